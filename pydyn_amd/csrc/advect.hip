// The three slopes advection operators (dynamu, dynamv, dynamw) with their CFL sub-cycle
// searches, and their mass-only forms for Check_CFL's rehearsal.
//
// x: one 64-lane workgroup per periodic (j, l) row.  The row (cells, faces, one tracer's four
//    moments and face fluxes) is staged in LDS; the search and the sub-steps run there, lanes
//    over faces then over cells, so a row costs one coalesced read and one write.
// y, z: closed lines.  A search kernel finds the sub-cycle count (per layer for y, one for the
//    whole region for z); a sub-step kernel stages a tile of lines (all cells of TI columns) in
//    LDS and runs every sub-step as two parallel phases, faces then cells.
//
// Tracers are processed one at a time over the same mass trajectory (recomputed per tracer,
// the same bits), so LDS does not grow with the number of tracers.
#include "tm5_lib.h"

namespace tm5 {

__device__ __forceinline__ bool skipped(const int *words, int mask)
{
    return (__hip_atomic_load(words + W_FLAGS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & mask) != 0;
}

// ------------------------------------------------------------------ x: dynamu
// pydyn/kernels.py:144-219 for one periodic row of n cells and n+1 faces (face k between cells
// k-1 and k, face 0 the same as face n).  The search (kernels.py:169-199): nloop = 1, 2, ...,
// each trial carrying every face's upwind cell through nloop mass-only sub-steps; on a failure
// every face is scaled by (nloop-1)/nloop in place.
template <bool MASS_ONLY>
__global__ void __launch_bounds__(64) k_x(int im, int jm, int lm, View3 m, Tracers q, View3 am, int limits,
                                          int nmax, int *words, int fail_bit, int skip_mask)
{
    if (skipped(words, skip_mask)) return;
    extern __shared__ double lds[];
    const int n = im;
    const int row = blockIdx.x;
    const int j = row % jm, l = row / jm;
    const int lane = threadIdx.x;
    double *M = lds;       // n   current mass
    double *M0 = M + n;    // n   the row's mass at the start
    double *A = M0 + n;    // n+1 faces
    double *R = A + n + 1; // 4n  rm, rxm, rym, rzm of one tracer
    double *F = R + 4 * n; // 4(n+1) f, pf, ft (=fy), fw (=fz)

    for (int i = lane; i < n; i += 64) M0[i] = M[i] = m(i, j, l);
    for (int k = lane; k <= n; k += 64) A[k] = am(k, j, l);
    __syncthreads();

    int nl = 1;
    for (;;) {
        bool ok = true;
        for (int k = lane; k <= n; k += 64) {
            const double ak = A[k];
            const int u = ak >= 0.0 ? (k == 0 ? n - 1 : k - 1) : (k == n ? 0 : k);
            const double alo = A[u], ahi = A[u + 1];
            double mx = M[u];
            for (int s = 0; s < nl; s++) {
                if (!(fabs(ak / mx) < 1.0)) {
                    ok = false;
                    break;
                }
                mx = (mx + alo) - ahi;
            }
        }
        if (__all(ok)) break;
        nl += 1;
        __syncthreads();
        for (int k = lane; k <= n; k += 64) A[k] = shrink(A[k], nl);
        __syncthreads();
        if (nl > nmax) {
            if (lane == 0) atomicOr(words + W_FLAGS, fail_bit);
            return;
        }
    }
    if (lane == 0) atomicMax(words + W_X_NLOOP, nl);

    if (MASS_ONLY) {  // each cell reads only itself and its two faces: in place
        for (int s = 0; s < nl; s++)
            for (int i = lane; i < n; i += 64) M[i] = (M[i] + A[i]) - A[i + 1];
        __syncthreads();
    } else {
        for (int t = 0; t < q.nt; t++) {
            if (t > 0) {
                for (int i = lane; i < n; i += 64) M[i] = M0[i];
            }
            for (int i = lane; i < n; i += 64) {
                const long long o = q.off(i, j, l, t);
                const double rm = q.q[0][o];
                R[i] = rm;
                R[n + i] = limits ? limit(q.q[1][o], rm) : q.q[1][o];
                R[2 * n + i] = q.q[2][o];
                R[3 * n + i] = q.q[3][o];
            }
            __syncthreads();
            for (int s = 0; s < nl; s++) {
                // faces 1..n; face 0 takes face n's fluxes (the periodic seam)
                for (int k = lane + 1; k <= n; k += 64) {
                    const double ak = A[k];
                    const int up = ak >= 0.0 ? k - 1 : (k == n ? 0 : k);
                    const double alpha = ak * (1.0 / M[up]);
                    const FaceFlux ff = face_flux(ak, alpha, R[up], R[n + up], R[2 * n + up], R[3 * n + up]);
                    F[k] = ff.f;
                    F[(n + 1) + k] = ff.pf;
                    F[2 * (n + 1) + k] = ff.ft;
                    F[3 * (n + 1) + k] = ff.fw;
                }
                __syncthreads();
                if (lane < 4) F[lane * (n + 1)] = F[lane * (n + 1) + n];
                __syncthreads();
                for (int i = lane; i < n; i += 64) {
                    const double m_old = M[i], a_lo = A[i], a_hi = A[i + 1];
                    const double m_new = (m_old + a_lo) - a_hi;
                    const double *f = F, *pf = F + (n + 1), *fy = F + 2 * (n + 1), *fz = F + 3 * (n + 1);
                    const double rm_new = R[i] - (f[i + 1] - f[i]);
                    double rx = slope_update(R[n + i], pf[i], pf[i + 1], a_lo, a_hi, rm_new, f[i], f[i + 1], m_old, m_new);
                    if (limits) rx = limit(rx, rm_new);
                    R[i] = rm_new;
                    R[n + i] = rx;
                    R[2 * n + i] = R[2 * n + i] - (fy[i + 1] - fy[i]);
                    R[3 * n + i] = R[3 * n + i] - (fz[i + 1] - fz[i]);
                    M[i] = m_new;
                }
                __syncthreads();
            }
            for (int i = lane; i < n; i += 64) {
                const long long o = q.off(i, j, l, t);
                q.q[0][o] = R[i];
                q.q[1][o] = R[n + i];
                q.q[2][o] = R[2 * n + i];
                q.q[3][o] = R[3 * n + i];
            }
            __syncthreads();
        }
    }
    for (int i = lane; i < n; i += 64) m(i, j, l) = M[i];
    for (int k = lane; k <= n; k += 64) am(k, j, l) = A[k] * (double)nl;  // "restore old am"
}

int launch_x(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 am, int limits, int nmax, int fail_bit,
             int skip_mask)
{
    if (im < 1 || im > 512 || jm < 1 || lm < 1) return TM5_EINVAL;
    const size_t lds = (size_t)(11 * im + 5) * sizeof(double);
    const dim3 grid(jm * lm), block(64);
    if (q.nt > 0)
        hipLaunchKernelGGL(k_x<false>, grid, block, lds, ctx->stream, im, jm, lm, m, q, am, limits, nmax, ctx->words,
                           fail_bit, skip_mask);
    else
        hipLaunchKernelGGL(k_x<true>, grid, block, lds, ctx->stream, im, jm, lm, m, q, am, limits, nmax, ctx->words,
                           fail_bit, skip_mask);
    return status(hipGetLastError());
}

// ------------------------------------------------------------------ y: dynamv's search
// pydyn/ops.py:315-341 with progressive scaling, one layer per workgroup: nloop = 1, 2, ...
// until every interior face of the layer passes its trial; on a failure the whole layer's
// faces are scaled by (nloop-1)/nloop in place (the reference scales bm(:,:,l) whole).  The
// polar outer faces must be zero (pydyn/kernels.py:233-234).
__global__ void __launch_bounds__(256) k_y_search(int im, int jm, int lm, View3 m, View3 bm, int nmax, int *nloop,
                                                  int *words, int fail_bit, int skip_mask)
{
    const int l = blockIdx.x;
    if (skipped(words, skip_mask)) return;
    const int tid = threadIdx.x, nth = blockDim.x;
    bool bad = false;
    for (int i = tid; i < im; i += nth) bad |= (bm(i, 0, l) != 0.0) || (bm(i, jm, l) != 0.0);
    if (__syncthreads_or(bad)) {
        if (tid == 0) {
            atomicOr(words + W_FLAGS, TM5_BAD_Y);
            nloop[l] = 0;
        }
        return;
    }
    const int nface = im * (jm - 1);
    int nl = 1;
    for (;;) {
        bool ok = true;
        for (int idx = tid; idx < nface; idx += nth) {
            const int i = idx % im, k = idx / im + 1;
            const double bk = bm(i, k, l);
            const int u = bk >= 0.0 ? k - 1 : k;
            const double blo = bm(i, u, l), bhi = bm(i, u + 1, l);
            double mx = m(i, u, l);
            for (int s = 0; s < nl; s++) {
                if (!(fabs(bk / mx) < 1.0)) {
                    ok = false;
                    break;
                }
                mx = (mx + blo) - bhi;
            }
        }
        if (__syncthreads_and(ok)) break;
        nl += 1;
        for (int idx = tid; idx < im * (jm + 1); idx += nth) {
            double &b = bm(idx % im, idx / im, l);
            b = shrink(b, nl);
        }
        __threadfence_block();
        __syncthreads();
        if (nl > nmax) {
            if (tid == 0) {
                atomicOr(words + W_FLAGS, fail_bit);
                nloop[l] = nl;
            }
            return;
        }
    }
    if (tid == 0) nloop[l] = nl;
}

// ------------------------------------------------------------------ y: dynamv's sub-steps
// pydyn/kernels.py:246-291 on a tile of TI columns of layer l (blockIdx.y), all jm cells.
// Along-flow slope rym; carried cross slopes rxm (ft) and rzm (fw).  Flow out of a pole cell
// carries no slope; the pole cells keep their rxm, rym.
template <bool MASS_ONLY>
__global__ void __launch_bounds__(256) k_y_steps(int im, int jm, int lm, int TI, View3 m, Tracers q, View3 bm,
                                                 const int *nloop, int limits, int *words, int skip_mask)
{
    if (skipped(words, skip_mask)) return;
    extern __shared__ double lds[];
    const int l = blockIdx.y, i0 = blockIdx.x * TI;
    const int ni = min(TI, im - i0);
    const int nl = nloop[l];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int ncell = jm * TI, nfc = (jm + 1) * TI;
    double *M = lds, *B = M + ncell;
    double *M0 = B + nfc, *R = M0 + ncell, *F = R + 4 * ncell;  // tracer part
    // index: cell (ti, j) at j * TI + ti; face (ti, k) at k * TI + ti
    for (int idx = tid; idx < ncell; idx += nth) {
        const int ti = idx % TI, j = idx / TI;
        if (ti < ni) M[idx] = m(i0 + ti, j, l);
    }
    for (int idx = tid; idx < nfc; idx += nth) {
        const int ti = idx % TI, k = idx / TI;
        if (ti < ni) B[idx] = bm(i0 + ti, k, l);
    }
    __syncthreads();

    if (MASS_ONLY) {
        for (int s = 0; s < nl; s++)
            for (int idx = tid; idx < ncell; idx += nth) M[idx] = (M[idx] + B[idx]) - B[idx + TI];
        __syncthreads();
    } else {
        for (int idx = tid; idx < ncell; idx += nth) M0[idx] = M[idx];
        for (int ti = tid; ti < TI; ti += nth) {  // the closed outer faces
            for (int a = 0; a < 4; a++) {
                F[a * nfc + ti] = 0.0;
                F[a * nfc + jm * TI + ti] = 0.0;
            }
        }
        for (int t = 0; t < q.nt; t++) {
            for (int idx = tid; idx < ncell; idx += nth) {
                const int ti = idx % TI, j = idx / TI;
                if (ti >= ni) continue;
                const long long o = q.off(i0 + ti, j, l, t);
                const double rm = q.q[0][o];
                R[idx] = rm;                                                    // rm
                R[ncell + idx] = limits ? limit(q.q[2][o], rm) : q.q[2][o];     // rym (along)
                R[2 * ncell + idx] = q.q[1][o];                                 // rxm
                R[3 * ncell + idx] = q.q[3][o];                                 // rzm
                M[idx] = M0[idx];
            }
            __syncthreads();
            for (int s = 0; s < nl; s++) {
                for (int idx = tid; idx < (jm - 1) * TI; idx += nth) {
                    const int ti = idx % TI, k = idx / TI + 1;
                    if (ti >= ni) continue;
                    const double bf = B[k * TI + ti];
                    const int u = (bf >= 0.0 ? k - 1 : k) * TI + ti;
                    const double beta = bf / M[u];
                    FaceFlux ff = face_flux(bf, beta, R[u], R[ncell + u], R[2 * ncell + u], R[3 * ncell + u]);
                    if (k == 1 && bf >= 0.0) {  // leaving the south pole cell
                        ff.f = beta * R[u];
                        ff.pf = (-3.0 * bf) * ff.f;
                        ff.ft = 0.0;
                        ff.fw = beta * R[3 * ncell + u];
                    } else if (k == jm - 1 && bf < 0.0) {  // leaving the north pole cell
                        ff.f = beta * R[u];
                        ff.pf = (-3.0 * bf) * ff.f;
                        ff.ft = 0.0;
                        ff.fw = beta * R[3 * ncell + u];
                    }
                    const int fo = k * TI + ti;
                    F[fo] = ff.f;
                    F[nfc + fo] = ff.pf;
                    F[2 * nfc + fo] = ff.ft;
                    F[3 * nfc + fo] = ff.fw;
                }
                __syncthreads();
                for (int idx = tid; idx < ncell; idx += nth) {
                    const int ti = idx % TI, j = idx / TI;
                    if (ti >= ni) continue;
                    const double b_lo = B[idx], b_hi = B[idx + TI];
                    const double m_old = M[idx];
                    const double m_new = (m_old + b_lo) - b_hi;
                    const double inv_mnew = 1.0 / m_new;
                    const double flow_conv = b_lo - b_hi;
                    const double flow_mean = 3.0 * (b_lo + b_hi);
                    const double mass_x3 = 3.0 * m_old;
                    const double *f = F, *pf = F + nfc, *fx = F + 2 * nfc, *fz = F + 3 * nfc;
                    const double rm_new = R[idx] - (f[idx + TI] - f[idx]);
                    const double rym = R[ncell + idx];
                    double ry = rym + ((((pf[idx] - pf[idx + TI]) - flow_conv * rym) + flow_mean * rm_new) -
                                       mass_x3 * (f[idx] + f[idx + TI])) *
                                          inv_mnew;
                    if (limits) ry = limit(ry, rm_new);
                    const double rx = R[2 * ncell + idx] - (fx[idx + TI] - fx[idx]);
                    R[idx] = rm_new;
                    R[3 * ncell + idx] = R[3 * ncell + idx] - (fz[idx + TI] - fz[idx]);
                    if (j != 0 && j != jm - 1) {
                        R[ncell + idx] = ry;
                        R[2 * ncell + idx] = rx;
                    }
                    M[idx] = m_new;
                }
                __syncthreads();
            }
            for (int idx = tid; idx < ncell; idx += nth) {
                const int ti = idx % TI, j = idx / TI;
                if (ti >= ni) continue;
                const long long o = q.off(i0 + ti, j, l, t);
                q.q[0][o] = R[idx];
                q.q[2][o] = R[ncell + idx];
                q.q[1][o] = R[2 * ncell + idx];
                q.q[3][o] = R[3 * ncell + idx];
            }
            __syncthreads();
        }
    }
    for (int idx = tid; idx < ncell; idx += nth) {
        const int ti = idx % TI, j = idx / TI;
        if (ti < ni) m(i0 + ti, j, l) = M[idx];
    }
    for (int idx = tid; idx < nfc; idx += nth) {
        const int ti = idx % TI, k = idx / TI;
        if (ti < ni) bm(i0 + ti, k, l) = B[idx] * (double)nl;  // "restore old bm"
    }
}

static int tile_width(int cells, int faces, int per_cell, int per_face, int cap)
{
    const long long per_line = (long long)cells * per_cell + (long long)faces * per_face;
    long long ti = (60 * 1024 / 8) / per_line;
    if (ti > cap) ti = cap;
    return (int)ti;
}

int launch_y(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 bm, int limits, int nmax, int fail_bit,
             int skip_mask)
{
    if (im < 1 || jm < 2 || lm < 1 || lm > ctx->y_cap) return TM5_EINVAL;
    const bool mass_only = q.nt == 0;
    const int TI = mass_only ? tile_width(jm, jm + 1, 1, 1, 16) : tile_width(jm, jm + 1, 6, 5, 8);
    if (TI < 1) return TM5_EINVAL;
    hipLaunchKernelGGL(k_y_search, dim3(lm), dim3(256), 0, ctx->stream, im, jm, lm, m, bm, nmax, ctx->y_nloop,
                       ctx->words, fail_bit, skip_mask);
    const size_t lds = (size_t)(mass_only ? (2 * jm + 1) * TI : (11 * jm + 5) * TI) * sizeof(double);
    const dim3 grid((im + TI - 1) / TI, lm);
    if (mass_only)
        hipLaunchKernelGGL(k_y_steps<true>, grid, dim3(256), lds, ctx->stream, im, jm, lm, TI, m, q, bm, ctx->y_nloop,
                           limits, ctx->words, skip_mask);
    else
        hipLaunchKernelGGL(k_y_steps<false>, grid, dim3(256), lds, ctx->stream, im, jm, lm, TI, m, q, bm,
                           ctx->y_nloop, limits, ctx->words, skip_mask);
    return status(hipGetLastError());
}

// ------------------------------------------------------------------ z: dynamw's search
// pydyn/ops.py:315-341 with fresh scaling (cm / n) over the whole region: the smallest n at which
// every column passes.  Each column finds its own smallest passing n >= 1; the largest of those
// is a candidate no smaller than the answer (every n below it fails some column).  The last
// workgroup to finish checks every column at the candidate and, if one fails there, raises the
// candidate to that column's next passing n, until all pass (the reference's sequential search,
// found as a fixpoint).
__device__ bool z_column_passes(const View3 &m, const View3 &cm, int i, int j, int lm, int n)
{
    const double dn = (double)n;
    for (int k = 1; k < lm; k++) {
        const double ck = cm(i, j, k) / dn;
        const int u = ck >= 0.0 ? k - 1 : k;
        const double clo = cm(i, j, u) / dn, chi = cm(i, j, u + 1) / dn;
        double mx = m(i, j, u);
        for (int s = 0; s < n; s++) {
            if (!(fabs(ck / mx) < 1.0)) return false;
            mx = (mx + clo) - chi;
        }
    }
    return true;
}

__device__ int z_column_next(const View3 &m, const View3 &cm, int i, int j, int lm, int from, int nmax)
{
    for (int n = from; n <= nmax; n++)
        if (z_column_passes(m, cm, i, j, lm, n)) return n;
    return nmax + 1;
}

__global__ void __launch_bounds__(256) k_z_search(int im, int jm, int lm, View3 m, View3 cm, int nmax, int *words,
                                                  int fail_bit, int skip_mask)
{
    // every workgroup takes its ticket, skipping or not, so that the last one always resets the
    // reduction words for the next launch
    const bool skip = skipped(words, skip_mask);
    const int tid = threadIdx.x;
    const int col = blockIdx.x * blockDim.x + tid;
    const int ncol = im * jm;
    int mine = 1;
    if (!skip && col < ncol) {
        const int i = col % im, j = col / im;
        if (cm(i, j, 0) != 0.0 || cm(i, j, lm) != 0.0) atomicOr(words + W_FLAGS, TM5_BAD_Z);
        mine = z_column_next(m, cm, i, j, lm, 1, nmax);
    }
    __shared__ int red[256];
    __shared__ bool last;
    red[tid] = mine;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] = max(red[tid], red[tid + s]);
        __syncthreads();
    }
    if (tid == 0) {
        atomicMax(words + W_Z_CAND, red[0]);
        __threadfence();
        last = atomicAdd(words + W_Z_TICKET, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    int cand = __hip_atomic_load(words + W_Z_CAND, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (!skip && cand > 1 && cand <= nmax) {
        int need = cand;
        for (int c = tid; c < ncol; c += blockDim.x) {
            const int i = c % im, j = c / im;
            if (!z_column_passes(m, cm, i, j, lm, cand)) need = max(need, z_column_next(m, cm, i, j, lm, cand + 1, nmax));
        }
        red[tid] = need;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red[tid] = max(red[tid], red[tid + s]);
            __syncthreads();
        }
        const int next = red[0];
        __syncthreads();
        if (next == cand) break;
        cand = next;
    }
    if (tid == 0) {
        if (!skip) {
            if (cand > nmax) atomicOr(words + W_FLAGS, fail_bit);
            words[W_Z_NADV] = cand;
        }
        words[W_Z_TICKET] = 0;
        words[W_Z_CAND] = 0;
    }
}

// ------------------------------------------------------------------ z: dynamw's sub-steps
// pydyn/kernels.py:324-346 on a tile of TI columns (i0.., j = blockIdx.y), all lm levels, with
// c = cm / n_advz; cm itself is not written.  Along-flow slope rzm; cross slopes rxm (ft), rym (fw).
template <bool MASS_ONLY>
__global__ void __launch_bounds__(256) k_z_steps(int im, int jm, int lm, int TI, View3 m, Tracers q, View3 cm,
                                                 int limits, int *words, int skip_mask)
{
    if (skipped(words, skip_mask)) return;
    extern __shared__ double lds[];
    const int j = blockIdx.y, i0 = blockIdx.x * TI;
    const int ni = min(TI, im - i0);
    const int nz = words[W_Z_NADV];
    const double dn = (double)nz;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int ncell = lm * TI, nfc = (lm + 1) * TI;
    double *M = lds, *C = M + ncell;
    double *M0 = C + nfc, *R = M0 + ncell, *F = R + 4 * ncell;
    for (int idx = tid; idx < ncell; idx += nth) {
        const int ti = idx % TI, l = idx / TI;
        if (ti < ni) M[idx] = m(i0 + ti, j, l);
    }
    for (int idx = tid; idx < nfc; idx += nth) {
        const int ti = idx % TI, k = idx / TI;
        if (ti < ni) C[idx] = cm(i0 + ti, j, k) / dn;
    }
    __syncthreads();

    if (MASS_ONLY) {
        for (int s = 0; s < nz; s++)
            for (int idx = tid; idx < ncell; idx += nth) M[idx] = (M[idx] + C[idx]) - C[idx + TI];
        __syncthreads();
    } else {
        for (int idx = tid; idx < ncell; idx += nth) M0[idx] = M[idx];
        for (int ti = tid; ti < TI; ti += nth) {
            for (int a = 0; a < 4; a++) {
                F[a * nfc + ti] = 0.0;
                F[a * nfc + lm * TI + ti] = 0.0;
            }
        }
        for (int t = 0; t < q.nt; t++) {
            for (int idx = tid; idx < ncell; idx += nth) {
                const int ti = idx % TI, l = idx / TI;
                if (ti >= ni) continue;
                const long long o = q.off(i0 + ti, j, l, t);
                const double rm = q.q[0][o];
                R[idx] = rm;                                                  // rm
                R[ncell + idx] = limits ? limit(q.q[3][o], rm) : q.q[3][o];   // rzm (along)
                R[2 * ncell + idx] = q.q[1][o];                               // rxm
                R[3 * ncell + idx] = q.q[2][o];                               // rym
                M[idx] = M0[idx];
            }
            __syncthreads();
            for (int s = 0; s < nz; s++) {
                for (int idx = tid; idx < (lm - 1) * TI; idx += nth) {
                    const int ti = idx % TI, k = idx / TI + 1;
                    if (ti >= ni) continue;
                    const double cf = C[k * TI + ti];
                    const int u = (cf >= 0.0 ? k - 1 : k) * TI + ti;
                    const double gamma = cf / M[u];
                    const FaceFlux ff = face_flux(cf, gamma, R[u], R[ncell + u], R[2 * ncell + u], R[3 * ncell + u]);
                    const int fo = k * TI + ti;
                    F[fo] = ff.f;
                    F[nfc + fo] = ff.pf;
                    F[2 * nfc + fo] = ff.ft;
                    F[3 * nfc + fo] = ff.fw;
                }
                __syncthreads();
                for (int idx = tid; idx < ncell; idx += nth) {
                    const int ti = idx % TI;
                    if (ti >= ni) continue;
                    const double c_lo = C[idx], c_hi = C[idx + TI];
                    const double m_old = M[idx];
                    const double m_new = (m_old + c_lo) - c_hi;
                    const double *f = F, *pf = F + nfc, *fx = F + 2 * nfc, *fy = F + 3 * nfc;
                    const double rm_new = (R[idx] + f[idx]) - f[idx + TI];  // advance: rm + f(l-1) - f(l)
                    double rz = slope_update(R[ncell + idx], pf[idx], pf[idx + TI], c_lo, c_hi, rm_new, f[idx],
                                             f[idx + TI], m_old, m_new);
                    if (limits) rz = limit(rz, rm_new);
                    R[idx] = rm_new;
                    R[ncell + idx] = rz;
                    R[2 * ncell + idx] = (R[2 * ncell + idx] + fx[idx]) - fx[idx + TI];
                    R[3 * ncell + idx] = (R[3 * ncell + idx] + fy[idx]) - fy[idx + TI];
                    M[idx] = m_new;
                }
                __syncthreads();
            }
            for (int idx = tid; idx < ncell; idx += nth) {
                const int ti = idx % TI, l = idx / TI;
                if (ti >= ni) continue;
                const long long o = q.off(i0 + ti, j, l, t);
                q.q[0][o] = R[idx];
                q.q[3][o] = R[ncell + idx];
                q.q[1][o] = R[2 * ncell + idx];
                q.q[2][o] = R[3 * ncell + idx];
            }
            __syncthreads();
        }
    }
    for (int idx = tid; idx < ncell; idx += nth) {
        const int ti = idx % TI, l = idx / TI;
        if (ti < ni) m(i0 + ti, j, l) = M[idx];
    }
}

int launch_z(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 cm, int limits, int nmax, int fail_bit,
             int skip_mask)
{
    if (im < 1 || jm < 1 || lm < 2) return TM5_EINVAL;
    const bool mass_only = q.nt == 0;
    const int TI = mass_only ? tile_width(lm, lm + 1, 1, 1, 32) : tile_width(lm, lm + 1, 6, 5, 16);
    if (TI < 1) return TM5_EINVAL;
    const int ncol = im * jm;
    hipLaunchKernelGGL(k_z_search, dim3((ncol + 255) / 256), dim3(256), 0, ctx->stream, im, jm, lm, m, cm, nmax,
                       ctx->words, fail_bit, skip_mask);
    const size_t lds = (size_t)(mass_only ? (2 * lm + 1) * TI : (11 * lm + 5) * TI) * sizeof(double);
    const dim3 grid((im + TI - 1) / TI, jm);
    if (mass_only)
        hipLaunchKernelGGL(k_z_steps<true>, grid, dim3(256), lds, ctx->stream, im, jm, lm, TI, m, q, cm, limits,
                           ctx->words, skip_mask);
    else
        hipLaunchKernelGGL(k_z_steps<false>, grid, dim3(256), lds, ctx->stream, im, jm, lm, TI, m, q, cm, limits,
                           ctx->words, skip_mask);
    return status(hipGetLastError());
}

}  // namespace tm5
