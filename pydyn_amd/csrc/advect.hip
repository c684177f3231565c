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
#include <type_traits>

#include "tm5_lib.h"

namespace tm5 {

__device__ __forceinline__ bool skipped(const int *words, int mask)
{
    return (words[W_FLAGS] & mask) != 0;  // written by earlier kernels of the stream
}

// ------------------------------------------------------------------ x: dynamu
// pydyn/kernels.py:144-219 for one periodic row of n cells and n+1 faces (face k between cells
// k-1 and k, face 0 the same as face n).  The search (kernels.py:169-199): nloop = 1, 2, ...,
// each trial carrying every face's upwind cell through nloop mass-only sub-steps; on a failure
// every face is scaled by (nloop-1)/nloop in place.
// With tracer ranges (tchunk > 0, grid y) several workgroups take the same row: each searches and
// moves the row's mass on its own, none writes m, am, rows_out or amax_out, and the caller runs
// the mass-only form afterwards.  MAKES: the step's fluxes made by the launch (fs, nd0), else read.
template <bool MASS_ONLY, bool MAKES>
__global__ void __launch_bounds__(64) k_x(int im, int jm, int lm, View3 m, Tracers q, View3 am, int limits,
                                          int nmax, int *words, int *rows_out, int fail_bit, int skip_mask,
                                          const int *imred, double *amax_out, int spec_k, double *bk, int tchunk,
                                          int hj0, int hj1, FluxSrc fs, int nd0, int keep_am)
{
    // rows j in hj0..hj1 also get their periodic halo, index -1 = index n - 1 and index n = index 0,
    // for m and every moment (a zoom tree's region 1, pydyn/zoom/advect.py:283-288; on a reduced
    // row the same cells of its reduced form, which the split back overwrites)
    // amax_out (optional): per row the largest |Courant| of its accepted trial (the reference's xi)
    // checked(spec_k): a search past nmax is a failed check of a checked step (spec_fail), not fail_bit
    // bk (whole rows only): the row's m and moments as read, i fastest, m first and then every
    // tracer's rm, rxm, rym, rzm (each im * jm * lm)
    // keep_am: am written back as it was read, not "restored" by nloop (a reduced-grid x, where
    // the reference puts back a copy of the faces taken before)
    // fs.u.p (whole periodic rows): the faces made from pu, am(k) = dt (a pu(k)) with face 0 the
    // same as face im, as dynam0 makes them, instead of read from am; the first nd0 workgroups (one
    // tracer range only) are dynam0's tiles making bm and cm (fs.bm.p): dispatched first, their
    // column sums (a chain of dependent adds a column) run while the rows keep the device busy
    if (skipped(words, skip_mask)) return;
    extern __shared__ double lds[];
    if (MAKES && (int)blockIdx.x < nd0) {  // a dynam0 tile: bm and cm of the step, am left to the rows
        const int ntx = (im + fs.TI - 1) / fs.TI;
        TimeSrc ts{};
        ts.a = fs.a;
        dynam0_tile(blockIdx.x % ntx, blockIdx.x / ntx, lds, im, jm, lm, fs.TI, fs.u, fs.v, ts, fs.bt, fs.dt, View3{},
                    fs.bm, fs.cm, 1, words, View3{}, View3{}, 0);
        return;
    }
    // rows renumbered XCD-contiguously, the tracer range slowest
    const int skip0 = MAKES ? nd0 : 0, nrx = gridDim.x - skip0;
    const int b = xcd_tile(blockIdx.y * nrx + (blockIdx.x - skip0), nrx * gridDim.y);
    const int row = b % nrx, bz = b / nrx;
    const int j = row % jm, l = row / jm;
    const int t0 = tchunk > 0 ? bz * tchunk : 0;
    const int t1 = tchunk > 0 && t0 + tchunk < q.nt ? t0 + tchunk : q.nt;
    const bool writes = tchunk == 0;  // m, am, rows_out, amax_out
    const int n = imred ? imred[j] : im;  // a reduced row has imred(j) cells
    const int lane = threadIdx.x;
    double *M = lds;       // n   current mass
    double *M0 = M + n;    // n   the row's mass at the start
    double *A = M0 + n;    // n+1 faces
    double *R = A + n + 1; // 4n  rm, rxm, rym, rzm of one tracer
    double *F = R + 4 * n; // 4(n+1) f, pf, ft (=fy), fw (=fz)
    double *A0 = F + 4 * (n + 1);  // n+1 (keep_am) the faces as read

    const long long ncell = (long long)im * jm * lm, brow = (long long)im * (j + (long long)jm * l);
    for (int i = lane; i < n; i += 64) {
        const double v = m(i, j, l);
        M0[i] = M[i] = v;
        if (bk) bk[brow + i] = v;
    }
    if (MAKES) {
        for (int k = lane; k <= n; k += 64) A[k] = fs.dt * (fs.a * fs.u(k == 0 ? im : k, j, l));
    } else {
        for (int k = lane; k <= n; k += 64) A[k] = am(k, j, l);
    }
    if (keep_am)
        for (int k = lane; k <= n; k += 64) A0[k] = A[k];
    // the first tracer's moments of the lane's (at most two) cells, in flight during the search
    constexpr int PRE = 2;
    const bool pre = !MASS_ONLY && t0 < t1 && n <= 64 * PRE;
    double qp[PRE][4];
#pragma unroll
    for (int c = 0; c < PRE; c++) {
        const int i = lane + 64 * c;
        if (pre && i < n) {
            const long long o = q.off(i, j, l, t0);
#pragma unroll
            for (int a = 0; a < 4; a++) qp[c][a] = q.q[a][o];
        }
    }
    __syncthreads();

    int nl = 1;
    double cmax = 0.0;
    for (;;) {
        bool ok = true;
        cmax = 0.0;
        for (int k = lane; k <= n; k += 64) {
            const double ak = A[k];
            const int u = ak >= 0.0 ? (k == 0 ? n - 1 : k - 1) : (k == n ? 0 : k);
            const double alo = A[u], ahi = A[u + 1];
            double mx = M[u];
            for (int s = 0; s < nl; s++) {
                const double c = fabs(ak / mx);
                if (!(c < 1.0)) {
                    ok = false;
                    break;
                }
                if (c > cmax) cmax = c;
                mx = (mx + alo) - ahi;
            }
        }
        if (__all(ok)) break;
        nl += 1;
        __syncthreads();
        for (int k = lane; k <= n; k += 64) A[k] = shrink(A[k], nl);
        __syncthreads();
        if (nl > nmax) {
            if (lane == 0) {
                if (checked(spec_k))
                    spec_fail(words, spec_k);
                else
                    atomicOr(words + W_FLAGS, fail_bit);
                if (writes) rows_out[row] = nl;
            }
            if (!MASS_ONLY && bk)  // the row's moments are not read below
                for (int t = 0; t < q.nt; t++)
                    for (int i = lane; i < n; i += 64) {
                        const long long o = q.off(i, j, l, t);
                        for (int a = 0; a < 4; a++) bk[(1 + 4 * t + a) * ncell + brow + i] = q.q[a][o];
                    }
            return;
        }
    }
    if (lane == 0 && writes) rows_out[row] = nl;  // one plain store per row: no contended atomic
    if (amax_out && writes) {  // the row's largest Courant number (a maximum: any order gives the same value)
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(cmax, off);
            if (o > cmax) cmax = o;
        }
        if (lane == 0) amax_out[row] = cmax;
    }

    if (MASS_ONLY) {  // each cell reads only itself and its two faces: in place
        for (int s = 0; s < nl; s++)
            for (int i = lane; i < n; i += 64) M[i] = (M[i] + A[i]) - A[i + 1];
        __syncthreads();
    } else {
        for (int t = t0; t < t1; t++) {
            if (t > t0) {
                for (int i = lane; i < n; i += 64) M[i] = M0[i];
            }
            for (int i = lane; i < n; i += 64) {
                const long long o = q.off(i, j, l, t);
                double rm, rx, ry, rz;
                if (pre && t == t0) {  // i = lane + 64 c with c < PRE
                    const int c = (i - lane) / 64;
                    rm = c == 0 ? qp[0][0] : qp[1][0];
                    rx = c == 0 ? qp[0][1] : qp[1][1];
                    ry = c == 0 ? qp[0][2] : qp[1][2];
                    rz = c == 0 ? qp[0][3] : qp[1][3];
                } else {
                    rm = q.q[0][o];
                    rx = q.q[1][o];
                    ry = q.q[2][o];
                    rz = q.q[3][o];
                }
                R[i] = rm;
                R[n + i] = limits ? limit(rx, rm) : rx;
                R[2 * n + i] = ry;
                R[3 * n + i] = rz;
                if (bk) {
                    double *b = bk + (1 + 4 * t) * ncell + brow + i;
                    b[0] = rm;
                    b[ncell] = rx;
                    b[2 * ncell] = ry;
                    b[3 * ncell] = rz;
                }
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
            if (lane < 8 && j >= hj0 && j <= hj1) {  // lane a < 4: q(-1) of moment a; 4 + a: q(n)
                const int a = lane & 3;
                q.q[a][q.off(lane < 4 ? -1 : n, j, l, t)] = R[a * n + (lane < 4 ? n - 1 : 0)];
            }
            __syncthreads();
        }
    }
    if (!writes) return;
    for (int i = lane; i < n; i += 64) m(i, j, l) = M[i];
    if (lane == 0 && j >= hj0 && j <= hj1) {
        m(-1, j, l) = M[n - 1];
        m(n, j, l) = M[0];
    }
    for (int k = lane; k <= n; k += 64) am(k, j, l) = keep_am ? A0[k] : A[k] * (double)nl;  // "restore old am"
}

// ------------------------------------------------------------------ x twice
// A checked step's last dynamu and the next checked step's first in one launch: they act on the
// same row one after the other (the step's closing m2phlb1, which only reads m, and the next step's
// Setup_MassFlow, which only makes fluxes, come between them in the reference's order and touch
// nothing either reads), so a workgroup takes its row through both.  Pass A takes the faces from
// am, pass B makes them from pu (FluxSrc) and stores them restored; the first nd0 workgroups make
// the next step's bm and cm (dynam0's tiles), as k_x's.

// k_x's search of a periodic row of n cells whose mass is M: nloop = 1, 2, ... with every face's
// upwind cell carried through nloop mass-only sub-steps, the faces A scaled by (nloop-1)/nloop in
// place after a failure; the accepted nloop, or nmax + 1 when none up to nmax passes
__device__ int x_row_search(int n, double *A, const double *M, int nmax, int lane)
{
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
        if (__all(ok)) return nl;
        nl += 1;
        __syncthreads();
        for (int k = lane; k <= n; k += 64) A[k] = shrink(A[k], nl);
        __syncthreads();
        if (nl > nmax) return nl;
    }
}

// k_x's nl sub-steps of one tracer on the row: faces 1..n (face 0 takes face n's fluxes), then
// every cell; M, R (rm, rxm, rym, rzm) in place, F the faces' fluxes
__device__ void x_row_substeps(int n, int nl, const double *A, double *M, double *R, double *F, int limits, int lane)
{
    for (int s = 0; s < nl; s++) {
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
}

// spec_a / spec_b: the checked steps' intervals (a search past nmax is a failed check, spec_fail_first);
// bk (null: none): pass B's input -- the row's m and moments after pass A, k_x's backup layout.  A
// row whose pass B fails its check still writes its part of bk (the host restores it whole).
__global__ void __launch_bounds__(64) k_xx(int im, int jm, int lm, View3 m, Tracers q, View3 am, FluxSrc fs, int nd0,
                                           int limits, int nmax, int *words, int *rows_out, int skip_mask,
                                           int spec_a, int spec_b, double *bk)
{
    if (skipped(words, skip_mask)) return;
    extern __shared__ double lds[];
    if ((int)blockIdx.x < nd0) {  // a dynam0 tile: the next step's bm and cm
        const int ntx = (im + fs.TI - 1) / fs.TI;
        TimeSrc ts{};
        ts.a = fs.a;
        dynam0_tile(blockIdx.x % ntx, blockIdx.x / ntx, lds, im, jm, lm, fs.TI, fs.u, fs.v, ts, fs.bt, fs.dt, View3{},
                    fs.bm, fs.cm, 1, words, View3{}, View3{}, 0);
        return;
    }
    const int row = xcd_tile(blockIdx.x - nd0, gridDim.x - nd0);  // rows renumbered XCD-contiguously
    const int j = row % jm, l = row / jm;
    const int n = im, lane = threadIdx.x;
    double *M = lds;            // n   current mass
    double *M0 = M + n;         // n   the row's mass at the start
    double *M1 = M0 + n;        // n   after pass A
    double *AA = M1 + n;        // n+1 pass A's faces
    double *AB = AA + n + 1;    // n+1 pass B's faces
    double *R = AB + n + 1;     // 4n  rm, rxm, rym, rzm of one tracer
    double *F = R + 4 * n;      // 4(n+1) f, pf, ft (=fy), fw (=fz)
    for (int i = lane; i < n; i += 64) M0[i] = M1[i] = m(i, j, l);
    for (int k = lane; k <= n; k += 64) {
        AA[k] = am(k, j, l);
        AB[k] = fs.dt * (fs.a * fs.u(k == 0 ? im : k, j, l));
    }
    // the first tracer's moments of the lane's (at most two) cells, in flight during the searches
    constexpr int PRE = 2;
    const bool pre = q.nt > 0 && n <= 64 * PRE;
    double qp[PRE][4];
#pragma unroll
    for (int c = 0; c < PRE; c++) {
        const int i = lane + 64 * c;
        if (pre && i < n) {
            const long long o = q.off(i, j, l, 0);
#pragma unroll
            for (int a = 0; a < 4; a++) qp[c][a] = q.q[a][o];
        }
    }
    __syncthreads();
    const int nla = x_row_search(n, AA, M0, nmax, lane);
    if (nla > nmax) {
        if (lane == 0) spec_fail_first(words, spec_a);
        return;
    }
    for (int s = 0; s < nla; s++)  // each cell reads only itself and its two faces: in place
        for (int i = lane; i < n; i += 64) M1[i] = (M1[i] + AA[i]) - AA[i + 1];
    __syncthreads();
    const int nlb = x_row_search(n, AB, M1, nmax, lane);
    const bool fail_b = nlb > nmax;
    if (lane == 0) {
        if (fail_b) spec_fail_first(words, spec_b);
        rows_out[row] = nlb;
    }
    if (fail_b && !bk) return;
    const long long ncell = (long long)im * jm * lm, brow = (long long)im * (j + (long long)jm * l);
    if (bk)
        for (int i = lane; i < n; i += 64) bk[brow + i] = M1[i];
    for (int t = 0; t < q.nt; t++) {
        for (int i = lane; i < n; i += 64) {
            M[i] = M0[i];
            const long long o = q.off(i, j, l, t);
            double rm, rx, ry, rz;
            if (pre && t == 0) {  // i = lane + 64 c with c < PRE
                const int c = (i - lane) / 64;
                rm = c == 0 ? qp[0][0] : qp[1][0];
                rx = c == 0 ? qp[0][1] : qp[1][1];
                ry = c == 0 ? qp[0][2] : qp[1][2];
                rz = c == 0 ? qp[0][3] : qp[1][3];
            } else {
                rm = q.q[0][o];
                rx = q.q[1][o];
                ry = q.q[2][o];
                rz = q.q[3][o];
            }
            R[i] = rm;
            R[n + i] = limits ? limit(rx, rm) : rx;
            R[2 * n + i] = ry;
            R[3 * n + i] = rz;
        }
        __syncthreads();
        x_row_substeps(n, nla, AA, M, R, F, limits, lane);
        // the tracer at the next step's start: into the backup, pass B's initial limiter
        for (int i = lane; i < n; i += 64) {
            if (bk) {
#pragma unroll
                for (int a = 0; a < 4; a++) bk[(1 + 4 * t + a) * ncell + brow + i] = R[a * n + i];
            }
            if (limits) R[n + i] = limit(R[n + i], R[i]);
        }
        __syncthreads();
        if (fail_b) continue;
        x_row_substeps(n, nlb, AB, M, R, F, limits, lane);
        for (int i = lane; i < n; i += 64) {
            const long long o = q.off(i, j, l, t);
#pragma unroll
            for (int a = 0; a < 4; a++) q.q[a][o] = R[a * n + i];
        }
        __syncthreads();
    }
    if (fail_b) return;
    if (q.nt == 0) {  // the mass alone through pass B
        for (int s = 0; s < nlb; s++)
            for (int i = lane; i < n; i += 64) M1[i] = (M1[i] + AB[i]) - AB[i + 1];
        __syncthreads();
    }
    const double *Mend = q.nt > 0 ? M : M1;
    for (int i = lane; i < n; i += 64) m(i, j, l) = Mend[i];
    for (int k = lane; k <= n; k += 64) am(k, j, l) = AB[k] * (double)nlb;  // "restore old am"
}

int launch_xx(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 am, int limits, int nmax, int skip_mask,
              int spec_a, int spec_b, double *backup, const FluxSrc &fs)
{
    if (im < 1 || im > 512 || jm < 1 || lm < 1 || jm * lm > ctx->x_cap || !checked(spec_a) || !checked(spec_b))
        return TM5_EINVAL;
    if (!fs.u.p || !fs.v.p || !fs.bm.p || !fs.cm.p || !fs.bt || fs.TI < 1 || lm * fs.TI > 8192) return TM5_EINVAL;
    size_t lds = (size_t)(13 * im + 6) * sizeof(double);
    if ((size_t)lm * fs.TI * sizeof(double) > lds) lds = (size_t)lm * fs.TI * sizeof(double);
    const int nd0 = ((im + fs.TI - 1) / fs.TI) * jm;
    ctx->x_nrows = jm * lm;
    TM5_LAUNCH(ctx, k_xx, dim3((unsigned)(nd0 + jm * lm)), dim3(64), lds, ctx->stream, im, jm, lm, m, q, am, fs, nd0,
               limits, nmax, ctx->words, ctx->x_rows, skip_mask, spec_a, spec_b, backup);
    return status(hipGetLastError());
}

int launch_x(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 am, int limits, int nmax, int fail_bit,
             int skip_mask, const int *imred, int *rows_out, double *amax_out, int spec_k, double *backup, int halo_j0,
             int halo_j1, const FluxSrc *flux, int keep_am)
{
    if (im < 1 || im > 512 || jm < 1 || lm < 1 || jm * lm > ctx->x_cap) return TM5_EINVAL;
    if ((backup || flux) && imred) return TM5_EINVAL;  // the backup holds whole rows; faces made from pu too
    const FluxSrc fs = flux ? *flux : FluxSrc{};
    size_t lds = (size_t)((keep_am ? 12 : 11) * im + (keep_am ? 6 : 5)) * sizeof(double);
    const dim3 grid(jm * lm), block(64);
    ctx->x_nrows = jm * lm;
    int *rows = rows_out ? rows_out : ctx->x_rows;
    int ranges = 1, tchunk = q.nt;
    if (q.nt >= 32 && !backup) {  // many tracers: tracer ranges over the grid, then the masses
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_x<false, false>, 64, lds) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        tchunk = tracer_chunk(grid.x, (long long)per_cu * (ctx->ncu > 0 ? ctx->ncu : 256), q.nt,
                              ctx->tile[TM5_TILE_RANGE_MIN]);
        ranges = (q.nt + tchunk - 1) / tchunk;
    }
    // the step's fluxes: am made by the rows and bm and cm by dynam0's tiles as extra workgroups of
    // the launch that takes every tracer (one range); with tracer ranges, a whole dynam0 launch of
    // their own before the x launches, which then read am
    unsigned nd0 = 0;
    if (fs.bm.p) {
        if (!fs.u.p || !fs.v.p || !fs.cm.p || !fs.bt || fs.TI < 1 || lm * fs.TI > 8192) return TM5_EINVAL;
        if (ranges > 1) {
            const int rc = launch_dynam0(ctx, im, jm, lm, fs.u, fs.v, fs.a, fs.bt, fs.dt, am, fs.bm, fs.cm, skip_mask, 1);
            if (rc) return rc;
        } else {
            nd0 = (unsigned)(((im + fs.TI - 1) / fs.TI) * jm);
            if ((size_t)lm * fs.TI * sizeof(double) > lds) lds = (size_t)lm * fs.TI * sizeof(double);
        }
    }
    // the tracers' launch, then (no tracers, or after tracer ranges) the masses' launch; the one
    // launch with every tracer or without tracers makes the fluxes when asked to
    auto go = [&](auto mass_only, auto makes, dim3 g, int tc, int nd) {
        constexpr bool MO = decltype(mass_only)::value, MK = decltype(makes)::value;
        static const char *const names[2][2] = {{"k_x<false, false>", "k_x<false, true>"},
                                                {"k_x<true, false>", "k_x<true, true>"}};
        TM5_LAUNCH_NAMED(ctx, names[MO][MK], (k_x<MO, MK>), g, block, lds, ctx->stream, im, jm, lm, m, q, am, limits,
                         nmax, ctx->words, rows, fail_bit, skip_mask, imred, amax_out, spec_k, backup, tc, halo_j0,
                         halo_j1, fs, nd, keep_am);
    };
    using T = std::true_type;
    using F = std::false_type;
    if (q.nt > 0) {
        if (nd0 && ranges == 1)
            go(F{}, T{}, dim3(grid.x + nd0), 0, (int)nd0);
        else
            go(F{}, F{}, dim3(grid.x, ranges), ranges > 1 ? tchunk : 0, 0);
    }
    if (q.nt == 0 && nd0)
        go(T{}, T{}, dim3(grid.x + nd0), 0, (int)nd0);
    else if (q.nt == 0 || ranges > 1)
        go(T{}, F{}, grid, 0, 0);
    return status(hipGetLastError());
}

// ------------------------------------------------------------------ y: dynamv's search
// pydyn/ops.py:315-341 with progressive scaling, per layer: nloop = 1, 2, ... until every
// interior face of the layer passes its trial; on a failure the whole layer's faces are scaled
// by (nloop-1)/nloop in place (the reference scales bm(:,:,l) whole).  The polar outer faces
// must be zero (pydyn/kernels.py:233-234).
constexpr int Y_FPT = 2;  // interior faces per thread in the y check at nloop = 1
constexpr int Y_NTH = 1024;  // threads of a y check workgroup

// grid (chunks, lm): every workgroup checks its chunk of layer l's faces at nloop = 1 (chunk 0
// also the polar faces) and or-s its verdict into the layer's word; the last workgroup of the
// layer to take its ticket reads the word, resets it and the ticket, and -- only where some
// face failed -- runs the sequential search for the layer.  ticket and verdict: lm ints each.
__global__ void __launch_bounds__(1024) k_y_search(int im, int jm, int lm, View3 m, View3 bm, int nmax, int *nloop,
                                                   int *ticket, int *verdict, int *words, int fail_bit, int skip_mask)
{
    const int l = blockIdx.y, chunk = blockIdx.x, chunks = gridDim.x;
    const bool skip = skipped(words, skip_mask);
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nface = im * (jm - 1);
    bool ok = true, bad = false;
    if (!skip) {
        if (chunk == 0)
            for (int i = tid; i < im; i += nth) bad |= (bm(i, 0, l) != 0.0) || (bm(i, jm, l) != 0.0);
        // nloop = 1: one Courant number per interior face, the loads independent of each other
#pragma unroll
        for (int r = 0; r < Y_FPT; r++) {
            const int idx = (chunk * Y_FPT + r) * nth + tid;
            if (idx < nface) {
                const int i = idx % im, k = idx / im + 1;
                const double bk = bm(i, k, l), mlo = m(i, k - 1, l), mhi = m(i, k, l);
                ok &= fabs(bk / (bk >= 0.0 ? mlo : mhi)) < 1.0;
            }
        }
    }
    ok = __syncthreads_and(ok);
    bad = __syncthreads_or(bad);
    __shared__ bool last;
    __shared__ int word;
    if (tid == 0) {
        if (!ok || bad) atomicOr(verdict + l, (ok ? 0 : 1) | (bad ? 2 : 0));
        __threadfence();
        last = atomicAdd(ticket + l, 1) == chunks - 1;
        if (last) {
            __threadfence();
            word = __hip_atomic_load(verdict + l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            verdict[l] = 0;
            ticket[l] = 0;
        }
    }
    __syncthreads();
    if (!last || skip) return;
    if (word & 2) {
        if (tid == 0) {
            atomicOr(words + W_FLAGS, TM5_BAD_Y);
            nloop[l] = 0;
        }
        return;
    }
    int nl = 1;
    if (word & 1) {
        // sub-cycling: the sequential search, the layer's faces scaled in place per retry
        __threadfence();
        for (;;) {
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
            ok = true;
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
        }
    }
    if (tid == 0) nloop[l] = nl;
}

// ------------------------------------------------------------------ y: dynamv's sub-steps
// pydyn/kernels.py:246-291 with the layer's nloop: one thread per cell of TI whole columns of
// layer l (blockIdx.y).  Each sub-step every thread reads its cell's neighbourhood (cells j-1,
// j, j+1 and faces j, j+1; independent loads), computes the fluxes of its two faces and its
// update, then -- after the workgroup's barrier, since the neighbours read the old values --
// writes it.  Along-flow slope rym; carried cross slopes rxm (ft) and rzm (fw).  Flow out of a
// pole cell carries no slope; the pole cells keep their rxm, rym.  The limiter of the first
// sub-step is taken on the values as they are read (it is the reference's initial limiter).
__device__ __forceinline__ FaceFlux y_face(double bf, double m_u, double rm, double rx, double ry, double rz, int k,
                                           int jm, bool lim0)
{
    // face k with flux bf and its upwind cell's mass and moments (ry as stored: the first
    // sub-step's limiter is taken here)
    FaceFlux ff{0.0, 0.0, 0.0, 0.0};
    if (k <= 0 || k >= jm) return ff;  // the closed polar faces
    const double beta = bf / m_u;
    if (lim0) ry = limit(ry, rm);
    ff = face_flux(bf, beta, rm, ry, rx, rz);
    if ((k == 1 && bf >= 0.0) || (k == jm - 1 && bf < 0.0)) {  // leaving a pole cell
        ff.f = beta * rm;
        ff.pf = (-3.0 * bf) * ff.f;
        ff.ft = 0.0;
        ff.fw = beta * rz;
    }
    return ff;
}

// With tracer ranges (tchunk > 0, grid z) several workgroups take the same lines: each moves the
// masses through the sub-steps on its own (in LDS), none writes them or restores bm, and the
// caller runs the mass-only form afterwards.
template <bool MASS_ONLY>
__global__ void __launch_bounds__(1024) k_y_steps(int im, int jm, int lm, int TI, View3 m, Tracers q, View3 bm,
                                                  const int *nloop, int limits, int *words, int skip_mask, int spec_k,
                                                  int tchunk)
{
    if (skipped(words, skip_mask)) return;
    extern __shared__ double yst[];  // 4 jm TI: one tracer's moments of the workgroup's cells; jm TI masses
    // tiles renumbered XCD-contiguously (as xcd_grid2), the tracer range slowest
    const int nbx = gridDim.x, nby = gridDim.y;
    const int b = xcd_tile((blockIdx.z * nby + blockIdx.y) * nbx + blockIdx.x, nbx * nby * gridDim.z);
    const int bx = b % nbx, l = (b / nbx) % nby, bz = b / (nbx * nby);
    const int t0 = tchunk > 0 ? bz * tchunk : 0;
    const int t1 = tchunk > 0 && t0 + tchunk < q.nt ? t0 + tchunk : q.nt;
    const bool writes_mass = MASS_ONLY || tchunk == 0;
    const int i0 = bx * TI;
    const int ti = threadIdx.x % TI, j = threadIdx.x / TI;
    const bool mine = j < jm && i0 + ti < im;
    const int i = mine ? i0 + ti : i0, jj = mine ? j : 0;
    const int nl = nloop ? nloop[l] : 1;  // no table: every layer at nloop 1
    if (checked(spec_k)) {
        bool ok = true, bad = false;
        if (mine) {
            const double b_lo = bm(i, jj, l);
            if (jj >= 1)
                ok = fabs(b_lo / (b_lo >= 0.0 ? m(i, jj - 1, l) : m(i, jj, l))) < 1.0;
            else
                bad = b_lo != 0.0;
            if (jj == jm - 1) bad = bad || bm(i, jm, l) != 0.0;
        }
        const bool any_bad = __any(bad), any_fail = __any(!ok);  // every lane votes
        if ((threadIdx.x & 63) == 0) {
            if (any_bad) atomicOr(words + W_FLAGS, TM5_BAD_Y);
            if (any_fail) spec_fail(words, spec_k);
        }
    }
    // the cells below and above (clamped at the poles, where their faces carry nothing)
    const int jlo = jj > 0 ? jj - 1 : 0, jhi = jj + 1 < jm ? jj + 1 : jm - 1;
    double m_cur = m(i, jj, l);
    for (int s = 0; s < nl; s++) {
        const double m_old = MASS_ONLY ? m(i, jj, l) : m_cur, b_lo = bm(i, jj, l), b_hi = bm(i, jj + 1, l);
        const double m_new = (m_old + b_lo) - b_hi;
        if (!MASS_ONLY) {
            const double inv_mnew = 1.0 / m_new;
            const double flow_conv = b_lo - b_hi;
            const double flow_mean = 3.0 * (b_lo + b_hi);
            const double mass_x3 = 3.0 * m_old;
            const bool lim0 = limits && s == 0;
            const bool pole = jj == 0 || jj == jm - 1;
            // both candidates of each face's upwind cell are read, independently of the flux signs:
            // the masses here, each tracer's moments from the workgroup's cells staged in LDS
            // ([moment][row][ti]; every cell read from HBM once)
            const int stride = jm * TI, me = jj * TI + ti, dn = jlo * TI + ti, up = jhi * TI + ti;
            double m_dn, m_up;
            if (tchunk > 0) {  // the line's masses at the sub-step's start, from LDS
                double *ym = yst + 4 * stride;
                if (mine) ym[me] = m_old;
                __syncthreads();
                m_dn = ym[dn];
                m_up = ym[up];
            } else {  // as written by the sub-step before
                m_dn = m(i, jlo, l);
                m_up = m(i, jhi, l);
            }
            for (int t = t0; t < t1; t++) {
                const long long o = q.off(i, jj, l, t);
                const double rm = q.q[0][o], rx0 = q.q[1][o], ry0 = q.q[2][o], rz0 = q.q[3][o];
                if (mine) {
                    yst[me] = rm;
                    yst[stride + me] = rx0;
                    yst[2 * stride + me] = ry0;
                    yst[3 * stride + me] = rz0;
                }
                __syncthreads();
                const double rm_d = yst[dn], rx_d = yst[stride + dn], ry_d = yst[2 * stride + dn], rz_d = yst[3 * stride + dn];
                const double rm_u = yst[up], rx_u = yst[stride + up], ry_u = yst[2 * stride + up], rz_u = yst[3 * stride + up];
                const bool lo_dn = b_lo >= 0.0, hi_own = b_hi >= 0.0;  // upwind: below / this cell
                const FaceFlux lo = y_face(b_lo, lo_dn ? m_dn : m_old, lo_dn ? rm_d : rm, lo_dn ? rx_d : rx0,
                                           lo_dn ? ry_d : ry0, lo_dn ? rz_d : rz0, jj, jm, lim0);
                const FaceFlux hi = y_face(b_hi, hi_own ? m_old : m_up, hi_own ? rm : rm_u, hi_own ? rx0 : rx_u,
                                           hi_own ? ry0 : ry_u, hi_own ? rz0 : rz_u, jj + 1, jm, lim0);
                double rym = ry0;
                if (lim0) rym = limit(rym, rm);
                const double rm_new = rm - (hi.f - lo.f);
                double ry = rym + ((((lo.pf - hi.pf) - flow_conv * rym) + flow_mean * rm_new) - mass_x3 * (lo.f + hi.f)) *
                                      inv_mnew;
                if (limits) ry = limit(ry, rm_new);
                const double rx = rx0 - (hi.ft - lo.ft);
                const double rz = rz0 - (hi.fw - lo.fw);
                if (mine) {  // no other thread reads the cell from memory
                    q.q[0][o] = rm_new;
                    q.q[3][o] = rz;
                    if (!pole) {
                        q.q[2][o] = ry;
                        q.q[1][o] = rx;
                    } else if (lim0) {
                        q.q[2][o] = rym;  // the pole keeps its (limited) slope
                    }
                }
                __syncthreads();  // the staged values are read before the next tracer's land
            }
            m_cur = m_new;
        }
        __syncthreads();
        if (mine && writes_mass) m(i, jj, l) = m_new;
        __threadfence_block();
        __syncthreads();
    }
    if (nl > 1 && mine && writes_mass) {  // "restore old bm": the search left the faces scaled
        bm(i, jj, l) = bm(i, jj, l) * (double)nl;
        if (jj == jm - 1) bm(i, jm, l) = bm(i, jm, l) * (double)nl;
    }
}

// columns (or z lines) of `len` cells per workgroup of at most 1024 threads
static int lines_per_block(int len, int cap) { return len > 1024 ? 0 : (1024 / len < cap ? 1024 / len : cap); }

static dim3 line_block(int TI, int len) { return dim3((unsigned)(((TI * len) + 63) / 64 * 64)); }

// known_one: the caller knows every layer passes at nloop 1 (a Check_CFL rehearsal at cap 1 on
// the same bits of m and bm passed), so the search is not launched
int launch_y(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 bm, int limits, int nmax, int fail_bit,
             int skip_mask, bool known_one, int spec_k)
{
    if (checked(spec_k)) known_one = true;
    if (im < 1 || jm < 2 || lm < 1 || lm > ctx->y_cap) return TM5_EINVAL;
    const bool mass_only = q.nt == 0;
    // many tracers: the widest tile, each tracer's rows read in the longest runs along i (the
    // context's tile is tuned for the one-tracer step)
    const int TI = lines_per_block(jm, ctx->tile[q.nt >= 8 ? TM5_TILE_Y_MANY : TM5_TILE_Y_LINES]);
    if (TI < 1) return TM5_EINVAL;
    const int chunks = (im * (jm - 1) + Y_NTH * Y_FPT - 1) / (Y_NTH * Y_FPT);
    if (!known_one)
        TM5_LAUNCH(ctx, k_y_search, dim3(chunks, lm), dim3(Y_NTH), 0, ctx->stream, im, jm, lm, m, bm, nmax,
                   ctx->y_nloop, ctx->y_nloop + ctx->y_cap, ctx->y_nloop + 2 * ctx->y_cap, ctx->words, fail_bit,
                   skip_mask);
    const int *nloop = known_one ? nullptr : ctx->y_nloop;
    const dim3 grid((im + TI - 1) / TI, lm), block = line_block(TI, jm);
    if (mass_only) {
        TM5_LAUNCH(ctx, k_y_steps<true>, grid, block, 0, ctx->stream, im, jm, lm, TI, m, q, bm, nloop, limits,
                   ctx->words, skip_mask, spec_k, 0);
        return status(hipGetLastError());
    }
    // many tracers: tracer ranges over the grid (one workgroup of these a CU), then the masses
    const int tchunk = q.nt >= 32 ? tracer_chunk((long long)grid.x * grid.y, ctx->ncu > 0 ? ctx->ncu : 256, q.nt,
                                                 ctx->tile[TM5_TILE_RANGE_MIN])
                                  : q.nt;
    const int ranges = (q.nt + tchunk - 1) / tchunk;
    TM5_LAUNCH(ctx, k_y_steps<false>, dim3(grid.x, grid.y, ranges), block, (size_t)5 * jm * TI * sizeof(double),
               ctx->stream, im, jm, lm, TI, m, q, bm, nloop, limits, ctx->words, skip_mask, spec_k,
               ranges > 1 ? tchunk : 0);
    if (ranges > 1)
        TM5_LAUNCH(ctx, k_y_steps<true>, grid, block, 0, ctx->stream, im, jm, lm, TI, m, q, bm, nloop, limits,
                   ctx->words, skip_mask, -1, 0);
    return status(hipGetLastError());
}

// ------------------------------------------------------------------ z: dynamw's search
// pydyn/ops.py:315-341 with fresh scaling (cm / n) over the whole region: the smallest n at which
// every column passes.  Each column finds its own smallest passing n >= 1; the largest of those
// is a candidate no smaller than the answer (every n below it fails some column).  The last
// workgroup to finish checks every column at the candidate and, if one fails there, raises the
// candidate to that column's next passing n, until all pass (the reference's sequential search,
// found as a fixpoint).
// `moves`: the trial advances the column's mass between sub-steps; a column outside a child's
// scope keeps its mass through the trial (pydyn/zoom/advect.py:415-441 advances the scope only).
__device__ bool z_column_passes(const View3 &m, const View3 &cm, int i, int j, int lm, int n, bool moves)
{
    const double dn = (double)n;
    for (int k = 1; k < lm; k++) {
        const double ck = cm(i, j, k) / dn;
        const int u = ck >= 0.0 ? k - 1 : k;
        const double clo = cm(i, j, u) / dn, chi = cm(i, j, u + 1) / dn;
        double mx = m(i, j, u);
        for (int s = 0; s < n; s++) {
            if (!(fabs(ck / mx) < 1.0)) return false;
            if (moves) mx = (mx + clo) - chi;
        }
    }
    return true;
}

__device__ int z_column_next(const View3 &m, const View3 &cm, int i, int j, int lm, int from, int nmax, bool moves)
{
    for (int n = from; n <= nmax; n++)
        if (z_column_passes(m, cm, i, j, lm, n, moves)) return n;
    return nmax + 1;
}

// the scope the trial advances, 0-based inclusive within the region's interior, less the columns
// of a hole (hi0..hi1, hj0..hj1; none by default) whose dynamw the caller runs elsewhere
struct ZScope {
    int i0, i1, j0, j1;
    int hi0 = 0, hi1 = -1, hj0 = 0, hj1 = -1;
    __device__ __forceinline__ bool has(int i, int j) const
    {
        return i >= i0 && i <= i1 && j >= j0 && j <= j1 && !(i >= hi0 && i <= hi1 && j >= hj0 && j <= hj1);
    }
};

constexpr int Z_FPT = 8;  // interior faces per thread in the z search

__global__ void __launch_bounds__(256) k_z_search(int im, int jm, int lm, View3 m, View3 cm, ZScope scope, int nmax,
                                                  int *ok_blocks, int *words, int fail_bit, int skip_mask)
{
    // every workgroup takes its ticket, skipping or not, so that the last one always resets it
    const bool skip = skipped(words, skip_mask);
    const int tid = threadIdx.x, nth = blockDim.x;
    bool ok = true;
    if (!skip) {
        // n = 1 (cm itself): one Courant number per interior face, the loads independent
        const int nface = im * jm * (lm - 1);  // < 2^31 (launch_z)
        const int base = blockIdx.x * nth * Z_FPT + tid;
#pragma unroll
        for (int r = 0; r < Z_FPT; r++) {
            const int idx = base + r * nth;
            if (idx < nface) {
                const int i = idx % im, rest = idx / im;
                const int j = rest % jm, k = rest / jm + 1;
                const double ck = cm(i, j, k), mlo = m(i, j, k - 1), mhi = m(i, j, k);
                ok &= fabs(ck / (ck >= 0.0 ? mlo : mhi)) < 1.0;
            }
        }
        const int col = blockIdx.x * nth + tid;  // the closed ends (pydyn/kernels.py:301-302)
        if (col < im * jm) {
            const int i = col % im, j = col / im;
            if (cm(i, j, 0) != 0.0 || cm(i, j, lm) != 0.0) atomicOr(words + W_FLAGS, TM5_BAD_Z);
        }
    }
    ok = __syncthreads_and(ok);
    __shared__ bool last;
    if (tid == 0) {
        __hip_atomic_store(ok_blocks + blockIdx.x, ok ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        last = atomicAdd(words + W_Z_TICKET, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    bool all = true;
    for (int b = tid; b < (int)gridDim.x; b += nth)
        all &= __hip_atomic_load(ok_blocks + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    all = __syncthreads_and(all);
    // n = 1 failed somewhere: the sequential search as a fixpoint over the columns (rare)
    __shared__ int red[256];
    int cand = all ? 1 : 2;
    const int ncol = im * jm;
    while (!skip && !all && cand <= nmax) {
        int need = cand;
        for (int c = tid; c < ncol; c += nth) {
            const int i = c % im, j = c / im;
            const bool moves = scope.has(i, j);
            if (!z_column_passes(m, cm, i, j, lm, cand, moves))
                need = max(need, z_column_next(m, cm, i, j, lm, cand + 1, nmax, moves));
        }
        red[tid] = need;
        __syncthreads();
        for (int s = nth / 2; s > 0; s >>= 1) {
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
    }
}

// ------------------------------------------------------------------ z: dynamw's sub-steps
// pydyn/kernels.py:324-346 with c = cm / n_advz (cm itself not written): one thread per cell of
// TI whole columns (i0.., j = blockIdx.y), the neighbourhood read into registers, the update
// written after the workgroup's barrier.  Along-flow slope rzm; cross slopes rxm (ft), rym (fw).
// A face's flux, its Courant number c / m(upwind) and upwind cell are the same for every tracer
// of a sub-step and are found once; each tracer's moments are staged in LDS for the neighbours.
// With tracer ranges (tchunk > 0, grid z) several workgroups take the same columns: each moves
// the masses through the sub-steps on its own (in LDS), none writes them, and the caller runs
// the mass-only form afterwards.
template <bool MASS_ONLY>
__global__ void __launch_bounds__(1024) k_z_steps(int im, int jm, int lm, int TI, View3 m, Tracers q, View3 cm,
                                                  int limits, int prelimit, int *words, int skip_mask, int fixed_n,
                                                  int spec_k, int tchunk, ZScope sc)
{
    // sc: the columns the sub-steps move (the grid's columns outside it are only tested, when checked)
    // prelimit: dynamw's limiter on rzm before the first sub-step (advect_z's, not the bare sub-steps')
    // fixed_n > 0: n_advz given by the caller, else the search's (the context's word)
    // checked(spec_k) (fixed_n 1): a checked step's dynamw, first testing its faces as k_z_mass1 does
    if (skipped(words, skip_mask)) return;
    extern __shared__ double zst[];  // 4 lm TI: one tracer's moments of the workgroup's cells; lm TI masses
    // tiles renumbered XCD-contiguously (as xcd_grid2), the tracer range slowest
    const int nbx = gridDim.x, nby = gridDim.y;
    const int b = xcd_tile((blockIdx.z * nby + blockIdx.y) * nbx + blockIdx.x, nbx * nby * gridDim.z);
    const int bx = b % nbx, j = (b / nbx) % nby, bz = b / (nbx * nby);
    const int t0 = tchunk > 0 ? bz * tchunk : 0;
    const int t1 = tchunk > 0 && t0 + tchunk < q.nt ? t0 + tchunk : q.nt;
    const int i0 = bx * TI;
    const int ti = threadIdx.x % TI, l = threadIdx.x / TI;
    const bool mine = l < lm && i0 + ti < im;
    const int i = mine ? i0 + ti : i0, ll = mine ? l : 0;
    const int nz = fixed_n > 0 ? fixed_n : words[W_Z_NADV];
    if (checked(spec_k)) {
        bool ok = true, bad = false;
        if (mine) {
            if (ll >= 1) {
                const double ck = cm(i, j, ll);
                ok = fabs(ck / (ck >= 0.0 ? m(i, j, ll - 1) : m(i, j, ll))) < 1.0;
            } else {
                bad = cm(i, j, 0) != 0.0 || cm(i, j, lm) != 0.0;
            }
        }
        const bool any_bad = __any(bad), any_fail = __any(!ok);  // every lane votes
        if ((threadIdx.x & 63) == 0) {
            if (any_bad) atomicOr(words + W_FLAGS, TM5_BAD_Z);
            if (any_fail) spec_fail(words, spec_k);
        }
    }
    if (j < sc.j0 || j > sc.j1 || i0 > sc.i1 || i0 + TI - 1 < sc.i0) return;  // a workgroup only tested
    if (j >= sc.hj0 && j <= sc.hj1 && i0 >= sc.hi0 && i0 + TI - 1 <= sc.hi1) return;  // inside the hole
    const bool act = sc.has(i0 + ti, j);
    const double dn = (double)nz;
    const double c_lo = cm(i, j, ll) / dn, c_hi = cm(i, j, ll + 1) / dn;
    // each face's upwind cell (the closed ends carry nothing)
    const bool has_lo = ll >= 1, has_hi = ll + 1 <= lm - 1;
    const int u_lo = c_lo >= 0.0 ? (has_lo ? ll - 1 : ll) : ll, u_hi = c_hi >= 0.0 ? ll : (has_hi ? ll + 1 : ll);
    double m_cur = m(i, j, ll);
    for (int s = 0; s < nz; s++) {
        const double m_old = MASS_ONLY ? m(i, j, ll) : m_cur;
        const double m_new = (m_old + c_lo) - c_hi;
        if (!MASS_ONLY) {
            const bool lim0 = prelimit && s == 0;
            // the Courant numbers of the sub-step, the same for every tracer, from the column's
            // masses at the sub-step's start (staged after the tracers' space)
            double *zm = zst + 4 * lm * TI;
            if (mine) zm[ll * TI + ti] = m_old;
            __syncthreads();
            const double g_lo = c_lo / zm[u_lo * TI + ti], g_hi = c_hi / zm[u_hi * TI + ti];
            // each tracer's moments of the workgroup's cells pass through LDS ([moment][level][ti]):
            // every cell is read from HBM once, its upwind neighbours from LDS, and the new values
            // go straight back (no other thread reads them from memory)
            const int stride = lm * TI, me = ll * TI + ti;
            // each tracer's moments are loaded while the tracer before it is worked on
            double nx[4] = {0.0, 0.0, 0.0, 0.0};
            if (t0 < t1) {
                const long long o0 = q.off(i, j, ll, t0);
                for (int k = 0; k < 4; k++) nx[k] = q.q[k][o0];
            }
            for (int t = t0; t < t1; t++) {
                const long long o = q.off(i, j, ll, t);
                const double rm = nx[0], rx0 = nx[1], ry0 = nx[2], rz0 = nx[3];
                if (t + 1 < t1) {
                    for (int k = 0; k < 4; k++) nx[k] = q.q[k][o + q.st];
                }
                if (mine) {
                    zst[me] = rm;
                    zst[stride + me] = rx0;
                    zst[2 * stride + me] = ry0;
                    zst[3 * stride + me] = rz0;
                }
                __syncthreads();
                FaceFlux lo{0.0, 0.0, 0.0, 0.0}, hi{0.0, 0.0, 0.0, 0.0};
                if (mine && has_lo) {
                    const int u = u_lo * TI + ti;
                    const double rmu = zst[u], rzu = zst[3 * stride + u];
                    lo = face_flux(c_lo, g_lo, rmu, lim0 ? limit(rzu, rmu) : rzu, zst[stride + u], zst[2 * stride + u]);
                }
                if (mine && has_hi) {
                    const int u = u_hi * TI + ti;
                    const double rmu = zst[u], rzu = zst[3 * stride + u];
                    hi = face_flux(c_hi, g_hi, rmu, lim0 ? limit(rzu, rmu) : rzu, zst[stride + u], zst[2 * stride + u]);
                }
                const double rzm = lim0 ? limit(rz0, rm) : rz0;
                const double rm_new = (rm + lo.f) - hi.f;  // advance: rm + f(l-1) - f(l)
                double rz = slope_update(rzm, lo.pf, hi.pf, c_lo, c_hi, rm_new, lo.f, hi.f, m_old, m_new);
                if (limits) rz = limit(rz, rm_new);
                if (mine && act) {
                    q.q[0][o] = rm_new;
                    q.q[1][o] = (rx0 + lo.ft) - hi.ft;
                    q.q[2][o] = (ry0 + lo.fw) - hi.fw;
                    q.q[3][o] = rz;
                }
                __syncthreads();  // the staged values are read before the next tracer's land
            }
            m_cur = m_new;
        }
        __syncthreads();
        if (mine && act && (MASS_ONLY || tchunk == 0)) m(i, j, ll) = m_new;
        __threadfence_block();
        __syncthreads();
    }
}

// The search over the whole region (the trial advancing only the scope's columns), then the
// sub-steps on the scope: region 1's scope is the region, a child's its block's.
__global__ void k_set_nadv(int *words, int n, int skip_mask)
{
    if (threadIdx.x == 0 && !skipped(words, skip_mask)) words[W_Z_NADV] = n;
}

// ------------------------------------------------------------------ Check_CFL's y and z at cap 1
// The rehearsal allows dynamvm and dynamwm no sub-cycling (max_nloop 1, pydyn/zoom/check_cfl.py:
// 258-290): a trial passes only if every interior face has |Courant| < 1 at nloop = 1, and then
// its one mass-only sub-step is the operator.  So the check needs no search: each face is tested
// where its column is updated, a failure raises the rehearsal's bit (the sub-steps already done
// are then discarded with the trial), and the closed ends are checked as the searches do.
// TI whole columns of a layer (y) or of a row (z) per workgroup: every read of the workgroup's
// cells precedes its writes.
__global__ void __launch_bounds__(1024) k_y_mass1(int im, int jm, int lm, int TI, View3 m, View3 bm, int *words,
                                                  int fail_bit, int skip_mask)
{
    if (skipped(words, skip_mask)) return;
    int bx, l;
    xcd_grid2(bx, l);
    const int i0 = bx * TI;
    const int ti = threadIdx.x % TI, j = threadIdx.x / TI;
    const bool mine = j < jm && i0 + ti < im;
    const int i = mine ? i0 + ti : i0, jj = mine ? j : 0;
    const double m_old = m(i, jj, l), b_lo = bm(i, jj, l), b_hi = bm(i, jj + 1, l);
    bool ok = true, bad = false;
    if (mine) {
        if (jj >= 1) {
            const double mlo = m(i, jj - 1, l);
            ok = fabs(b_lo / (b_lo >= 0.0 ? mlo : m_old)) < 1.0;
        } else {
            bad = b_lo != 0.0;
        }
        if (jj == jm - 1) bad = bad || b_hi != 0.0;
    }
    const double m_new = (m_old + b_lo) - b_hi;
    __syncthreads();
    if (mine) m(i, jj, l) = m_new;
    const bool any_bad = __any(bad), any_fail = __any(!ok);  // every lane votes
    if ((threadIdx.x & 63) == 0) {
        if (any_bad) atomicOr(words + W_FLAGS, TM5_BAD_Y);
        if (any_fail) atomicOr(words + W_FLAGS, fail_bit);
    }
}

__global__ void __launch_bounds__(1024) k_z_mass1(int im, int jm, int lm, int TI, View3 m, View3 cm, int *words,
                                                  int fail_bit, int skip_mask)
{
    if (skipped(words, skip_mask)) return;
    int bx, j;
    xcd_grid2(bx, j);
    const int i0 = bx * TI;
    const int ti = threadIdx.x % TI, l = threadIdx.x / TI;
    const bool mine = l < lm && i0 + ti < im;
    const int i = mine ? i0 + ti : i0, ll = mine ? l : 0;
    const double m_old = m(i, j, ll);
    const double c_lo = cm(i, j, ll) / 1.0, c_hi = cm(i, j, ll + 1) / 1.0;  // cm / n_advz at n = 1
    bool ok = true, bad = false;
    if (mine) {
        if (ll >= 1) {
            const double ck = cm(i, j, ll), mlo = m(i, j, ll - 1);
            ok = fabs(ck / (ck >= 0.0 ? mlo : m_old)) < 1.0;
        } else {
            bad = cm(i, j, 0) != 0.0 || cm(i, j, lm) != 0.0;
        }
    }
    const double m_new = (m_old + c_lo) - c_hi;
    __syncthreads();
    if (mine) m(i, j, ll) = m_new;
    const bool any_bad = __any(bad), any_fail = __any(!ok);  // every lane votes
    if ((threadIdx.x & 63) == 0) {
        if (any_bad) atomicOr(words + W_FLAGS, TM5_BAD_Z);
        if (any_fail) atomicOr(words + W_FLAGS, fail_bit);
    }
}

int launch_y_mass1(tm5_ctx *ctx, int im, int jm, int lm, View3 m, View3 bm, int fail_bit, int skip_mask)
{
    if (im < 1 || jm < 2 || lm < 1) return TM5_EINVAL;
    const int TI = lines_per_block(jm, ctx->tile[TM5_TILE_Y_LINES]);
    if (TI < 1) return TM5_EINVAL;
    TM5_LAUNCH(ctx, k_y_mass1, dim3((im + TI - 1) / TI, lm), line_block(TI, jm), 0, ctx->stream, im, jm, lm, TI, m,
               bm, ctx->words, fail_bit, skip_mask);
    return status(hipGetLastError());
}

int launch_z_mass1(tm5_ctx *ctx, int im, int jm, int lm, View3 m, View3 cm, int fail_bit, int skip_mask)
{
    if (im < 1 || jm < 1 || lm < 2) return TM5_EINVAL;
    const int TI = lines_per_block(lm, ctx->tile[TM5_TILE_Z_LINES]);
    if (TI < 1) return TM5_EINVAL;
    TM5_LAUNCH(ctx, k_z_mass1, dim3((im + TI - 1) / TI, jm), line_block(TI, lm), 0, ctx->stream, im, jm, lm, TI, m,
               cm, ctx->words, fail_bit, skip_mask);
    return status(hipGetLastError());
}

int launch_z_scoped(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 cm, int i0, int i1, int j0, int j1,
                    int limits, int nmax, int fail_bit, int skip_mask, bool known_one, int spec_k, int hi0, int hi1,
                    int hj0, int hj1)
{
    if (checked(spec_k)) known_one = true;
    if (im < 1 || jm < 1 || lm < 2 || (long long)im * jm * (lm + 1) >= (1LL << 31)) return TM5_EINVAL;
    if (i0 < 0 || j0 < 0 || i1 >= im || j1 >= jm || i1 < i0 || j1 < j0) return TM5_EINVAL;
    const bool mass_only = q.nt == 0;
    // with a hole, tiles of three columns: the workgroups inside it return at once
    const bool holed = hi0 <= hi1 && hj0 <= hj1;
    const int TI = lines_per_block(lm, holed ? 3 : ctx->tile[TM5_TILE_Z_LINES]);
    if (TI < 1) return TM5_EINVAL;
    const long long nface = (long long)im * jm * (lm - 1);
    long long nb = (nface + 256LL * Z_FPT - 1) / (256LL * Z_FPT);
    const long long nb_cols = ((long long)im * jm + 255) / 256;
    if (nb < nb_cols) nb = nb_cols;
    if (nb > ctx->z_cap) return TM5_EINVAL;
    if (holed && !known_one) return TM5_EINVAL;  // a hole: n_advz 1 known
    const ZScope scope{i0, i1, j0, j1, hi0, hi1, hj0, hj1};
    if (!known_one)  // as launch_y's known_one: n_advz 1 without a search
        TM5_LAUNCH(ctx, k_z_search, dim3((unsigned)nb), dim3(256), 0, ctx->stream, im, jm, lm, m, cm, scope, nmax,
                   ctx->z_ok, ctx->words, fail_bit, skip_mask);
    const int fixed_n = known_one ? 1 : 0;
    // the scope's views, or -- a checked step on a scope smaller than the region, whose test covers
    // every column of the region as the rehearsal's trial does -- the region's with the scope
    const bool whole = checked(spec_k) && (i0 > 0 || j0 > 0 || i1 < im - 1 || j1 < jm - 1);
    const int si = whole ? im : i1 - i0 + 1, sj = whole ? jm : j1 - j0 + 1;
    const int o0 = whole ? 0 : i0, o1 = whole ? 0 : j0;
    View3 ms{m.p + o0 * m.si + o1 * m.sj, m.si, m.sj, m.sl}, cs{cm.p + o0 * cm.si + o1 * cm.sj, cm.si, cm.sj, cm.sl};
    Tracers qs = q;
    for (int a = 0; a < 4 && q.nt > 0; a++) qs.q[a] = q.q[a] + (o0 * q.si + o1 * q.sj);
    const ZScope moved = whole ? scope : ZScope{0, si - 1, 0, sj - 1, hi0 - o0, hi1 - o0, hj0 - o1, hj1 - o1};
    const dim3 grid((si + TI - 1) / TI, sj), block = line_block(TI, lm);
    // the tiles that move columns (the kernel's others, only tested or inside the hole, return
    // at once)
    long long tiles = 0;
    for (int by = moved.j0; by <= moved.j1 && by < (int)grid.y; by++) {
        for (int bx = 0; bx < (int)grid.x; bx++) {
            const int t0 = bx * TI, t1 = t0 + TI - 1;
            if (t0 > moved.i1 || t1 < moved.i0) continue;
            if (by >= moved.hj0 && by <= moved.hj1 && t0 >= moved.hi0 && t1 <= moved.hi1) continue;
            tiles++;
        }
    }
    // a hole over the whole scope moves nothing: no launch unless it carries a checked step's test
    if (tiles == 0 && !checked(spec_k)) return status(hipGetLastError());
    if (mass_only) {
        TM5_LAUNCH(ctx, k_z_steps<true>, grid, block, 0, ctx->stream, si, sj, lm, TI, ms, qs, cs, limits, limits,
                   ctx->words, skip_mask, fixed_n, spec_k, 0, moved);
        return status(hipGetLastError());
    }
    // many tracers: tracer ranges over the grid (one workgroup of these a CU; of a hole's
    // three-column tiles twelve: two waves each at the kernel's six waves a SIMD), then the
    // masses.  The ranges are sized for the tiles that move columns; with none (a checked step's
    // test alone) one range.
    const long long slots = (long long)(ctx->ncu > 0 ? ctx->ncu : 256) * (holed ? 12 : 1);
    const int tchunk = q.nt >= 32 && tiles > 0 ? tracer_chunk(tiles, slots, q.nt, ctx->tile[TM5_TILE_RANGE_MIN])
                                               : q.nt;
    const int ranges = (q.nt + tchunk - 1) / tchunk;
    TM5_LAUNCH(ctx, k_z_steps<false>, dim3(grid.x, grid.y, ranges), block, (size_t)(4 * lm * TI + lm * TI) * sizeof(double),
               ctx->stream, si, sj, lm, TI, ms, qs, cs, limits, limits, ctx->words, skip_mask, fixed_n, spec_k,
               ranges > 1 ? tchunk : 0, moved);
    if (ranges > 1)
        TM5_LAUNCH(ctx, k_z_steps<true>, grid, block, 0, ctx->stream, si, sj, lm, TI, ms, qs, cs, limits, limits,
                   ctx->words, skip_mask, fixed_n, SPEC_NONE, 0, moved);
    return status(hipGetLastError());
}

int launch_z(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 cm, int limits, int nmax, int fail_bit,
             int skip_mask, bool known_one, int spec_k)
{
    return launch_z_scoped(ctx, im, jm, lm, m, q, cm, 0, im - 1, 0, jm - 1, limits, nmax, fail_bit, skip_mask,
                           known_one, spec_k);
}

// dynamw's sub-steps alone with a given n_advz (left in the context's word for the kernel)
int launch_z_substeps(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 cm, int n_advz, int limits,
                      int prelimit)
{
    if (im < 1 || jm < 1 || lm < 2 || n_advz < 1 || q.nt < 1) return TM5_EINVAL;
    const int TI = lines_per_block(lm, ctx->tile[TM5_TILE_Z_LINES]);
    if (TI < 1) return TM5_EINVAL;
    TM5_LAUNCH(ctx, k_set_nadv, dim3(1), dim3(64), 0, ctx->stream, ctx->words, n_advz, TM5_ERRORS);
    TM5_LAUNCH(ctx, k_z_steps<false>, dim3((im + TI - 1) / TI, jm), line_block(TI, lm),
               (size_t)(4 * lm * TI + lm * TI) * sizeof(double), ctx->stream, im, jm, lm, TI,
               m, q, cm, limits, prelimit, ctx->words, TM5_ERRORS, 0, SPEC_NONE, 0, ZScope{0, im - 1, 0, jm - 1});
    return status(hipGetLastError());
}

}  // namespace tm5
