// The 'v' operator: per column, the convection + Kzz diffusion matrix conv1, its explicit
// inverse C by Crout LU without pivoting and substitution against the identity, slopes times
// diag(C), rm <- C rm (pydyn/kernels.py:350-457).
//
// One column per 64-lane wavefront, four wavefronts per workgroup, the column's matrices in
// LDS.  Lane c owns matrix column c while the matrix is built (the up/downdraft recursions are
// sequential in the level, their scalars computed redundantly by every lane), the LU is
// right-looking with lanes over the trailing columns, and the substitutions run with lane c
// solving right-hand side e_c.  Every element still receives its updates in increasing k, the
// reference's order, so the factors are bit-identical to its left-looking loops.
#include "tm5_lib.h"

namespace tm5 {

__device__ __forceinline__ bool skipped_c(const int *words, int mask)
{
    return (words[W_FLAGS] & mask) != 0;  // written by earlier kernels of the stream
}

static inline int lead(int n) { return (n % 2 == 0) ? n + 1 : n; }
static inline size_t wave_lds(int n) { return (size_t)(2 * (n + 1) * lead(n) + 7 * n + 2 * (n + 1)) * sizeof(double); }

__global__ void __launch_bounds__(256) k_convec(int im, int jm, int lmax, int wpb, View3 m, Tracers q, View3 entu,
                                                View3 detu, View3 entd, View3 detd, IView2 top, IView2 lfs, View3 dkg,
                                                double dt, int *words, int skip_mask)
{
    if (skipped_c(words, skip_mask)) return;
    extern __shared__ double lds[];
    const int n = lmax, ld = (n % 2 == 0) ? n + 1 : n;
    const int wave = threadIdx.x >> 6, c = threadIdx.x & 63;
    const int ncol = im * jm;
    int col = blockIdx.x * wpb + wave;
    const bool active = col < ncol;
    if (!active) col = ncol - 1;
    const int i = col % im, j = col / im;
    const size_t per_wave = 2 * (size_t)(n + 1) * ld + 7 * n + 2 * (n + 1);
    double *U = lds + wave * per_wave;   // fu, then conv1 and its LU
    double *D = U + (n + 1) * ld;         // fd, then f, then Y and C
    double *mc = D + (n + 1) * ld;        // the column's m, entu, detu, entd, detd, dkg
    double *eu = mc + n, *du = eu + n, *ed = du + n, *dd = ed + n, *dk = dd + n;
    double *old = dk + n;                 // a tracer's rm before the update
    double *amu = old + n, *amd = amu + (n + 1);

    if (c < n) {
        mc[c] = m(i, j, c);
        eu[c] = entu(i, j, c);
        du[c] = detu(i, j, c);
        ed[c] = entd(i, j, c);
        dd[c] = detd(i, j, c);
        dk[c] = dkg(i, j, c);
    }
    for (int x = c; x < (n + 1) * ld; x += 64) U[x] = D[x] = 0.0;
    if (c <= n) amu[c] = amd[c] = 0.0;
    __syncthreads();

    const int ctop = top(i, j), clfs = lfs(i, j);
    // updraft, k = 1..n: fu(k, c) for the column c this lane owns (kernels.py:366-374)
    {
        double run = 0.0, amu_prev = 0.0;
        for (int k = 1; k <= n; k++) {
            const bool on = ctop >= k;
            const double inflow = amu_prev + eu[k - 1];
            const double a = inflow - du[k - 1];
            const bool rising = a > 0.0;
            const double keep = rising ? max0(1.0 - du[k - 1] / inflow) : 0.0;
            const double amu_k = (on && rising) ? a : 0.0;
            if (c < n) {
                double v = 0.0;
                if (c < k - 1)
                    v = on ? run * keep : 0.0;
                else if (c == k - 1)
                    v = on ? eu[k - 1] / mc[k - 1] * keep : 0.0;
                U[k * ld + c] = v;
                run = v;
            }
            if (c == 0) amu[k] = amu_k;
            amu_prev = amu_k;
        }
    }
    // downdraft, k = n..2: fd(k-1, c) (kernels.py:376-383)
    {
        double run = 0.0, amd_prev = 0.0;
        for (int k = n; k >= 2; k--) {
            const bool on = clfs >= k;
            const double inflow = amd_prev - ed[k - 1];
            const double a = inflow + dd[k - 1];
            const bool sinking = a < 0.0;
            const double keep = sinking ? max0(1.0 + dd[k - 1] / inflow) : 0.0;
            const double amd_k = (on && sinking) ? a : 0.0;
            if (c < n) {
                double v = 0.0;
                if (c >= k)
                    v = on ? run * keep : 0.0;
                else if (c == k - 1)
                    v = on ? -ed[k - 1] / mc[k - 1] * keep : 0.0;
                D[(k - 1) * ld + c] = v;
                run = v;
            }
            if (c == 0) amd[k - 1] = amd_k;
            amd_prev = amd_k;
        }
    }
    __syncthreads();
    // f = fu + fd with subsidence and diffusion through the interfaces k = 1..n-1 (:385-391)
    if (c < n) {
        for (int k = 1; k < n; k++) {
            double v = U[k * ld + c] + D[k * ld + c];
            if (c == k) v = v - amu[k] / mc[k];
            if (c == k - 1) v = v - amd[k] / mc[k - 1];
            if (c == k - 1) v = v + dk[k - 1] / mc[k - 1];
            if (c == k) v = v - dk[k - 1] / mc[k];
            D[k * ld + c] = v;
        }
    }
    __syncthreads();
    // conv1 (:392-395), into U
    if (c < n) {
        for (int k = 0; k < n; k++) {
            double v = (-dt) * (D[k * ld + c] - D[(k + 1) * ld + c]);
            if (c == k) v = v + 1.0;
            U[k * ld + c] = v;
        }
    }
    // Crout LU without pivoting, right-looking (:404-407 reordered, same per-element sequence)
    for (int k = 0; k < n; k++) {
        __syncthreads();
        const double piv = U[k * ld + k];
        if (c > k && c < n) U[c * ld + k] = U[c * ld + k] / piv;
        __syncthreads();
        if (c > k && c < n) {
            const double ukc = U[k * ld + c];
            for (int r = k + 1; r < n; r++) U[r * ld + c] = U[r * ld + c] - U[r * ld + k] * ukc;
        }
    }
    __syncthreads();
    // forward and back substitution of e_c, lane c's own column of D (:409-419)
    if (c < n) {
        for (int r = 0; r < n; r++) D[r * ld + c] = (r == c) ? 1.0 : 0.0;
        for (int r = 1; r < n; r++) {
            double acc = D[r * ld + c];
            for (int s = 0; s < r; s++) acc = acc - U[r * ld + s] * D[s * ld + c];
            D[r * ld + c] = acc;
        }
        for (int r = n - 1; r >= 0; r--) {
            double acc = D[r * ld + c];
            for (int s = r + 1; s < n; s++) acc = acc - U[r * ld + s] * D[s * ld + c];
            D[r * ld + c] = acc / U[r * ld + r];
        }
    }
    __syncthreads();
    // slopes times diag(C); rm <- C rm accumulated in k from zero (:448-457)
    for (int t = 0; t < q.nt; t++) {
        if (c < n) old[c] = q.q[0][q.off(i, j, c, t)];
        __syncthreads();
        if (c < n && active) {
            double acc = 0.0;
            for (int k = 0; k < n; k++) acc = acc + D[c * ld + k] * old[k];
            const double d = D[c * ld + c];
            const long long o = q.off(i, j, c, t);
            q.q[0][o] = acc;
            q.q[1][o] = d * q.q[1][o];
            q.q[2][o] = d * q.q[2][o];
            q.q[3][o] = d * q.q[3][o];
        }
        __syncthreads();
    }
}

int launch_convec(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 entu, View3 detu, View3 entd,
                  View3 detd, IView2 top, IView2 lfs, View3 dkg, double dt, int lmax)
{
    if (im < 1 || jm < 1 || lmax < 1 || lmax > lm || lmax > 48) return TM5_EINVAL;
    const size_t per = wave_lds(lmax);
    int wpb = (int)((60 * 1024) / per);
    if (wpb > 4) wpb = 4;
    if (wpb < 1) return TM5_EINVAL;
    const int ncol = im * jm;
    const dim3 grid((ncol + wpb - 1) / wpb), block(64 * wpb);
    TM5_LAUNCH(ctx, k_convec, grid, block, per * wpb, ctx->stream, im, jm, lmax, wpb, m, q, entu, detu, entd, detd,
                       top, lfs, dkg, dt, ctx->words, TM5_ERRORS);
    return status(hipGetLastError());
}

}  // namespace tm5
