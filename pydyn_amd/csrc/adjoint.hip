// The adjoint of the advection operators' tracer map (dynamu, dynamv, dynamw).
//
// Given the air-mass trajectory, each sub-step maps the tracers (rm, and the slopes along and
// across the line) linearly, except the slope limiter clip(s, -|rm|, |rm|), whose derivative is
// ds where it passes s and +-sign(rm) drm where it clips.  One 64-lane workgroup per line (an x
// row, a y column of a layer, a z column): the line's forward input is staged in LDS, and for
// each sub-step, last to first, the forward is replayed bit-exactly from the input up to that
// sub-step (its fluxes, Courant numbers and limiter branches are what the transpose needs), then
// the transposed sub-step is applied to the lambdas.  Sub-cycle counts are small, so replaying
// costs less than storing every sub-step's state.
//
// Along a line the tracer arrays are (rm, s, t, w): s the slope along the line, t and w the two
// carried across it -- x: (rm, rxm, rym, rzm); y: (rm, rym, rxm, rzm); z: (rm, rzm, rxm, rym).
// Cell c lies between faces c and c+1.  x lines are periodic (face 0 is face n), y and z lines
// closed (faces 0 and n carry nothing); y's pole cells keep their s and t, and the flux leaving
// a pole cell carries no slope (pydyn/kernels.py:246-291).
#include <string.h>

#include "tm5_lib.h"

namespace tm5 {

__device__ __forceinline__ bool skipped_adj(const int *words, int mask) { return (words[W_FLAGS] & mask) != 0; }

enum { LX = 0, LY = 1, LZ = 2 };

struct Line {
    double *q[4];      // forward input tracers, along-line order, element of cell 0
    double *lam[4];    // lambdas (in: after the operator, out: before), same addressing
    const double *m;   // forward input mass of cell 0
    const double *a;   // faces before the operator, face 0
    long long cs, ls, ms, fs;  // strides (elements): tracer cells, lambda cells, mass cells, faces
};

// the arrays of one line: lines are x rows (j, l), y columns (i, l), z columns (i, j)
__device__ __forceinline__ Line line_of(int kind, int line, int im, int jm, View3 m, Tracers q, Tracers lam, View3 flux,
                                        int t)
{
    int i = 0, j = 0, l = 0;
    if (kind == LX) {
        j = line % jm, l = line / jm;
    } else if (kind == LY) {
        i = line % im, l = line / im;
    } else {
        i = line % im, j = line / im;
    }
    static const int order[3][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}, {0, 3, 1, 2}};
    Line L;
    const long long o = q.off(i, j, l, t), ol = lam.off(i, j, l, t);
    for (int k = 0; k < 4; k++) {
        L.q[k] = q.q[order[kind][k]] + o;
        L.lam[k] = lam.q[order[kind][k]] + ol;
    }
    L.m = &m(i, j, l);
    L.a = &flux(i, j, l);
    L.cs = kind == LX ? q.si : kind == LY ? q.sj : q.sl;
    L.ls = kind == LX ? lam.si : kind == LY ? lam.sj : lam.sl;
    L.ms = kind == LX ? m.si : kind == LY ? m.sj : m.sl;
    L.fs = kind == LX ? flux.si : kind == LY ? flux.sj : flux.sl;
    return L;
}

// The upwind cell and the Courant number of face k, as the forward computes them: x alpha =
// a (1/m), y beta = b / m, z gamma = c / m.
__device__ __forceinline__ void face_coeff(int kind, int n, int k, const double *A, const double *M, int &u, double &c)
{
    const double ak = A[k];
    if (kind == LX) {
        u = ak >= 0.0 ? k - 1 : (k == n ? 0 : k);
        c = ak * (1.0 / M[u]);
    } else {
        u = ak >= 0.0 ? k - 1 : k;
        c = ak / M[u];
    }
}

// One forward sub-step of the line in LDS (the forward kernels' arithmetic).  analyze: keep the
// state, record per face the Courant number (AL) and per cell m_new (MN), rm_new (RN) and the
// limiter's branch (CODE); else advance M and Q.
__device__ void fwd_substep(int kind, int n, bool limits, bool analyze, double *M, double *Q, const double *A, double *F,
                            double *AL, double *MN, double *RN, double *CODE)
{
    const int lane = threadIdx.x;
    const int nf = n + 1;
    const int k0 = 1, k1 = kind == LX ? n : n - 1;
    for (int k = k0 + lane; k <= k1; k += 64) {
        int u;
        double c;
        face_coeff(kind, n, k, A, M, u, c);
        FaceFlux ff = face_flux(A[k], c, Q[u], Q[n + u], Q[2 * n + u], Q[3 * n + u]);
        if (kind == LY && ((k == 1 && A[k] >= 0.0) || (k == n - 1 && A[k] < 0.0))) {  // leaving a pole cell
            ff.f = c * Q[u];
            ff.pf = (-3.0 * A[k]) * ff.f;
            ff.ft = 0.0;
            ff.fw = c * Q[3 * n + u];
        }
        F[k] = ff.f;
        F[nf + k] = ff.pf;
        F[2 * nf + k] = ff.ft;
        F[3 * nf + k] = ff.fw;
        AL[k] = c;
    }
    __syncthreads();
    if (lane < 4) {
        if (kind == LX) {
            F[lane * nf] = F[lane * nf + n];  // the periodic seam
        } else {
            F[lane * nf] = 0.0;  // the closed ends
            F[lane * nf + n] = 0.0;
        }
    }
    __syncthreads();
    const double *f = F, *pf = F + nf, *ft = F + 2 * nf, *fw = F + 3 * nf;
    for (int i = lane; i < n; i += 64) {
        const double m_old = M[i], a_lo = A[i], a_hi = A[i + 1];
        const double m_new = (m_old + a_lo) - a_hi;
        const double rm = Q[i], s = Q[n + i], t = Q[2 * n + i], w = Q[3 * n + i];
        double rm_new, s_raw, t_new, w_new;
        if (kind == LZ) {
            rm_new = (rm + f[i]) - f[i + 1];
            s_raw = slope_update(s, pf[i], pf[i + 1], a_lo, a_hi, rm_new, f[i], f[i + 1], m_old, m_new);
            t_new = (t + ft[i]) - ft[i + 1];
            w_new = (w + fw[i]) - fw[i + 1];
        } else if (kind == LX) {
            rm_new = rm - (f[i + 1] - f[i]);
            s_raw = slope_update(s, pf[i], pf[i + 1], a_lo, a_hi, rm_new, f[i], f[i + 1], m_old, m_new);
            t_new = t - (ft[i + 1] - ft[i]);
            w_new = w - (fw[i + 1] - fw[i]);
        } else {
            const double inv = 1.0 / m_new;
            const double conv = a_lo - a_hi, mean = 3.0 * (a_lo + a_hi), mx3 = 3.0 * m_old;
            rm_new = rm - (f[i + 1] - f[i]);
            s_raw = s + ((((pf[i] - pf[i + 1]) - conv * s) + mean * rm_new) - mx3 * (f[i] + f[i + 1])) * inv;
            t_new = t - (ft[i + 1] - ft[i]);
            w_new = w - (fw[i + 1] - fw[i]);
        }
        const double s_new = limits ? limit(s_raw, rm_new) : s_raw;
        const bool pole = kind == LY && (i == 0 || i == n - 1);
        if (analyze) {
            MN[i] = m_new;
            RN[i] = rm_new;
            double code = 0.0;
            if (limits) {
                const double hi = fabs(rm_new), lo = -hi;
                const bool low = !(s_raw > lo);
                const double tt = low ? lo : s_raw;
                code = !(tt < hi) ? 2.0 : (low ? 1.0 : 0.0);
            }
            CODE[i] = code;
        } else {
            // every cell reads only its own values and the fluxes: in place after the barrier
            M[i] = m_new;
            Q[i] = rm_new;
            Q[3 * n + i] = w_new;
            if (!pole) {
                Q[n + i] = s_new;
                Q[2 * n + i] = t_new;
            }
        }
    }
    __syncthreads();
}

// the limiter's transpose: (lambda of s, lambda added to rm) from lambda of the clipped value
__device__ __forceinline__ void lim_ad(double code, double lam, double r, double &ls, double &lr)
{
    const double sg = r > 0.0 ? 1.0 : (r < 0.0 ? -1.0 : 0.0);
    ls = code == 0.0 ? lam : 0.0;
    lr = code == 1.0 ? -sg * lam : (code == 2.0 ? sg * lam : 0.0);
}

// The transpose of one sub-step, applied to the lambdas L (4n) in place; the forward quantities
// of the sub-step (M = m_old, A, AL, MN, RN, CODE) from fwd_substep(analyze).  LF, CF: face
// scratch (4 nf each); QV, LRN, LS: cell scratch.
__device__ void ad_substep(int kind, int n, bool limits, const double *M, const double *A, const double *AL,
                           const double *MN, const double *RN, const double *CODE, double *L, double *LF, double *CF,
                           double *QV, double *LRN, double *KEEP)
{
    const int lane = threadIdx.x;
    const int nf = n + 1;
    // cells: the limiter, the slope update's coefficients
    for (int i = lane; i < n; i += 64) {
        double ls_new = L[n + i], lt = L[2 * n + i];
        const bool pole = kind == LY && (i == 0 || i == n - 1);
        KEEP[2 * i] = pole ? ls_new : 0.0;  // a pole keeps its s and t
        KEEP[2 * i + 1] = pole ? lt : 0.0;
        if (pole) ls_new = lt = 0.0;
        double ls_raw = ls_new, lr = 0.0;
        if (limits) lim_ad(CODE[i], ls_new, RN[i], ls_raw, lr);
        const double a_lo = A[i], a_hi = A[i + 1];
        const double qv = kind == LY ? ls_raw * (1.0 / MN[i]) : ls_raw / MN[i];
        const double lrn = (L[i] + lr) + 3.0 * (a_lo + a_hi) * qv;
        QV[i] = qv;
        LRN[i] = lrn;
        L[n + i] = ls_raw - (a_lo - a_hi) * qv;  // s's own share
        L[2 * n + i] = lt;
    }
    __syncthreads();
    // faces: the lambdas of the fluxes from the two cells around them
    for (int k = lane; k < nf; k += 64) {
        double lf = 0.0, lpf = 0.0, lft = 0.0, lfw = 0.0;
        if (k < n) {  // face k is the lower face of cell k
            lf = lf + (LRN[k] - 3.0 * M[k] * QV[k]);
            lpf = lpf + QV[k];
            lft = lft + L[2 * n + k];
            lfw = lfw + L[3 * n + k];
        }
        if (k >= 1) {  // and the upper face of cell k-1
            lf = lf + (-LRN[k - 1] - 3.0 * M[k - 1] * QV[k - 1]);
            lpf = lpf - QV[k - 1];
            lft = lft - L[2 * n + k - 1];
            lfw = lfw - L[3 * n + k - 1];
        }
        LF[k] = lf;
        LF[nf + k] = lpf;
        LF[2 * nf + k] = lft;
        LF[3 * nf + k] = lfw;
    }
    __syncthreads();
    if (kind == LX && lane < 4) {  // face 0 carries face n's fluxes
        LF[lane * nf + n] += LF[lane * nf];
        LF[lane * nf] = 0.0;
    }
    __syncthreads();
    // faces: the flux computations' transposes, into the face's upwind cell
    const int k0 = 1, k1 = kind == LX ? n : n - 1;
    for (int k = lane; k < nf; k += 64) {
        double crm = 0.0, cs = 0.0, ct = 0.0, cw = 0.0;
        if (k >= k0 && k <= k1) {
            const double ak = A[k], c = AL[k];
            const double lpf = LF[nf + k], lf = LF[k] + (-3.0 * ak) * lpf;
            const bool special = kind == LY && ((k == 1 && ak >= 0.0) || (k == n - 1 && ak < 0.0));
            if (special) {
                crm = c * lf;
                cw = c * LF[3 * nf + k];
            } else {
                const double g = ak >= 0.0 ? c * (1.0 - c) : -(c * (1.0 + c));
                crm = c * lf;
                cs = ak * (c * c) * lpf + g * lf;
                ct = c * LF[2 * nf + k];
                cw = c * LF[3 * nf + k];
            }
        }
        CF[k] = crm;
        CF[nf + k] = cs;
        CF[2 * nf + k] = ct;
        CF[3 * nf + k] = cw;
    }
    __syncthreads();
    // cells: gather from the faces whose upwind cell they are
    for (int i = lane; i < n; i += 64) {
        double add[4] = {0.0, 0.0, 0.0, 0.0};
        // the face above (i+1) takes cell i upwind when its flow is >= 0
        const int kh = i + 1;
        if (kh <= k1 && A[kh] >= 0.0)
            for (int v = 0; v < 4; v++) add[v] = add[v] + CF[v * nf + kh];
        // the face below takes it when its flow is < 0 (x: cell 0's lower face is face n)
        const int kl = (kind == LX && i == 0) ? n : i;
        if (kl >= k0 && A[kl] < 0.0)
            for (int v = 0; v < 4; v++) add[v] = add[v] + CF[v * nf + kl];
        L[i] = LRN[i] + add[0];
        L[n + i] = (L[n + i] + add[1]) + KEEP[2 * i];
        L[2 * n + i] = (L[2 * n + i] + add[2]) + KEEP[2 * i + 1];
        L[3 * n + i] = L[3 * n + i] + add[3];
    }
    __syncthreads();
}

// The adjoint of one operator over lines of n cells.  nloop: per x row, per y layer, or the
// single z count; the faces the forward used are the operator's input faces scaled by each
// retry of the search ((k-1)/k, k = 2..nloop; x and y) or divided by n_advz (z).
template <int KIND>
__global__ void __launch_bounds__(64) k_line_ad(int im, int jm, int lm, View3 m, Tracers q, Tracers lam, View3 flux,
                                                const int *nloop, int limits, int *words, int skip_mask)
{
    if (skipped_adj(words, skip_mask)) return;
    extern __shared__ double lds[];
    const int n = KIND == LX ? im : KIND == LY ? jm : lm;
    const int nf = n + 1;
    const int line = blockIdx.x, lane = threadIdx.x;
    const int nl = KIND == LX ? nloop[line] : KIND == LY ? nloop[line / im] : nloop[0];
    double *M0 = lds, *M = M0 + n, *MN = M + n, *RN = MN + n, *CODE = RN + n, *QV = CODE + n, *LRN = QV + n;
    double *R0 = LRN + n, *C0 = R0 + n, *KEEP = C0 + n;  // KEEP: 2n
    double *Q0 = KEEP + 2 * n, *Q = Q0 + 4 * n, *L = Q + 4 * n;
    double *A = L + 4 * n, *AL = A + nf, *F = AL + nf, *LF = F + 4 * nf, *CF = LF + 4 * nf;

    const Line ln0 = line_of(KIND, line, im, jm, m, q, lam, flux, 0);
    for (int i = lane; i < n; i += 64) M0[i] = ln0.m[i * ln0.ms];
    for (int k = lane; k < nf; k += 64) {
        double a = ln0.a[k * ln0.fs];
        if (KIND == LZ) {
            a = a / (double)nl;
        } else {
            for (int r = 2; r <= nl; r++) a = shrink(a, r);
        }
        A[k] = a;
        AL[k] = 0.0;
    }
    __syncthreads();
    for (int t = 0; t < q.nt; t++) {
        const Line ln = line_of(KIND, line, im, jm, m, q, lam, flux, t);
        for (int i = lane; i < n; i += 64) {
            const double rm = ln.q[0][i * ln.cs], s = ln.q[1][i * ln.cs];
            R0[i] = rm;
            double code = 0.0, s_lim = s;
            if (limits) {
                const double hi = fabs(rm), lo = -hi;
                const bool low = !(s > lo);
                const double tt = low ? lo : s;
                code = !(tt < hi) ? 2.0 : (low ? 1.0 : 0.0);
                s_lim = limit(s, rm);
            }
            C0[i] = code;
            Q0[i] = rm;
            Q0[n + i] = s_lim;
            Q0[2 * n + i] = ln.q[2][i * ln.cs];
            Q0[3 * n + i] = ln.q[3][i * ln.cs];
            for (int v = 0; v < 4; v++) L[v * n + i] = ln.lam[v][i * ln.ls];
        }
        __syncthreads();
        for (int sb = nl - 1; sb >= 0; sb--) {
            for (int i = lane; i < n; i += 64) {
                M[i] = M0[i];
                for (int v = 0; v < 4; v++) Q[v * n + i] = Q0[v * n + i];
            }
            __syncthreads();
            for (int s = 0; s < sb; s++) fwd_substep(KIND, n, limits, false, M, Q, A, F, AL, MN, RN, CODE);
            fwd_substep(KIND, n, limits, true, M, Q, A, F, AL, MN, RN, CODE);
            ad_substep(KIND, n, limits, M, A, AL, MN, RN, CODE, L, LF, CF, QV, LRN, KEEP);
        }
        // the initial limiter's transpose, and the lambdas before the operator
        for (int i = lane; i < n; i += 64) {
            double ls = L[n + i], lr = 0.0;
            if (limits) lim_ad(C0[i], L[n + i], R0[i], ls, lr);
            ln.lam[0][i * ln.ls] = L[i] + lr;
            ln.lam[1][i * ln.ls] = ls;
            ln.lam[2][i * ln.ls] = L[2 * n + i];
            ln.lam[3][i * ln.ls] = L[3 * n + i];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_copy_view3(int n0, int n1, int n2, View3 src, View3 dst);

static size_t line_ad_lds(int n) { return (size_t)(23 * n + 14 * (n + 1)) * sizeof(double); }

// The operator's adjoint: the sub-cycle counts of its forward found again on copies (x per row,
// y per layer, z for the region; the search kernels scale the copies, not the inputs), then one
// line per workgroup.  work: device scratch of (im+1)(jm+1)(lm+1) + im jm lm doubles.
int launch_advect_adjoint(tm5_ctx *ctx, int axis, int im, int jm, int lm, View3 m, Tracers q, View3 flux, Tracers lam,
                          int limits, int max_nloop, double *work)
{
    if (im < 1 || jm < 2 || lm < 2 || axis < 0 || axis > 2 || !work || q.nt != lam.nt || q.nt < 1) return TM5_EINVAL;
    const int n = axis == 0 ? im : axis == 1 ? jm : lm;
    if (line_ad_lds(n) > 64 * 1024) return TM5_EINVAL;
    const long long fi = axis == 0 ? im + 1 : im, fj = axis == 1 ? jm + 1 : jm, fl = axis == 2 ? lm + 1 : lm;
    View3 wm{work, 1, im, (long long)im * jm}, wf{work + (long long)im * jm * lm, 1, fi, fi * fj};
    TM5_LAUNCH(ctx, k_copy_view3, dim3((unsigned)(((long long)im * jm * lm + 255) / 256)), dim3(256), 0, ctx->stream,
               im, jm, lm, m, wm);
    TM5_LAUNCH(ctx, k_copy_view3, dim3((unsigned)((fi * fj * fl + 255) / 256)), dim3(256), 0, ctx->stream, (int)fi,
               (int)fj, (int)fl, flux, wf);
    const int *nloop = nullptr;
    int lines = 0;
    if (axis == 0) {
        if (launch_x(ctx, im, jm, lm, wm, no_tracers(), wf, 0, max_nloop - 1, TM5_FAIL_X, TM5_ERRORS)) return TM5_EINVAL;
        nloop = ctx->x_rows;
        lines = jm * lm;
    } else if (axis == 1) {
        if (launch_y(ctx, im, jm, lm, wm, no_tracers(), wf, 0, max_nloop - 1, TM5_FAIL_Y, TM5_ERRORS)) return TM5_EINVAL;
        nloop = ctx->y_nloop;
        lines = im * lm;
    } else {
        if (launch_z(ctx, im, jm, lm, wm, no_tracers(), wf, 0, max_nloop, TM5_FAIL_Z, TM5_ERRORS)) return TM5_EINVAL;
        nloop = ctx->words + W_Z_NADV;
        lines = im * jm;
    }
    const size_t lds = line_ad_lds(n);
    if (axis == 0)
        TM5_LAUNCH(ctx, k_line_ad<LX>, dim3((unsigned)lines), dim3(64), lds, ctx->stream, im, jm, lm, m, q, lam, flux,
                   nloop, limits, ctx->words, TM5_ERRORS);
    else if (axis == 1)
        TM5_LAUNCH(ctx, k_line_ad<LY>, dim3((unsigned)lines), dim3(64), lds, ctx->stream, im, jm, lm, m, q, lam, flux,
                   nloop, limits, ctx->words, TM5_ERRORS);
    else
        TM5_LAUNCH(ctx, k_line_ad<LZ>, dim3((unsigned)lines), dim3(64), lds, ctx->stream, im, jm, lm, m, q, lam, flux,
                   nloop, limits, ctx->words, TM5_ERRORS);
    return status(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_copy_view3(int n0, int n1, int n2, View3 src, View3 dst)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n0 * n1 * n2) return;
    const int i = (int)(idx % n0), j = (int)((idx / n0) % n1), l = (int)(idx / ((long long)n0 * n1));
    dst(i, j, l) = src(i, j, l);
}

}  // namespace tm5

using namespace tm5;

extern "C" int tm5_advect_adjoint(tm5_ctx *c, int axis, int im, int jm, int lm, const tm5_view *m,
                                  const tm5_tracers *q, const tm5_view *flux, const tm5_tracers *lam, int limits,
                                  int max_nloop, double *work)
{
    if (!c || !m || !q || !flux || !lam || !m->data || !flux->data) return TM5_EINVAL;
    return launch_advect_adjoint(c, axis, im, jm, lm, view3(m), tracers(q), view3(flux), tracers(lam), limits,
                                 max_nloop, work);
}

extern "C" int tm5_column_adjoint(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *lam,
                                  const tm5_zphysics *p, const char *ops, const tm5_view *lam_emis)
{
    if (!c || !m || !lam || !p || !ops || !lam_emis || !column_supported(p->lmax) || lm > 64 || p->lmax > lm)
        return TM5_EINVAL;
    ColumnArgs a{};
    if (!strcmp(ops, "v")) {
    } else if (!strcmp(ops, "vsccsv") || !strcmp(ops, "vscMcsv")) {
        a.twice = 1;
    } else if (!strcmp(ops, "csv")) {
        a.pre = 1;
    } else if (!strcmp(ops, "vsc")) {
        a.post = 1;
    } else {
        return TM5_EINVAL;
    }
    a.im = im;
    a.jm = jm;
    a.lm = lm;
    a.m = view3(m);
    a.q = tracers(lam);
    a.entu = view3(&p->entu);
    a.detu = view3(&p->detu);
    a.entd = view3(&p->entd);
    a.detd = view3(&p->detd);
    a.dkg = view3(&p->dkg);
    a.top = iview2(&p->cloud_top);
    a.lfs = iview2(&p->cloud_lfs);
    a.dt = p->dt_conv;
    a.emis = view3(&p->emis);
    a.factor = p->factor;
    a.ratio = p->ratio;
    a.efrac = p->efrac;
    a.level = p->level - 1;
    a.tracer = p->tracer - 1;
    a.escale = p->escale;
    a.loss = view3(&p->loss);
    a.loss_sr = p->loss_stride;
    a.dt_chem = p->dt_chem;
    a.adjoint = 1;
    a.lam_emis = view3(lam_emis);
    return launch_column(c, a, p->lmax, TM5_ERRORS);
}
