// dynam0, m2phlb1, dkg, the CH4 injection and decay, and a scale: per-column or per-cell
// kernels, one thread per column/cell with i the fastest index (coalesced for i-fastest views).
#include "tm5_lib.h"

namespace tm5 {

__device__ __forceinline__ bool skipped_p(const int *words, int mask)
{
    return (words[W_FLAGS] & mask) != 0;  // written by earlier kernels of the stream
}

// ------------------------------------------------------------------ dynam0
// (the tile of dynam0 is tm5_lib.h's dynam0_tile, shared with the x kernel's fused form)
__global__ void __launch_bounds__(256) k_dynam0(int im, int jm, int lm, int TI, View3 pu, View3 pv, TimeSrc ts,
                                                const double *bt, double dt, View3 am, View3 bm, View3 cm, int cyclic,
                                                int *words, int skip_mask, View3 pu_out, View3 pv_out, int am_halo)
{
    if (skipped_p(words, skip_mask)) return;
    extern __shared__ double conv[];  // [lm][TI]
    dynam0_tile(blockIdx.x, blockIdx.y, conv, im, jm, lm, TI, pu, pv, ts, bt, dt, am, bm, cm, cyclic, words, pu_out,
                pv_out, am_halo);
}

// ------------------------------------------------------------------ m2phlb1
// pydyn/zoom/massflow.py:282-290 on a tile of TI columns: the column masses staged in LDS,
// sp = seqsum(m over l) g / dxyp(j) per column, then phlb = at + bt sp in parallel.
// tile (bx, j) of m2phlb1 (sh: its LDS, (lm + 1) * TI doubles); mcopy (when mcopy.p is set): a
// copy of m on the way
__device__ __forceinline__ void m2phlb1_tile(int bx, int j, double *sh, int im, int lm, int TI, View3 m,
                                             const double *dxyp, const double *at, const double *bt, View3 sp,
                                             View3 phlb, View3 mcopy)
{
    double *M = sh, *P = sh + lm * TI;  // M [lm][TI], P [TI]
    const int i0 = bx * TI;
    const int ni = min(TI, im - i0);
    const int tid = threadIdx.x, nth = blockDim.x;
#pragma unroll 8
    for (int idx = tid; idx < lm * TI; idx += nth) {
        const int ti = idx % TI, l = idx / TI;
        if (ti < ni) {
            const double v = m(i0 + ti, j, l);
            M[idx] = v;
            if (mcopy.p) mcopy(i0 + ti, j, l) = v;
        }
    }
    __syncthreads();
    for (int ti = tid; ti < ni; ti += nth) {
        double p = M[ti];
        for (int l = 1; l < lm; l++) p = p + M[l * TI + ti];
        p = p * 9.80665 / dxyp[j];
        P[ti] = p;
        sp(i0 + ti, j, 0) = p;
    }
    __syncthreads();
    for (int idx = tid; idx < (lm + 1) * TI; idx += nth) {
        const int ti = idx % TI, l = idx / TI;
        if (ti < ni) phlb(i0 + ti, j, l) = at[l] + bt[l] * P[ti];
    }
}

__global__ void __launch_bounds__(256) k_m2phlb1(int im, int jm, int lm, int TI, View3 m, const double *dxyp,
                                                 const double *at, const double *bt, View3 sp, View3 phlb, int *words,
                                                 int skip_mask, View3 mcopy)
{
    // TI columns i per workgroup (TM5_TILE_M2PHLB)
    if (skipped_p(words, skip_mask)) return;
    extern __shared__ double sh[];
    m2phlb1_tile(blockIdx.x, blockIdx.y, sh, im, lm, TI, m, dxyp, at, bt, sp, phlb, mcopy);
}

// ------------------------------------------------------------------ cluster sums
// The reduced grid's seqsum over clusters (pydyn/redgrid.py:56-62 through ops.seqsum): out(k, a,
// b) = q(k c, a, b) + q(k c + 1, a, b) + ... + q(k c + c - 1, a, b), summed from the first term.
__global__ void __launch_bounds__(256) k_cluster_sum(int n, int c, int r1, int r2, View3 q, View3 out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * r1 * r2) return;
    const int k = (int)(t % n), a = (int)((t / n) % r1), b = (int)(t / ((long long)n * r1));
    double acc = q(k * c, a, b);
    for (int s = 1; s < c; s++) acc = acc + q(k * c + s, a, b);
    out(k, a, b) = acc;
}

int launch_cluster_sum(tm5_ctx *ctx, int n, int c, int r1, int r2, View3 q, View3 out)
{
    if (n < 1 || c < 1 || r1 < 1 || r2 < 1) return TM5_EINVAL;
    const long long total = (long long)n * r1 * r2;
    TM5_LAUNCH(ctx, k_cluster_sum, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, n, c, r1, r2, q,
               out);
    return status(hipGetLastError());
}

// ------------------------------------------------------------------ dkg
// pydyn/diffusion.py:256-263: dkg(l) = max(0, kvh(l)) 2 (m(l+1) + m(l)) / dz^2, dz = gph(l+2) -
// gph(l), for l < lmax - 1; zero at lmax - 1.
__global__ void __launch_bounds__(256) k_dkg(int im, int jm, int lmax, View3 kvh, View3 m, View3 gph, View3 dkg)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= im * jm * lmax) return;
    const int i = cell % im, j = (cell / im) % jm, l = cell / (im * jm);
    if (l == lmax - 1) {
        dkg(i, j, l) = 0.0;
        return;
    }
    const double dz = gph(i, j, l + 2) - gph(i, j, l);
    dkg(i, j, l) = max0(kvh(i, j, l)) * 2.0 * (m(i, j, l + 1) + m(i, j, l)) / (dz * dz);
}

// ------------------------------------------------------------------ sources: do_add_2d
// pydyn/emis.py:52-54 inject, pydyn/zoom/physics.py:516-525: rm += x, rzm -= x.
__global__ void __launch_bounds__(256) k_add_2d(int im, int jm, View3 emis, double factor, double ratio,
                                                double efrac, Tracers q, int lev, int t, int *words, int skip_mask)
{
    if (skipped_p(words, skip_mask)) return;
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= im * jm) return;
    const int i = col % im, j = col / im;
    const double x = ((emis(i, j, 0) * factor) * ratio) * efrac;
    const long long o = q.off(i, j, lev, t);
    q.q[0][o] = q.q[0][o] + x;
    q.q[3][o] = q.q[3][o] - x;
}

// ------------------------------------------------------------------ chemistry: decay
// pydyn/chem.py:168-178: factor = exp(-(((L1 + L2) + L3) + L4) dt), rm..rzm of the tracers times it.
__global__ void __launch_bounds__(256) k_decay(int im, int jm, int lm, View3 loss, long long sr, double dt, Tracers q,
                                               int t0, int count, int *words, int skip_mask)
{
    if (skipped_p(words, skip_mask)) return;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= im * jm * lm) return;
    const int i = cell % im, j = (cell / im) % jm, l = cell / (im * jm);
    const double *L = &loss(i, j, l);
    const double total = ((L[0] + L[sr]) + L[2 * sr]) + L[3 * sr];
    const double f = fdlibm_exp(-(total * dt));
    for (int t = t0; t < t0 + count; t++) {
        const long long o = q.off(i, j, l, t);
        for (int a = 0; a < 4; a++) q.q[a][o] = q.q[a][o] * f;
    }
}

__global__ void __launch_bounds__(256) k_scale(int n0, int n1, int n2, View3 a, double factor)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n0 * n1 * n2) return;
    const int i = (int)(idx % n0), j = (int)((idx / n0) % n1), l = (int)(idx / ((long long)n0 * n1));
    a(i, j, l) = a(i, j, l) * factor;
}

static dim3 blocks(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// ------------------------------------------------------------------ the libm route
__global__ void __launch_bounds__(256) k_libm(int fn, long long n, const double *x, const double *y, double *out)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double v = x[k];
    out[k] = fn == TM5_LIBM_EXP ? fdlibm_exp(v) : fn == TM5_LIBM_LOG ? fdlibm_log(v) : fdlibm_pow(v, y[k]);
}

int launch_libm(tm5_ctx *ctx, int fn, long long n, const double *x, const double *y, double *out)
{
    if (n < 0 || n > (1LL << 40) || !x || !out || fn < TM5_LIBM_EXP || fn > TM5_LIBM_POW || (fn == TM5_LIBM_POW && !y))
        return TM5_EINVAL;
    if (n == 0) return 0;
    TM5_LAUNCH(ctx, k_libm, blocks(n), dim3(256), 0, ctx->stream, fn, n, x, y, out);
    return status(hipGetLastError());
}

int launch_dynam0(tm5_ctx *ctx, int im, int jm, int lm, View3 pu, View3 pv, double a, const double *bt, double dt,
                  View3 am, View3 bm, View3 cm, int skip_mask, int cyclic, const Dynam0Extra *ex)
{
    const int TI = ctx->tile[TM5_TILE_DYNAM0];
    if (im < 1 || jm < 1 || lm < 1 || lm * TI > 8192 || !bt) return TM5_EINVAL;
    const dim3 grid((im + TI - 1) / TI, jm);
    Dynam0Extra x{};
    if (ex) x = *ex;
    else x.ts.a = a;
    TM5_LAUNCH(ctx, k_dynam0, grid, dim3(256), (size_t)lm * TI * sizeof(double), ctx->stream, im, jm, lm, TI, pu, pv,
               x.ts, bt, dt, am, bm, cm, cyclic, ctx->words, skip_mask, x.pu_out, x.pv_out, x.am_halo);
    return status(hipGetLastError());
}

// the regions of a k_dynam0_regions launch: region k's tiles are workgroups first[k] .. first[k+1]-1
struct Dynam0Set {
    int n, TI;
    int first[D0_REGIONS + 1];
    const double *bt;
    Dynam0Region r[D0_REGIONS];
};

__global__ void __launch_bounds__(256) k_dynam0_regions(Dynam0Set set, int *words, int skip_mask)
{
    if (skipped_p(words, skip_mask)) return;
    extern __shared__ double conv[];  // [lm][TI] of the workgroup's region
    int k = set.n - 1;
    while (k > 0 && (int)blockIdx.x < set.first[k]) k--;
    const Dynam0Region &g = set.r[k];
    const int t = (int)blockIdx.x - set.first[k], ntx = (g.im + set.TI - 1) / set.TI;
    dynam0_tile(t % ntx, t / ntx, conv, g.im, g.jm, g.lm, set.TI, g.pu, g.pv, g.ex.ts, set.bt, g.dt, g.am, g.bm, g.cm,
                g.cyclic, words, g.ex.pu_out, g.ex.pv_out, g.ex.am_halo);
}

int launch_dynam0_regions(tm5_ctx *ctx, int n, const Dynam0Region *regions, const double *bt, int skip_mask)
{
    const int TI = ctx->tile[TM5_TILE_DYNAM0];
    if (n < 0 || !bt || (n > 0 && !regions)) return TM5_EINVAL;
    for (int k0 = 0; k0 < n; k0 += D0_REGIONS) {  // D0_REGIONS regions a launch
        Dynam0Set set{};
        set.n = n - k0 < D0_REGIONS ? n - k0 : D0_REGIONS;
        set.TI = TI;
        set.bt = bt;
        int lm_max = 1;
        for (int k = 0; k < set.n; k++) {
            const Dynam0Region &g = regions[k0 + k];
            if (g.im < 1 || g.jm < 1 || g.lm < 1 || g.lm * TI > 8192) return TM5_EINVAL;
            set.r[k] = g;
            if (!g.ex.ts.t0) set.r[k].ex.ts.a = g.a;
            set.first[k + 1] = set.first[k] + ((g.im + TI - 1) / TI) * g.jm;
            lm_max = g.lm > lm_max ? g.lm : lm_max;
        }
        TM5_LAUNCH(ctx, k_dynam0_regions, dim3((unsigned)set.first[set.n]), dim3(256), (size_t)lm_max * TI * sizeof(double),
                   ctx->stream, set, ctx->words, skip_mask);
        const int rc = status(hipGetLastError());
        if (rc) return rc;
    }
    return 0;
}

int launch_m2phlb1(tm5_ctx *ctx, int im, int jm, int lm, View3 m, const double *dxyp, const double *at,
                   const double *bt, View3 sp, View3 phlb, const View3 *mcopy)
{
    const int TI = ctx->tile[TM5_TILE_M2PHLB];
    if (im < 1 || jm < 1 || lm < 1 || (lm + 1) * TI > 8192 || !dxyp || !at || !bt) return TM5_EINVAL;
    const dim3 grid((im + TI - 1) / TI, jm);
    const View3 none{};
    TM5_LAUNCH(ctx, k_m2phlb1, grid, dim3(256), (size_t)(lm + 1) * TI * sizeof(double), ctx->stream, im, jm, lm, TI,
               m, dxyp, at, bt, sp, phlb, ctx->words, TM5_ERRORS, mcopy ? *mcopy : none);
    return status(hipGetLastError());
}

int launch_dkg(tm5_ctx *ctx, int im, int jm, View3 kvh, View3 m, View3 gph, int lmax, View3 dkg)
{
    if (im < 1 || jm < 1 || lmax < 1) return TM5_EINVAL;
    TM5_LAUNCH(ctx, k_dkg, blocks((long long)im * jm * lmax), dim3(256), 0, ctx->stream, im, jm, lmax, kvh, m, gph,
                       dkg);
    return status(hipGetLastError());
}

int launch_add_2d(tm5_ctx *ctx, int im, int jm, View3 emis, double factor, double ratio, double efrac, Tracers q,
                  int level, int tracer)
{
    if (im < 1 || jm < 1 || level < 1 || tracer < 1 || tracer > q.nt) return TM5_EINVAL;
    TM5_LAUNCH(ctx, k_add_2d, blocks((long long)im * jm), dim3(256), 0, ctx->stream, im, jm, emis, factor, ratio,
                       efrac, q, level - 1, tracer - 1, ctx->words, TM5_ERRORS);
    return status(hipGetLastError());
}

int launch_decay(tm5_ctx *ctx, int im, int jm, int lm, View3 loss, long long sr, double dt, Tracers q, int t0,
                 int count)
{
    if (im < 1 || jm < 1 || lm < 1 || t0 < 0 || count < 0 || t0 + count > q.nt) return TM5_EINVAL;
    TM5_LAUNCH(ctx, k_decay, blocks((long long)im * jm * lm), dim3(256), 0, ctx->stream, im, jm, lm, loss, sr, dt,
                       q, t0, count, ctx->words, TM5_ERRORS);
    return status(hipGetLastError());
}

// ------------------------------------------------------------------ CH4 loss rates
// pydyn/chem.py:129-170 (ch4_loss_kernel, then decay) for one column (i, j) per thread: the zonal
// columns' rebin onto the model layers by the Fortran loop of rebin_levels(ip_mix) -- each layer's
// overlaps from zero in order, a running j_start, the exit at the first input layer above it, the
// mean where the layer has thickness; the layers below the first edge take the first value -- the
// tropospheric Cl likewise on its own edges (pl_trop, 0), the Spivakovsky OH and Cl rates by the
// fdlibm exp and pow, the choice at the tropopause, the soil sink at the surface, and the decay
// factor exp(-(((L1 + L2) + L3) + L4) dt).  Layers that are not strictly descending raise
// TM5_BAD_LEVELS (the reference's ValueError).
__global__ void __launch_bounds__(256) k_loss_rates(LossArgs a, int *words)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= a.im * a.jm) return;
    const int i = col % a.im, j = col / a.im, lm = a.lm, ns = a.nstrat, nt = a.ntrop;
    const bool trop = a.pl.p != nullptr;
    bool bad = false;
    for (int l = 0; l < lm; l++) bad = bad || !(a.phlb(i, j, l + 1) < a.phlb(i, j, l));
    if (trop) {
        for (int k = 0; k + 1 < nt; k++) bad = bad || !(a.pl(i, j, k + 1) < a.pl(i, j, k));
        bad = bad || !(0.0 < a.pl(i, j, nt - 1));
    }
    if (bad) {
        atomicOr(words + W_FLAGS, TM5_BAD_LEVELS);
        return;
    }
    const double *E = a.edges + (size_t)j * (ns + 1);
    const double *SOH = a.s_oh + (size_t)j * ns, *SCL = a.s_cl + (size_t)j * ns, *SOD = a.s_od + (size_t)j * ns;
    const double OH_A = 2.45e-12, OH_EA = 1775.0, CL_A = 2.36e-12, CL_T0 = 298.0, CL_N = 1.37, CL_EA = 939.0;
    int js = 0, jt = 0;
    for (int l = 0; l < lm; l++) {
        const double op_bot = a.phlb(i, j, l), op_top = a.phlb(i, j, l + 1), width = op_bot - op_top;
        while (js < ns && E[js + 1] >= op_bot) js++;
        double s_oh = 0.0, s_cl = 0.0, s_od = 0.0;
        for (int k = js; k < ns; k++) {
            const double ip_bot = E[k], ip_top = E[k + 1];
            if (ip_bot <= op_top) break;
            const double thick = (op_bot < ip_bot ? op_bot : ip_bot) - (op_top > ip_top ? op_top : ip_top);
            if (thick > 0.0) {
                s_oh = s_oh + SOH[k] * thick;
                s_cl = s_cl + SCL[k] * thick;
                s_od = s_od + SOD[k] * thick;
            }
        }
        if (width > 0.0) {
            s_oh = s_oh / width;
            s_cl = s_cl / width;
            s_od = s_od / width;
        }
        if (op_bot > E[0]) {  // below the zonal columns' first edge
            s_oh = SOH[0];
            s_cl = SCL[0];
            s_od = SOD[0];
        }
        const double T = a.temper(i, j, l);
        double cl_col = 0.0;
        if (trop) {
            while (jt < nt && (jt + 1 < nt ? a.pl(i, j, jt + 1) : 0.0) >= op_bot) jt++;
            double acc = 0.0;
            for (int k = jt; k < nt; k++) {
                const double ip_bot = a.pl(i, j, k), ip_top = k + 1 < nt ? a.pl(i, j, k + 1) : 0.0;
                if (ip_bot <= op_top) break;
                const double thick = (op_bot < ip_bot ? op_bot : ip_bot) - (op_top > ip_top ? op_top : ip_top);
                if (thick > 0.0) acc = acc + a.cl(i, j, k) * thick;
            }
            if (width > 0.0) acc = acc / width;
            if (op_bot > a.pl(i, j, 0)) acc = a.cl(i, j, 0);
            const double krate = (CL_A * fdlibm_pow(T / CL_T0, CL_N)) * fdlibm_exp(-CL_EA / T);
            cl_col = (acc * krate) * a.cl_scale;
        }
        const double oh_spiv = ((OH_A * fdlibm_exp(-OH_EA / T)) * a.oh3d(i, j, l)) * a.oh_scale;
        const bool strato = op_bot < a.trop[j];
        const double L0 = strato ? s_oh : oh_spiv;
        const double L1 = strato ? s_cl : s_cl + cl_col;
        const double L2 = s_od * a.o1d_scale;
        const double L3 = (l == 0 && a.soil.p) ? -(a.soil_rate * a.soil(i, j, 0)) : 0.0;
        double *L = &a.loss(i, j, l);
        L[0] = L0;
        L[a.loss_sr] = L1;
        L[2 * a.loss_sr] = L2;
        L[3 * a.loss_sr] = L3;
        if (a.decay.p) a.decay(i, j, l) = fdlibm_exp(-((((L0 + L1) + L2) + L3) * a.dt));
    }
}

int launch_loss_rates(tm5_ctx *ctx, const LossArgs &a)
{
    if (a.im < 1 || a.jm < 1 || a.lm < 1 || a.nstrat < 1 || !a.edges || !a.trop || !a.s_oh || !a.s_cl || !a.s_od ||
        !a.temper.p || !a.phlb.p || !a.oh3d.p || !a.loss.p || (a.pl.p != nullptr) != (a.cl.p != nullptr) ||
        (a.pl.p && a.ntrop < 1))
        return TM5_EINVAL;
    TM5_LAUNCH(ctx, k_loss_rates, blocks((long long)a.im * a.jm), dim3(256), 0, ctx->stream, a, ctx->words);
    return status(hipGetLastError());
}

// ------------------------------------------------------------------ Calc_Kzz
// pydyn/diffusion.py:71-276 for one column (i, j) per thread, lanes along i: the friction velocity,
// the interface heights (written to gph and read back), then three sweeps up the column that
// recompute each level's quantities from the inputs (the same operations, so the same bits): the
// first bulk-Richardson sweep, the second with the convective excess, and the interfaces' free
// and boundary-layer diffusivities with dkg.  Every `if` of the Fortran is a 0/1 mask (jqif)
// multiplying both branches, max / min propagate NaN as NumPy's, log / pow / exp are fdlibm.
__device__ __forceinline__ double np_max(double a, double b) { return (a >= b || a != a) ? a : b; }
__device__ __forceinline__ double np_min(double a, double b) { return (a <= b || a != a) ? a : b; }
__device__ __forceinline__ double jqif(double x, double y) { return __double_as_longlong(x - y) < 0 ? 0.0 : 1.0; }

__global__ void __launch_bounds__(256) k_calc_kzz(KzzArgs a)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= a.im * a.jm) return;
    const int i = col % a.im, j = col / a.im, lm = a.lm;
    const double GRAV = 9.80665, VKARMAN = 0.4, CP_AIR = 1004.0, LVAP = 2.5e6, RHO_AIR = 1.2, CKMIN = 1.0e-15;
    const double SFFRAC = 0.1, ONET = 1.0 / 3.0, APZERO = 100000.0, BETAH = 15.0, FAK = 8.5, FAKN = 7.2;
    const double RAIR = 8.3144 * 1.0e3 / 28.94;
    const double EPSILO = RAIR / 461.51, ZKAPPA = RAIR / CP_AIR, ZCRDQ = 1.0 / EPSILO - 1.0;
    const double CCON = SFFRAC * VKARMAN, BINM = 15.0 * SFFRAC;
    const double TINY = 1.0e-9, ZACB = 100.0, RICR = 0.3, SSFRAC = 0.1, KVHMIN = 0.1;
    auto S = [&](int k) { return a.surf[k](i, j, 0); };  // oro lsm sr srols u10m v10m sshf slhf ewss nsss
    const double spm = a.spm(i, j, 0), sp = a.sp(i, j, 0), dxyp = a.dxyp[j], dxx = a.dxx[j], dyy = a.dyy;

    // ustar (pydyn/diffusion.py:76-91)
    double ust;
    {
        const double xland = S(1) / 100.0;
        const double tau = sqrt(S(8) * S(8) + S(9) * S(9));
        const double ustar_sea = np_max(0.001, sqrt(tau / RHO_AIR));
        const double wind10m = sqrt(S(4) * S(4) + S(5) * S(5));
        const double sr_land = np_max(1e-2, S(3)), sr_help = np_min(S(2), 0.03);
        double land = ((VKARMAN * wind10m) / fdlibm_log(10.0 / sr_help)) * fdlibm_log(75.0 / sr_help);
        land = land / fdlibm_log(10.0 / sr_land);
        const double ustar_land = xland > 0.0 ? land : 0.0;
        ust = xland * ustar_land + (1.0 - xland) * ustar_sea;
    }
    if (a.ustar.p) a.ustar(i, j, 0) = ust;
    // the interface heights (pydyn/diffusion.py:94-106)
    {
        double g = S(0) / GRAV;
        a.gph(i, j, 0) = g;
        for (int k = 0; k + 1 < lm; k++) {
            const double tv = a.t(i, j, k) * (1.0 + 0.608 * a.q(i, j, k));
            const double ratio = (a.at[k] + a.bt[k] * sp) / (a.at[k + 1] + a.bt[k + 1] * sp);
            g = g + ((tv * 287.05) * fdlibm_log(ratio)) / GRAV;
            a.gph(i, j, k + 1) = g;
        }
        a.gph(i, j, lm) = 200000.0;
    }
    const double G0 = a.gph(i, j, 0);
    auto PH = [&](int l) { return a.at[l] + a.bt[l] * spm; };
    auto M = [&](int l) { return ((PH(l) - PH(l + 1)) * dxyp) / GRAV; };
    auto PF = [&](int l) { return (PH(l) + PH(l + 1)) * 0.5; };
    auto WTHM = [&](int l) { return a.t(i, j, l) * fdlibm_pow(APZERO / PF(l), ZKAPPA); };
    auto ZTHVK = [&](int l) { return WTHM(l) * (1.0 + ZCRDQ * a.q(i, j, l)); };
    auto WHM = [&](int l) {
        const double lo = PH(l), pf = PF(l), intfac = (lo - pf) / (lo - PH(l + 1));
        return (a.gph(i, j, l) - G0) * (1.0 - intfac) + (a.gph(i, j, l + 1) - G0) * intfac;
    };
    auto UW = [&](int l) { return ((dxx * (a.pu(i + 1, j, l) + a.pu(i, j, l))) * 0.5) / M(l); };
    auto VW = [&](int l) { return ((dyy * (a.pv(i, j + 1, l) + a.pv(i, j, l))) * 0.5) / M(l); };

    // the surface (pydyn/diffusion.py:160-166, 203-206)
    const double dz1 = a.gph(i, j, 1) - a.gph(i, j, 0), dp1 = PH(0) - PH(1);
    const double wheat = ((((-S(6)) * dz1) * GRAV) / dp1) / CP_AIR;
    const double wqflx = ((((-S(7)) * dz1) * GRAV) / dp1) / LVAP;
    const double ustr = np_max(ust, 0.01);
    const double wthm0 = WTHM(0), uw0 = UW(0), vw0 = VW(0), whm0 = WHM(0);
    const double wtseff = wthm0 * (1.0 + ZCRDQ * a.q(i, j, 0));
    const double wheatv = wheat + (ZCRDQ * wthm0) * wqflx;
    const double cube = (ustr * ustr) * ustr;
    const double wobkl = ((-wtseff) * cube) / ((GRAV * VKARMAN) * (wheatv + copysign(1.0e-10, wheatv)));
    // the two bulk-Richardson sweeps (pydyn/diffusion.py:127-155)
    auto sweep = [&](double pblh, bool first, double reference) {
        double jwork = 1.0, wrino = 0.0;
        for (int k = first ? 0 : 1; k < lm; k++) {
            const double whm_k = k == 0 ? whm0 : WHM(k);
            const double du = UW(k) - uw0, dv = VW(k) - vw0;
            const double vvk = np_max((du * du + dv * dv) + (ZACB * ustr) * ustr, TINY);
            const double tkv = WTHM(k) * (1.0 + ZCRDQ * a.q(i, j, k));
            const double dz = whm_k - whm0;
            double zrino = first ? ((GRAV * (tkv - wtseff)) * dz) / (wtseff * vvk)
                                 : ((GRAV * (tkv - reference)) * dz) / (vvk * wtseff);
            zrino = jwork * (zrino + copysign(1.0e-10, zrino));
            const double jq = jqif(RICR, zrino);
            const double frac = (RICR - wrino) / (zrino - wrino);
            const double below = k == 0 ? 0.0 : WHM(k - 1);
            const double at_ricr = k == 0 ? frac * whm0 : below + frac * (whm_k - below);
            pblh = jq * pblh + (1.0 - jq) * at_ricr;
            jwork = jq * jwork;
            wrino = zrino + (1.0 - jwork) * 0.1;
        }
        const double top = WHM(lm - 1);
        return (1.0 - jwork) * pblh + (first ? top * jwork : jwork * top);
    };
    double pblh = sweep(0.0, true, 0.0);
    {
        const double jq = jqif(wheatv, 0.0);
        const double fmt = fdlibm_pow(jq * (1.0 - (BINM * pblh) / wobkl) + (1.0 - jq), ONET);
        const double zexcess = (wheatv * FAK) / (ustr * fmt);
        const double thvparc = wtseff + jq * zexcess;
        pblh = np_max(sweep(pblh, false, thvparc), 100.0);
    }
    if (a.pblh.p) a.pblh(i, j, 0) = pblh;
    // the interfaces l + 1 (pydyn/diffusion.py:178-253) and dkg (:256-263)
    const double P = pblh, W = wobkl, HV = wheatv, U = ustr, TS = wtseff;
    const int n = a.lmax - 1;
    for (int l = 0; l + 1 < lm; l++) {
        const double z = a.gph(i, j, l + 1) - G0;
        double jq = jqif(1000.0, z);
        const double zlambdac = jq * 300.0 + (1.0 - jq) * (30.0 + 270.0 * fdlibm_exp(1.0 - z / 1000.0));
        double cml2 = 1.0 / (1.0 / zlambdac + 1.0 / (VKARMAN * z));
        cml2 = cml2 * cml2;
        const double du = UW(l + 1) - UW(l), dv = VW(l + 1) - VW(l);
        const double whm_l = WHM(l), whm_u = WHM(l + 1);
        const double zdz = whm_u - whm_l;
        const double zdup2 = np_max(du * du + dv * dv, 1.0e-10) / (zdz * zdz);
        const double log_pf0 = fdlibm_log(PF(l)), log_pf1 = fdlibm_log(PF(l + 1));
        const double arg = (fdlibm_log(PH(l + 1)) - log_pf0) / (log_pf1 - log_pf0);
        const double th0 = ZTHVK(l), dthv = ZTHVK(l + 1) - th0;
        const double zthva = th0 + arg * dthv;
        const double zrinub = (((GRAV / zthva) * dthv) / zdz) / zdup2;
        const double zfunst = sqrt(np_max(1.0 - 18.0 * zrinub, 0.0));
        double zfstabh = 1.0 / (1.0 + (10.0 * zrinub) * (1.0 + 8.0 * zrinub));
        const double zkvn = cml2 * sqrt(zdup2);
        jq = jqif(zrinub, 0.0);
        const double wkvf = jq * np_max(CKMIN, zkvn * zfstabh) + (1.0 - jq) * np_max(CKMIN, zkvn * zfunst);

        jq = jqif(P, z);
        const double zh = (jq * z) / P, zl = (jq * z) / W + (1.0 - jq);
        const double zzh = jq * ((1.0 - zh) * (1.0 - zh));
        double jq1 = jqif(HV, 0.0);
        const double jq2 = jq * (1.0 - jq1), jq3 = jq * jq1;
        const double jq6 = (jq1 * jqif(P, whm_l)) * jqif(whm_u, P);
        jq1 = jqif(SFFRAC, zh);
        const double jq4 = jq3 * jq1, jq5 = jq3 * (1.0 - jq1);
        cml2 = 1.0 / (1.0 / (VKARMAN * z) + 1.0 / 450.0);
        cml2 = cml2 * cml2;
        const double jqq = jqif(zrinub, 0.0);
        zfstabh = jqq / (1.0 + (10.0 * zrinub) * sqrt(1.0 + jqq * zrinub)) + (1.0 - jqq);
        double zpblk = (cml2 * sqrt(zdup2)) * zfstabh;
        const double zkvh = jq2 * np_max(zpblk, KVHMIN) + (1.0 - jq2) * wkvf;
        double zslask = jq4 * (1.0 - BETAH * zl) + (1.0 - jq4);
        zpblk = (1.0 - jq4) * zkvh + jq4 * ((((U * VKARMAN) * z) * zzh) * fdlibm_pow(zslask, ONET));
        zslask = jq5 * (1.0 - ((SSFRAC * BETAH) * P) / W) + (1.0 - jq5);
        const double wsc = U * fdlibm_pow(zslask, ONET);
        zpblk = (1.0 - jq5) * zpblk + (((jq5 * wsc) * VKARMAN) * z) * zzh;
        const double zwstr = fdlibm_pow(((fabs(HV) * GRAV) * P) / TS, ONET);
        const double obukov = jq3 * W - (1.0 - jq3);
        const double stab = 1.0 - ((SSFRAC * BETAH) * P) / obukov;
        const double term1 = fdlibm_pow(stab, ONET);
        const double term3 = ((CCON * FAKN) * zwstr) / (U * term1);
        const double pr = jq3 * (term1 / sqrt(stab) + term3) + (1.0 - jq3);
        double kvh = jq3 * np_max(zpblk / pr, KVHMIN) + (1.0 - jq3) * zkvh;
        const double kvhentr = (0.2 * HV) / (dthv / zdz);
        kvh = jq6 * kvhentr + (1.0 - jq6) * kvh;
        if (a.kvh.p) a.kvh(i, j, l) = kvh;
        if (l < n) {
            const double dz = a.gph(i, j, l + 2) - a.gph(i, j, l);
            a.dkg(i, j, l) = ((np_max(0.0, kvh) * 2.0) * (M(l + 1) + M(l))) / (dz * dz);
        }
    }
    a.dkg(i, j, n) = 0.0;
}

int launch_calc_kzz(tm5_ctx *ctx, const KzzArgs &a)
{
    if (a.im < 1 || a.jm < 1 || a.lm < 2 || a.lmax < 1 || a.lmax > a.lm || !a.at || !a.bt || !a.dxyp || !a.dxx ||
        !a.gph.p || !a.dkg.p)
        return TM5_EINVAL;
    for (int k = 0; k < 10; k++)
        if (!a.surf[k].p) return TM5_EINVAL;
    TM5_LAUNCH(ctx, k_calc_kzz, blocks((long long)a.im * a.jm), dim3(256), 0, ctx->stream, a);
    return status(hipGetLastError());
}

int launch_scale(tm5_ctx *ctx, int n0, int n1, int n2, View3 a, double factor)
{
    if (n0 < 1 || n1 < 1 || n2 < 1) return TM5_EINVAL;
    TM5_LAUNCH(ctx, k_scale, blocks((long long)n0 * n1 * n2), dim3(256), 0, ctx->stream, n0, n1, n2, a, factor);
    return status(hipGetLastError());
}

}  // namespace tm5
