// Host/device declarations shared by the translation units of libtm5hip.
#pragma once

#include <hip/hip_runtime.h>

#include "tm5_device.h"
#include "tm5hip.h"

// Device scratch words of a context (ints): the flag word, the sub-cycle results the
// operators leave, and the z search's grid-wide reduction state.
enum {
    W_FLAGS = 0,
    W_X_NLOOP = 1,  // max nloop over the x rows of the last tm5_advect_x / tm5_mass_x
    W_Z_NADV = 2,   // n_advz of the last z operator
    W_Z_TICKET = 3, // blocks of the z search that have finished (reset by the last one)
    W_Z_CAND = 4,   // the z search's running candidate
    W_SPEC_AT = 5,  // tm5_region_run: the first queued interval whose rehearsal failed
    W_SPEC_KIND = 6, // 1: that interval's checked step failed a check (its state moved), 0: a trial
    // a queued run's clock, kept on the device so that a step captured once can be replayed for any
    // interval: the interval (counted from the run's start) and the step within it
    W_CLOCK_K = 8,
    W_CLOCK_S = 9,
    W_WORDS = 16,
};

// spec_k of the step kernels: a checked step of queued interval spec_k (>= 0), of the interval the
// clock names (SPEC_CLOCK), or not a checked step (SPEC_NONE); SPEC_REHEARSAL: Check_CFL's
// mass-only rehearsal at its caps, whose failed test raises TM5_REHEARSAL (the verdict) instead
constexpr int SPEC_NONE = -1;
constexpr int SPEC_CLOCK = -2;
constexpr int SPEC_REHEARSAL = -3;
__host__ __device__ inline bool checked(int spec_k) { return spec_k != SPEC_NONE; }

struct tm5_prof;  // the launches timed between tm5_profile_begin and tm5_profile_end (api.hip)

struct tm5_ctx {
    hipStream_t stream;
    int device;
    int *words;     // device, W_WORDS ints
    int *hwords;    // pinned host mirror
    int *y_nloop;   // device, per-layer nloop of the last y operator, then y_cap tickets, y_cap verdicts
    int y_cap;      // its length
    int *x_rows;    // device, per-row nloop of the last x operator
    int x_cap;      // its length
    int x_nrows;    // rows the last x operator wrote
    int *z_ok;      // device, per-block verdicts of the z check at n = 1
    int z_cap;      // its length
    int tile[TM5_TILES];  // tile shapes (tm5_ctx_set_tile)
    int ncu;        // the device's compute units
    double *cbuf;   // device, the split column form's matrices (grown on demand)
    size_t cbuf_n;  // its length in doubles
    tm5_prof *prof; // non-null while profiling (then prof == prof_store)
    tm5_prof *prof_store;  // owned: the events and the last totals
    double *noalloc;       // tm5_convec_noalloc's static device buffer on this context's device
    int skip_or;           // bits every zoom-tree launcher adds to its skip mask (a queued tree run's TM5_SPEC)
};

namespace tm5 {

// Profiling: events recorded on the context's stream around a launch while profiling is on.
void prof_mark(tm5_ctx *ctx, const char *name, bool stop);
struct Span {
    tm5_ctx *ctx;
    const char *name;
    Span(tm5_ctx *c, const char *n) : ctx(c), name(n)
    {
        if (ctx->prof) prof_mark(ctx, name, false);
    }
    ~Span()
    {
        if (ctx->prof) prof_mark(ctx, name, true);
    }
};

// every kernel of the library is launched through these, on the context's stream
#define TM5_LAUNCH_NAMED(ctx_, name, kernel, ...)    \
    do {                                             \
        ::tm5::Span span_((ctx_), (name));           \
        hipLaunchKernelGGL(kernel, __VA_ARGS__);     \
    } while (0)
#define TM5_LAUNCH(ctx_, kernel, ...) TM5_LAUNCH_NAMED(ctx_, #kernel, kernel, __VA_ARGS__)

struct View3 {
    double *p;
    long long si, sj, sl;
    __device__ __forceinline__ double &operator()(int i, int j, int l) const
    {
        return p[(long long)i * si + (long long)j * sj + (long long)l * sl];
    }
};

struct Tracers {
    double *q[4];  // rm, rxm, rym, rzm
    long long si, sj, sl, st;
    int nt;
    __device__ __forceinline__ long long off(int i, int j, int l, int t) const
    {
        return (long long)i * si + (long long)j * sj + (long long)l * sl + (long long)t * st;
    }
};

struct IView2 {
    int *p;
    long long si, sj;
    __device__ __forceinline__ int operator()(int i, int j) const { return p[(long long)i * si + (long long)j * sj]; }
};

// A view of a halo'd array addressed by Fortran indices: element (i, j, l[, t]) at
// p + (i - l0) s0 + (j - l1) s1 + (l - l2) s2 + (t - l3) s3, p the element at the lower bounds.
struct HV {
    double *p;
    long long s0, s1, s2, s3;
    int l0, l1, l2, l3;
    __device__ __forceinline__ double &operator()(int i, int j, int l) const
    {
        return p[(long long)(i - l0) * s0 + (long long)(j - l1) * s1 + (long long)(l - l2) * s2];
    }
    __device__ __forceinline__ double &operator()(int i, int j, int l, int t) const
    {
        return p[(long long)(i - l0) * s0 + (long long)(j - l1) * s1 + (long long)(l - l2) * s2 +
                 (long long)(t - l3) * s3];
    }
};

inline HV hv(const tm5_hview &v)
{
    return HV{v.data, v.stride[0], v.stride[1], v.stride[2], v.stride[3], v.lower[0], v.lower[1], v.lower[2], v.lower[3]};
}

// A region of the zoom tree on the device (tm5_zregion with its views as HV).
struct ZR {
    int im, jm, lm, nt;
    int xr, yr, zr, tr, tref;
    int ibeg, iend, jbeg, jend, isr, ier, jsr, jer;
    int xcyc, touch_sp, touch_np, whole;
    HV m, q[4], am, bm, cm, pu, pv, sp, phlb;
    const double *dxyp;
    const int *owned, *edge;
    int nred;
    const int *jred, *imred;
};

inline ZR zr(const tm5_zregion &d)
{
    ZR r;
    r.im = d.im, r.jm = d.jm, r.lm = d.lm, r.nt = d.nt;
    r.xr = d.xref, r.yr = d.yref, r.zr = d.zref, r.tr = d.tref, r.tref = d.tref_abs;
    r.ibeg = d.ibeg, r.iend = d.iend, r.jbeg = d.jbeg, r.jend = d.jend;
    r.isr = d.isr, r.ier = d.ier, r.jsr = d.jsr, r.jer = d.jer;
    r.xcyc = d.xcyc, r.touch_sp = d.touch_sp, r.touch_np = d.touch_np, r.whole = d.whole;
    r.m = hv(d.m), r.q[0] = hv(d.rm), r.q[1] = hv(d.rxm), r.q[2] = hv(d.rym), r.q[3] = hv(d.rzm);
    r.am = hv(d.am), r.bm = hv(d.bm), r.cm = hv(d.cm), r.pu = hv(d.pu), r.pv = hv(d.pv);
    r.sp = hv(d.sp), r.phlb = hv(d.phlb);
    r.dxyp = d.dxyp, r.owned = d.owned, r.edge = d.edge;
    r.nred = d.nred, r.jred = d.jred, r.imred = d.imred;
    return r;
}

// the halo-free view of a halo'd array from Fortran (i0, j0, l0)
inline View3 sub3(const HV &v, int i0, int j0, int l0)
{
    return View3{v.p + (long long)(i0 - v.l0) * v.s0 + (long long)(j0 - v.l1) * v.s1 + (long long)(l0 - v.l2) * v.s2,
                 v.s0, v.s1, v.s2};
}

inline Tracers subq(const ZR &r, int i0, int j0, int l0)
{
    Tracers t;
    for (int a = 0; a < 4; a++) {
        const HV &v = r.q[a];
        t.q[a] = v.p + (long long)(i0 - v.l0) * v.s0 + (long long)(j0 - v.l1) * v.s1 + (long long)(l0 - v.l2) * v.s2;
    }
    t.si = r.q[0].s0, t.sj = r.q[0].s1, t.sl = r.q[0].s2, t.st = r.q[0].s3;
    t.nt = r.nt;
    return t;
}

inline View3 view3(const tm5_view *v) { return View3{v->data, v->stride[0], v->stride[1], v->stride[2]}; }

inline Tracers tracers(const tm5_tracers *q)
{
    Tracers t;
    t.q[0] = q->rm;
    t.q[1] = q->rxm;
    t.q[2] = q->rym;
    t.q[3] = q->rzm;
    t.si = q->stride[0];
    t.sj = q->stride[1];
    t.sl = q->stride[2];
    t.st = q->stride[3];
    t.nt = q->nt;
    return t;
}

inline Tracers no_tracers()
{
    Tracers t{};
    t.nt = 0;
    return t;
}

inline IView2 iview2(const tm5_iview *v) { return IView2{v->data, v->stride[0], v->stride[1]}; }

inline int status(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }

// A checked step's failed test (spec_k: its interval in the queue): the kernels queued after it
// skip, and the host restores the interval's state from the backup before redoing it.
__device__ __forceinline__ void spec_fail(int *words, int spec_k)
{
    if (spec_k == SPEC_REHEARSAL) {
        atomicOr(words + W_FLAGS, TM5_REHEARSAL);
        return;
    }
    words[W_SPEC_AT] = spec_k >= 0 ? spec_k : words[W_CLOCK_K];  // every raiser of one launch writes the same
    words[W_SPEC_KIND] = 1;
    __threadfence();
    atomicOr(words + W_FLAGS, TM5_SPEC);
}

// Setup_MassFlow's fluxes made by a step's first x launch instead of a dynam0 launch (a region's
// queued steps), with dynam0's operations: each row makes the am it uses, am = dt (a pu) with
// am(0) = am(im) (a cyclic row), and stores it; with bm.p set, extra workgroups of the launch make
// bm and cm (dynam0's tiles of TI columns without am) for the step's later kernels.  u.p null: am
// is read from its array.
struct FluxSrc {
    View3 u, v;        // pu, pv (the meteo's, before the amplitude)
    double a, dt;      // the amplitude at the step's middle; ndyn / 2
    const double *bt;  // the hybrid b at the interfaces (cm)
    View3 bm, cm;      // where the extra workgroups store bm and cm (bm.p null: none)
    int TI;            // columns of a dynam0 tile
};

// spec_fail for a launch that checks two queued intervals (k_xx: a step's last x and the next
// step's first): the earlier failing interval is the one the host redoes, so the smallest is kept
// (the host sets W_SPEC_AT to INT_MAX before it queues)
__device__ __forceinline__ void spec_fail_first(int *words, int spec_k)
{
    atomicMin(words + W_SPEC_AT, spec_k);
    words[W_SPEC_KIND] = 1;
    __threadfence();
    atomicOr(words + W_FLAGS, TM5_SPEC);
}

// launchers (advect.hip)
// rows_out (default: the context's): per row nloop; amax_out (optional): per row the largest Courant number
// halo_j0..halo_j1 (0-based rows): those rows also get their periodic halo cells, index -1 and n of
// the views (a zoom tree's region 1; a reduced row's n is its imred); keep_am: am written back as read
int launch_x(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 am, int limits, int nmax, int fail_bit,
             int skip_mask, const int *imred = nullptr, int *rows_out = nullptr, double *amax_out = nullptr,
             int spec_k = -1, double *backup = nullptr, int halo_j0 = 0, int halo_j1 = -1,
             const FluxSrc *flux = nullptr, int keep_am = 0);
// A checked step's last x (interval spec_a, faces from am) and the next checked step's first
// (interval spec_b, faces made from pu with fs, its bm and cm by extra workgroups) in one launch, a
// row taking both; backup (null: none): the next step's input, as launch_x's backup
int launch_xx(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 am, int limits, int nmax, int skip_mask,
              int spec_a, int spec_b, double *backup, const FluxSrc &fs);
// known_one: the caller knows the sub-cycle count is 1 everywhere (no search launched); spec_k >= 0:
// a checked step of queued interval spec_k, which tests Check_CFL's cap 1 itself
int launch_y(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 bm, int limits, int nmax, int fail_bit,
             int skip_mask, bool known_one = false, int spec_k = -1);
int launch_z(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 cm, int limits, int nmax, int fail_bit,
             int skip_mask, bool known_one = false, int spec_k = -1);
// The column physics of one launch (column.hip): convec with, around it, the sources (s), the
// chemistry (c) and m2phlb1 (M) in the sequence the flags give:
//   plain "v";  pre "csv";  post "vsc";  twice "vsccsv";  twice + mid "vscMcsv".
struct ColumnArgs {
    int im, jm, lm;
    View3 m;
    Tracers q;
    View3 entu, detu, entd, detd, dkg;
    IView2 top, lfs;
    double dt;  // convection's dt = ndyn / (2 tref)
    int pre, twice, mid, post;
    const int *owned;  // null: every column; else (im, jm) i fastest, 1 where the region acts
    View3 emis;
    double factor, ratio, efrac;
    int level, tracer;      // 0-based
    const double *escale;   // null, or per tracer: the emission field's scale (0: no emission)
    View3 loss;
    long long loss_sr;
    double dt_chem;
    const double *dxyp, *at, *bt;
    View3 sp, phlb;
    int adjoint;      // the transposed sequence on lambdas in q; the emission's lambda into lam_emis
    View3 lam_emis;   // (im, jm)
    // zfused: dynamw at n_advz 1 before and after the sequence ("z v s c M c s v z"), on cm with
    // the limiter flag zlimits; spec_k >= 0: a checked step's, testing both at cap 1 first
    int zfused, zlimits, spec_k;
    View3 cm;
    // zpost: the dynamw that follows the sequence in the step, at n_advz 1 on cm with zlimits,
    // moving the columns i in [zi0, zi1], j in [zj0, zj1] (0-based; a column the region does not
    // own is moved but skips the sequence); spec_k >= 0: a checked step's, testing every column's
    // faces at cap 1 first
    int zpost, zi0, zi1, zj0, zj1;
    // zpre: the dynamw before the sequence on the columns in [zpi0, zpi1] x
    // [zpj0, zpj1] outside the hole [zhi0, zhi1] x [zhj0, zhj1] (0-based), at n_advz 1 on cm with
    // zlimits; the caller has moved the rest of its scope and tested every column
    int zpre, zpi0, zpi1, zpj0, zpj1, zhi0, zhi1, zhj0, zhj1;
    // non-null: the factor pass of the split form (many tracers), which stores each column's rows
    // of C there ((im, jm) columns, i fastest, of N x N row-major) and leaves the tracers alone
    double *cbuf;
    // the LAPACK kind (pydyn/lapack.py): the inverse by DGETRF2 + DGETRI (k_lapack_inverse)
    // instead of the Crout form; build_only: the factor pass stores conv1's rows, not C's
    int lapack, build_only;
};
// The tracer ranges a many-tracer kernel splits nt tracers over (grid z): enough for `tiles`
// workgroups a range to fill the device's `slots` workgroup slots in nearly whole rounds, each
// range at least min_chunk tracers.  Returns the tracers a range (nt for a single range).
int tracer_chunk(long long tiles, long long slots, int nt, int min_chunk);
bool column_supported(int n);
// dynam0's meteo amplitude a(t): `a` as given (t0 null), or, for a step of a queued run, from the
// context's clock words and the run's start *t0 (device) with t = ((t0 + K nread) + S ndyn) +
// ndyn / 2 -- the host's expression, exact in doubles for whole seconds -- through the triangle
// wave of the meteo's amplitude (x = t / period, frac = x - floor(x), 1 + amplitude (1 - 4 |frac -
// 1/2|): the host's double)
struct TimeSrc {
    double a;
    const double *t0;
    double nread, ndyn, period, amplitude;
};
__device__ __forceinline__ double time_amplitude(const TimeSrc &s, const int *words)
{
    if (!s.t0) return s.a;
    const double t = ((*s.t0 + (double)words[W_CLOCK_K] * s.nread) + (double)words[W_CLOCK_S] * s.ndyn) + s.ndyn / 2.0;
    const double x = t / s.period;
    const double frac = x - floor(x);
    return 1.0 + s.amplitude * (1.0 - 4.0 * fabs(frac - 0.5));
}
// What a zoom tree's dynam0 adds: the amplitude's source, the fluxes a pu, a pv stored into pu_out,
// pv_out (p null: not stored) as the reference's interpolated meteo, and am(im) of a cyclic
// region's halo (am_halo: am(im + 1) = am(1), index im + 1 of the am view)
struct Dynam0Extra {
    TimeSrc ts;
    View3 pu_out, pv_out;
    int am_halo;
};
// pydyn/kernels.py:102-125 + pydyn/zoom/massflow.py:446-461 on a tile of TI columns (i0.., j =
// blockIdx.y), all levels: the fluxes are the meteo's times the amplitude a (rounded, as the
// interpolated meteo the reference stores), am = dt pu, bm = dt pv with am(0) = am(im), the
// convergence of every (i, l) computed in parallel into LDS; then per column cm = -dt sd with
// pit summed from the surface and sd accumulated from the top, in the reference's order.
// tile (bx, j) of dynam0 (the workgroup's share; conv: its LDS, lm * TI doubles)
__device__ __forceinline__ void dynam0_tile(int bx, int j, double *conv, int im, int jm, int lm, int TI, View3 pu,
                                            View3 pv, const TimeSrc &ts, const double *bt, double dt, View3 am,
                                            View3 bm, View3 cm, int cyclic, const int *words, View3 pu_out,
                                            View3 pv_out, int am_halo)
{
    // TI columns i per workgroup (the context's TM5_TILE_DYNAM0)
    const int i0 = bx * TI;
    const int ni = min(TI, im - i0);
    const int tid = threadIdx.x, nth = blockDim.x;
    const double a = time_amplitude(ts, words);
#pragma unroll 4
    for (int idx = tid; idx < lm * TI; idx += nth) {
        const int ti = idx % TI, l = idx / TI;
        if (ti >= ni) continue;
        const int i = i0 + ti;
        const double p0 = a * pu(i, j, l), p1 = a * pu(i + 1, j, l);
        const double v0 = a * pv(i, j, l), v1 = a * pv(i, j + 1, l);
        conv[idx] = ((p0 - p1) + v0) - v1;
        const double a0 = dt * (i == 0 && cyclic ? a * pu(im, j, l) : p0);  // a cyclic region: am(0) = am(im)
        const double a1 = dt * p1, b0 = dt * v0, b1 = dt * v1;
        if (am.p) {  // null: am is made by the x kernel that uses it
            am(i, j, l) = a0;
            if (i == im - 1) am(im, j, l) = a1;
            if (am_halo && i == 1) am(im + 1, j, l) = a0;
        }
        bm(i, j, l) = b0;
        if (j == jm - 1) bm(i, jm, l) = b1;
        if (pu_out.p) {
            pu_out(i, j, l) = p0;
            if (i == im - 1) pu_out(im, j, l) = p1;
        }
        if (pv_out.p) {
            pv_out(i, j, l) = v0;
            if (j == jm - 1) pv_out(i, jm, l) = v1;
        }
    }
    __syncthreads();
    for (int ti = tid; ti < ni; ti += nth) {
        const int i = i0 + ti;
        // (unrolled so that the LDS and bt loads of the next levels issue ahead of the chain of sums)
        double pit = conv[ti];
#pragma unroll 8
        for (int l = 1; l < lm; l++) pit = pit + conv[l * TI + ti];
        cm(i, j, 0) = 0.0;
        cm(i, j, lm) = 0.0;
        if (lm > 1) {
            double sd = conv[(lm - 1) * TI + ti] - (bt[lm - 1] - bt[lm]) * pit;
            cm(i, j, lm - 1) = (-dt) * sd;
#pragma unroll 8
            for (int l = lm - 2; l >= 1; l--) {
                sd = (sd + conv[l * TI + ti]) - (bt[l] - bt[l + 1]) * pit;
                cm(i, j, l) = (-dt) * sd;
            }
        }
    }
}

int launch_dynam0(tm5_ctx *ctx, int im, int jm, int lm, View3 pu, View3 pv, double a, const double *bt, double dt,
                  View3 am, View3 bm, View3 cm, int skip_mask, int cyclic = 1, const Dynam0Extra *ex = nullptr);
// dynam0 of several regions in one launch (a zoom tree's Setup_MassFlow: no region's fluxes depend
// on another's): each region's tiles as launch_dynam0's, the regions' workgroups one after another
constexpr int D0_REGIONS = 8;
struct Dynam0Region {
    int im, jm, lm, cyclic;
    View3 pu, pv, am, bm, cm;
    double a, dt;  // a: the amplitude when ex.ts has no clock
    Dynam0Extra ex;
};
int launch_dynam0_regions(tm5_ctx *ctx, int n, const Dynam0Region *regions, const double *bt, int skip_mask);
// mcopy: receives a copy of m
int launch_m2phlb1(tm5_ctx *ctx, int im, int jm, int lm, View3 m, const double *dxyp, const double *at,
                   const double *bt, View3 sp, View3 phlb, const View3 *mcopy = nullptr);
// ext (non-null): the split form's buffer of ext_n doubles in place of the context's (nothing is
// allocated; too small is TM5_ESCRATCH)
int launch_column(tm5_ctx *ctx, const ColumnArgs &args, int lmax, int skip_mask, double *ext = nullptr,
                  size_t ext_n = 0);
size_t column_matrix_doubles(int lmax);  // the split form's buffer per column
int column_split_nt();                   // the split form from this many tracers on
// ncol row-major n x n matrices to their inverses, LAPACK kind (lapack.hip); info: a device int
// or null (then TM5_SINGULAR in the flag word)
int launch_lapack_inverse(tm5_ctx *ctx, long long ncol, int n, const double *A, double *Y, int *info, int skip_mask);
int launch_libm(tm5_ctx *ctx, int fn, long long n, const double *x, const double *y, double *out);
// The CH4 loss rates of tm5_loss_rates (physics.hip)
struct LossArgs {
    int im, jm, lm, nstrat, ntrop;
    View3 temper, phlb, oh3d;             // (im, jm, lm), phlb (im, jm, lm + 1)
    const double *trop, *edges;           // (jm), (jm, nstrat + 1) contiguous
    const double *s_oh, *s_cl, *s_od;     // (jm, nstrat) contiguous
    View3 pl, cl, soil;                   // (im, jm, ntrop) twice, (im, jm, 1); p null: not applied
    double oh_scale, cl_scale, o1d_scale, soil_rate, dt;
    View3 loss;                           // (im, jm, lm) of reaction 0 ...
    long long loss_sr;                    // ... and the reaction axis' stride
    View3 decay;                          // (im, jm, lm); p null: not written
};
int launch_loss_rates(tm5_ctx *ctx, const LossArgs &a);
// Calc_Kzz of tm5_calc_kzz (physics.hip)
struct KzzArgs {
    int im, jm, lm, lmax;
    View3 spm, sp;                // (im, jm)
    View3 pu, pv, t, q;           // (im + 1, jm, lm), (im, jm + 1, lm), (im, jm, lm) twice
    View3 surf[10];               // oro lsm sr srols u10m v10m sshf slhf ewss nsss, (im, jm)
    const double *at, *bt;        // lm + 1
    const double *dxyp, *dxx;     // jm: the cells' areas and zonal widths
    double dyy;                   // the cells' meridional width
    View3 gph, dkg;               // (im, jm, lm + 1), (im, jm, lmax)
    View3 ustar, pblh, kvh;       // (im, jm), (im, jm), (im, jm, lm - 1); p null: not written
};
int launch_calc_kzz(tm5_ctx *ctx, const KzzArgs &a);

// dynamw on the columns i0..i1, j0..j1 (0-based) less the hole hi0..hi1, hj0..hj1 (none by default;
// a hole needs n_advz 1 known: known_one or a checked step)
int launch_z_scoped(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 cm, int i0, int i1, int j0, int j1,
                    int limits, int nmax, int fail_bit, int skip_mask, bool known_one = false, int spec_k = -1,
                    int hi0 = 0, int hi1 = -1, int hj0 = 0, int hj1 = -1);
int launch_cluster_sum(tm5_ctx *ctx, int n, int c, int r1, int r2, View3 q, View3 out);

// the zoom tree's launchers (zoom.hip); every one adds ctx->skip_or to its skip mask.  tracers = 0:
// the mass-only form of Check_CFL's rehearsal (masks with TM5_REHEARSAL)
int zoom_put_x(tm5_ctx *ctx, const ZR &P, const ZR &C, int tracers);
int zoom_put_y(tm5_ctx *ctx, const ZR &P, const ZR &C, int tracers);
int zoom_compress(tm5_ctx *ctx, const ZR &C, int axis, int a0, int a1, int l0, int l1, int tracers, bool un);
int zoom_mix_edges(tm5_ctx *ctx, const ZR &C, int xedges, int tracers);
// a zoom region's column pass with the step's dynamw around it fused, at n_advz 1 with the limiter
// flag (0-based columns): post, the one after it on its scope i0..i1, j0..j1 (spec_k >= 0: a
// checked step's, tested at cap 1); pre, the one before it on pi0..pi1, pj0..pj1
// less the hole hi0..hi1, hj0..hj1, the rest of whose scope has moved and been tested
struct ZPost {
    int post = 0, i0 = 0, i1 = -1, j0 = 0, j1 = -1;
    int pre = 0, pi0 = 0, pi1 = -1, pj0 = 0, pj1 = -1, hi0 = 0, hi1 = -1, hj0 = 0, hj1 = -1;
    int limits = 1, spec_k = -1;
};
int zoom_physics(tm5_ctx *c, const tm5_zregion *r, const tm5_zphysics *p, const char *ops, const ZPost *zp);
int zoom_update_parent(tm5_ctx *ctx, const ZR &P, const ZR &C, int tracers);
// update_parent (with tracers) and the child's m2phlb, one launch
int zoom_update_m2phlb(tm5_ctx *ctx, const ZR &P, const ZR &C, const double *at, const double *bt, int skip_mask);
int zoom_coarsen(tm5_ctx *ctx, const ZR &P, const ZR &C, int skip_mask);
int zoom_coarsen_ambm(tm5_ctx *ctx, const ZR &P, const ZR &C, int skip_mask);
// clock_n > 0: the kernel also moves the context's clock on by a step of an interval of clock_n
int zoom_m2phlb(tm5_ctx *ctx, const ZR &R, const double *pdxyp, const double *at, const double *bt, int skip_mask,
                int clock_n);
int zoom_x_halo(tm5_ctx *ctx, const ZR &R, int js, int je, int tracers, int skip_mask);
int zoom_slab(tm5_ctx *ctx, const HV &f, int axis, int k, int n0, int n1, double *save, int restore, int skip_mask);
int zoom_am_halo(tm5_ctx *ctx, const ZR &R, int skip_mask);
constexpr int FUSE_COMPRESS = 1, FUSE_OUTER = 2;  // zoom_x_child / zoom_y_child's fuse bits
// fuse: 1 the compress before and the uncompress after made in the kernel, 2 the outer interface faces
// taken as zero (the kernel's rows / columns only; for y with known_one only)
int zoom_x_child(tm5_ctx *ctx, const ZR &C, int js, int je, int ls, int le, int limits, int nmax, int mass_only,
                 int fail_bit, int skip_mask, int spec_k, int fuse);
int zoom_y_child(tm5_ctx *ctx, const ZR &C, int is, int ie, int ls, int le, int limits, int nmax, int mass_only,
                 int fail_bit, int skip_mask, bool known_one, int spec_k, int fuse);

}  // namespace tm5

// region 1's dynamu on its reduced grid (zoom.hip)
int zoom_x_reduced(tm5_ctx *c, const tm5_zregion &d, int js, int je, int limits, int max_nloop, int mass_only, int fail,
                   int skip, int spec_k);

namespace tm5 {
int launch_z_substeps(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 cm, int n_advz, int limits,
                      int prelimit);
// Check_CFL's mass-only y and z with no sub-cycling allowed: the check at nloop 1 and its step
int launch_y_mass1(tm5_ctx *ctx, int im, int jm, int lm, View3 m, View3 bm, int fail_bit, int skip_mask);
int launch_z_mass1(tm5_ctx *ctx, int im, int jm, int lm, View3 m, View3 cm, int fail_bit, int skip_mask);

}  // namespace tm5
