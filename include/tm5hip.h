/* libtm5hip: the TM5 transport step on AMD Instinct MI355X (gfx950), FP64, bit-identical to
   pydyn's NumPy reference.

   This is the C ABI that replaces pydyn.gpu's backend boundary (Backend.launch of the PD_GLOBAL
   kernels, pydyn/gpu/backend.py:317-338, driven by the operator layer pydyn/gpu/kernels.py and
   the body table pydyn/gpu/zoom.py:38-74).  One entry point per operator; every call is
   asynchronous on the context's stream and returns 0 or an error code (a hipError_t value, or
   TM5_EINVAL); nothing throws across the ABI.

   Arrays are caller-owned device memory described by tm5_view: the address of element (0,0,0,0)
   and the element strides of the axes (i, j, l, t) -- the halo-free pydyn layout
   (pydyn/kernels.py:7-16): cells (im, jm, lm[, nt]), am (im+1, jm, lm), bm (im, jm+1, lm),
   cm (im, jm, lm+1).  Any strides work; i-fastest strides give coalesced sweeps.  The four
   tracer arrays rm, rxm, rym, rzm share one stride set (tm5_tracers), as in the reference
   (pydyn/gpu/kernels.py:48-52).  Kernels write in place, as the reference's do.

   Errors found on the device (a sub-cycle cap exceeded, non-zero outer faces) are recorded in
   the context's flag word, like the reference's deferred device flags (pydyn/gpu/backend.py:
   173-181), and read with tm5_ctx_flags.  Once an error bit is set, later operators on the
   context do nothing until the flags are cleared. */
#ifndef TM5HIP_H
#define TM5HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tm5_ctx tm5_ctx;

typedef struct tm5_view {
    double *data;      /* device address of element (0, 0, 0, 0) */
    int64_t stride[4]; /* element strides of i, j, l, t (t unused for 3-D arrays) */
} tm5_view;

typedef struct tm5_tracers {
    double *rm, *rxm, *rym, *rzm; /* tracer mass and its x, y, z slopes, (im, jm, lm, nt) */
    int64_t stride[4];            /* shared strides of i, j, l, t */
    int nt;                       /* number of tracers */
} tm5_tracers;

/* A halo'd array with its Fortran bounds (the zoom tree's layout, pydyn/layout.py:26-36):
   element (i, j, l, t) at data + sum((index - lower) * stride), data the element at the lower
   bounds. */
typedef struct tm5_hview {
    double *data;
    int64_t stride[4];
    int32_t lower[4];
} tm5_hview;

typedef struct tm5_iview {
    int32_t *data;
    int64_t stride[2]; /* i, j */
} tm5_iview;

/* flag word bits */
#define TM5_FAIL_X 1     /* dynamu: nloop too high (CFLError) */
#define TM5_FAIL_Y 2     /* dynamv: nloop too high (CFLError) */
#define TM5_FAIL_Z 4     /* dynamw: too many CFL violations (CFLError) */
#define TM5_BAD_Y 8      /* non-zero flux through the polar outer faces (ValueError) */
#define TM5_BAD_Z 16     /* non-zero halo interfaces in cm (ValueError) */
#define TM5_REHEARSAL 32 /* a Check_CFL mass-only operator exceeded its cap (the verdict) */
#define TM5_FAIL_XC 64   /* a child's dynamu: nloop too high (CFLError "dynuamu: nloop too high") */
#define TM5_FAIL_YC 128  /* a child's dynamv: nloop too high (CFLError "nloop too high in dynamv") */
#define TM5_ERRORS 1759  /* any error bit (all but TM5_REHEARSAL and TM5_SPEC) */
#define TM5_SPEC 256     /* tm5_region_run: a queued interval's trial or checked step disagreed with its plan;
                            the kernels queued after it did nothing (cleared before tm5_region_run returns) */
#define TM5_BAD_LEVELS 512 /* tm5_loss_rates: layers not strictly descending (ValueError) */
#define TM5_SINGULAR 1024  /* the LAPACK kind's DGETRF / DGETRI met an exactly zero pivot (ValueError) */

#define TM5_EINVAL (-1)  /* bad arguments (shapes beyond the kernels' limits, null pointers) */
#define TM5_EFLAGS (-2)  /* the region driver stopped: an error bit is set in the flag word */
#define TM5_ESCRATCH (-3) /* tm5_convec_noalloc: one row of columns needs more than its static buffer holds */

/* ------------------------------------------------------------------ context */
/* A context: the stream every call is queued on (0 = the default stream) and the device scratch
   the operators use (flag word, search results).  Not thread-safe; one context per stream.
   Replaces pydyn/gpu/backend.py:460-541 (stream, flags, scratch). */
int tm5_ctx_create(uint64_t stream, tm5_ctx **out);
int tm5_ctx_destroy(tm5_ctx *ctx);
/* Wait for the stream, then the flag word (and clear it when clear != 0).  The reference's
   Backend.check (pydyn/gpu/backend.py:173-181, 674-689). */
int tm5_ctx_flags(tm5_ctx *ctx, int clear, int *flags);
/* The last sub-cycle counts found on the device (valid after tm5_ctx_flags or a stream sync):
   x: max over rows; y: per layer into nloop_layers[lm] (may be NULL); z: n_advz. */
int tm5_ctx_nloop(tm5_ctx *ctx, int *x_max, int *y_max, int *nloop_layers, int lm, int *z_nadv);
/* Clear the bits of `mask` in the flag word from the stream (a kernel: no wait, capturable). */
int tm5_ctx_clear(tm5_ctx *ctx, int mask);
/* Tile shapes of the line and column kernels, set per context (pydyn_amd.tune measures them;
   they change speed only, never results).  No counterpart in the reference, whose launch shapes
   are fixed in its kernel source. */
#define TM5_TILE_Y_LINES 0 /* columns i per y-line workgroup (dynamv, dynamvm), 1..16, at most 1024 / jm; 16 */
#define TM5_TILE_Z_LINES 1 /* columns i per z-line workgroup (dynamw, dynamwm), 1..32, at most 1024 / lm; 32 */
#define TM5_TILE_DYNAM0 2  /* columns i per dynam0 workgroup, 1..64, lm * tile <= 8192; 32 */
#define TM5_TILE_M2PHLB 3  /* columns i per m2phlb1 workgroup, 1..64, (lm + 1) * tile <= 8192; 32 */
/* the many-tracer forms (nt >= 8: the batched members x tracers) */
#define TM5_TILE_Y_MANY 4     /* columns i per y-line workgroup with nt >= 8, 1..16, at most 1024 / jm; 16 */
#define TM5_TILE_RANGE_MIN 5  /* fewest tracers in a tracer range of the sweeps and the column apply, 1..64; 16 */
#define TM5_TILE_APPLY 6      /* columns per workgroup of the split column form's apply: 8, 16 or 32; 16 */
#define TM5_TILES 7
int tm5_ctx_set_tile(tm5_ctx *ctx, int which, int value);
int tm5_ctx_get_tile(tm5_ctx *ctx, int which, int *value);
/* Per-kernel GPU time: from tm5_profile_begin every launch on the context is bracketed by
   events on its stream; tm5_profile_end waits for the stream and totals them per kernel, and
   tm5_profile_entry reads the idx-th total (TM5_EINVAL past the last).  Profiling never fails a
   launch: if an event cannot be made the span stops recording and tm5_profile_end returns
   TM5_ENOMEM with no totals.  The reference's
   Backend.profile_begin / profile_end (pydyn/gpu/backend.py:633-649). */
int tm5_profile_begin(tm5_ctx *ctx);
int tm5_profile_end(tm5_ctx *ctx);
int tm5_profile_entry(tm5_ctx *ctx, int idx, const char **kernel, long long *count, double *seconds);

/* ------------------------------------------------------------------ advection */
/* dynamu: periodic x rows, each (j, l) row with its own nloop < max_nloop; am left scaled by
   the search and multiplied back by nloop.  pydyn/gpu/kernels.py:64 advect_x_rows
   (pydyn/kernels.py:129-219). */
int tm5_advect_x(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                 const tm5_view *am, int limits, int max_nloop);
/* dynamv: y columns with polar caps, one nloop < max_nloop per layer, bm of the layer scaled
   by the search and multiplied back.  pydyn/gpu/kernels.py:103 advect_y (pydyn/kernels.py:
   223-291). */
int tm5_advect_y(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                 const tm5_view *bm, int limits, int max_nloop);
/* dynamw: z columns, one n_advz <= iter_limit for the region, cm / n_advz fresh, cm unchanged.
   pydyn/gpu/kernels.py:227 advect_z (pydyn/kernels.py:295-346). */
int tm5_advect_z(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                 const tm5_view *cm, int limits, int iter_limit);

/* The mass-only operators of Check_CFL's rehearsal: a cap exceeded sets TM5_REHEARSAL (not an
   error), and they do nothing once it is set.  pydyn/gpu/kernels.py:420 mass_x_rows, :447
   mass_y, :497 mass_z (pydyn/cfl.py:140-194, pydyn/zoom/check_cfl.py:206-216). */
int tm5_mass_x(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_view *am, int max_nloop);
int tm5_mass_y(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_view *bm, int max_nloop);
int tm5_mass_z(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_view *cm, int iter_limit);

/* ------------------------------------------------------------------ mass fluxes */
/* dynam0: am = dt (a pu), bm = dt (a pv), cm from the convergence with bt (device, lm+1), the
   periodic face am(0) = am(im); a is the meteo amplitude (1.0: pu, pv as given).
   pydyn/gpu/kernels.py:300 fluxes_from_meteo + pydyn/zoom/massflow.py:446-461 dynam0. */
int tm5_dynam0(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *pu, const tm5_view *pv, double a,
               const double *bt, double dt, const tm5_view *am, const tm5_view *bm, const tm5_view *cm);
/* m2phlb1: sp = seqsum(m) g / dxyp, phlb = at + bt sp (dxyp (jm), at, bt (lm+1) on the device).
   pydyn/zoom/massflow.py:282-290 (device pd_m2phlb1, pydyn/gpu/csrc/mass.cuh:44-79). */
int tm5_m2phlb1(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const double *dxyp, const double *at,
                const double *bt, const tm5_view *sp, const tm5_view *phlb);

/* ------------------------------------------------------------------ physics */
/* convec on every column: the convection + Kzz diffusion matrix, its Crout inverse, slopes
   times diag, rm = C rm.  pydyn/gpu/kernels.py:351 convec (pydyn/kernels.py:350-457).
   With nt >= 8 (and in tm5_zoom_physics) the inverses pass through a context-owned buffer of
   im jm lmax^2 doubles, allocated on first use and grown (after waiting for the stream) when a
   larger grid needs more. */
int tm5_convec(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
               const tm5_view *entu, const tm5_view *detu, const tm5_view *entd, const tm5_view *detd,
               const tm5_iview *cloud_top, const tm5_iview *cloud_lfs, const tm5_view *dkg, double dt, int lmax);
/* convec with the LAPACK kind's inverse: reference LAPACK's DGETRF (recursive DGETRF2) and DGETRI
   (lwork = n) per column with reference BLAS loop orders -- pydyn/gpu/kernels.py:351 convec(...,
   inverse="lapack") (pydyn/gpu/csrc/convec.cuh:17-206, pydyn/lapack.py:43-66).  mask: null (every
   column) or as tm5_convec_masked's.  Register-form lmax only (4, 5, 8, 10, 15, 20, 25, 26);
   the inverses pass through the context's buffer.  A zero pivot sets TM5_SINGULAR. */
int tm5_convec_lapack(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                      const tm5_view *entu, const tm5_view *detu, const tm5_view *entd, const tm5_view *detd,
                      const tm5_iview *cloud_top, const tm5_iview *cloud_lfs, const tm5_view *dkg, double dt, int lmax,
                      const int *mask);
/* tm5_convec that allocates nothing.  No counterpart in the reference: it is for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc).  Same arguments, same bits as tm5_convec.  With nt >= 8 tm5_convec
   keeps every column's inverse in a context buffer it allocates and grows; this one uses a
   static device buffer of TM5_NOALLOC_DOUBLES doubles (32 MiB, one per GPU, held until the
   process ends) and runs the columns in bands of whole j rows that fit it.  A row of im columns
   needing more than the buffer (im lmax^2 > TM5_NOALLOC_DOUBLES) is TM5_ESCRATCH.  Every call on
   one GPU shares that buffer: calls that overlap on one GPU (two streams, two threads) race. */
#define TM5_NOALLOC_DOUBLES (4ull << 20)
int tm5_convec_noalloc(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                       const tm5_view *entu, const tm5_view *detu, const tm5_view *entd, const tm5_view *detd,
                       const tm5_iview *cloud_top, const tm5_iview *cloud_lfs, const tm5_view *dkg, double dt,
                       int lmax);
/* The LAPACK kind's inverse alone: ncol row-major n x n device matrices A to Y (A == Y allowed),
   n in 4, 5, 8, 10, 15, 20, 25, 26; info (device int, or null: TM5_SINGULAR) set to 1 on a zero
   pivot.  pydyn/gpu/kernels.py:403 inverse_rows (csrc/convec.cuh pd_inverse_rows). */
int tm5_inverse_rows(tm5_ctx *ctx, long long ncol, int n, const double *A, double *Y, int *info);
/* convec on the columns mask (im, jm, i fastest, nonzero = act) marks: pydyn/gpu/kernels.py:351
   convec's `cols`, the flat i * jm + j list, as a mask; register form only (lmax in 4, 5, 8, 10,
   15, 20, 25, 26), else TM5_EINVAL. */
int tm5_convec_masked(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                      const tm5_view *entu, const tm5_view *detu, const tm5_view *entd, const tm5_view *detd,
                      const tm5_iview *cloud_top, const tm5_iview *cloud_lfs, const tm5_view *dkg, double dt, int lmax,
                      const int *mask);
/* advect_x_rows (pydyn/gpu/kernels.py:64): every (j, l) row of nrow[j] cells (NULL: im; a reduced
   row's cells first), per row nloop and the largest |Courant| of the accepted trial into the
   device arrays nloop (int, jm lm) and amax (double, jm lm; NULL: not kept), row j + jm l. */
int tm5_advect_x_rows(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                      const tm5_view *am, const int *nrow, int limits, int max_nloop, int *nloop, double *amax);
/* mass_x_rows (pydyn/gpu/kernels.py:420) with reduced rows: TM5_REHEARSAL when a row exceeds its cap. */
int tm5_mass_x_rows(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_view *am, const int *nrow,
                    int max_nloop);
/* advect_z_substeps (pydyn/gpu/kernels.py:262): n_advz sub-steps with cm / n_advz (cm not
   written); prelimit applies dynamw's limiter to rzm first. */
int tm5_advect_z_substeps(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                          const tm5_view *cm, int n_advz, int limits, int prelimit);
/* cluster_sum (pydyn/gpu/kernels.py:327): out (n, r1, r2) = seqsum over clusters of c cells of
   q (n c, r1, r2) along its first axis, from each cluster's first term. */
int tm5_cluster_sum(tm5_ctx *ctx, int n, int c, int r1, int r2, const tm5_view *q, const tm5_view *out);

/* dkg (im, jm, lmax) from kvh (im, jm, lm-1), m and gph (im, jm, lm+1).
   pydyn/diffusion.py:256-263 (device part of pd_calc_kzz, pydyn/gpu/csrc/kzz.cuh:91-242). */
int tm5_dkg(tm5_ctx *ctx, int im, int jm, const tm5_view *kvh, const tm5_view *m, const tm5_view *gph, int lmax,
            const tm5_view *dkg);
/* do_add_2d of x = ((emis factor) ratio) efrac at the 1-based level and tracer, every cell.
   pydyn/zoom/physics.py:516-525 sources_region with pydyn/emis.py:52-54 inject. */
int tm5_add_2d(tm5_ctx *ctx, int im, int jm, const tm5_view *emis, double factor, double ratio, double efrac,
               const tm5_tracers *q, int level, int tracer);
/* CH4 decay: factor = exp(-(((L1 + L2) + L3) + L4) dt) (fdlibm exp), rm..rzm of tracers
   [t0, t0 + count) times factor.  loss: (im, jm, lm, 4), stride[3] the reaction axis.
   pydyn/chem.py:168-178 + pydyn/zoom/physics.py:528-539 (device pd_scale_owned_all). */
int tm5_decay(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *loss, double dt, const tm5_tracers *q, int t0,
              int count);
/* a = a * factor on a 3-D view of n0 x n1 x n2.  pydyn/zoom/host.py:226-230 scale. */
int tm5_scale(tm5_ctx *ctx, int n0, int n1, int n2, const tm5_view *a, double factor);
/* The device's libm over contiguous device arrays of n doubles: out = exp(x), log(x) or
   pow(x, y) (fn TM5_LIBM_*; y only for pow), by the fdlibm methods the kernels use.  The
   reference's libm route ops.libm.exp/log/pow (pydyn/ops.py:165-232) with its device twin
   (pydyn/gpu/csrc/pmath.cuh). */
/* The CH4 loss rates and decay of one month / step (pydyn/chem.py:129-170, ch4_loss_kernel; the
   reference's device pydyn.gpu.kernels.loss_rates, gpu/kernels.py:545-645 with
   csrc/chem.cuh): loss (im, jm, lm, 4; stride[3] the reaction axis) and, when decay is not
   null, decay (im, jm, lm) = exp(-(((L1 + L2) + L3) + L4) dt).  temper, oh_3d (im, jm, lm),
   phlb (im, jm, lm + 1) descending; trop (jm) the tropopause pressures and edges
   (jm, nstrat + 1) the zonal columns' edges in Pa, both computed by the caller on the host as
   the reference does (pydyn/chem.py:106-119); strat_oh/cl/od (jm, nstrat) contiguous; pl_trop and
   cl_trop (im, jm, ntrop) the tropospheric Cl on its descending pressures, both null for none;
   soil (im, jm) or null.  Layers not strictly descending set TM5_BAD_LEVELS. */
int tm5_loss_rates(tm5_ctx *ctx, int im, int jm, int lm, int nstrat, int ntrop, const tm5_view *temper,
                   const tm5_view *phlb, const double *trop, const double *edges, const double *strat_oh,
                   const double *strat_cl, const double *strat_od, const tm5_view *oh_3d, const tm5_view *pl_trop,
                   const tm5_view *cl_trop, const tm5_view *soil, double oh_scale, double cl_scale,
                   double o1d_scale, double soil_rate, double dt, const tm5_view *loss, const tm5_view *decay);
/* Calc_Kzz of one dynamic step (pydyn/diffusion.py:266-276 calc_kzz; the reference's device
   pydyn.gpu.diffusion.calc_kzz, gpu/diffusion.py:38 with csrc/kzz.cuh): dkg (im, jm, lmax) and the
   interface heights gph (im, jm, lm + 1); ustar (im, jm), pblh (im, jm) and kvh (im, jm, lm - 1)
   when not null.  spm, sp (im, jm) the surface pressure at the middle and the start of the step;
   pu (im + 1, jm, lm), pv (im, jm + 1, lm), t, q (im, jm, lm); surface[10] the (im, jm) fields oro,
   lsm, sr, srols, u10m, v10m, sshf, slhf, ewss, nsss; at, bt (lm + 1), dxyp (jm) device arrays;
   dxx (jm) the cells' zonal widths AE dx GTOR cos(lat GTOR) and dyy = AE dy GTOR, computed by the
   caller on the host (the reference evaluates that cos on the host, gpu/diffusion.py:55-57). */
int tm5_calc_kzz(tm5_ctx *ctx, int im, int jm, int lm, int lmax, const tm5_view *spm, const tm5_view *sp,
                 const tm5_view *pu, const tm5_view *pv, const tm5_view *t, const tm5_view *q,
                 const tm5_view *surface, const double *at, const double *bt, const double *dxyp, const double *dxx,
                 double dyy, const tm5_view *gph, const tm5_view *dkg, const tm5_view *ustar, const tm5_view *pblh,
                 const tm5_view *kvh);
#define TM5_LIBM_EXP 0
#define TM5_LIBM_LOG 1
#define TM5_LIBM_POW 2
int tm5_libm(tm5_ctx *ctx, int fn, int64_t n, const double *x, const double *y, double *out);

/* ------------------------------------------------------------------ the region driver */
/* One global region resident on the device, run interval by interval by a native driver:
   Check_CFL (rehearsal on device scratch, the verdict read once per attempt), then the n steps
   of ndyn, each Setup_MassFlow (dynam0 of a(t) pu0, a(t) pv0 at the middle of the step) and
   the operators "xyzvsc" m2phlb1 "csvzyx" m2phlb1 with sources (level 1 of tracer 1) and
   chemistry (decay of every tracer).  pydyn/zoom/steps.py:259-318, pydyn/zoom/check_cfl.py:
   258-290, pydyn/forward.py:388-451; the meteo amplitude a(t) is the triangle wave of
   pydyn_amd.meteo (mod_period, mod_amplitude). */
typedef struct tm5_region_desc {
    int im, jm, lm, lmax, limits;
    tm5_view m, am, bm, cm, pu0, pv0, sp, phlb;
    tm5_tracers q;
    tm5_view entu, detu, entd, detd, dkg, emis, loss;
    tm5_iview cloud_top, cloud_lfs;
    const double *at, *bt, *dxyp; /* device, lm+1, lm+1, jm */
    double mod_period, mod_amplitude, sec_month, mol_ratio;
} tm5_region_desc;

typedef struct tm5_region tm5_region;

int tm5_region_create(tm5_ctx *ctx, const tm5_region_desc *desc, tm5_region **out);
int tm5_region_destroy(tm5_region *region);
/* One interval of nread seconds from t0: returns the (ndyn, n) Check_CFL chose. */
int tm5_region_interval(tm5_region *region, double t0, int nread, int ndyn_max, int *ndyn, int *n);
/* The operators of a step one by one, as tm5_region_interval runs them: 'D' Setup_MassFlow at
   time t, 'x' 'y' 'z' advection, 'V' the middle "vsc m2phlb1 csv", 'M' m2phlb1; a step of ndyn
   from t0 is "D" at t0 + ndyn/2, then "xyzVzyxM". */
int tm5_region_ops(tm5_region *region, double t, int ndyn, const char *ops);
/* count intervals of nread seconds from t0, as count calls of tm5_region_interval would run
   them (interval k starts at t0 + k nread), with the host reading the device once per batch
   instead of once per interval.  Every interval is queued on a planned step (the previous
   interval's): the Check_CFL trials the plan expects to fail are rehearsed and their verdicts
   checked on the device, and the planned steps run as checked steps, which make the passing
   rehearsal's tests themselves on the same bits of m (x rows at nloop <= 9, y and z faces at
   nloop 1) and keep the interval's starting m and moments in a backup.  When a verdict or a
   check disagrees with the plan, the kernels queued after it skip; the interval's state comes
   back from the backup, the run resumes from it through tm5_region_interval, then queues the
   rest again.  ndyn[k], n[k]: what Check_CFL chose for interval k.  The reference's queued
   submission is Pipeline.submit (pydyn/gpu/zoom.py:944-952).  The backup (m and four moments
   per tracer, im jm lm doubles each) is allocated on the first call. */
int tm5_region_run(tm5_region *region, double t0, int nread, int ndyn_max, int count, int *ndyn, int *n);


/* ------------------------------------------------------------------ the zoom tree */
/* The bodies the zoom driver (pydyn/zoom/steps.py, check_cfl.py, massflow.py) runs through
   ZoomState.lib for a state on a device: pydyn.gpu.zoom's names (pydyn/gpu/zoom.py:38-74, the
   list pydyn/zoom/host.py:42-48).  A region is described by tm5_zregion: its sizes, its
   refinement relative to its parent, its footprint in the parent, its interior scope, and its
   halo'd arrays (tm5_hview, declared bounds m (-1:im+2, -1:jm+2, 1:lm), rm..rzm (.., 1:nt),
   am bm cm (0:im+1, 0:jm+1, 0:lm+1), pu (0:im, 1:jm, 1:lm), pv (1:im, 1:jm+1, 1:lm),
   sp (-1:im+2, -1:jm+2, 1:1), phlb (-1:im+2, -1:jm+2, 1:lm+1)).  `tracers` = 0 is the
   mass-only form of Check_CFL's rehearsal (putm_xedges, mixm_edges, ...): those kernels skip
   once TM5_REHEARSAL is set. */
typedef struct tm5_zregion {
    int im, jm, lm, nt;
    int xref, yref, zref, tref;      /* refinement relative to the parent (1 for region 1) */
    int tref_abs;                    /* time refinement relative to the globe: dt = ndyn / (2 tref_abs) */
    int ibeg, iend, jbeg, jend;      /* footprint in the parent's cells (children) */
    int isr, ier, jsr, jer;          /* interior scope */
    int xcyc, touch_sp, touch_np;    /* region 1: 1 1 1 */
    int whole;                       /* im / xref spans the globe (no x walls) */
    tm5_hview m, rm, rxm, rym, rzm, am, bm, cm, pu, pv, sp, phlb;
    const double *dxyp;              /* device, jm */
    const int32_t *owned;            /* device (im, jm), i fastest: 1 where convec, sources, chemistry act */
    const int32_t *edge;             /* device (im, jm), i fastest: blank_zoom_region's edge map */
    /* region 1's reduced grid (pydyn/redgrid.py): nred rows jred[] (device, 1-based), each row's
       cell count imred[] (device, jm; im on a uniform row); nred = 0: uniform */
    int nred;
    const int32_t *jred;
    const int32_t *imred;
    double *scratch;                 /* device, >= (elements of m) + (elements of am) doubles when nred > 0 */
} tm5_zregion;

/* put_xedges / put_yedges: the parent's cells at each wall of the child into the child's halo
   and interface cells (pydyn/zoom/exchange.py:17-105). */
int tm5_zoom_put_xedges(tm5_ctx *ctx, const tm5_zregion *parent, const tm5_zregion *child, int tracers);
int tm5_zoom_put_yedges(tm5_ctx *ctx, const tm5_zregion *parent, const tm5_zregion *child, int tracers);
/* _compress_x/yedges and _uncompress_x/yedges over rows js..je (x) or columns is..ie (y) and
   levels ls..le (pydyn/zoom/exchange.py:109-177). */
int tm5_zoom_compress_xedges(tm5_ctx *ctx, const tm5_zregion *r, int js, int je, int ls, int le, int tracers);
int tm5_zoom_uncompress_xedges(tm5_ctx *ctx, const tm5_zregion *r, int js, int je, int ls, int le, int tracers);
int tm5_zoom_compress_yedges(tm5_ctx *ctx, const tm5_zregion *r, int is, int ie, int ls, int le, int tracers);
int tm5_zoom_uncompress_yedges(tm5_ctx *ctx, const tm5_zregion *r, int is, int ie, int ls, int le, int tracers);
/* mix_edges: the child's ring blocks set to their means (pydyn/zoom/exchange.py:181-231). */
int tm5_zoom_mix_edges(tm5_ctx *ctx, const tm5_zregion *r, int tracers);
/* update_parent: the parent's m, rm over the footprint from the child's sums
   (pydyn/zoom/exchange.py:235-258). */
int tm5_zoom_update_parent(tm5_ctx *ctx, const tm5_zregion *parent, const tm5_zregion *child, int tracers);
/* _coarsen_block and coarsen_ambm (pydyn/zoom/massflow.py:79-177). */
int tm5_zoom_coarsen_block(tm5_ctx *ctx, const tm5_zregion *parent, const tm5_zregion *child);
int tm5_zoom_coarsen_ambm(tm5_ctx *ctx, const tm5_zregion *parent, const tm5_zregion *child);
/* m2phlb1 (parent = NULL, region 1) or a child's m2phlb (pydyn/zoom/massflow.py:24-58); at, bt
   device arrays of lm+1. */
int tm5_zoom_m2phlb(tm5_ctx *ctx, const tm5_zregion *r, const tm5_zregion *parent, const double *at,
                    const double *bt);
/* dynam0 of a region from its pu, pv: dt = ndyn / (2 tref_abs), a cyclic region's halo faces
   (pydyn/zoom/massflow.py:185-203).  The halos of am, bm, cm are never written otherwise, so the
   reference's zero-fill of the whole arrays leaves them as they are. */
int tm5_zoom_dynam0(tm5_ctx *ctx, const tm5_zregion *r, const double *bt, int ndyn, int rehearsal);
/* outer_x/y_faces_zeroed: a child's outer interface faces saved into `save` and zeroed
   (restore = 0), or put back (restore = 1); axis 0 = am, 1 = bm.  save holds
   2 (jm+2)(lm+2) (x) or 2 (im+2)(lm+2) (y) doubles (pydyn/zoom/advect.py:121-157). */
int tm5_zoom_outer_faces(tm5_ctx *ctx, const tm5_zregion *r, int axis, int restore, double *save, int rehearsal);
/* dynamu / dynamv / dynamw of a region on the block's scope: region 1 on its interior (x then
   writes its periodic halos on rows js..je), a child with its ends held; mass_only = the
   rehearsal's dynamum/dynamvm/dynamwm, whose failure sets TM5_REHEARSAL
   (pydyn/zoom/advect.py:209-472, pydyn/zoom/check_cfl.py:101-226). */
int tm5_zoom_advect_x(tm5_ctx *ctx, const tm5_zregion *r, int js, int je, int limits, int max_nloop, int mass_only);
int tm5_zoom_advect_y(tm5_ctx *ctx, const tm5_zregion *r, int is, int ie, int limits, int max_nloop, int mass_only);
int tm5_zoom_advect_z(tm5_ctx *ctx, const tm5_zregion *r, int is, int ie, int js, int je, int limits,
                      int iter_limit, int mass_only);

/* sources_region's do_add_2d (pydyn/zoom/physics.py:55-64): rm + x, rzm - x at 1-based (level,
   tracer) on the cells the region owns; x (im, jm) in kg. */
int tm5_zoom_inject(tm5_ctx *ctx, const tm5_zregion *r, const tm5_view *x, int level, int tracer);
/* chemistry_region (pydyn/zoom/physics.py:67-78): rm and the slopes of 1-based tracers t0 ..
   t0+count-1 times decay (the scope's factors, (ier-isr+1, jer-jsr+1, lm)) on the owned cells. */
int tm5_zoom_decay(tm5_ctx *ctx, const tm5_zregion *r, const tm5_view *decay, int t0, int count);

/* The column physics of a region on the cells it owns: convec (convection and Kzz diffusion,
   dt_conv), do_add_2d (x = ((emis factor) ratio) efrac at level/tracer) and the CH4 decay
   (exp(-(sum of the four loss rates) dt_chem)).  `ops` is the sequence run in one pass over
   each column: "v", "csv", "vsc", "vsccsv", or "vscMcsv" (M = m2phlb1 of every column, owned
   or not) -- pydyn/zoom/physics.py:34-78, pydyn/zoom/steps.py:31-69.  Between two v's the air
   mass does not change, so the convection matrix is built and inverted once. */
typedef struct tm5_zphysics {
    tm5_view entu, detu, entd, detd, dkg; /* (im, jm, >= lmax) kg/s; dkg (im, jm, lmax) */
    tm5_iview cloud_top, cloud_lfs;
    int lmax;
    double dt_conv;
    tm5_view emis;                        /* (im, jm) */
    double factor, ratio, efrac;
    int level, tracer;                    /* 1-based */
    const double *escale;                 /* device, nt: tracer t gets ((emis escale[t]) factor) ratio) efrac where
                                             escale[t] != 0 (an ensemble); NULL: tracer `tracer` gets it unscaled */
    tm5_view loss;                        /* (im, jm, lm) with the 4 rates at stride loss_stride */
    int64_t loss_stride;
    double dt_chem;
    const double *at, *bt;                /* for M */
} tm5_zphysics;

int tm5_zoom_physics(tm5_ctx *ctx, const tm5_zregion *r, const tm5_zphysics *p, const char *ops);

/* ------------------------------------------------------------------ the zoom tree's native driver */
/* A zoom tree resident on the device, run interval by interval by a native driver: the
   reference's zoom driver (pydyn/zoom/steps.py:15-88 do_steps / proces_region / step,
   pydyn/zoom/massflow.py:62-193 Setup_MassFlow with coarsen_region, pydyn/zoom/check_cfl.py:
   228-290 Check_CFL of the tree) over the bodies above, with every region's convection, sources
   (its emission, factor month2dt of the half step ndyn / (2 tref), into `tracer` at `level`, or
   every tracer's with escale) and chemistry (its decay over the half step) on the cells it owns;
   the fluxes of a step are the regions' steady pu0, pv0 times the meteo's amplitude a(t) at the
   step's middle (also stored into each region's pu, pv, as the reference stores the interpolated
   meteo).  The regions' arrays must be contiguous with i fastest in their declared bounds. */
typedef struct tm5_tree_desc {
    int nregions;
    const tm5_zregion *regions;   /* nregions, region 1 (the globe) first */
    const int *parent;            /* per region: its parent's 1-based index (0 for region 1) */
    const tm5_zphysics *physics;  /* per region: the column-physics inputs (dt_conv, factor, dt_chem set per step) */
    const tm5_view *pu0, *pv0;    /* per region: pu (0:im, 1:jm, 1:lm) and pv (1:im, 1:jm+1, 1:lm) at amplitude 1 */
    const double *at, *bt;        /* device, lm + 1 */
    double mod_period, mod_amplitude, sec_month;
    int limits;
} tm5_tree_desc;

typedef struct tm5_tree tm5_tree;

#define TM5_ENOMEM (-4)  /* host memory exhausted (nothing throws across the ABI) */

int tm5_tree_create(tm5_ctx *ctx, const tm5_tree_desc *desc, tm5_tree **out);
int tm5_tree_destroy(tm5_tree *tree);
/* One interval of nread seconds from t0 the reference's way: Check_CFL (the host reads its verdict
   once per attempt), then its n steps of ndyn; returns the (ndyn, n) Check_CFL chose. */
int tm5_tree_interval(tm5_tree *tree, double t0, int nread, int ndyn_max, int *ndyn, int *n);
/* count intervals from t0, as count calls of tm5_tree_interval would run them, queued: each on the
   previous interval's step, the trials that plan expects to fail rehearsed and verified on the
   device, the planned steps run as checked steps (the rehearsal's tests made inside the step's own
   kernels: x rows at nloop <= 9, y and z faces at nloop 1, so no search runs) with every region's
   m and moments backed up at the interval's start, the host reading the device once per batch.
   A step of a given ndyn is captured into a HIP graph the second time it is queued and replayed
   after that (its time comes from a device clock).  A miss restores the interval's state from the
   backup and redoes it through tm5_tree_interval.  The reference's queued, graph-replayed
   submission is Pipeline.submit (pydyn/gpu/zoom.py:944-1041). */
int tm5_tree_run(tm5_tree *tree, double t0, int nread, int ndyn_max, int count, int *ndyn, int *n);
/* Options: TM5_TREE_GRAPHS (1: capture and replay steps, the default; 0: issue every launch),
   TM5_TREE_BATCH (most intervals queued per read of the device, 1..64; 16), TM5_TREE_MODE: how a
   queued interval is verified -- TM5_TREE_CHECKED (checked steps, a backup of every region's m
   and moments per interval), TM5_TREE_REHEARSED (the planned trial rehearsed mass-only and its
   verdict checked on the device before plain steps; no backup) or TM5_TREE_AUTO (checked while
   the backup is at most 512 MB, the default). */
#define TM5_TREE_GRAPHS 0
#define TM5_TREE_BATCH 1
#define TM5_TREE_MODE 2
#define TM5_TREE_AUTO 0
#define TM5_TREE_CHECKED 1
#define TM5_TREE_REHEARSED 2
int tm5_tree_option(tm5_tree *tree, int which, int value);
/* (queued intervals kept, redone, steps replayed from graphs, graphs captured) since creation */
int tm5_tree_tally(tm5_tree *tree, long long *kept, long long *redone, long long *replayed, long long *captured);

/* ------------------------------------------------------------------ adjoints */
/* The adjoint of an advection operator's tracer map (axis 0 dynamu, 1 dynamv, 2 dynamw) at the
   forward trajectory: m, q, flux the operator's forward inputs (unchanged), lam the lambdas after
   it, overwritten with those before it.  The limiter's branches come from the recomputed forward
   sub-steps.  work: device scratch of (im+1)(jm+1)(lm+1) + im jm lm doubles.  No counterpart in
   the reference (it has no adjoint); the forward operators are pydyn/kernels.py:129-346. */
int tm5_advect_adjoint(tm5_ctx *ctx, int axis, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                       const tm5_view *flux, const tm5_tracers *lam, int limits, int max_nloop, double *work);
/* The adjoint of a column-physics sequence ("v", "csv", "vsc", "vsccsv"/"vscMcsv") of a single
   region on lambdas lam, with m the air mass the forward saw; the emission's lambda is added
   into lam_emis (im, jm).  p as for tm5_zoom_physics (at, bt unused). */
int tm5_column_adjoint(tm5_ctx *ctx, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *lam,
                       const tm5_zphysics *p, const char *ops, const tm5_view *lam_emis);

/* save / restore / scale of whole arrays (pydyn/zoom/host.py:52-71) */
int tm5_zoom_copy(tm5_ctx *ctx, const double *src, double *dst, int64_t n);
int tm5_zoom_scale(tm5_ctx *ctx, double *a, int64_t n, double factor);

#ifdef __cplusplus
}
#endif
#endif
