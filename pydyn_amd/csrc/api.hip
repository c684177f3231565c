// The extern "C" boundary of libtm5hip (include/tm5hip.h): the context, one entry point per
// operator, and the native driver of a single global region.
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "tm5_lib.h"

// tm5_convec_noalloc's buffer: one copy in every GPU's memory while the library is loaded.  Host
// code never passes `scratch` to a kernel or a copy (in host code the name is a host-side stand-in,
// not a device address): its device address comes from hipGetSymbolAddress, once per device, with
// that device current, and every use goes through that address.
static __device__ __align__(256) double scratch[TM5_NOALLOC_DOUBLES];
static constexpr int MAX_DEVICES = 64;
static double *g_scratch[MAX_DEVICES];  // the device address of `scratch` per device (null: not yet)
static std::mutex g_scratch_mu;

// the address of the current device's `scratch` (device: the current device)
static hipError_t scratch_address(int device, double **out)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    try {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch[device]) {
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return e;
            if (!p) return hipErrorInvalidSymbol;
            g_scratch[device] = (double *)p;
        }
        *out = g_scratch[device];
    } catch (...) {  // the mutex could not be locked: nothing may throw across the C ABI
        return hipErrorUnknown;
    }
    return hipSuccess;
}

// events of the launches timed while profiling, and the totals of the last profile_end
struct tm5_prof {
    std::vector<hipEvent_t> events;                  // created on demand, reused
    size_t used = 0;                                 // events recorded so far
    std::vector<const char *> names;                 // per span: its kernel (spans use events 2s, 2s+1)
    std::vector<const char *> total_names;
    std::vector<long long> counts;
    std::vector<double> seconds;
    bool failed = false;  // an event or a record could not be made: profiling stopped, no totals
};

namespace tm5 {

// Never throws and never fails the launch it brackets: when an event cannot be created or
// recorded (or a list cannot grow) profiling stops for the rest of the span, and the next
// tm5_profile_end reports TM5_ENOMEM instead of totals that would pair the wrong events.
void prof_mark(tm5_ctx *ctx, const char *name, bool stop)
{
    tm5_prof *p = ctx->prof;
    try {
        if (p->used == p->events.size()) {
            p->events.reserve(p->events.size() + 1);  // may throw, before the event exists
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) throw std::bad_alloc();
            p->events.push_back(e);
        }
        if (!stop) p->names.push_back(name);
    } catch (...) {
        p->failed = true;
        ctx->prof = nullptr;
        return;
    }
    if (hipEventRecord(p->events[p->used], ctx->stream) != hipSuccess) {
        p->failed = true;
        ctx->prof = nullptr;
        return;
    }
    p->used++;
}

int launch_convec(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 entu, View3 detu, View3 entd,
                  View3 detd, IView2 top, IView2 lfs, View3 dkg, double dt, int lmax);
int launch_dkg(tm5_ctx *ctx, int im, int jm, View3 kvh, View3 m, View3 gph, int lmax, View3 dkg);
int launch_add_2d(tm5_ctx *ctx, int im, int jm, View3 emis, double factor, double ratio, double efrac, Tracers q,
                  int level, int tracer);
int launch_decay(tm5_ctx *ctx, int im, int jm, int lm, View3 loss, long long sr, double dt, Tracers q, int t0,
                 int count);
int launch_scale(tm5_ctx *ctx, int n0, int n1, int n2, View3 a, double factor);



// a word of the context's scratch set from the stream (a kernel, so that it can be captured)
__global__ void k_set_word(int *words, int index, int value, int mask)
{
    if (threadIdx.x == 0) words[index] = (words[index] & ~mask) | (value & mask);
}

// A queued trial's verdict against the plan (expect_fail: the plan cuts the step after this
// trial): a different verdict makes every kernel queued after it skip.  The trial's bit is cleared.
__global__ void k_spec_check(int *words, int k, int expect_fail)
{
    if (threadIdx.x != 0) return;
    const int f = words[W_FLAGS];
    const bool failed = (f & TM5_REHEARSAL) != 0;
    int out = f & ~TM5_REHEARSAL;
    if (!(f & TM5_SPEC) && failed != (expect_fail != 0)) {
        out |= TM5_SPEC;
        words[W_SPEC_AT] = k;
    }
    words[W_FLAGS] = out;
}

__global__ void __launch_bounds__(256) k_copy3(int n0, int n1, int n2, View3 src, View3 dst)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n0 * n1 * n2) return;
    const int i = (int)(idx % n0), j = (int)((idx / n0) % n1), l = (int)(idx / ((long long)n0 * n1));
    dst(i, j, l) = src(i, j, l);
}

// the region's m and moments back from a checked step's backup (k_x's layout: i fastest, m first,
// then every tracer's rm, rxm, rym, rzm)
__global__ void __launch_bounds__(256) k_restore(int im, int jm, int lm, const double *bk, View3 m, Tracers q)
{
    const long long ncell = (long long)im * jm * lm;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncell) return;
    const int i = (int)(idx % im), j = (int)((idx / im) % jm), l = (int)(idx / ((long long)im * jm));
    m(i, j, l) = bk[idx];
    for (int t = 0; t < q.nt; t++) {
        const long long o = q.off(i, j, l, t);
        for (int a = 0; a < 4; a++) q.q[a][o] = bk[(1 + 4 * t + a) * ncell + idx];
    }
}

static int copy3(tm5_ctx *ctx, int n0, int n1, int n2, View3 src, View3 dst)
{
    const long long n = (long long)n0 * n1 * n2;
    TM5_LAUNCH(ctx, k_copy3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n0, n1, n2, src, dst);
    return status(hipGetLastError());
}

static int set_word(tm5_ctx *ctx, int index, int value, int mask)
{
    TM5_LAUNCH(ctx, k_set_word, dim3(1), dim3(64), 0, ctx->stream, ctx->words, index, value, mask);
    return status(hipGetLastError());
}

static int read_words(tm5_ctx *ctx)
{
    hipError_t e = hipMemcpyAsync(ctx->hwords, ctx->words, W_WORDS * sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return status(e);
}

static bool bad_view(const tm5_view *v) { return v == nullptr || v->data == nullptr; }
static bool bad_tracers(const tm5_tracers *q)
{
    return q == nullptr || q->nt < 0 || (q->nt > 0 && (!q->rm || !q->rxm || !q->rym || !q->rzm));
}

}  // namespace tm5

using namespace tm5;

#define TRY(x)                  \
    do {                        \
        int rc_ = (x);          \
        if (rc_ != 0) return rc_; \
    } while (0)

// ------------------------------------------------------------------ context
extern "C" int tm5_ctx_create(uint64_t stream, tm5_ctx **out)
{
    if (!out) return TM5_EINVAL;
    tm5_ctx *c = (tm5_ctx *)calloc(1, sizeof(tm5_ctx));
    if (!c) return TM5_EINVAL;
    c->stream = (hipStream_t)stream;
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    c->y_cap = 256;
    c->x_cap = 1 << 17;
    c->z_cap = 1 << 13;
    c->tile[TM5_TILE_Y_LINES] = 16;
    c->tile[TM5_TILE_Z_LINES] = 32;
    c->tile[TM5_TILE_DYNAM0] = 32;
    c->tile[TM5_TILE_M2PHLB] = 32;
    c->tile[TM5_TILE_Y_MANY] = 16;
    c->tile[TM5_TILE_RANGE_MIN] = 16;
    c->tile[TM5_TILE_APPLY] = 16;
    if (e == hipSuccess) e = hipMalloc((void **)&c->words, W_WORDS * sizeof(int));
    // per layer: nloop, the y check's tickets and verdicts (zero between launches)
    if (e == hipSuccess) e = hipMalloc((void **)&c->y_nloop, 3 * c->y_cap * sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->y_nloop, 0, 3 * c->y_cap * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&c->x_rows, c->x_cap * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&c->z_ok, c->z_cap * sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->hwords, W_WORDS * sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = scratch_address(c->device, &c->noalloc);  // the device is current here
    if (e != hipSuccess) {
        tm5_ctx_destroy(c);
        return (int)e;
    }
    memset(c->hwords, 0, W_WORDS * sizeof(int));
    e = hipMemcpyAsync(c->words, c->hwords, W_WORDS * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        tm5_ctx_destroy(c);
        return (int)e;
    }
    *out = c;
    return 0;
}

extern "C" int tm5_ctx_destroy(tm5_ctx *c)
{
    if (!c) return 0;
    if (c->stream || c->words) (void)hipStreamSynchronize(c->stream);
    if (c->prof_store) {
        for (hipEvent_t e : c->prof_store->events) (void)hipEventDestroy(e);
        delete c->prof_store;
    }
    if (c->words) (void)hipFree(c->words);
    if (c->y_nloop) (void)hipFree(c->y_nloop);
    if (c->x_rows) (void)hipFree(c->x_rows);
    if (c->z_ok) (void)hipFree(c->z_ok);
    if (c->cbuf) (void)hipFree(c->cbuf);
    if (c->hwords) (void)hipHostFree(c->hwords);
    free(c);
    return 0;
}

extern "C" int tm5_ctx_flags(tm5_ctx *c, int clear, int *flags)
{
    if (!c) return TM5_EINVAL;
    TRY(read_words(c));
    if (flags) *flags = c->hwords[W_FLAGS];
    if (clear) TRY(set_word(c, W_FLAGS, 0, ~0));
    return 0;
}

extern "C" int tm5_ctx_clear(tm5_ctx *c, int mask)
{
    if (!c) return TM5_EINVAL;
    return set_word(c, W_FLAGS, 0, mask);
}

extern "C" int tm5_profile_begin(tm5_ctx *c)
{
    if (!c) return TM5_EINVAL;
    TRY(status(hipStreamSynchronize(c->stream)));
    if (!c->prof_store) c->prof_store = new (std::nothrow) tm5_prof();
    if (!c->prof_store) return TM5_ENOMEM;
    c->prof = c->prof_store;
    c->prof->used = 0;
    c->prof->failed = false;
    c->prof->names.clear();
    return 0;
}

extern "C" int tm5_profile_end(tm5_ctx *c)
{
    if (!c || !c->prof_store || (!c->prof && !c->prof_store->failed)) return TM5_EINVAL;  // not profiling
    tm5_prof *p = c->prof_store;
    c->prof = nullptr;  // stop recording; the totals stay until the next begin
    TRY(status(hipStreamSynchronize(c->stream)));
    p->total_names.clear();
    p->counts.clear();
    p->seconds.clear();
    if (p->failed) {
        p->failed = false;
        return TM5_ENOMEM;
    }
    try {
        for (size_t s = 0; s < p->names.size() && 2 * s + 1 < p->used; s++) {
            float ms = 0.0f;
            TRY(status(hipEventElapsedTime(&ms, p->events[2 * s], p->events[2 * s + 1])));
            size_t k = 0;
            while (k < p->total_names.size() && strcmp(p->total_names[k], p->names[s]) != 0) k++;
            if (k == p->total_names.size()) {
                p->total_names.push_back(p->names[s]);
                p->counts.push_back(0);
                p->seconds.push_back(0.0);
            }
            p->counts[k] += 1;
            p->seconds[k] += ms * 1e-3;
        }
    } catch (const std::bad_alloc &) {
        p->total_names.clear();
        p->counts.clear();
        p->seconds.clear();
        return TM5_ENOMEM;
    }
    return 0;
}

extern "C" int tm5_profile_entry(tm5_ctx *c, int idx, const char **kernel, long long *count, double *seconds)
{
    if (!c) return TM5_EINVAL;
    tm5_prof *p = c->prof_store;
    if (!p || idx < 0 || (size_t)idx >= p->total_names.size()) return TM5_EINVAL;
    if (kernel) *kernel = p->total_names[idx];
    if (count) *count = p->counts[idx];
    if (seconds) *seconds = p->seconds[idx];
    return 0;
}

extern "C" int tm5_ctx_set_tile(tm5_ctx *c, int which, int value)
{
    static const int hi[TM5_TILES] = {16, 32, 64, 64, 16, 64, 32};
    if (!c || which < 0 || which >= TM5_TILES || value < 1 || value > hi[which]) return TM5_EINVAL;
    if (which == TM5_TILE_APPLY && value != 8 && value != 16 && value != 32) return TM5_EINVAL;
    c->tile[which] = value;
    return 0;
}

extern "C" int tm5_ctx_get_tile(tm5_ctx *c, int which, int *value)
{
    if (!c || !value || which < 0 || which >= TM5_TILES) return TM5_EINVAL;
    *value = c->tile[which];
    return 0;
}

extern "C" int tm5_ctx_nloop(tm5_ctx *c, int *x_max, int *y_max, int *nloop_layers, int lm, int *z_nadv)
{
    if (!c || lm < 0 || lm > c->y_cap) return TM5_EINVAL;
    TRY(read_words(c));
    if (x_max) {
        *x_max = 0;
        if (c->x_nrows > 0) {
            int *rows = (int *)malloc(c->x_nrows * sizeof(int));
            if (!rows) return TM5_EINVAL;
            hipError_t e = hipMemcpy(rows, c->x_rows, c->x_nrows * sizeof(int), hipMemcpyDeviceToHost);
            if (e == hipSuccess)
                for (int r = 0; r < c->x_nrows; r++) *x_max = rows[r] > *x_max ? rows[r] : *x_max;
            free(rows);
            if (e != hipSuccess) return (int)e;
        }
    }
    if (z_nadv) *z_nadv = c->hwords[W_Z_NADV];
    if (lm > 0 && (nloop_layers || y_max)) {
        int buf[256];
        hipError_t e = hipMemcpy(buf, c->y_nloop, lm * sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return (int)e;
        int mx = 0;
        for (int l = 0; l < lm; l++) mx = buf[l] > mx ? buf[l] : mx;
        if (nloop_layers) memcpy(nloop_layers, buf, lm * sizeof(int));
        if (y_max) *y_max = mx;
    }
    return 0;
}

// ------------------------------------------------------------------ operators
extern "C" int tm5_advect_x(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                            const tm5_view *am, int limits, int max_nloop)
{
    if (!c || bad_view(m) || bad_tracers(q) || bad_view(am)) return TM5_EINVAL;
    return launch_x(c, im, jm, lm, view3(m), tracers(q), view3(am), limits, max_nloop - 1, TM5_FAIL_X, TM5_ERRORS);
}

extern "C" int tm5_advect_y(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                            const tm5_view *bm, int limits, int max_nloop)
{
    if (!c || bad_view(m) || bad_tracers(q) || bad_view(bm)) return TM5_EINVAL;
    return launch_y(c, im, jm, lm, view3(m), tracers(q), view3(bm), limits, max_nloop - 1, TM5_FAIL_Y, TM5_ERRORS);
}

extern "C" int tm5_advect_z(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                            const tm5_view *cm, int limits, int iter_limit)
{
    if (!c || bad_view(m) || bad_tracers(q) || bad_view(cm)) return TM5_EINVAL;
    return launch_z(c, im, jm, lm, view3(m), tracers(q), view3(cm), limits, iter_limit, TM5_FAIL_Z, TM5_ERRORS);
}

static const int REHEARSAL_SKIP = TM5_ERRORS | TM5_REHEARSAL;

extern "C" int tm5_mass_x(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_view *am, int max_nloop)
{
    if (!c || bad_view(m) || bad_view(am)) return TM5_EINVAL;
    return launch_x(c, im, jm, lm, view3(m), no_tracers(), view3(am), 0, max_nloop - 1, TM5_REHEARSAL, REHEARSAL_SKIP);
}

extern "C" int tm5_mass_y(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_view *bm, int max_nloop)
{
    if (!c || bad_view(m) || bad_view(bm)) return TM5_EINVAL;
    return launch_y(c, im, jm, lm, view3(m), no_tracers(), view3(bm), 0, max_nloop, TM5_REHEARSAL, REHEARSAL_SKIP);
}

extern "C" int tm5_mass_z(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_view *cm, int iter_limit)
{
    if (!c || bad_view(m) || bad_view(cm)) return TM5_EINVAL;
    return launch_z(c, im, jm, lm, view3(m), no_tracers(), view3(cm), 0, iter_limit, TM5_REHEARSAL, REHEARSAL_SKIP);
}

extern "C" int tm5_dynam0(tm5_ctx *c, int im, int jm, int lm, const tm5_view *pu, const tm5_view *pv, double a,
                          const double *bt, double dt, const tm5_view *am, const tm5_view *bm, const tm5_view *cm)
{
    if (!c || bad_view(pu) || bad_view(pv) || bad_view(am) || bad_view(bm) || bad_view(cm)) return TM5_EINVAL;
    return launch_dynam0(c, im, jm, lm, view3(pu), view3(pv), a, bt, dt, view3(am), view3(bm), view3(cm), TM5_ERRORS);
}

extern "C" int tm5_m2phlb1(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const double *dxyp,
                           const double *at, const double *bt, const tm5_view *sp, const tm5_view *phlb)
{
    if (!c || bad_view(m) || bad_view(sp) || bad_view(phlb)) return TM5_EINVAL;
    return launch_m2phlb1(c, im, jm, lm, view3(m), dxyp, at, bt, view3(sp), view3(phlb));
}

extern "C" int tm5_convec(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                          const tm5_view *entu, const tm5_view *detu, const tm5_view *entd, const tm5_view *detd,
                          const tm5_iview *cloud_top, const tm5_iview *cloud_lfs, const tm5_view *dkg, double dt,
                          int lmax)
{
    if (!c || bad_view(m) || bad_tracers(q) || bad_view(entu) || bad_view(detu) || bad_view(entd) || bad_view(detd) ||
        bad_view(dkg) || !cloud_top || !cloud_lfs || !cloud_top->data || !cloud_lfs->data)
        return TM5_EINVAL;
    if (column_supported(lmax) && lm <= 64) {
        ColumnArgs a{};
        a.im = im;
        a.jm = jm;
        a.lm = lm;
        a.m = view3(m);
        a.q = tracers(q);
        a.entu = view3(entu);
        a.detu = view3(detu);
        a.entd = view3(entd);
        a.detd = view3(detd);
        a.dkg = view3(dkg);
        a.top = iview2(cloud_top);
        a.lfs = iview2(cloud_lfs);
        a.dt = dt;
        return launch_column(c, a, lmax, TM5_ERRORS);
    }
    return launch_convec(c, im, jm, lm, view3(m), tracers(q), view3(entu), view3(detu), view3(entd), view3(detd),
                         iview2(cloud_top), iview2(cloud_lfs), view3(dkg), dt, lmax);
}

// tm5_convec with the split form's matrices in the static `scratch` (through its device address
// c->noalloc), band by band of whole j rows: a column's convec reads and writes its own column
// only, so bands are independent.  Nothing is allocated.
extern "C" int tm5_convec_noalloc(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                                  const tm5_view *entu, const tm5_view *detu, const tm5_view *entd,
                                  const tm5_view *detd, const tm5_iview *cloud_top, const tm5_iview *cloud_lfs,
                                  const tm5_view *dkg, double dt, int lmax)
{
    if (!c || bad_view(m) || bad_tracers(q) || bad_view(entu) || bad_view(detu) || bad_view(entd) || bad_view(detd) ||
        bad_view(dkg) || !cloud_top || !cloud_lfs || !cloud_top->data || !cloud_lfs->data || im < 1 || jm < 1)
        return TM5_EINVAL;
    if (!column_supported(lmax) || lm > 64)  // the LDS form: no buffer
        return launch_convec(c, im, jm, lm, view3(m), tracers(q), view3(entu), view3(detu), view3(entd), view3(detd),
                             iview2(cloud_top), iview2(cloud_lfs), view3(dkg), dt, lmax);
    if (!c->noalloc) return TM5_EINVAL;
    ColumnArgs a{};
    a.im = im;
    a.lm = lm;
    a.dt = dt;
    const View3 vm = view3(m), ve[5] = {view3(entu), view3(detu), view3(entd), view3(detd), view3(dkg)};
    const Tracers vq = tracers(q);
    const IView2 vt = iview2(cloud_top), vl = iview2(cloud_lfs);
    if (vq.nt < column_split_nt()) {  // the one-kernel form: no buffer at all
        a.jm = jm;
        a.m = vm, a.q = vq, a.entu = ve[0], a.detu = ve[1], a.entd = ve[2], a.detd = ve[3], a.dkg = ve[4];
        a.top = vt, a.lfs = vl;
        return launch_column(c, a, lmax, TM5_ERRORS);
    }
    const size_t per_row = (size_t)im * column_matrix_doubles(lmax);
    if (per_row > TM5_NOALLOC_DOUBLES) return TM5_ESCRATCH;
    const int band = (int)(TM5_NOALLOC_DOUBLES / per_row < (size_t)jm ? TM5_NOALLOC_DOUBLES / per_row : (size_t)jm);
    for (int j0 = 0; j0 < jm; j0 += band) {
        a.jm = jm - j0 < band ? jm - j0 : band;
        a.m = View3{vm.p + (long long)j0 * vm.sj, vm.si, vm.sj, vm.sl};
        View3 *dst[5] = {&a.entu, &a.detu, &a.entd, &a.detd, &a.dkg};
        for (int k = 0; k < 5; k++) *dst[k] = View3{ve[k].p + (long long)j0 * ve[k].sj, ve[k].si, ve[k].sj, ve[k].sl};
        a.q = vq;
        for (int k = 0; k < 4; k++) a.q.q[k] = vq.q[k] + (long long)j0 * vq.sj;
        a.top = IView2{vt.p + (long long)j0 * vt.sj, vt.si, vt.sj};
        a.lfs = IView2{vl.p + (long long)j0 * vl.sj, vl.si, vl.sj};
        TRY(launch_column(c, a, lmax, TM5_ERRORS, c->noalloc, (size_t)TM5_NOALLOC_DOUBLES));
    }
    return 0;
}

// convec with the LAPACK-kind inverse (DGETRF2 + DGETRI per column), on every column or (mask
// non-null) on the columns the mask marks
extern "C" int tm5_convec_lapack(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                                 const tm5_view *entu, const tm5_view *detu, const tm5_view *entd,
                                 const tm5_view *detd, const tm5_iview *cloud_top, const tm5_iview *cloud_lfs,
                                 const tm5_view *dkg, double dt, int lmax, const int *mask)
{
    if (!c || bad_view(m) || bad_tracers(q) || bad_view(entu) || bad_view(detu) || bad_view(entd) || bad_view(detd) ||
        bad_view(dkg) || !cloud_top || !cloud_lfs || !cloud_top->data || !cloud_lfs->data)
        return TM5_EINVAL;
    if (!column_supported(lmax) || lm > 64) return TM5_EINVAL;
    ColumnArgs a{};
    a.im = im;
    a.jm = jm;
    a.lm = lm;
    a.m = view3(m);
    a.q = tracers(q);
    a.entu = view3(entu);
    a.detu = view3(detu);
    a.entd = view3(entd);
    a.detd = view3(detd);
    a.dkg = view3(dkg);
    a.top = iview2(cloud_top);
    a.lfs = iview2(cloud_lfs);
    a.dt = dt;
    a.owned = mask;
    a.lapack = 1;
    return launch_column(c, a, lmax, TM5_ERRORS);
}

// the LAPACK-kind inverse of ncol row-major n x n device matrices A into Y; info: a device int
// that a zero pivot sets to 1 (null: TM5_SINGULAR in the flag word)
extern "C" int tm5_inverse_rows(tm5_ctx *c, long long ncol, int n, const double *A, double *Y, int *info)
{
    if (!c) return TM5_EINVAL;
    return launch_lapack_inverse(c, ncol, n, A, Y, info, TM5_ERRORS);
}

// convec on the columns an (im, jm) mask (i fastest) marks; the others keep their tracers
extern "C" int tm5_convec_masked(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                                 const tm5_view *entu, const tm5_view *detu, const tm5_view *entd,
                                 const tm5_view *detd, const tm5_iview *cloud_top, const tm5_iview *cloud_lfs,
                                 const tm5_view *dkg, double dt, int lmax, const int *mask)
{
    if (!c || bad_view(m) || bad_tracers(q) || bad_view(entu) || bad_view(detu) || bad_view(entd) || bad_view(detd) ||
        bad_view(dkg) || !cloud_top || !cloud_lfs || !cloud_top->data || !cloud_lfs->data || !mask)
        return TM5_EINVAL;
    if (!column_supported(lmax) || lm > 64) return TM5_EINVAL;
    ColumnArgs a{};
    a.im = im;
    a.jm = jm;
    a.lm = lm;
    a.m = view3(m);
    a.q = tracers(q);
    a.entu = view3(entu);
    a.detu = view3(detu);
    a.entd = view3(entd);
    a.detd = view3(detd);
    a.dkg = view3(dkg);
    a.top = iview2(cloud_top);
    a.lfs = iview2(cloud_lfs);
    a.dt = dt;
    a.owned = mask;
    return launch_column(c, a, lmax, TM5_ERRORS);
}

extern "C" int tm5_advect_x_rows(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                                 const tm5_view *am, const int *nrow, int limits, int max_nloop, int *nloop,
                                 double *amax)
{
    if (!c || bad_view(m) || bad_tracers(q) || bad_view(am)) return TM5_EINVAL;
    return launch_x(c, im, jm, lm, view3(m), tracers(q), view3(am), limits, max_nloop - 1, TM5_FAIL_X, TM5_ERRORS, nrow,
                    nloop, amax);
}

extern "C" int tm5_mass_x_rows(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_view *am,
                               const int *nrow, int max_nloop)
{
    if (!c || bad_view(m) || bad_view(am)) return TM5_EINVAL;
    return launch_x(c, im, jm, lm, view3(m), no_tracers(), view3(am), 0, max_nloop - 1, TM5_REHEARSAL, REHEARSAL_SKIP,
                    nrow);
}

extern "C" int tm5_advect_z_substeps(tm5_ctx *c, int im, int jm, int lm, const tm5_view *m, const tm5_tracers *q,
                                     const tm5_view *cm, int n_advz, int limits, int prelimit)
{
    if (!c || bad_view(m) || bad_tracers(q) || bad_view(cm)) return TM5_EINVAL;
    return launch_z_substeps(c, im, jm, lm, view3(m), tracers(q), view3(cm), n_advz, limits, prelimit);
}

extern "C" int tm5_cluster_sum(tm5_ctx *c, int n, int cs, int r1, int r2, const tm5_view *q, const tm5_view *out)
{
    if (!c || bad_view(q) || bad_view(out)) return TM5_EINVAL;
    return launch_cluster_sum(c, n, cs, r1, r2, view3(q), view3(out));
}

extern "C" int tm5_dkg(tm5_ctx *c, int im, int jm, const tm5_view *kvh, const tm5_view *m, const tm5_view *gph,
                       int lmax, const tm5_view *dkg)
{
    if (!c || bad_view(kvh) || bad_view(m) || bad_view(gph) || bad_view(dkg)) return TM5_EINVAL;
    return launch_dkg(c, im, jm, view3(kvh), view3(m), view3(gph), lmax, view3(dkg));
}

extern "C" int tm5_add_2d(tm5_ctx *c, int im, int jm, const tm5_view *emis, double factor, double ratio, double efrac,
                          const tm5_tracers *q, int level, int tracer)
{
    if (!c || bad_view(emis) || bad_tracers(q)) return TM5_EINVAL;
    return launch_add_2d(c, im, jm, view3(emis), factor, ratio, efrac, tracers(q), level, tracer);
}

extern "C" int tm5_decay(tm5_ctx *c, int im, int jm, int lm, const tm5_view *loss, double dt, const tm5_tracers *q,
                         int t0, int count)
{
    if (!c || bad_view(loss) || bad_tracers(q)) return TM5_EINVAL;
    return launch_decay(c, im, jm, lm, view3(loss), loss->stride[3], dt, tracers(q), t0, count);
}

extern "C" int tm5_loss_rates(tm5_ctx *c, int im, int jm, int lm, int nstrat, int ntrop, const tm5_view *temper,
                              const tm5_view *phlb, const double *trop, const double *edges, const double *strat_oh,
                              const double *strat_cl, const double *strat_od, const tm5_view *oh_3d,
                              const tm5_view *pl_trop, const tm5_view *cl_trop, const tm5_view *soil,
                              double oh_scale, double cl_scale, double o1d_scale, double soil_rate, double dt,
                              const tm5_view *loss, const tm5_view *decay)
{
    if (!c || bad_view(temper) || bad_view(phlb) || bad_view(oh_3d) || bad_view(loss)) return TM5_EINVAL;
    LossArgs a{};
    a.im = im;
    a.jm = jm;
    a.lm = lm;
    a.nstrat = nstrat;
    a.ntrop = ntrop;
    a.temper = view3(temper);
    a.phlb = view3(phlb);
    a.oh3d = view3(oh_3d);
    a.trop = trop;
    a.edges = edges;
    a.s_oh = strat_oh;
    a.s_cl = strat_cl;
    a.s_od = strat_od;
    if (pl_trop && pl_trop->data) a.pl = view3(pl_trop);
    if (cl_trop && cl_trop->data) a.cl = view3(cl_trop);
    if (soil && soil->data) a.soil = view3(soil);
    a.oh_scale = oh_scale;
    a.cl_scale = cl_scale;
    a.o1d_scale = o1d_scale;
    a.soil_rate = soil_rate;
    a.dt = dt;
    a.loss = view3(loss);
    a.loss_sr = loss->stride[3];
    if (decay && decay->data) a.decay = view3(decay);
    return launch_loss_rates(c, a);
}

extern "C" int tm5_calc_kzz(tm5_ctx *c, int im, int jm, int lm, int lmax, const tm5_view *spm, const tm5_view *sp,
                            const tm5_view *pu, const tm5_view *pv, const tm5_view *t, const tm5_view *q,
                            const tm5_view *surface, const double *at, const double *bt, const double *dxyp,
                            const double *dxx, double dyy, const tm5_view *gph, const tm5_view *dkg,
                            const tm5_view *ustar, const tm5_view *pblh, const tm5_view *kvh)
{
    if (!c || bad_view(spm) || bad_view(sp) || bad_view(pu) || bad_view(pv) || bad_view(t) || bad_view(q) ||
        !surface || bad_view(gph) || bad_view(dkg))
        return TM5_EINVAL;
    KzzArgs a{};
    a.im = im;
    a.jm = jm;
    a.lm = lm;
    a.lmax = lmax;
    a.spm = view3(spm);
    a.sp = view3(sp);
    a.pu = view3(pu);
    a.pv = view3(pv);
    a.t = view3(t);
    a.q = view3(q);
    for (int k = 0; k < 10; k++) {
        if (bad_view(surface + k)) return TM5_EINVAL;
        a.surf[k] = view3(surface + k);
    }
    a.at = at;
    a.bt = bt;
    a.dxyp = dxyp;
    a.dxx = dxx;
    a.dyy = dyy;
    a.gph = view3(gph);
    a.dkg = view3(dkg);
    if (ustar && ustar->data) a.ustar = view3(ustar);
    if (pblh && pblh->data) a.pblh = view3(pblh);
    if (kvh && kvh->data) a.kvh = view3(kvh);
    return launch_calc_kzz(c, a);
}

extern "C" int tm5_libm(tm5_ctx *c, int fn, int64_t n, const double *x, const double *y, double *out)
{
    if (!c) return TM5_EINVAL;
    return launch_libm(c, fn, (long long)n, x, y, out);
}

extern "C" int tm5_scale(tm5_ctx *c, int n0, int n1, int n2, const tm5_view *a, double factor)
{
    if (!c || bad_view(a)) return TM5_EINVAL;
    return launch_scale(c, n0, n1, n2, view3(a), factor);
}

// ------------------------------------------------------------------ the region driver
struct tm5_region {
    tm5_ctx *ctx;
    tm5_region_desc d;
    int skip;         // the step kernels' skip mask: TM5_ERRORS, with TM5_SPEC while queuing ahead
    bool ms_fresh;    // ms holds m (the next rehearsal need not copy it)
    int plan;         // tm5_region_run: the step the last interval settled on (0: none yet)
    int batch;        // tm5_region_run: intervals it queues per read of the device
    double *scratch;  // the rehearsal's m, am, bm, cm (i-fastest, contiguous)
    View3 ms, ams, bms, cms;
    double *backup;   // tm5_region_run: a queued interval's m and moments at its start (k_x's layout),
                      // two of them (even and odd intervals)
    size_t backup_n;  // the length of each in doubles
};

static View3 packed(double *p, int n0, int n1) { return View3{p, 1, n0, (long long)n0 * n1}; }

extern "C" int tm5_region_create(tm5_ctx *ctx, const tm5_region_desc *d, tm5_region **out)
{
    if (!ctx || !d || !out || d->im < 1 || d->jm < 2 || d->lm < 2 || d->lmax < 1 || d->lmax > d->lm) return TM5_EINVAL;
    tm5_region *r = (tm5_region *)calloc(1, sizeof(tm5_region));
    if (!r) return TM5_EINVAL;
    r->ctx = ctx;
    r->d = *d;
    r->skip = TM5_ERRORS;
    const int im = d->im, jm = d->jm, lm = d->lm;
    const size_t nm = (size_t)im * jm * lm, na = (size_t)(im + 1) * jm * lm, nb = (size_t)im * (jm + 1) * lm,
                 nc = (size_t)im * jm * (lm + 1);
    hipError_t e = hipMalloc((void **)&r->scratch, (nm + na + nb + nc) * sizeof(double));
    if (e != hipSuccess) {
        free(r);
        return (int)e;
    }
    r->ms = packed(r->scratch, im, jm);
    r->ams = packed(r->scratch + nm, im + 1, jm);
    r->bms = packed(r->scratch + nm + na, im, jm + 1);
    r->cms = packed(r->scratch + nm + na + nb, im, jm);
    *out = r;
    return 0;
}

extern "C" int tm5_region_destroy(tm5_region *r)
{
    if (!r) return 0;
    (void)hipStreamSynchronize(r->ctx->stream);
    (void)hipFree(r->scratch);
    if (r->backup) (void)hipFree(r->backup);
    free(r);
    return 0;
}

// pydyn_amd.meteo's amplitude(t): the same operations, so the same double
static double amplitude(const tm5_region_desc &d, double t)
{
    const double x = t / d.mod_period;
    const double frac = x - floor(x);
    return 1.0 + d.mod_amplitude * (1.0 - 4.0 * fabs(frac - 0.5));
}

// pydyn/cfl.py:208-213
static int new_valid_timestep(int dtime, int nread, int cfl_outputstep, int tref_max)
{
    for (int dt = dtime - 1; dt > 0; dt--)
        if (nread % dt == 0 && cfl_outputstep % dt == 0 && dt % (2 * tref_max) == 0) return dt;
    return -1;
}

// Fortran nint (pydyn/ops.py:78-80)
static int nint(double x) { return (int)(floor(fabs(x) + 0.5) * (x >= 0 ? 1.0 : -1.0)); }

static int dynam0_at(tm5_region *r, double t, int ndyn, View3 am, View3 bm, View3 cm)
{
    const tm5_region_desc &d = r->d;
    const double a = amplitude(d, t);
    return launch_dynam0(r->ctx, d.im, d.jm, d.lm, view3(&d.pu0), view3(&d.pv0), a, d.bt, (double)ndyn / 2.0, am, bm,
                         cm, r->skip, 1);
}

// The kernels of Check_CFL's rehearsal of the n steps of ndyn on the scratch copy.
static int rehearse_queue(tm5_region *r, double t0, int ndyn, int n, bool clear = true)
{
    tm5_ctx *c = r->ctx;
    const tm5_region_desc &d = r->d;
    const int im = d.im, jm = d.jm, lm = d.lm;
    const int skip = REHEARSAL_SKIP | (r->skip & TM5_SPEC);
    if (clear) TRY(set_word(c, W_FLAGS, 0, TM5_REHEARSAL));
    if (!r->ms_fresh) TRY(copy3(c, im, jm, lm, view3(&d.m), r->ms));
    r->ms_fresh = false;  // the trial moves ms
    for (int i = 1; i <= n; i++) {
        TRY(dynam0_at(r, t0 + ((double)i - 0.5) * (double)ndyn, ndyn, r->ams, r->bms, r->cms));
        const char *order = "xyzzyx";
        for (const char *op = order; *op; op++) {
            if (*op == 'x')
                TRY(launch_x(c, im, jm, lm, r->ms, no_tracers(), r->ams, 0, 10 - 1, TM5_REHEARSAL, skip));
            else if (*op == 'y')  // caps 1: the check at nloop 1 and its step
                TRY(launch_y_mass1(c, im, jm, lm, r->ms, r->bms, TM5_REHEARSAL, skip));
            else
                TRY(launch_z_mass1(c, im, jm, lm, r->ms, r->cms, TM5_REHEARSAL, skip));
        }
    }
    return 0;
}

// Check_CFL's rehearsal of the n steps of ndyn on the scratch copy (pydyn/zoom/check_cfl.py:
// 258-290 on one region): the verdict comes back in the rehearsal bit.
static int rehearse(tm5_region *r, double t0, int ndyn, int n, bool *ok)
{
    tm5_ctx *c = r->ctx;
    TRY(rehearse_queue(r, t0, ndyn, n));
    TRY(read_words(c));
    if (c->hwords[W_FLAGS] & TM5_ERRORS) return TM5_EFLAGS;
    *ok = (c->hwords[W_FLAGS] & TM5_REHEARSAL) == 0;
    return 0;
}

// A step after a passing Check_CFL rehearsal of its ndyn (known_one), or a checked step that makes
// the rehearsal's tests itself (spec_k >= 0: queued interval spec_k; a failed test raises
// TM5_SPEC).  Both rest on one fact: the rehearsal's mass-only operators -- dynamum at cap 10,
// dynamvm and dynamwm at cap 1, in the order x y z z y x -- move m through the same bits as the
// step's operators, given the same fluxes (the mass-only x sub-steps add the same faces as
// dynamu's; convection, sources, chemistry and m2phlb1 leave m alone).  So after a passing
// rehearsal every y layer and the z region pass at 1, and a step whose x rows pass at nloop <= 9
// and whose y and z faces pass at 1 is exactly a step whose rehearsal passes.
// backup (checked x only): where the x kernel leaves the m and moments it reads.
// flux (x): the step's fluxes made by the launch (am by its rows, bm and cm by extra workgroups)
// instead of a dynam0 launch before it
static int operator_(tm5_region *r, char op, int ndyn, bool known_one = false, int spec_k = -1,
                     double *backup = nullptr, const FluxSrc *flux = nullptr)
{
    tm5_ctx *c = r->ctx;
    const tm5_region_desc &d = r->d;
    const int im = d.im, jm = d.jm, lm = d.lm;
    const Tracers q = tracers(&d.q);
    switch (op) {
    case 'x':
        if (spec_k >= 0)  // the rehearsal's cap (its search stops past 10 - 1)
            return launch_x(c, im, jm, lm, view3(&d.m), q, view3(&d.am), d.limits, 10 - 1, TM5_FAIL_X, r->skip,
                            nullptr, nullptr, nullptr, spec_k, backup, 0, -1, flux);
        return launch_x(c, im, jm, lm, view3(&d.m), q, view3(&d.am), d.limits, 50 - 1, TM5_FAIL_X, r->skip,
                        nullptr, nullptr, nullptr, -1, nullptr, 0, -1, flux);
    case 'y':
        return launch_y(c, im, jm, lm, view3(&d.m), q, view3(&d.bm), d.limits, 10 - 1, TM5_FAIL_Y, r->skip,
                        known_one, spec_k);
    case 'z':
        return launch_z(c, im, jm, lm, view3(&d.m), q, view3(&d.cm), d.limits, 15, TM5_FAIL_Z, r->skip, known_one,
                        spec_k);
    case 'v':
        return launch_convec(c, im, jm, lm, view3(&d.m), q, view3(&d.entu), view3(&d.detu), view3(&d.entd),
                             view3(&d.detd), iview2(&d.cloud_top), iview2(&d.cloud_lfs), view3(&d.dkg),
                             (double)ndyn / 2.0, d.lmax);
    case 's': {
        const double factor = ((double)ndyn / 2.0) / d.sec_month;  // month2dt, pydyn/emis.py:65-68
        return launch_add_2d(c, im, jm, view3(&d.emis), factor, d.mol_ratio, 1.0, q, 1, 1);
    }
    case 'c':
        return launch_decay(c, im, jm, lm, view3(&d.loss), d.loss.stride[3], (double)ndyn / 2.0, q, 0, q.nt);
    }
    return TM5_EINVAL;
}

// "v s c m2phlb1 c s v" in one kernel when lmax has a register form (column.hip), else the
// operators one by one.  with_z: a step's "z ... z" around it, at n_advz 1 (operator_'s
// known_one and spec_k), fused into the column kernel where it has one.
static int middle(tm5_region *r, int ndyn, bool with_z = false, int spec_k = -1)
{
    const tm5_region_desc &d = r->d;
    if (!column_supported(d.lmax) || d.lm > 64) {
        if (with_z) TRY(operator_(r, 'z', ndyn, true, spec_k));
        for (const char *op = "vsc"; *op; op++) TRY(operator_(r, *op, ndyn));
        TRY(launch_m2phlb1(r->ctx, d.im, d.jm, d.lm, view3(&d.m), d.dxyp, d.at, d.bt, view3(&d.sp), view3(&d.phlb)));
        for (const char *op = "csv"; *op; op++) TRY(operator_(r, *op, ndyn));
        if (with_z) TRY(operator_(r, 'z', ndyn, true, spec_k));
        return 0;
    }
    ColumnArgs a{};
    a.im = d.im;
    a.jm = d.jm;
    a.lm = d.lm;
    a.m = view3(&d.m);
    a.q = tracers(&d.q);
    a.entu = view3(&d.entu);
    a.detu = view3(&d.detu);
    a.entd = view3(&d.entd);
    a.detd = view3(&d.detd);
    a.dkg = view3(&d.dkg);
    a.top = iview2(&d.cloud_top);
    a.lfs = iview2(&d.cloud_lfs);
    a.dt = (double)ndyn / 2.0;
    a.twice = 1;
    a.mid = 1;
    a.emis = view3(&d.emis);
    a.factor = ((double)ndyn / 2.0) / d.sec_month;  // month2dt, pydyn/emis.py:65-68
    a.ratio = d.mol_ratio;
    a.efrac = 1.0;
    a.level = 0;
    a.tracer = 0;
    a.loss = view3(&d.loss);
    a.loss_sr = d.loss.stride[3];
    a.dt_chem = (double)ndyn / 2.0;
    a.dxyp = d.dxyp;
    a.at = d.at;
    a.bt = d.bt;
    a.sp = view3(&d.sp);
    a.phlb = view3(&d.phlb);
    a.zfused = with_z ? 1 : 0;
    a.zlimits = d.limits;
    a.spec_k = spec_k;
    a.cm = view3(&d.cm);
    return launch_column(r->ctx, a, d.lmax, r->skip);
}

// the fluxes of a step of ndyn whose middle is at t, made by its first x launch (Setup_MassFlow's
// dynam0 at t with pydyn_amd.meteo's amplitude: am by the rows, bm and cm by extra workgroups)
static FluxSrc step_flux(const tm5_region *r, double t, int ndyn)
{
    const tm5_region_desc &d = r->d;
    FluxSrc f{};
    f.u = view3(&d.pu0);
    f.v = view3(&d.pv0);
    f.a = amplitude(d, t);
    f.dt = (double)ndyn / 2.0;
    f.bt = d.bt;
    f.bm = view3(&d.bm);
    f.cm = view3(&d.cm);
    f.TI = r->ctx->tile[TM5_TILE_DYNAM0];
    return f;
}

// one step of ndyn after Check_CFL's rehearsal of it passed
static int step(tm5_region *r, int ndyn)
{
    const tm5_region_desc &d = r->d;
    TRY(operator_(r, 'x', ndyn, true));
    TRY(operator_(r, 'y', ndyn, true));
    TRY(middle(r, ndyn, true));  // "z v s c m2phlb1 c s v z"
    TRY(operator_(r, 'y', ndyn, true));
    TRY(operator_(r, 'x', ndyn, true));
    TRY(launch_m2phlb1(r->ctx, d.im, d.jm, d.lm, view3(&d.m), d.dxyp, d.at, d.bt, view3(&d.sp), view3(&d.phlb)));
    r->ms_fresh = false;
    return 0;
}

// A checked step of queued interval k (the rehearsal's tests made by its own kernels; backup: the
// first of its interval writes its input there) whose fluxes its first x launch makes, no dynam0
// launched.  Its last x is left pending (*pend: the interval it checks): a pending x and the next
// step's first run as one launch (k_xx), and flush_x launches one alone.  Its closing m2phlb1 is
// left to close_step: the column kernel's middle m2phlb1 overwrites sp and phlb before anything
// reads them, so only the last step before the state is seen (a trial's rehearsal, the end of the
// batch) is closed.
static int queued_step(tm5_region *r, int ndyn, int k, double *backup, const FluxSrc &fl, int *pend)
{
    const tm5_region_desc &d = r->d;
    if (*pend >= 0)
        TRY(launch_xx(r->ctx, d.im, d.jm, d.lm, view3(&d.m), tracers(&d.q), view3(&d.am), d.limits, 10 - 1, r->skip,
                      *pend, k, backup, fl));
    else
        TRY(operator_(r, 'x', ndyn, true, k, backup, &fl));
    *pend = -1;
    TRY(operator_(r, 'y', ndyn, true, k));
    TRY(middle(r, ndyn, true, k));  // "z v s c m2phlb1 c s v z"
    TRY(operator_(r, 'y', ndyn, true, k));
    *pend = k;
    return 0;
}

// the pending last x of a queued step, alone
static int flush_x(tm5_region *r, int *pend)
{
    if (*pend < 0) return 0;
    const int k = *pend;
    *pend = -1;
    return operator_(r, 'x', 0, true, k);
}

// An open queued step's m2phlb1, with the copy of m for the next rehearsal
static int close_step(tm5_region *r)
{
    const tm5_region_desc &d = r->d;
    TRY(launch_m2phlb1(r->ctx, d.im, d.jm, d.lm, view3(&d.m), d.dxyp, d.at, d.bt, view3(&d.sp), view3(&d.phlb),
                       &r->ms));
    r->ms_fresh = true;
    return 0;
}

extern "C" int tm5_region_interval(tm5_region *r, double t0, int nread, int ndyn_max, int *ndyn_out, int *n_out)
{
    if (!r || nread < 1 || ndyn_max < 1 || nread % ndyn_max != 0) return TM5_EINVAL;
    r->ms_fresh = false;  // the caller may have changed m since the last call
    const tm5_region_desc &d = r->d;
    int ndyn = ndyn_max, n = nread / ndyn_max;
    bool ok = false;
    for (int iteration = 1; iteration <= 100 && !ok; iteration++) {
        TRY(rehearse(r, t0, ndyn, n, &ok));
        if (ok) break;
        const int ndyn_old = ndyn;
        ndyn = new_valid_timestep(ndyn, nread, nread, 1);
        if (ndyn < 1) return TM5_EINVAL;
        n = nint((double)(n * ndyn_old) / (double)ndyn);
        const double fraction = (double)ndyn / (double)ndyn_old;
        TRY(launch_scale(r->ctx, d.im + 1, d.jm, d.lm, view3(&d.am), fraction));
        TRY(launch_scale(r->ctx, d.im, d.jm + 1, d.lm, view3(&d.bm), fraction));
        TRY(launch_scale(r->ctx, d.im, d.jm, d.lm + 1, view3(&d.cm), fraction));
    }
    if (!ok) return TM5_EINVAL;
    for (int s = 0; s < n; s++) {
        TRY(dynam0_at(r, (t0 + (double)(s * ndyn)) + (double)ndyn / 2.0, ndyn, view3(&d.am), view3(&d.bm),
                      view3(&d.cm)));
        TRY(step(r, ndyn));
    }
    if (ndyn_out) *ndyn_out = ndyn;
    if (n_out) *n_out = n;
    return 0;
}

extern "C" int tm5_region_ops(tm5_region *r, double t, int ndyn, const char *ops)
{
    if (!r || !ops || ndyn < 1) return TM5_EINVAL;
    r->ms_fresh = false;
    const tm5_region_desc &d = r->d;
    for (const char *op = ops; *op; op++) {
        switch (*op) {
        case 'D': TRY(dynam0_at(r, t, ndyn, view3(&d.am), view3(&d.bm), view3(&d.cm))); break;
        case 'x': case 'y': case 'z': TRY(operator_(r, *op, ndyn)); break;
        case 'V': TRY(middle(r, ndyn)); break;
        case 'M':
            TRY(launch_m2phlb1(r->ctx, d.im, d.jm, d.lm, view3(&d.m), d.dxyp, d.at, d.bt, view3(&d.sp), view3(&d.phlb)));
            break;
        default: return TM5_EINVAL;
        }
    }
    return 0;
}

// new_valid_timestep's chain from ndyn_max down to ndyn: the trials Check_CFL makes when it
// settles on ndyn, each with its n; false if ndyn is not on the chain
static bool trial_chain(int nread, int ndyn_max, int ndyn, int *nd, int *nn, int *count)
{
    int a = ndyn_max, n = nread / ndyn_max, k = 0;
    for (;;) {
        if (k >= 100) return false;
        nd[k] = a;
        nn[k] = n;
        k++;
        if (a == ndyn) break;
        const int b = new_valid_timestep(a, nread, nread, 1);
        if (b < 1) return false;
        n = nint((double)(n * a) / (double)b);
        a = b;
    }
    *count = k;
    return true;
}

extern "C" int tm5_region_run(tm5_region *r, double t0, int nread, int ndyn_max, int count, int *ndyn_out, int *n_out)
{
    if (!r || !ndyn_out || !n_out || count < 0 || nread < 1 || ndyn_max < 1 || nread % ndyn_max != 0)
        return TM5_EINVAL;
    tm5_ctx *c = r->ctx;
    const tm5_region_desc &d = r->d;
    r->ms_fresh = false;  // the caller may have changed m since the last call
    // queuing ahead needs every step kernel to honour the skip mask (the register column form)
    if (!column_supported(d.lmax) || d.lm > 64) {
        for (int k = 0; k < count; k++)
            TRY(tm5_region_interval(r, t0 + (double)k * (double)nread, nread, ndyn_max, ndyn_out + k, n_out + k));
        return 0;
    }
    // two backups, by the parity of the interval: a k_xx launch writes interval k's while it still
    // checks interval k - 1, whose backup a failure there restores
    const size_t ncell = (size_t)d.im * d.jm * d.lm, nb = ncell * (1 + 4 * (size_t)d.q.nt);
    if (r->backup_n < nb) {
        if (r->backup) TRY(status(hipFree(r->backup)));
        r->backup = nullptr;
        r->backup_n = 0;
        TRY(status(hipMalloc((void **)&r->backup, 2 * nb * sizeof(double))));
        r->backup_n = nb;
    }
    auto backup_of = [&](int k) { return r->backup + (size_t)(k & 1) * nb; };
    TRY(set_word(c, W_FLAGS, 0, TM5_REHEARSAL | TM5_SPEC));
    TRY(set_word(c, W_SPEC_KIND, 0, -1));
    TRY(set_word(c, W_SPEC_AT, INT_MAX, -1));  // k_xx keeps the smallest failing interval
    // intervals queued per read of the device: doubled after a batch that held to its plan, back
    // to 2 after one that did not (the kernels queued after a miss only skip, but each costs a
    // dispatch)
    int &batch = r->batch;
    if (batch < 2) batch = 2;
    int &plan = r->plan;  // the step Check_CFL is predicted to settle on: the previous interval's
    if (plan < 1 || plan > ndyn_max) plan = ndyn_max;
    int i = 0, miss_at = -1;  // miss_at: the interval of the last miss
    while (i < count) {
        int nd[100], nn[100], trials = 0;
        if (!trial_chain(nread, ndyn_max, plan, nd, nn, &trials)) {
            plan = ndyn_max;
            trial_chain(nread, ndyn_max, plan, nd, nn, &trials);
        }
        const int end = i + batch < count ? i + batch : count;
        r->skip = TM5_ERRORS | TM5_SPEC;
        int rc = 0;
        bool open = false;  // a queued step whose m2phlb1 is yet to be launched
        int pend = -1;      // the interval of the queued step whose last x is yet to be launched
        for (int k = i; k < end && rc == 0; k++) {
            const double t = t0 + (double)k * (double)nread;
            // the trials the plan expects to fail are rehearsed; the last one is the steps
            // themselves, checked
            if (trials > 1 && open) {
                rc = flush_x(r, &pend);
                if (rc == 0) rc = close_step(r);
                open = false;
            }
            for (int q = 0; q + 1 < trials && rc == 0; q++) {
                rc = rehearse_queue(r, t, nd[q], nn[q], false);  // k_spec_check clears the bit
                if (rc == 0) {
                    TM5_LAUNCH(c, k_spec_check, dim3(1), dim3(64), 0, c->stream, c->words, k, 1);
                    rc = status(hipGetLastError());
                }
            }
            const int ndyn = nd[trials - 1], n = nn[trials - 1];
            for (int s = 0; s < n && rc == 0; s++) {
                const double ts = (t + (double)(s * ndyn)) + (double)ndyn / 2.0;
                rc = queued_step(r, ndyn, k, s == 0 ? backup_of(k) : nullptr, step_flux(r, ts, ndyn), &pend);
                open = rc == 0;
            }
        }
        if (rc == 0) rc = flush_x(r, &pend);
        if (rc == 0 && open) rc = close_step(r);
        r->skip = TM5_ERRORS;
        TRY(rc);
        TRY(read_words(c));
        if (c->hwords[W_FLAGS] & TM5_ERRORS) return TM5_EFLAGS;
        const int f = (c->hwords[W_FLAGS] & TM5_SPEC) ? c->hwords[W_SPEC_AT] : end;
        for (int k = i; k < f; k++) {
            ndyn_out[k] = nd[trials - 1];
            n_out[k] = nn[trials - 1];
        }
        if (f >= end) {
            i = end;
            batch = batch < 16 ? 2 * batch : 16;
            continue;
        }
        batch = 2;
        // interval f's trials or checks disagreed with the plan: nothing after them ran.  A failed
        // check left the interval's state part-stepped: it comes back from the backup.
        const bool check_failed = c->hwords[W_SPEC_KIND] != 0;
        if (check_failed) {
            const long long nc = (long long)ncell;
            TM5_LAUNCH(c, k_restore, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->stream, d.im, d.jm, d.lm,
                       backup_of(f), view3(&d.m), tracers(&d.q));
            TRY(status(hipGetLastError()));
        }
        TRY(set_word(c, W_FLAGS, 0, TM5_SPEC | TM5_REHEARSAL));
        TRY(set_word(c, W_SPEC_KIND, 0, -1));
        TRY(set_word(c, W_SPEC_AT, INT_MAX, -1));
        r->ms_fresh = false;
        // The interval is queued again on the step its miss points to -- after a failed check of
        // the planned step the next shorter one of new_valid_timestep's chain, after a trial that
        // passed against the plan the longest -- as checked steps again, so a wrong guess is only
        // another miss.  A second miss on the same interval redoes it the reference's way, and the
        // next ones are planned on its step.
        if (f != miss_at) {
            miss_at = f;
            plan = check_failed ? new_valid_timestep(nd[trials - 1], nread, nread, 1) : ndyn_max;
            if (plan >= 1) {
                i = f;
                continue;
            }
        }
        TRY(tm5_region_interval(r, t0 + (double)f * (double)nread, nread, ndyn_max, ndyn_out + f, n_out + f));
        plan = ndyn_out[f];
        i = f + 1;
    }
    return 0;
}
