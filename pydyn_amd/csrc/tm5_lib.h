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
    W_WORDS = 8,
};

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
    tm5_prof *prof; // non-null while profiling (then prof == prof_store)
    tm5_prof *prof_store;  // owned: the events and the last totals
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

// launchers (advect.hip)
int launch_x(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 am, int limits, int nmax, int fail_bit,
             int skip_mask);
int launch_y(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 bm, int limits, int nmax, int fail_bit,
             int skip_mask);
int launch_z(tm5_ctx *ctx, int im, int jm, int lm, View3 m, Tracers q, View3 cm, int limits, int nmax, int fail_bit,
             int skip_mask);

}  // namespace tm5
