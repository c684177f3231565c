// The native driver of a zoom tree (include/tm5hip.h, tm5_tree_*): the reference's zoom driver --
// do_steps / Proces_Region / the two half steps (pydyn/zoom/steps.py:15-88), Setup_MassFlow with
// coarsen_region (pydyn/zoom/massflow.py:62-193) and the tree's Check_CFL (pydyn/zoom/
// check_cfl.py:228-290) -- issuing the library's zoom bodies (zoom.hip) from C++.
//
// Two ways to run an interval:
//   tm5_tree_interval  the reference's: Check_CFL's mass-only rehearsal of the interval's steps
//                      (m, am, bm, cm saved and restored), its verdict read by the host once per
//                      attempt, the step cut on a failure; then the steps.  After a passing
//                      rehearsal every y layer and z region passes at nloop 1 (see below), so the
//                      steps launch no y or z search.
//   tm5_tree_run       intervals queued without reading the device between them: each planned
//                      on the previous interval's step, its steps run as checked steps.
// Both rest on one fact: the rehearsal's mass-only operators -- in the order and on the scopes
// the CFL split orders give ("xyzzyx" in blocks of 3, each child's copies of its parent's blocks)
// -- move every region's m through the same bits as the step's operators ("xyzvsccsvzyx" in
// blocks of 6) given the same fluxes: the two orders interleave the regions' advection operators
// identically, their scopes and outer-face zeroing agree block for block, and convection, sources,
// chemistry, m2phlb and the step's extra put_edges at block ends (overwritten before a child
// reads them) leave the masses alone.  So a step whose x rows pass at nloop <= 9 and whose y and z
// faces pass at nloop 1 (every column of a region for z, as the rehearsal's whole-region trial)
// is exactly a step whose rehearsal passes.  A checked step makes those tests in its kernels; a
// failure raises TM5_SPEC (every later kernel skips) and the host redoes the interval from the
// backup the interval's first step took.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "tm5_lib.h"

using namespace tm5;

namespace {

const char SPLITORDER[] = "xyzvsccsvzyx";  // dims.F90's splitorder (pydyn/zoom/tree.py:18)
constexpr int NOPS = 6;
const char CFL_SPLITORDER[] = "xyzzyx";  // advectm_cfl.F90's own order and block size
constexpr int CFL_NOPS = 3;
constexpr int CFL_CAP[3] = {10, 1, 1};  // cfl_max_nloop_{x,y,z}
constexpr int MAX_SEGS = 32;

// define_splitorderzoom, zoom_mode 1 (pydyn/zoom/tree.py:160-174): each block of the parent's
// order repeated tref times, the even copies reversed
std::string child_order(const std::string &parent, int nops, int tref)
{
    std::string out;
    for (size_t j0 = 0; j0 < parent.size(); j0 += nops) {
        const std::string blk = parent.substr(j0, nops);
        for (int i = 1; i <= tref; i++) out += (i % 2 == 1) ? blk : std::string(blk.rbegin(), blk.rend());
    }
    return out;
}

// pydyn/cfl.py:208-213
int new_valid_timestep(int dtime, int nread, int cfl_outputstep, int tref_max)
{
    for (int dt = dtime - 1; dt > 0; dt--)
        if (nread % dt == 0 && cfl_outputstep % dt == 0 && dt % (2 * tref_max) == 0) return dt;
    return -1;
}

// Fortran nint (pydyn/ops.py:78-80)
int nint(double x) { return (int)(floor(fabs(x) + 0.5) * (x >= 0 ? 1.0 : -1.0)); }

// pydyn_amd.meteo's amplitude(t): the same operations, so the same double
double amplitude(double period, double amp, double t)
{
    const double x = t / period;
    const double frac = x - floor(x);
    return 1.0 + amp * (1.0 - 4.0 * fabs(frac - 0.5));
}

// Flat copies of whole arrays, dst[s] = src[s] for every segment s (grid y), skipped under the
// mask; first_step: only at a queued interval's first step (the clock's step 0).
struct Segs {
    const double *src[MAX_SEGS];
    double *dst[MAX_SEGS];
    long long n[MAX_SEGS];
    int count;
};

__global__ void __launch_bounds__(256) k_copy_segs(Segs s, int *words, int skip_mask, int first_step)
{
    if ((words[W_FLAGS] & skip_mask) != 0) return;
    if (first_step && words[W_CLOCK_S] != 0) return;
    const int k = blockIdx.y;
    const double *src = s.src[k];
    double *dst = s.dst[k];
    const long long n = s.n[k];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// the clock of a queued run: interval K (from the run's start t0), step 0
__global__ void k_clock_set(int *words, int k, double *t0p, double t0)
{
    if (threadIdx.x != 0) return;
    words[W_CLOCK_K] = k;
    words[W_CLOCK_S] = 0;
    *t0p = t0;
}

// the flag word's bits in mask set to value (a kernel: capturable, ordered on the stream)
__global__ void k_tree_word(int *words, int index, int value, int mask)
{
    if (threadIdx.x == 0) words[index] = (words[index] & ~mask) | (value & mask);
}

// A queued trial's verdict against the plan (expect_fail: the plan cuts the step after it): another
// verdict makes every kernel queued after it skip.  The rehearsal bit is cleared.
__global__ void k_tree_spec_check(int *words, int k, int expect_fail)
{
    if (threadIdx.x != 0) return;
    const int f = words[W_FLAGS];
    int out = f & ~TM5_REHEARSAL;
    if (!(f & TM5_SPEC) && ((f & TM5_REHEARSAL) != 0) != (expect_fail != 0)) {
        out |= TM5_SPEC;
        words[W_SPEC_AT] = k;
        words[W_SPEC_KIND] = 0;
    }
    words[W_FLAGS] = out;
}

struct TReg {
    tm5_zregion d;
    ZR R;
    int parent = 0;
    std::vector<int> children;
    std::string order, cfl_order;
    tm5_zphysics phys;
    tm5_view pu0, pv0;
    long long nm = 0, nq = 0, nf = 0;  // elements of m, of each tracer array, of am / bm / cm
    double *slab[2] = {nullptr, nullptr};
};

struct Graph {
    int ndyn = 0, nread = 0, checked = 0, uses = 0;
    hipGraphExec_t exec = nullptr;
};

}  // namespace

struct tm5_tree {
    tm5_ctx *ctx = nullptr;
    int n = 0;
    std::vector<TReg> reg;  // 1-based
    const double *at = nullptr, *bt = nullptr;
    double period = 0.0, amp = 0.0, sec_month = 0.0;
    int limits = 1, tref_max = 1;
    // the step being issued
    int spec = SPEC_NONE;    // SPEC_CLOCK: a checked step
    bool known_one = false;  // y and z at nloop 1 without a search
    int clock_n = 0;         // > 0: the step's last kernel moves the clock on (intervals of clock_n steps)
    TimeSrc ts{};            // dynam0's amplitude
    std::vector<int> status, done, cstatus;
    std::vector<ZPost> zdefer;  // a region's dynamw deferred to its next column pass (pre set)
    bool m2phlb1_done = false;
    int half = 0;
    // buffers
    double *save = nullptr;    // Check_CFL's m, am, bm, cm of every region
    double *backup = nullptr;  // a queued interval's m and moments of every region at its start
    size_t backup_n = 0;
    double *t0dev = nullptr;   // the queued run's start (the clock's origin)
    // the queue
    int plan = 0, batch = 2, batch_max = 16, graphs_on = 1;
    int mode = TM5_TREE_AUTO;  // TM5_TREE_MODE: checked steps with a backup, or rehearsed plans
    long long kept = 0, redone = 0, replayed = 0, captured = 0;
    std::vector<Graph> graphs;
    hipStream_t cap = nullptr;  // the capture stream
};

namespace {

int mask(const tm5_tree *t, bool rehearsal)
{
    return (rehearsal ? (TM5_ERRORS | TM5_REHEARSAL) : TM5_ERRORS) | t->ctx->skip_or;
}

const ZR &R(const tm5_tree *t, int r) { return t->reg[r].R; }

int word(tm5_tree *t, int index, int value, int m)
{
    tm5_ctx *c = t->ctx;
    TM5_LAUNCH(c, k_tree_word, dim3(1), dim3(64), 0, c->stream, c->words, index, value, m);
    return status(hipGetLastError());
}

int read_words(tm5_ctx *c)
{
    hipError_t e = hipMemcpyAsync(c->hwords, c->words, W_WORDS * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return status(e);
}

#define TTRY(x)                   \
    do {                          \
        int rc_ = (x);            \
        if (rc_ != 0) return rc_; \
    } while (0)

// flat copies of the segments (in groups of MAX_SEGS)
int copy_segs(tm5_tree *t, const std::vector<const double *> &src, const std::vector<double *> &dst,
              const std::vector<long long> &n, int skip_mask, int first_step)
{
    tm5_ctx *c = t->ctx;
    for (size_t s0 = 0; s0 < src.size(); s0 += MAX_SEGS) {
        Segs s{};
        long long most = 0;
        for (size_t k = s0; k < src.size() && k < s0 + MAX_SEGS; k++) {
            s.src[s.count] = src[k];
            s.dst[s.count] = dst[k];
            s.n[s.count] = n[k];
            most = n[k] > most ? n[k] : most;
            s.count++;
        }
        long long blocks = (most + 255) / 256;
        if (blocks > 2048) blocks = 2048;
        if (blocks < 1) blocks = 1;
        TM5_LAUNCH(c, k_copy_segs, dim3((unsigned)blocks, (unsigned)s.count), dim3(256), 0, c->stream, s, c->words,
                   skip_mask, first_step);
        TTRY(status(hipGetLastError()));
    }
    return 0;
}

// Check_CFL's save / restore of every region's m, am, bm, cm
int save_restore(tm5_tree *t, bool restore)
{
    std::vector<const double *> src;
    std::vector<double *> dst;
    std::vector<long long> n;
    double *p = t->save;
    for (int r = 1; r <= t->n; r++) {
        const TReg &g = t->reg[r];
        double *arr[4] = {g.d.m.data, g.d.am.data, g.d.bm.data, g.d.cm.data};
        const long long len[4] = {g.nm, g.nf, g.nf, g.nf};
        for (int a = 0; a < 4; a++) {
            src.push_back(restore ? p : arr[a]);
            dst.push_back(restore ? arr[a] : p);
            n.push_back(len[a]);
            p += len[a];
        }
    }
    return copy_segs(t, src, dst, n, t->ctx->skip_or, 0);
}

// a queued interval's backup of every region's m and moments (at its first step), or its restore
int backup_restore(tm5_tree *t, bool restore)
{
    std::vector<const double *> src;
    std::vector<double *> dst;
    std::vector<long long> n;
    double *p = t->backup;
    for (int r = 1; r <= t->n; r++) {
        const TReg &g = t->reg[r];
        double *arr[5] = {g.d.m.data, g.d.rm.data, g.d.rxm.data, g.d.rym.data, g.d.rzm.data};
        const long long len[5] = {g.nm, g.nq, g.nq, g.nq, g.nq};
        for (int a = 0; a < 5; a++) {
            src.push_back(restore ? p : arr[a]);
            dst.push_back(restore ? arr[a] : p);
            n.push_back(len[a]);
            p += len[a];
        }
    }
    return copy_segs(t, src, dst, n, restore ? 0 : (TM5_ERRORS | t->ctx->skip_or), restore ? 0 : 1);
}

// ------------------------------------------------------------------ scopes (pydyn/zoom/advect.py:47-118)
void block_of(const tm5_tree *t, int r, int st, const std::string &order, int nops, std::string &blk, bool &second)
{
    const int tr = t->reg[r].d.tref;
    const int start = (st / (nops * tr)) * nops * tr;
    blk = order.substr(start, nops);
    second = st % (nops * tr) >= nops;
}

// the j range of dynamu: the whole column when y follows x in the block, else the rows inside
// the y refinement
void x_scope(const tm5_tree *t, int r, const std::string &blk, int &js, int &je)
{
    const tm5_zregion &g = t->reg[r].d;
    const int yr = r == 1 ? 0 : g.yref;
    bool seen = false;
    js = 1, je = g.jm;
    for (char d : blk) {
        if (d == 'x') {
            seen = true;
        } else if (d == 'y') {
            if (!seen) {
                js = 1, je = g.jm;
            } else {
                js = g.touch_sp == 1 ? 1 : yr + 1;
                je = g.touch_np == 1 ? g.jm : g.jm - yr;
            }
        }
    }
}

// the i range of dynamv
void y_scope(const tm5_tree *t, int r, const std::string &blk, int &is, int &ie)
{
    const tm5_zregion &g = t->reg[r].d;
    const int xr = r == 1 ? 0 : g.xref;
    bool seen = false;
    is = 1, ie = g.im;
    for (char d : blk) {
        if (d == 'x') {
            if (!seen || g.xcyc == 1) {
                is = 1, ie = g.im;
            } else {
                is = xr + 1, ie = g.im - xr;
            }
        } else if (d == 'y') {
            seen = true;
        }
    }
}

// the (i, j) scope of dynamw
void z_scope(const tm5_tree *t, int r, const std::string &blk, int &is, int &ie, int &js, int &je)
{
    const tm5_zregion &g = t->reg[r].d;
    const int xr = r == 1 ? 0 : g.xref, yr = r == 1 ? 0 : g.yref;
    bool seen = false;
    is = 1, ie = g.im, js = 1, je = g.jm;
    for (char d : blk) {
        if (d == 'x') {
            if (!seen || g.xcyc == 1) {
                is = 1, ie = g.im;
            } else {
                is = xr + 1, ie = g.im - xr;
            }
        } else if (d == 'y') {
            if (!seen) {
                js = 1, je = g.jm;
            } else {
                js = g.touch_sp == 1 ? 1 : yr + 1;
                je = g.touch_np == 1 ? g.jm : g.jm - yr;
            }
        } else if (d == 'z') {
            seen = true;
        }
    }
}

// ------------------------------------------------------------------ the step's bodies
// where compress / uncompress apply (as tm5_zoom_compress_x/yedges decide)
bool xedges_on(const tm5_zregion &d) { return !(d.xref == 1 || d.whole); }
bool yedges_on(const tm5_zregion &d)
{
    return !(d.yref == 1 || (d.xcyc == 1 && d.touch_sp == 1 && d.touch_np == 1));
}

int put_edges(tm5_tree *t, int r, int axis, int tracers)
{
    for (int ch : t->reg[r].children) {
        if (axis == 0)
            TTRY(zoom_put_x(t->ctx, R(t, r), R(t, ch), tracers));
        else
            TTRY(zoom_put_y(t->ctx, R(t, r), R(t, ch), tracers));
    }
    return 0;
}

// A region's deferred dynamw, moved now on the columns its column pass was to take (before any
// other operator of the region, or the end of its Proces_Region, reads them).
int flush_z(tm5_tree *t, int r)
{
    ZPost &zd = t->zdefer[r];
    if (!zd.pre) return 0;
    zd.pre = 0;
    const ZR &Rr = t->reg[r].R;
    return launch_z_scoped(t->ctx, Rr.im, Rr.jm, Rr.lm, sub3(Rr.m, 1, 1, 1), subq(Rr, 1, 1, 1), sub3(Rr.cm, 1, 1, 0),
                           zd.pi0, zd.pi1, zd.pj0, zd.pj1, t->limits, 15, TM5_FAIL_Z, mask(t, false), true, t->spec,
                           zd.hi0, zd.hi1, zd.hj0, zd.hj1);
}

// advect_x_zoom (pydyn/zoom/advect.py:160-176): put_xedges, dynamu on the block's scope with the
// outer faces zeroed in the second half
int advect_x(tm5_tree *t, int r)
{
    tm5_ctx *c = t->ctx;
    TReg &g = t->reg[r];
    const ZR &Rr = g.R;
    TTRY(flush_z(t, r));
    TTRY(put_edges(t, r, 0, 1));
    std::string blk;
    bool second;
    block_of(t, r, t->status[r], g.order, NOPS, blk, second);
    int js, je;
    x_scope(t, r, blk, js, je);
    const bool zero = second && !g.d.whole;
    const bool chk = checked(t->spec);
    const int mk = mask(t, false);
    if (r == 1) {
        if (g.d.nred == 0) {  // the periodic halo of rows js..je written by the x kernel itself
            TTRY(launch_x(c, Rr.im, Rr.jm, Rr.lm, sub3(Rr.m, 1, 1, 1), subq(Rr, 1, 1, 1), sub3(Rr.am, 0, 1, 1),
                          t->limits, chk ? CFL_CAP[0] - 1 : 50 - 1, TM5_FAIL_X, mk, nullptr, nullptr, nullptr, t->spec,
                          nullptr, js - 1, je - 1));
        } else {
            TTRY(zoom_x_reduced(c, g.d, js, je, t->limits, chk ? CFL_CAP[0] : 50, 0, TM5_FAIL_X, mk, t->spec));
        }
        return 0;
    }
    // a child: the compress / uncompress and the outer faces' zeroing inside the x kernel
    const int fuse = (xedges_on(g.d) ? FUSE_COMPRESS : 0) | (zero ? FUSE_OUTER : 0);
    TTRY(zoom_x_child(c, Rr, js, je, 1, Rr.lm, t->limits, chk ? CFL_CAP[0] - 1 : 50 - 1, 0, TM5_FAIL_XC, mk, t->spec,
                      fuse));
    TTRY(tm5_zoom_mix_edges(c, &g.d, 1));
    return 0;
}

// advect_y_zoom (pydyn/zoom/advect.py:179-194)
int advect_y(tm5_tree *t, int r)
{
    tm5_ctx *c = t->ctx;
    TReg &g = t->reg[r];
    const ZR &Rr = g.R;
    TTRY(flush_z(t, r));
    TTRY(put_edges(t, r, 1, 1));
    std::string blk;
    bool second;
    block_of(t, r, t->status[r], g.order, NOPS, blk, second);
    int is, ie;
    y_scope(t, r, blk, is, ie);
    const bool zero = second && (g.d.touch_sp != 1 || g.d.touch_np != 1);
    const int mk = mask(t, false);
    if (r == 1)
        return launch_y(c, Rr.im, Rr.jm, Rr.lm, sub3(Rr.m, 1, 1, 1), subq(Rr, 1, 1, 1), sub3(Rr.bm, 1, 1, 1),
                        t->limits, 10 - 1, TM5_FAIL_Y, mk, t->known_one, t->spec);
    if (t->known_one || checked(t->spec)) {
        // no search: the compress / uncompress and the outer faces' zeroing (a child touches
        // neither pole: both faces) inside the sub-step kernel
        const int fuse = (yedges_on(g.d) ? FUSE_COMPRESS : 0) | (zero ? FUSE_OUTER : 0);
        if (zero && (g.d.touch_sp == 1 || g.d.touch_np == 1)) return TM5_EINVAL;
        TTRY(zoom_y_child(c, Rr, is, ie, 1, Rr.lm, t->limits, 10 - 1, 0, TM5_FAIL_YC, mk, true, t->spec, fuse));
        return tm5_zoom_mix_edges(c, &g.d, 1);
    }
    if (zero) TTRY(tm5_zoom_outer_faces(c, &g.d, 1, 0, g.slab[1], 0));
    TTRY(tm5_zoom_compress_yedges(c, &g.d, is, ie, 1, Rr.lm, 1));
    TTRY(zoom_y_child(c, Rr, is, ie, 1, Rr.lm, t->limits, 10 - 1, 0, TM5_FAIL_YC, mk, false, t->spec, 0));
    TTRY(tm5_zoom_uncompress_yedges(c, &g.d, is, ie, 1, Rr.lm, 1));
    TTRY(tm5_zoom_mix_edges(c, &g.d, 1));
    if (zero) TTRY(tm5_zoom_outer_faces(c, &g.d, 1, 1, g.slab[1], 0));
    return 0;
}

// advect_z_zoom (pydyn/zoom/advect.py:197-206)
//
// The dynamw before a column pass (z after y, the pass its block's next operator), at n_advz 1,
// moves now only the columns something reads or rewrites before that pass, and the pass (the
// column kernel, or the split form's factor pass and apply) moves the rest: for region 1, whose
// child runs in between and reads none of its columns (the edge puts come in x and y), the
// child's footprint and a column around it (update_parent rewrites the footprint's m and rm); for
// a child without children, its ring of xref columns and yref rows at every wall (mix_edges,
// next, mixes blocks of it).  The launch now tests every column in a checked step, as the whole
// dynamw's would.
int advect_z(tm5_tree *t, int r)
{
    tm5_ctx *c = t->ctx;
    TReg &g = t->reg[r];
    const ZR &Rr = g.R;
    TTRY(flush_z(t, r));
    std::string blk;
    bool second;
    const int s = t->status[r];
    block_of(t, r, s, g.order, NOPS, blk, second);
    int is, ie, js, je;
    z_scope(t, r, blk, is, ie, js, je);
    const bool defer = (t->known_one || checked(t->spec)) && s + 1 < (int)g.order.size() && g.order[s + 1] == 'v' && s >= 1 && g.order[s - 1] == 'y' &&
                       (r == 1 ? g.children.size() == 1 : g.children.empty());
    auto scoped = [&](int i0, int i1, int j0, int j1, int hi0, int hi1, int hj0, int hj1) {
        return launch_z_scoped(c, Rr.im, Rr.jm, Rr.lm, sub3(Rr.m, 1, 1, 1), subq(Rr, 1, 1, 1), sub3(Rr.cm, 1, 1, 0), i0,
                               i1, j0, j1, t->limits, 15, TM5_FAIL_Z, mask(t, false), t->known_one, t->spec, hi0, hi1,
                               hj0, hj1);
    };
    ZPost zd;
    zd.pre = 1;
    if (defer && r == 1) {
        const tm5_zregion &ch = t->reg[g.children[0]].d;  // footprint ibeg..iend, jbeg..jend (1-based)
        const int fi0 = std::max(ch.ibeg - 2, is - 1), fi1 = std::min(ch.iend, ie - 1);
        const int fj0 = std::max(ch.jbeg - 2, js - 1), fj1 = std::min(ch.jend, je - 1);
        if (fi0 <= fi1 && fj0 <= fj1) {
            TTRY(scoped(fi0, fi1, fj0, fj1, 0, -1, 0, -1));
            zd.pi0 = is - 1, zd.pi1 = ie - 1, zd.pj0 = js - 1, zd.pj1 = je - 1;
            zd.hi0 = fi0, zd.hi1 = fi1, zd.hj0 = fj0, zd.hj1 = fj1;
            t->zdefer[r] = zd;
            return 0;
        }
    } else if (defer) {
        // the inner rectangle: every column at least xref from an x wall and yref from a y wall
        const int ii0 = std::max(g.d.xref, is - 1), ii1 = std::min(Rr.im - 1 - g.d.xref, ie - 1);
        const int ij0 = std::max(g.d.yref, js - 1), ij1 = std::min(Rr.jm - 1 - g.d.yref, je - 1);
        if (ii0 <= ii1 && ij0 <= ij1) {
            // many tracers: the ring in up to three launches, the scope's rows below and above the
            // inner rectangle (whole rows, the context's tiles), then its rows beside it (a hole,
            // three-column tiles); the first launch tests every column in a checked step.  Few
            // tracers: one holed launch over the scope (each launch is mostly its fixed cost).
            if (subq(Rr, 1, 1, 1).nt < column_split_nt()) {
                TTRY(scoped(is - 1, ie - 1, js - 1, je - 1, ii0, ii1, ij0, ij1));
                zd.pi0 = ii0, zd.pi1 = ii1, zd.pj0 = ij0, zd.pj1 = ij1;
                t->zdefer[r] = zd;
                return tm5_zoom_mix_edges(c, &g.d, 1);
            }
            bool tested = false;
            auto part = [&](int j0, int j1, bool hole) {
                if (j0 > j1) return 0;
                const int spec = tested ? SPEC_NONE : t->spec;
                tested = true;
                return launch_z_scoped(c, Rr.im, Rr.jm, Rr.lm, sub3(Rr.m, 1, 1, 1), subq(Rr, 1, 1, 1),
                                       sub3(Rr.cm, 1, 1, 0), is - 1, ie - 1, j0, j1, t->limits, 15, TM5_FAIL_Z,
                                       mask(t, false), true, spec, hole ? ii0 : 0, hole ? ii1 : -1,
                                       hole ? ij0 : 0, hole ? ij1 : -1);
            };
            TTRY(part(js - 1, ij0 - 1, false));
            TTRY(part(ij1 + 1, je - 1, false));
            TTRY(part(ij0, ij1, true));
            zd.pi0 = ii0, zd.pi1 = ii1, zd.pj0 = ij0, zd.pj1 = ij1;
            t->zdefer[r] = zd;
            return tm5_zoom_mix_edges(c, &g.d, 1);
        }
    }
    TTRY(scoped(is - 1, ie - 1, js - 1, je - 1, 0, -1, 0, -1));
    if (r != 1) TTRY(tm5_zoom_mix_edges(c, &g.d, 1));
    return 0;
}

// convection, sources and chemistry of a region in one column pass (tm5_zoom_physics); zp: with the
// step's next dynamw fused
int column(tm5_tree *t, int r, int ndyn, const char *ops, const ZPost *zp)
{
    TReg &g = t->reg[r];
    tm5_zphysics p = g.phys;
    const double half = (double)ndyn / (2.0 * (double)g.d.tref_abs);
    p.dt_conv = half;
    p.factor = half / t->sec_month;  // month2dt (pydyn/emis.py:65-68)
    p.dt_chem = half;
    p.at = t->at;
    p.bt = t->bt;
    return zoom_physics(t->ctx, &g.d, &p, ops, zp);
}

// The column pass starting at region r's current status: the longest sequence the order allows
// ("vsc" + "csv" as one pass when the air mass cannot change between them: region 1 across its
// two half steps with the m2phlb1 between them folded in, a child within one Proces_Region).
int column_pass(tm5_tree *t, int r, int ndyn)
{
    const std::string &order = t->reg[r].order;
    const int s = t->status[r], tr = t->reg[r].d.tref;
    const char *ops;
    int n;
    if (order.compare(s, 6, "vsccsv") == 0 && (r == 1 || s / (NOPS * tr) == (s + 5) / (NOPS * tr))) {
        ops = r == 1 ? "vscMcsv" : "vsccsv";
        n = 6;
        if (r == 1) t->m2phlb1_done = true;
    } else if (order.compare(s, 3, "vsc") == 0) {
        ops = "vsc", n = 3;
    } else if (order.compare(s, 3, "csv") == 0) {
        ops = "csv", n = 3;
    } else if (order[s] == 'v') {
        ops = "v", n = 1;
    } else {
        return TM5_EINVAL;  // sources or chemistry apart from convection
    }
    // The operator after the pass a dynamw at n_advz 1 (a checked step's, or one its rehearsal
    // passed at 1): it runs in the pass (the column kernel, or the split form's apply for many
    // tracers), on the scope of its block.  Nothing comes between them in the order (the children run after a z only when
    // y precedes it, and region 1's "vscMcsv" ends its first half with nothing launched); a
    // child's mix_edges follows, as after its z.
    ZPost zp = t->zdefer[r];  // the dynamw before the pass, deferred to it (pre set)
    t->zdefer[r].pre = 0;
    zp.limits = t->limits;
    zp.spec_k = t->spec;
    const int zs = s + n;
    const bool post = zs < (int)order.size() && order[zs] == 'z' && (t->known_one || checked(t->spec));
    if (post) {
        std::string blk;
        bool second;
        block_of(t, r, zs, order, NOPS, blk, second);
        int is, ie, js, je;
        z_scope(t, r, blk, is, ie, js, je);
        zp.post = 1;
        zp.i0 = is - 1, zp.i1 = ie - 1, zp.j0 = js - 1, zp.j1 = je - 1;
    }
    TTRY(column(t, r, ndyn, ops, zp.pre || zp.post ? &zp : nullptr));
    t->done[r] = post ? zs + 1 : s + n;
    if (post && r != 1) TTRY(tm5_zoom_mix_edges(t->ctx, &t->reg[r].d, 1));
    return 0;
}

// do_steps (pydyn/zoom/steps.py:31-69): the next operators of region r up to a block boundary or a
// z right after y.  The reference also puts the edges to the children at a block boundary unless
// the last operator was x; those puts are not launched: the children run only after a z that
// follows y or at the end of a block ending in x, so the x and y of that same block have put
// their edges again since, over the same cells, before any child reads them.
int do_steps(tm5_tree *t, int r, int ndyn)
{
    const std::string &order = t->reg[r].order;
    for (int k = 0; k < NOPS; k++) {
        const int s = t->status[r];
        const char op = order[s];
        if (s >= t->done[r]) {
            if (op == 'x')
                TTRY(advect_x(t, r));
            else if (op == 'y')
                TTRY(advect_y(t, r));
            else if (op == 'z')
                TTRY(advect_z(t, r));
            else if (op == 'v' || op == 's' || op == 'c')
                TTRY(column_pass(t, r, ndyn));
            else
                return TM5_EINVAL;
        }
        t->status[r] += 1;
        const int now = t->status[r];
        if (now % NOPS == 0) return 0;
        if (op == 'z' && order[now - 2] == 'y') return 0;
    }
    return 0;
}

// Proces_Region (pydyn/zoom/steps.py:72-82): tref rounds of do_steps, each followed by the
// children; then update_parent and m2phlb (m2phlb1 for region 1)
int proces_region(tm5_tree *t, int r, int ndyn)
{
    tm5_ctx *c = t->ctx;
    TReg &g = t->reg[r];
    for (int k = 0; k < g.d.tref; k++) {
        TTRY(do_steps(t, r, ndyn));
        for (int ch : g.children) TTRY(proces_region(t, ch, ndyn));
        if (t->status[r] % NOPS != 0) TTRY(do_steps(t, r, ndyn));
    }
    TTRY(flush_z(t, r));
    const int mk = mask(t, false);
    if (r != 1) {  // update_parent and m2phlb: disjoint arrays, one launch
        TTRY(zoom_update_m2phlb(c, R(t, g.parent), g.R, t->at, t->bt, mk));
    } else if (t->m2phlb1_done) {
        t->m2phlb1_done = false;  // done inside the fused column pass
    } else {
        // the step's last kernel: in a queued run it moves the clock on
        TTRY(zoom_m2phlb(c, g.R, nullptr, t->at, t->bt, mk, t->half == 1 ? t->clock_n : 0));
    }
    return 0;
}

// one ndyn step (fluxes set): every status reset, Proces_Region(1) for each half
int step(tm5_tree *t, int ndyn)
{
    for (int r = 1; r <= t->n; r++) {
        t->status[r] = t->done[r] = 0;
        t->zdefer[r].pre = 0;
    }
    t->m2phlb1_done = false;
    for (t->half = 0; t->half < 2; t->half++) TTRY(proces_region(t, 1, ndyn));
    for (int r = 1; r <= t->n; r++) TTRY(flush_z(t, r));  // none is left: every one has its pass
    if (t->m2phlb1_done) return TM5_EINVAL;  // the step's second half must end with region 1's m2phlb1
    return 0;
}

// ------------------------------------------------------------------ Setup_MassFlow
// region r's dynam0 (pydyn/zoom/massflow.py:181-186 through the meteo interpolation): the meteo's
// fluxes at the step's amplitude into its pu, pv, and am, bm, cm from them
Dynam0Region dynam0_of(tm5_tree *t, int r, int ndyn)
{
    TReg &g = t->reg[r];
    const ZR &Rr = g.R;
    Dynam0Region d{};
    d.im = Rr.im, d.jm = Rr.jm, d.lm = Rr.lm, d.cyclic = Rr.xcyc;
    d.pu = view3(&g.pu0);
    d.pv = view3(&g.pv0);
    d.am = sub3(Rr.am, 0, 1, 1);
    d.bm = sub3(Rr.bm, 1, 1, 1);
    d.cm = sub3(Rr.cm, 1, 1, 0);
    d.a = t->ts.a;
    d.dt = (double)ndyn / (2.0 * (double)g.d.tref_abs);
    d.ex.ts = t->ts;
    d.ex.ts.ndyn = (double)ndyn;
    d.ex.pu_out = sub3(Rr.pu, 0, 1, 1);
    d.ex.pv_out = sub3(Rr.pv, 1, 1, 1);
    d.ex.am_halo = Rr.xcyc == 1;
    return d;
}

int coarsen_region(tm5_tree *t, int r)
{
    const int p = t->reg[r].parent;
    if (p <= 0) return 0;
    tm5_ctx *c = t->ctx;
    const int mk = mask(t, false);
    TTRY(zoom_coarsen(c, R(t, p), R(t, r), mk));
    const int pp = t->reg[p].parent;
    TTRY(zoom_m2phlb(c, R(t, p), pp > 0 ? t->reg[pp].d.dxyp : nullptr, t->at, t->bt, mk, 0));
    return coarsen_region(t, p);
}

// pydyn/zoom/massflow.py:181-193: dynam0 and coarsen_region region by region, then coarsen_ambm.
// Every region's dynam0 runs first, in one launch: coarsen_region(r) touches only r and its
// ancestors, which come before r, so no dynam0 after it in the reference's order is one it reads
// or writes.
int setup_massflow(tm5_tree *t, int ndyn, bool rehearsal)
{
    std::vector<Dynam0Region> regs;
    try {
        regs.reserve((size_t)t->n);
    } catch (...) {
        return TM5_ENOMEM;
    }
    for (int r = 1; r <= t->n; r++) regs.push_back(dynam0_of(t, r, ndyn));
    TTRY(launch_dynam0_regions(t->ctx, t->n, regs.data(), t->bt, mask(t, rehearsal)));
    for (int r = 1; r <= t->n; r++) TTRY(coarsen_region(t, r));
    for (int r = 1; r <= t->n; r++)
        for (int ch : t->reg[r].children) TTRY(zoom_coarsen_ambm(t->ctx, R(t, r), R(t, ch), mask(t, false)));
    return 0;
}

// ------------------------------------------------------------------ Check_CFL (pydyn/zoom/check_cfl.py)
int mass_op(tm5_tree *t, int r, char op)
{
    tm5_ctx *c = t->ctx;
    TReg &g = t->reg[r];
    std::string blk;
    bool second;
    block_of(t, r, t->cstatus[r], g.cfl_order, CFL_NOPS, blk, second);
    // The mass-only operators at their caps as single kernels: x at cap 10 with its own search (and
    // region 1's halo, a child's compress / uncompress and outer faces inside the kernel), y and z
    // at cap 1 as the test of every face at nloop 1 and the one sub-step (SPEC_REHEARSAL: a failed
    // test raises the verdict; what the failing kernel still moves is discarded with the trial).
    const ZR &Rr = g.R;
    const int mk = mask(t, true);
    if (op == 'x') {
        TTRY(put_edges(t, r, 0, 0));
        int js, je;
        x_scope(t, r, blk, js, je);
        const bool zero = second && !g.d.whole;
        if (r == 1 && g.d.nred == 0)
            return launch_x(c, Rr.im, Rr.jm, Rr.lm, sub3(Rr.m, 1, 1, 1), no_tracers(), sub3(Rr.am, 0, 1, 1), 0,
                            CFL_CAP[0] - 1, TM5_REHEARSAL, mk, nullptr, nullptr, nullptr, SPEC_NONE, nullptr, js - 1,
                            je - 1);
        if (r == 1) return tm5_zoom_advect_x(c, &g.d, js, je, 0, CFL_CAP[0], 1);
        const int fuse = (xedges_on(g.d) ? FUSE_COMPRESS : 0) | (zero ? FUSE_OUTER : 0);
        TTRY(zoom_x_child(c, Rr, js, je, 1, Rr.lm, 0, CFL_CAP[0] - 1, 1, TM5_REHEARSAL, mk, SPEC_NONE, fuse));
        return tm5_zoom_mix_edges(c, &g.d, 0);
    }
    if (op == 'y') {
        TTRY(put_edges(t, r, 1, 0));
        int is, ie;
        y_scope(t, r, blk, is, ie);
        const bool zero = second && (g.d.touch_sp != 1 || g.d.touch_np != 1);
        if (r == 1) return launch_y_mass1(c, Rr.im, Rr.jm, Rr.lm, sub3(Rr.m, 1, 1, 1), sub3(Rr.bm, 1, 1, 1), TM5_REHEARSAL, mk);
        if (zero && (g.d.touch_sp == 1 || g.d.touch_np == 1)) return TM5_EINVAL;
        const int fuse = (yedges_on(g.d) ? FUSE_COMPRESS : 0) | (zero ? FUSE_OUTER : 0);
        TTRY(zoom_y_child(c, Rr, is, ie, 1, Rr.lm, 0, CFL_CAP[1], 1, TM5_REHEARSAL, mk, true, SPEC_REHEARSAL, fuse));
        return tm5_zoom_mix_edges(c, &g.d, 0);
    }
    int is, ie, js, je;
    z_scope(t, r, blk, is, ie, js, je);
    if (r == 1) return launch_z_mass1(c, Rr.im, Rr.jm, Rr.lm, sub3(Rr.m, 1, 1, 1), sub3(Rr.cm, 1, 1, 0), TM5_REHEARSAL, mk);
    TTRY(launch_z_scoped(c, Rr.im, Rr.jm, Rr.lm, sub3(Rr.m, 1, 1, 1), no_tracers(), sub3(Rr.cm, 1, 1, 0), is - 1, ie - 1,
                         js - 1, je - 1, 0, CFL_CAP[2], TM5_REHEARSAL, mk, true, SPEC_REHEARSAL));
    return tm5_zoom_mix_edges(c, &g.d, 0);
}

int cfl_do_steps(tm5_tree *t, int r)
{
    for (int k = 0; k < CFL_NOPS; k++) {
        TTRY(mass_op(t, r, t->reg[r].cfl_order[t->cstatus[r]]));
        t->cstatus[r] += 1;
        if (t->cstatus[r] % CFL_NOPS == 0) return 0;
    }
    return 0;
}

// determine_cfl_iter (pydyn/zoom/check_cfl.py:248-266): after a failure every kernel is a no-op
int determine_cfl_iter(tm5_tree *t, int r)
{
    TReg &g = t->reg[r];
    for (int k = 0; k < g.d.tref; k++) {
        TTRY(cfl_do_steps(t, r));
        for (int ch : g.children) TTRY(determine_cfl_iter(t, ch));
        if (t->cstatus[r] % CFL_NOPS != 0) TTRY(cfl_do_steps(t, r));
    }
    if (r != 1) TTRY(zoom_update_parent(t->ctx, R(t, g.parent), g.R, 0));
    return 0;
}

// the kernels of one Check_CFL attempt: the n steps of ndyn from t0 rehearsed on the state, m, am,
// bm, cm restored; the verdict is the rehearsal bit (cleared first when clear)
int rehearse(tm5_tree *t, double t0, int ndyn, int n, bool clear)
{
    if (clear) TTRY(word(t, W_FLAGS, 0, TM5_REHEARSAL));
    TTRY(save_restore(t, false));
    const TimeSrc keep = t->ts;
    for (int i = 1; i <= n; i++) {
        t->ts = TimeSrc{};
        t->ts.a = amplitude(t->period, t->amp, t0 + ((double)i - 0.5) * (double)ndyn);
        int rc = setup_massflow(t, ndyn, true);
        for (int r = 1; r <= t->n && rc == 0; r++) t->cstatus[r] = 0;
        if (rc == 0) rc = determine_cfl_iter(t, 1);
        if (rc == 0) rc = determine_cfl_iter(t, 1);
        if (rc != 0) {
            t->ts = keep;
            return rc;
        }
    }
    t->ts = keep;
    return save_restore(t, true);
}

int scale_fluxes(tm5_tree *t, double fraction)
{
    for (int r = 1; r <= t->n; r++) {
        const TReg &g = t->reg[r];
        TTRY(tm5_zoom_scale(t->ctx, g.d.am.data, g.nf, fraction));
        TTRY(tm5_zoom_scale(t->ctx, g.d.bm.data, g.nf, fraction));
        TTRY(tm5_zoom_scale(t->ctx, g.d.cm.data, g.nf, fraction));
    }
    return 0;
}

// check_cfl_zoom (pydyn/zoom/check_cfl.py:269-290): ndyn, n in and out
int check_cfl(tm5_tree *t, double t0, int nread, int &ndyn, int &n)
{
    tm5_ctx *c = t->ctx;
    for (int iteration = 1; iteration <= 100; iteration++) {
        TTRY(rehearse(t, t0, ndyn, n, true));
        TTRY(read_words(c));
        if (c->hwords[W_FLAGS] & TM5_ERRORS) return TM5_EFLAGS;
        if (!(c->hwords[W_FLAGS] & TM5_REHEARSAL)) return 0;
        const int ndyn_old = ndyn;
        ndyn = new_valid_timestep(ndyn, nread, nread, t->tref_max);
        if (ndyn < 1) return TM5_EINVAL;
        n = nint((double)(n * ndyn_old) / (double)ndyn);
        TTRY(scale_fluxes(t, (double)ndyn / (double)ndyn_old));
    }
    return TM5_EINVAL;  // "check_cfl_zoom: exceeded maximum number of time steps"
}

// a planned step of a queued interval: checked (the backup at its first step, the rehearsal's tests
// in the kernels) or, when the interval's plan was rehearsed and verified, plain at nloop 1;
// Setup_MassFlow at the clock's time, the step; its last kernel moves the clock on
int checked_step(tm5_tree *t, int ndyn, int nread, bool checked_ = true)
{
    t->spec = checked_ ? SPEC_CLOCK : SPEC_NONE;
    t->known_one = true;
    t->clock_n = nread / ndyn;
    t->ts = TimeSrc{};
    t->ts.t0 = t->t0dev;
    t->ts.nread = (double)nread;
    t->ts.period = t->period;
    t->ts.amplitude = t->amp;
    int rc = checked_ ? backup_restore(t, false) : 0;
    if (rc == 0) rc = setup_massflow(t, ndyn, false);
    if (rc == 0) rc = step(t, ndyn);
    t->spec = SPEC_NONE;
    t->known_one = false;
    t->clock_n = 0;
    t->ts = TimeSrc{};
    return rc;
}

// a checked step from its graph: captured the second time its (ndyn, nread) is queued (the first
// issues every launch, which also sizes every buffer the launches use), replayed after that
int graph_step(tm5_tree *t, int ndyn, int nread, bool chk)
{
    tm5_ctx *c = t->ctx;
    if (!t->graphs_on || c->prof) return checked_step(t, ndyn, nread, chk);
    Graph *g = nullptr;
    for (Graph &x : t->graphs)
        if (x.ndyn == ndyn && x.nread == nread && x.checked == (int)chk) g = &x;
    if (!g) {
        t->graphs.push_back(Graph{ndyn, nread, (int)chk, 0, nullptr});
        g = &t->graphs.back();
    }
    if (g->exec) {
        TTRY(status(hipGraphLaunch(g->exec, c->stream)));
        t->replayed++;
        return 0;
    }
    if (g->uses++ == 0) return checked_step(t, ndyn, nread, chk);
    hipStream_t keep = c->stream;
    c->stream = t->cap;
    hipError_t e = hipStreamBeginCapture(t->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        c->stream = keep;
        return status(e);
    }
    const int rc = checked_step(t, ndyn, nread, chk);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(t->cap, &graph);
    c->stream = keep;
    if (rc != 0 || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != 0 ? rc : status(e);
    }
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g->exec = nullptr;
        return status(e);
    }
    t->captured++;
    TTRY(status(hipGraphLaunch(g->exec, c->stream)));
    t->replayed++;
    return 0;
}

// new_valid_timestep's chain from ndyn_max down to ndyn: the trials Check_CFL makes when it
// settles on ndyn, each with its n; false if ndyn is not on the chain
bool trial_chain(int nread, int ndyn_max, int ndyn, int tref_max, int *nd, int *nn, int *count)
{
    int a = ndyn_max, n = nread / ndyn_max, k = 0;
    for (;;) {
        if (k >= 100) return false;
        nd[k] = a;
        nn[k] = n;
        k++;
        if (a == ndyn) break;
        const int b = new_valid_timestep(a, nread, nread, tref_max);
        if (b < 1) return false;
        n = nint((double)(n * a) / (double)b);
        a = b;
    }
    *count = k;
    return true;
}

int tree_interval(tm5_tree *t, double t0, int nread, int ndyn_max, int *ndyn_out, int *n_out)
{
    if (nread < 1 || ndyn_max < 1 || nread % ndyn_max != 0) return TM5_EINVAL;
    int ndyn = ndyn_max, n = nread / ndyn_max;
    TTRY(check_cfl(t, t0, nread, ndyn, n));
    t->known_one = true;  // the rehearsal at cap 1 passed: every y layer and z region at nloop 1
    for (int s = 0; s < n; s++) {
        t->ts = TimeSrc{};
        t->ts.a = amplitude(t->period, t->amp, (t0 + (double)(s * ndyn)) + (double)ndyn / 2.0);
        int rc = setup_massflow(t, ndyn, false);
        if (rc == 0) rc = step(t, ndyn);
        if (rc != 0) {
            t->known_one = false;
            return rc;
        }
    }
    t->known_one = false;
    if (ndyn_out) *ndyn_out = ndyn;
    if (n_out) *n_out = n;
    return 0;
}

// The queue's mode: checked steps need a backup of every region's m and moments per interval, a
// rehearsed plan a mass-only rehearsal (whose cost does not grow with the tracers): the backup up
// to BACKUP_MAX doubles, the rehearsal beyond.
constexpr size_t BACKUP_MAX = (size_t)64 << 20;  // 512 MB

int tree_run(tm5_tree *t, double t0, int nread, int ndyn_max, int count, int *ndyn_out, int *n_out)
{
    tm5_ctx *c = t->ctx;
    size_t need = 0;
    for (int r = 1; r <= t->n; r++) need += (size_t)t->reg[r].nm + 4 * (size_t)t->reg[r].nq;
    const bool chk = t->mode == TM5_TREE_CHECKED || (t->mode == TM5_TREE_AUTO && need <= BACKUP_MAX);
    if (chk && t->backup_n < need) {
        TTRY(status(hipStreamSynchronize(c->stream)));
        if (t->backup) TTRY(status(hipFree(t->backup)));
        t->backup = nullptr;
        t->backup_n = 0;
        TTRY(status(hipMalloc((void **)&t->backup, need * sizeof(double))));
        t->backup_n = need;
    }
    TTRY(word(t, W_FLAGS, 0, TM5_REHEARSAL | TM5_SPEC));
    TTRY(word(t, W_SPEC_KIND, 0, -1));
    if (t->batch < 2) t->batch = 2;
    if (t->plan < 1 || t->plan > ndyn_max) t->plan = ndyn_max;
    int i = 0;
    while (i < count) {
        int nd[100], nn[100], trials = 0;
        if (!trial_chain(nread, ndyn_max, t->plan, t->tref_max, nd, nn, &trials)) {
            t->plan = ndyn_max;
            if (!trial_chain(nread, ndyn_max, t->plan, t->tref_max, nd, nn, &trials)) return TM5_EINVAL;
        }
        const int end = i + t->batch < count ? i + t->batch : count;
        TM5_LAUNCH(c, k_clock_set, dim3(1), dim3(64), 0, c->stream, c->words, i, t->t0dev, t0);
        TTRY(status(hipGetLastError()));
        c->skip_or = TM5_SPEC;
        int rc = 0;
        for (int k = i; k < end && rc == 0; k++) {
            const double tk = t0 + (double)k * (double)nread;
            // the trials the plan expects to fail are rehearsed and their verdicts checked; without
            // checked steps the planned trial too, expected to pass
            for (int q = 0; q < trials - (chk ? 1 : 0) && rc == 0; q++) {
                rc = rehearse(t, tk, nd[q], nn[q], false);  // k_tree_spec_check clears the bit
                if (rc == 0) {
                    TM5_LAUNCH(c, k_tree_spec_check, dim3(1), dim3(64), 0, c->stream, c->words, k, q + 1 < trials);
                    rc = status(hipGetLastError());
                }
            }
            for (int s = 0; s < nn[trials - 1] && rc == 0; s++) rc = graph_step(t, nd[trials - 1], nread, chk);
        }
        c->skip_or = 0;
        TTRY(rc);
        TTRY(read_words(c));
        if (c->hwords[W_FLAGS] & TM5_ERRORS) return TM5_EFLAGS;
        const int f = (c->hwords[W_FLAGS] & TM5_SPEC) ? c->hwords[W_SPEC_AT] : end;
        if (f < i || f > end) return TM5_EINVAL;
        for (int k = i; k < f; k++) {
            ndyn_out[k] = nd[trials - 1];
            n_out[k] = nn[trials - 1];
        }
        t->kept += f - i;
        if (f >= end) {
            i = end;
            t->batch = t->batch < t->batch_max ? 2 * t->batch : t->batch_max;
            if (t->batch > t->batch_max) t->batch = t->batch_max;
            continue;
        }
        t->batch = 2 < t->batch_max ? 2 : t->batch_max;
        // interval f's trials or checks disagreed with the plan: nothing after them ran.  A failed
        // check left the interval's state part-stepped: it comes back from the backup.  Then the
        // interval is redone the reference's way, and the next ones planned on its step.
        if (c->hwords[W_SPEC_KIND]) TTRY(backup_restore(t, true));
        TTRY(word(t, W_FLAGS, 0, TM5_SPEC | TM5_REHEARSAL));
        TTRY(word(t, W_SPEC_KIND, 0, -1));
        TTRY(tree_interval(t, t0 + (double)f * (double)nread, nread, ndyn_max, ndyn_out + f, n_out + f));
        t->redone++;
        t->plan = ndyn_out[f];
        i = f + 1;
    }
    return 0;
}

bool contiguous(const tm5_hview &v, long long n0, long long n1, long long n2, int l0, int l1, int l2)
{
    return v.data && v.stride[0] == 1 && v.stride[1] == n0 && v.stride[2] == n0 * n1 && v.lower[0] == l0 &&
           v.lower[1] == l1 && v.lower[2] == l2 && n2 > 0;
}

void free_tree(tm5_tree *t)
{
    if (!t) return;
    if (t->ctx) (void)hipStreamSynchronize(t->ctx->stream);
    for (Graph &g : t->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (TReg &g : t->reg)
        for (double *s : g.slab)
            if (s) (void)hipFree(s);
    if (t->save) (void)hipFree(t->save);
    if (t->backup) (void)hipFree(t->backup);
    if (t->t0dev) (void)hipFree(t->t0dev);
    if (t->cap) (void)hipStreamDestroy(t->cap);
    delete t;
}

int tree_create(tm5_ctx *ctx, const tm5_tree_desc *d, tm5_tree **out)
{
    if (!ctx || !d || !out || d->nregions < 1 || !d->regions || !d->parent || !d->physics || !d->pu0 || !d->pv0 ||
        !d->at || !d->bt || d->mod_period == 0.0 || d->sec_month == 0.0)
        return TM5_EINVAL;
    tm5_tree *t = new tm5_tree();
    t->ctx = ctx;
    t->n = d->nregions;
    t->at = d->at;
    t->bt = d->bt;
    t->period = d->mod_period;
    t->amp = d->mod_amplitude;
    t->sec_month = d->sec_month;
    t->limits = d->limits;
    t->reg.resize(t->n + 1);
    t->status.assign(t->n + 1, 0);
    t->done.assign(t->n + 1, 0);
    t->cstatus.assign(t->n + 1, 0);
    t->zdefer.assign(t->n + 1, ZPost{});
    int rc = 0;
    size_t save_n = 0;
    for (int r = 1; r <= t->n && rc == 0; r++) {
        TReg &g = t->reg[r];
        g.d = d->regions[r - 1];
        g.R = zr(g.d);
        g.parent = d->parent[r - 1];
        g.phys = d->physics[r - 1];
        g.pu0 = d->pu0[r - 1];
        g.pv0 = d->pv0[r - 1];
        const int im = g.d.im, jm = g.d.jm, lm = g.d.lm, nt = g.d.nt;
        g.nm = (long long)(im + 4) * (jm + 4) * lm;
        g.nq = g.nm * nt;
        g.nf = (long long)(im + 2) * (jm + 2) * (lm + 2);
        const tm5_hview *q[4] = {&g.d.rm, &g.d.rxm, &g.d.rym, &g.d.rzm};
        bool ok = (r == 1) == (g.parent == 0) && g.parent < r && nt >= 1 && g.d.tref >= 1 && g.d.tref_abs >= 1 &&
                  g.d.lm == d->regions[0].lm && contiguous(g.d.m, im + 4, jm + 4, lm, -1, -1, 1) &&
                  contiguous(g.d.am, im + 2, jm + 2, lm + 2, 0, 0, 0) && contiguous(g.d.bm, im + 2, jm + 2, lm + 2, 0, 0, 0) &&
                  contiguous(g.d.cm, im + 2, jm + 2, lm + 2, 0, 0, 0) && g.pu0.data && g.pv0.data && g.d.edge &&
                  g.d.owned && g.d.dxyp;
        for (int a = 0; a < 4 && ok; a++)
            ok = contiguous(*q[a], im + 4, jm + 4, lm, -1, -1, 1) && q[a]->stride[3] == g.nm && q[a]->lower[3] == 1;
        if (!ok) rc = TM5_EINVAL;
        if (r == 1 && rc == 0 && !(g.d.xcyc == 1 && g.d.touch_sp == 1 && g.d.touch_np == 1)) rc = TM5_EINVAL;
        if (rc == 0 && r > 1) t->reg[g.parent].children.push_back(r);
        save_n += (size_t)g.nm + 3 * (size_t)g.nf;
        if (g.d.tref_abs > t->tref_max) t->tref_max = g.d.tref_abs;
        const size_t ns = 2 * (size_t)((im > jm ? im : jm) + 2) * (lm + 2);
        for (int a = 0; a < 2 && rc == 0; a++) rc = status(hipMalloc((void **)&g.slab[a], ns * sizeof(double)));
    }
    if (rc == 0) {  // the split orders (pydyn/zoom/tree.py:160-174), parents before children
        t->reg[1].order = SPLITORDER;
        t->reg[1].cfl_order = CFL_SPLITORDER;
        for (int r = 2; r <= t->n; r++) {
            const TReg &p = t->reg[t->reg[r].parent];
            const int tr = t->reg[r].d.tref;
            if (tr > 4 || tr == 3) rc = TM5_EINVAL;  // "define_splitorderzoom: tref_ not 1, 2 or 4"
            t->reg[r].order = child_order(p.order, NOPS, tr);
            t->reg[r].cfl_order = child_order(p.cfl_order, CFL_NOPS, tr);
        }
    }
    if (rc == 0) rc = status(hipMalloc((void **)&t->save, save_n * sizeof(double)));
    if (rc == 0) rc = status(hipMalloc((void **)&t->t0dev, sizeof(double)));
    if (rc == 0) rc = status(hipStreamCreateWithFlags(&t->cap, hipStreamNonBlocking));
    if (rc != 0) {
        free_tree(t);
        return rc;
    }
    *out = t;
    return 0;
}

}  // namespace

// ------------------------------------------------------------------ the C ABI
extern "C" int tm5_tree_create(tm5_ctx *ctx, const tm5_tree_desc *desc, tm5_tree **out)
{
    try {
        return tree_create(ctx, desc, out);
    } catch (const std::bad_alloc &) {
        return TM5_ENOMEM;
    } catch (...) {
        return TM5_EINVAL;
    }
}

extern "C" int tm5_tree_destroy(tm5_tree *t)
{
    free_tree(t);
    return 0;
}

extern "C" int tm5_tree_interval(tm5_tree *t, double t0, int nread, int ndyn_max, int *ndyn, int *n)
{
    if (!t) return TM5_EINVAL;
    try {
        return tree_interval(t, t0, nread, ndyn_max, ndyn, n);
    } catch (const std::bad_alloc &) {
        return TM5_ENOMEM;
    } catch (...) {
        return TM5_EINVAL;
    }
}

extern "C" int tm5_tree_run(tm5_tree *t, double t0, int nread, int ndyn_max, int count, int *ndyn, int *n)
{
    if (!t || !ndyn || !n || count < 0 || nread < 1 || ndyn_max < 1 || nread % ndyn_max != 0) return TM5_EINVAL;
    try {
        const int rc = tree_run(t, t0, nread, ndyn_max, count, ndyn, n);
        t->ctx->skip_or = 0;
        return rc;
    } catch (const std::bad_alloc &) {
        t->ctx->skip_or = 0;
        return TM5_ENOMEM;
    } catch (...) {
        t->ctx->skip_or = 0;
        return TM5_EINVAL;
    }
}

extern "C" int tm5_tree_option(tm5_tree *t, int which, int value)
{
    if (!t) return TM5_EINVAL;
    switch (which) {
    case TM5_TREE_GRAPHS:
        t->graphs_on = value != 0;
        return 0;
    case TM5_TREE_BATCH:
        if (value < 1 || value > 64) return TM5_EINVAL;
        t->batch_max = value;
        if (t->batch > value) t->batch = value;
        return 0;
    case TM5_TREE_MODE:
        if (value != TM5_TREE_AUTO && value != TM5_TREE_CHECKED && value != TM5_TREE_REHEARSED) return TM5_EINVAL;
        t->mode = value;
        return 0;
    default:
        return TM5_EINVAL;
    }
}

extern "C" int tm5_tree_tally(tm5_tree *t, long long *kept, long long *redone, long long *replayed, long long *captured)
{
    if (!t) return TM5_EINVAL;
    if (kept) *kept = t->kept;
    if (redone) *redone = t->redone;
    if (replayed) *replayed = t->replayed;
    if (captured) *captured = t->captured;
    return 0;
}
