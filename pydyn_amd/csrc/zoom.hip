// The zoom tree on the device: a global region with nested children (pydyn/zoom/*.py).
//
// A region's arrays keep the Fortran bounds the model declares (m (-1:im+2, -1:jm+2, 1:lm),
// am bm cm (0:im+1, 0:jm+1, 0:lm+1), ...): the exchange between a parent and a child is halo
// arithmetic, so the kernels index with Fortran indices through HV, a view that knows its lower
// bounds.  Region 1 runs the single-region kernels on its interior views; this file has what a
// tree adds:
//   exchange   put_x/yedges (parent cells into the child's halo and interface cells),
//              compress/uncompress_x/yedges, mix_edges (the child's ring of parent-cell blocks
//              set to their means), update_parent (the parent's footprint from the child's sums)
//   mass flow  coarsen_block (the parent's fluxes and mass over the footprint from the
//              child's), coarsen_ambm (the child's outer-interface fluxes from the parent's),
//              m2phlb for a child (the ring's pressure rescaled to the parent cell's area)
//   advection  the child's x rows and y columns with their end cells held (they are the
//              parent's), their sub-cycle searches, and the halos region 1's x writes
//   helpers    the outer interface faces zeroed and restored, whole-array copy and scale.
// Every kernel is one thread per output element (or per line for the advection), each doing
// the reference's operations in its order; the sums that the reference accumulates from zero
// start from 0.0 here too (0.0 + x can change the sign of a zero).
#include <string.h>

#include <type_traits>

#include "tm5_lib.h"

namespace tm5 {

__device__ __forceinline__ bool zskip(const int *words, int mask) { return (words[W_FLAGS] & mask) != 0; }

static dim3 zgrid(long long n, int per = 256) { return dim3((unsigned)((n + per - 1) / per)); }

// a launcher's skip mask with the bits the context adds (a queued tree run's TM5_SPEC)
static inline int zm(const tm5_ctx *ctx, int mask) { return mask | ctx->skip_or; }

// ------------------------------------------------------------------ exchange: put_x/yedges
// pydyn/zoom/exchange.py:17-60: for every child row jc and level lc, the parent cells ib-1, ib,
// ie, ie+1 of the row jbeg + (jc-1)/yref into the child's cells 0, 1..xref, im+1-xref..im,
// im+1, scaled by 1/(yref zref) (outer) or 1/(xref yref zref) (inner; the slopes take the
// outer factor for their inner cells too).  f = 0 is m, f = 1 + a nt + (t-1) tracer array a.
__global__ void __launch_bounds__(256) k_put_x(ZR P, ZR C, int nf, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n2 = (long long)C.jm * C.lm;
    if (idx >= n2 * nf) return;
    const int jc = (int)(idx % C.jm) + 1, lc = (int)((idx / C.jm) % C.lm) + 1, f = (int)(idx / n2);
    const int jp = C.jbeg + (jc - 1) / C.yr, lp = 1 + (lc - 1) / C.zr;
    const double yz = 1.0 / (double)(C.yr * C.zr), xyz = 1.0 / (double)(C.xr * C.yr * C.zr);
    const int ib = C.ibeg, ie = C.iend;
    if (f == 0) {
        const double west = P.m(ib - 1, jp, lp) * yz, west_in = P.m(ib, jp, lp) * xyz;
        const double east_in = P.m(ie, jp, lp) * xyz, east = P.m(ie + 1, jp, lp) * yz;
        C.m(0, jc, lc) = west;
        C.m(C.im + 1, jc, lc) = east;
        for (int i = 1; i <= C.xr; i++) {
            C.m(i, jc, lc) = west_in;
            C.m(C.im + 1 - i, jc, lc) = east_in;
        }
        return;
    }
    const int a = (f - 1) / C.nt, t = (f - 1) % C.nt + 1;
    const HV &src = P.q[a], &dst = C.q[a];
    const double inner = a == 0 ? xyz : yz;
    const double west = src(ib - 1, jp, lp, t) * yz, west_in = src(ib, jp, lp, t) * inner;
    const double east_in = src(ie, jp, lp, t) * inner, east = src(ie + 1, jp, lp, t) * yz;
    dst(0, jc, lc, t) = west;
    dst(C.im + 1, jc, lc, t) = east;
    for (int i = 1; i <= C.xr; i++) {
        dst(i, jc, lc, t) = west_in;
        dst(C.im + 1 - i, jc, lc, t) = east_in;
    }
}

// pydyn/zoom/exchange.py:63-105: the same across the y-walls, column ic of the child from the
// parent column mod(ibeg - 1 + (ic-1)/xref, im) + 1.
__global__ void __launch_bounds__(256) k_put_y(ZR P, ZR C, int nf, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n2 = (long long)C.im * C.lm;
    if (idx >= n2 * nf) return;
    const int ic = (int)(idx % C.im) + 1, lc = (int)((idx / C.im) % C.lm) + 1, f = (int)(idx / n2);
    const int ip = (C.ibeg - 1 + (ic - 1) / C.xr) % P.im + 1, lp = 1 + (lc - 1) / C.zr;
    const double xz = 1.0 / (double)(C.xr * C.zr), xyz = 1.0 / (double)(C.xr * C.yr * C.zr);
    const int jb = C.jbeg, je = C.jend;
    if (f == 0) {
        const double south = P.m(ip, jb - 1, lp) * xz, south_in = P.m(ip, jb, lp) * xyz;
        const double north_in = P.m(ip, je, lp) * xyz, north = P.m(ip, je + 1, lp) * xz;
        C.m(ic, 0, lc) = south;
        C.m(ic, C.jm + 1, lc) = north;
        for (int j = 1; j <= C.yr; j++) {
            C.m(ic, j, lc) = south_in;
            C.m(ic, C.jm + 1 - j, lc) = north_in;
        }
        return;
    }
    const int a = (f - 1) / C.nt, t = (f - 1) % C.nt + 1;
    const HV &src = P.q[a], &dst = C.q[a];
    const double inner = a == 0 ? xyz : xz;
    const double south = src(ip, jb - 1, lp, t) * xz, south_in = src(ip, jb, lp, t) * inner;
    const double north_in = src(ip, je, lp, t) * inner, north = src(ip, je + 1, lp, t) * xz;
    dst(ic, 0, lc, t) = south;
    dst(ic, C.jm + 1, lc, t) = north;
    for (int j = 1; j <= C.yr; j++) {
        dst(ic, j, lc, t) = south_in;
        dst(ic, C.jm + 1 - j, lc, t) = north_in;
    }
}

// ------------------------------------------------------------------ compress / uncompress
// pydyn/zoom/exchange.py:109-129 (x) and :147-165 (y) for one line (j, l) or (i, l): the
// interface cells 1..xref summed in order into cell xref (and im-xref+1..im into im-xref+1),
// the halo cell moved next to them; the sum for m and rm, the move for every array.  f = 0 is
// m, f = t (1..nt) the tracer arrays of tracer t.  `axis` 0 = x (lines (j, l)), 1 = y.
__device__ __forceinline__ double &zat(const HV &q, int axis, int k, int a, int l, int t)
{
    return axis == 0 ? q(k, a, l, t) : q(a, k, l, t);
}

__global__ void __launch_bounds__(256) k_compress(ZR C, int axis, int a0, int a1, int l0, int l1, int nf, int *words,
                                                  int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const int na = a1 - a0 + 1, nl = l1 - l0 + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)na * nl * nf) return;
    const int a = a0 + (int)(idx % na), l = l0 + (int)((idx / na) % nl), f = (int)(idx / ((long long)na * nl));
    const int r = axis == 0 ? C.xr : C.yr, n = axis == 0 ? C.im : C.jm;
    if (f == 0) {
        const HV &m = C.m;
        double s = zat(m, axis, 1, a, l, 1);
        for (int k = 2; k <= r; k++) s = s + zat(m, axis, k, a, l, 1);
        zat(m, axis, r, a, l, 1) = s;
        zat(m, axis, r - 1, a, l, 1) = zat(m, axis, 0, a, l, 1);
        s = zat(m, axis, n - r + 1, a, l, 1);
        for (int k = n - r + 2; k <= n; k++) s = s + zat(m, axis, k, a, l, 1);
        zat(m, axis, n - r + 1, a, l, 1) = s;
        zat(m, axis, n - r + 2, a, l, 1) = zat(m, axis, n + 1, a, l, 1);
        return;
    }
    const int t = f;
    const HV &rm = C.q[0];
    double s = zat(rm, axis, 1, a, l, t);
    for (int k = 2; k <= r; k++) s = s + zat(rm, axis, k, a, l, t);
    zat(rm, axis, r, a, l, t) = s;
    s = zat(rm, axis, n - r + 1, a, l, t);
    for (int k = n - r + 2; k <= n; k++) s = s + zat(rm, axis, k, a, l, t);
    zat(rm, axis, n - r + 1, a, l, t) = s;
    for (int q = 0; q < 4; q++) {
        zat(C.q[q], axis, r - 1, a, l, t) = zat(C.q[q], axis, 0, a, l, t);
        zat(C.q[q], axis, n - r + 2, a, l, t) = zat(C.q[q], axis, n + 1, a, l, t);
    }
}

// pydyn/zoom/exchange.py:132-144 (x), :168-177 (y): the edge cell's value over r shared evenly
// over the interface cells 1..r (and n-r+1..n) again.
__global__ void __launch_bounds__(256) k_uncompress(ZR C, int axis, int a0, int a1, int l0, int l1, int nf, int *words,
                                                    int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const int na = a1 - a0 + 1, nl = l1 - l0 + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)na * nl * nf) return;
    const int a = a0 + (int)(idx % na), l = l0 + (int)((idx / na) % nl), f = (int)(idx / ((long long)na * nl));
    const int r = axis == 0 ? C.xr : C.yr, n = axis == 0 ? C.im : C.jm;
    const int t = f == 0 ? 1 : f;
    const int q0 = f == 0 ? -1 : 0, q1 = f == 0 ? -1 : 3;
    for (int q = q0; q <= q1; q++) {
        const HV &v = q < 0 ? C.m : C.q[q];
        const double lo = zat(v, axis, r, a, l, t) / (double)r;
        for (int k = 1; k <= r; k++) zat(v, axis, k, a, l, t) = lo;
        const double hi = zat(v, axis, n - r + 1, a, l, t) / (double)r;
        for (int k = n - r + 1; k <= n; k++) zat(v, axis, k, a, l, t) = hi;
    }
}

// ------------------------------------------------------------------ mix_edges
// pydyn/zoom/exchange.py:181-231: every parent-cell block (xref x yref child cells) of the
// child's ring (edge == -1) replaced by the mean of its ring cells, accumulated from zero j-outer,
// i-inner; blocks at the x walls only when x edges are mixed, at a y wall unless the region
// touches that pole.  One thread per (ring block, level, array).
// The blocks that are spread, enumerated: the x-wall columns of blocks (ib = 0 and imp-1, every
// jb) when x edges are mixed, then the y-wall rows (jb = 0 unless the region touches the south
// pole, jmp-1 unless it touches the north pole) over the columns the x walls did not take.
struct RingBlocks {
    int imp, jmp, nx, ib0, ib1, ny_s, ny_n;
    __host__ __device__ int count() const { return nx + (ny_s + ny_n) * (ib1 - ib0 + 1); }
    __device__ void at(int k, int &ib, int &jb) const
    {
        if (k < nx) {
            jb = k % jmp;
            ib = k < jmp ? 0 : imp - 1;
            return;
        }
        k -= nx;
        const int w = ib1 - ib0 + 1;
        ib = ib0 + k % w;
        jb = (k / w == 0 && ny_s) ? 0 : jmp - 1;
    }
};

static RingBlocks ring_blocks(const ZR &C, int xedges)
{
    RingBlocks r;
    r.imp = C.im / C.xr, r.jmp = C.jm / C.yr;
    r.nx = xedges ? (r.imp > 1 ? 2 : 1) * r.jmp : 0;
    r.ib0 = xedges ? 1 : 0;
    r.ib1 = xedges ? r.imp - 2 : r.imp - 1;
    r.ny_s = C.touch_sp != 1 ? 1 : 0;
    r.ny_n = (C.touch_np != 1 && !(r.ny_s && r.jmp == 1)) ? 1 : 0;
    if (r.ib1 < r.ib0) r.ny_s = r.ny_n = 0;
    return r;
}

__global__ void __launch_bounds__(256) k_mix_edges(ZR C, RingBlocks ring, int nf, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const long long nb = ring.count();
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb * C.lm * nf) return;
    int ib, jb;
    ring.at((int)(idx % nb), ib, jb);
    const int l = (int)((idx / nb) % C.lm) + 1, f = (int)(idx / (nb * C.lm));
    const HV &v = f == 0 ? C.m : C.q[(f - 1) / C.nt];
    const int t = f == 0 ? 1 : (f - 1) % C.nt + 1;
    const double xyr = (double)(C.xr * C.yr);
    double acc = 0.0;
    for (int b = 0; b < C.yr; b++)
        for (int a = 0; a < C.xr; a++) {
            const int i = ib * C.xr + a, j = jb * C.yr + b;  // 0-based within the region
            const double term = C.edge[(long long)j * C.im + i] == -1 ? v(i + 1, j + 1, l, t) / xyr : 0.0;
            acc = acc + term;
        }
    for (int b = 0; b < C.yr; b++)
        for (int a = 0; a < C.xr; a++) v(ib * C.xr + a + 1, jb * C.yr + b + 1, l, t) = acc;
}

// The same mixing with a workgroup per (level, field): the ring's cells are read in block order
// (a block's cells i fastest, then its rows), so neighbouring lanes take neighbouring cells of a
// row instead of one lane a block; each block's mean is summed in LDS in the order above (j outer,
// i inner, from zero) and written back to its cells.
__global__ void __launch_bounds__(256) k_mix_edges_lds(ZR C, RingBlocks ring, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    extern __shared__ double mix[];  // every cell's term, then every block's mean
    const int nb = ring.count(), per = C.xr * C.yr, ncell = nb * per;
    const int l = (int)(blockIdx.x % C.lm) + 1, f = (int)(blockIdx.x / C.lm);
    const HV &v = f == 0 ? C.m : C.q[(f - 1) / C.nt];
    const int t = f == 0 ? 1 : (f - 1) % C.nt + 1;
    const double xyr = (double)(C.xr * C.yr);
    double *term = mix, *mean = mix + ncell;
    auto cell = [&](int c, int &i, int &j) {  // 0-based within the region
        int ib, jb;
        ring.at(c / per, ib, jb);
        const int w = c % per;
        i = ib * C.xr + w % C.xr;
        j = jb * C.yr + w / C.xr;
    };
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        int i, j;
        cell(c, i, j);
        term[c] = C.edge[(long long)j * C.im + i] == -1 ? v(i + 1, j + 1, l, t) / xyr : 0.0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        double acc = 0.0;
        for (int w = 0; w < per; w++) acc = acc + term[k * per + w];
        mean[k] = acc;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        int i, j;
        cell(c, i, j);
        v(i + 1, j + 1, l, t) = mean[c / per];
    }
}

// ------------------------------------------------------------------ update_parent
// pydyn/zoom/exchange.py:235-258: the parent's m and rm over the footprint from the child's
// block sums (from zero, i outer, j inner); f = 0 m, f = t rm of tracer t.
__device__ __forceinline__ void update_parent_at(const ZR &P, const ZR &C, int nf, long long idx)
{
    const int imp = C.im / C.xr, jmp = C.jm / C.yr;
    const long long nb = (long long)imp * jmp;
    if (idx >= nb * C.lm * nf) return;
    const int ip = (int)(idx % imp), jp = (int)((idx / imp) % jmp);
    const int l = (int)((idx / nb) % C.lm) + 1, f = (int)(idx / (nb * C.lm));
    const HV &src = f == 0 ? C.m : C.q[0];
    const HV &dst = f == 0 ? P.m : P.q[0];
    const int t = f == 0 ? 1 : f;
    double acc = 0.0;
    for (int a = 0; a < C.xr; a++)
        for (int b = 0; b < C.yr; b++) acc = acc + src(ip * C.xr + a + 1, jp * C.yr + b + 1, l, t);
    dst(C.ibeg + ip, C.jbeg + jp, l, t) = acc;
}

__global__ void __launch_bounds__(256) k_update_parent(ZR P, ZR C, int nf, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    update_parent_at(P, C, nf, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------------ coarsen_block
// pydyn/zoom/massflow.py:79-147: the parent's am, bm, cm, m over the footprint from the child's,
// fluxes times tref_ (relative).  am: child faces xref ip summed over each block's rows (j
// ascending); bm: child faces 1 + (jp-1) yref summed over each block's columns (i ascending);
// cm: over the block i outer / j inner; m: j outer / i inner.  All from zero.
__global__ void __launch_bounds__(256) k_coarsen(ZR P, ZR C, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const int imp = C.im / C.xr, jmp = C.jm / C.yr, lm = C.lm;
    const long long na = (long long)(imp + 1) * jmp * lm, nbm = (long long)imp * (jmp + 1) * lm;
    const long long nc = (long long)imp * jmp * lm;
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const double tr = (double)C.tr;
    const int ib = C.ibeg, jb = C.jbeg;
    auto pcol = [&](int k) { return (ib - 2 + k) % P.im + 1; };  // parent i of child block k = 0..imp
    if (idx < na) {
        const int ipf = (int)(idx % (imp + 1)), jp = (int)((idx / (imp + 1)) % jmp), l = (int)(idx / ((long long)(imp + 1) * jmp)) + 1;
        double acc = 0.0;
        for (int b = 0; b < C.yr; b++) acc = acc + C.am(C.xr * ipf, jp * C.yr + b + 1, l) * tr;
        P.am(pcol(ipf), jb + jp, l) = acc;
        return;
    }
    idx -= na;
    if (idx < nbm) {
        const int ip = (int)(idx % imp), jpf = (int)((idx / imp) % (jmp + 1)), l = (int)(idx / ((long long)imp * (jmp + 1))) + 1;
        double acc = 0.0;
        for (int a = 0; a < C.xr; a++) acc = acc + C.bm(ip * C.xr + a + 1, 1 + C.yr * jpf, l) * tr;
        P.bm(pcol(ip + 1), jb + jpf, l) = acc;
        return;
    }
    idx -= nbm;
    if (idx < 2 * nc) {
        const bool mass = idx >= nc;
        if (mass) idx -= nc;
        const int ip = (int)(idx % imp), jp = (int)((idx / imp) % jmp), l = (int)(idx / ((long long)imp * jmp)) + 1;
        double acc = 0.0;
        if (!mass) {
            for (int a = 0; a < C.xr; a++)
                for (int b = 0; b < C.yr; b++) acc = acc + C.cm(ip * C.xr + a + 1, jp * C.yr + b + 1, l) * tr;
            P.cm(pcol(ip + 1), jb + jp, l) = acc;
        } else {
            for (int b = 0; b < C.yr; b++)
                for (int a = 0; a < C.xr; a++) acc = acc + C.m(ip * C.xr + a + 1, jp * C.yr + b + 1, l);
            P.m(pcol(ip + 1), jb + jp, l) = acc;
        }
    }
}

// ------------------------------------------------------------------ coarsen_ambm
// pydyn/zoom/massflow.py:150-177: the child's outer-interface fluxes from the parent's face
// flux divided over the child faces it covers.  First (jm lm) threads: the x walls (unless the
// child spans the globe); then (im lm): the y walls of a child touching no pole.
__global__ void __launch_bounds__(256) k_coarsen_ambm(ZR P, ZR C, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nx = (long long)C.jm * C.lm, ny = (long long)C.im * C.lm;
    if (idx < nx) {
        if (C.whole) return;
        const int jc = (int)(idx % C.jm) + 1, lc = (int)(idx / C.jm) + 1;
        const int j = C.jbeg + (jc - 1) / C.yr, lev = 1 + (lc - 1) / C.zr;
        const double div = (double)(C.yr * C.zr);
        C.am(C.xr - 1, jc, lc) = P.am(C.ibeg - 1, j, lev) / div;
        C.am(C.im - C.xr + 1, jc, lc) = P.am(C.iend, j, lev) / div;
        return;
    }
    idx -= nx;
    if (idx < ny) {
        const int ic = (int)(idx % C.im) + 1, lc = (int)(idx / C.im) + 1;
        const int i = C.ibeg + (ic - 1) / C.xr, lev = 1 + (lc - 1) / C.zr;
        const double div = (double)(C.xr * C.zr);
        if (C.touch_sp != 1) C.bm(ic, C.yr, lc) = P.bm(i, C.jbeg, lev) / div;
        if (C.touch_np != 1) C.bm(ic, C.jm - C.yr + 2, lc) = P.bm(i, C.jend + 1, lev) / div;
    }
}

// ------------------------------------------------------------------ m2phlb / m2phlb1
// pydyn/zoom/massflow.py:24-58: sp = seqsum(m over l) g / dxyp(j) per column, a child's ring
// cells rescaled p dxyp(j) xref yref / dxyp_parent(jpar); phlb = at + bt sp.  The sp halo stays
// zero (it is never written).  pdxyp null: region 1 (m2phlb1).
// clock_n > 0: the last kernel of a queued run's step, which moves the context's clock on to the
// next step (clock_n steps an interval; thread 0: nothing else in the kernel reads the clock)
// One workgroup's tile: TI columns i0 + 1 .. (0-based bx TI ..) of row j + 1, as k_m2phlb1 tiles
// them: the columns' masses staged in LDS (lm TI doubles, then TI for sp), each column's sum from
// level 1 up by one thread, then phlb of every level in parallel.
__device__ __forceinline__ void m2phlb_tile(const ZR &R, const double *pdxyp, const double *at, const double *bt,
                                            int bx, int j, int TI, double *sh)
{
    double *M = sh, *P = sh + R.lm * TI;
    const int i0 = bx * TI, lm = R.lm;
    const int ni = min(TI, R.im - i0);
    const int tid = threadIdx.x, nth = blockDim.x;
    for (int idx = tid; idx < lm * TI; idx += nth) {
        const int ti = idx % TI, l = idx / TI;
        if (ti < ni) M[idx] = R.m(i0 + ti + 1, j + 1, l + 1);
    }
    __syncthreads();
    for (int ti = tid; ti < ni; ti += nth) {
        const int i = i0 + ti + 1, jj = j + 1;  // Fortran
        double p = M[ti];
        for (int l = 1; l < lm; l++) p = p + M[l * TI + ti];
        p = p * 9.80665 / R.dxyp[jj - 1];
        if (pdxyp) {
            const int xr = R.xr, yr = R.yr;
            const bool inner_i = xr < i && i < R.im - xr + 1, inner_j = yr < jj && jj < R.jm - yr + 1;
            const bool keep = (inner_i && inner_j) || (R.xcyc == 1 && inner_j) || (R.touch_np == 1 && jj > R.jm - yr) ||
                              (R.touch_sp == 1 && jj < yr + 1);
            if (!keep) {
                const int jpar = R.jbeg + (jj - 1) / yr;
                p = p * R.dxyp[jj - 1] * (double)xr * (double)yr / pdxyp[jpar - 1];
            }
        }
        R.sp(i, jj, 1) = p;
        P[ti] = p;
    }
    __syncthreads();
    for (int idx = tid; idx < (lm + 1) * TI; idx += nth) {
        const int ti = idx % TI, l = idx / TI;
        if (ti < ni) R.phlb(i0 + ti + 1, j + 1, l + 1) = at[l] + bt[l] * P[ti];
    }
}

// grid (tiles of TI columns, jm); clock_n > 0: the last kernel of a queued run's step, which moves
// the context's clock on to the next step (clock_n steps an interval; thread 0: nothing else in the
// kernel reads the clock)
__global__ void __launch_bounds__(256) k_m2phlb_zoom(ZR R, const double *pdxyp, const double *at, const double *bt,
                                                     int TI, int *words, int skip_mask, int clock_n)
{
    if (zskip(words, skip_mask)) return;
    extern __shared__ double msh[];
    if (clock_n > 0 && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        const int s = words[W_CLOCK_S] + 1;
        words[W_CLOCK_S] = s < clock_n ? s : 0;
        if (s >= clock_n) words[W_CLOCK_K] += 1;
    }
    m2phlb_tile(R, pdxyp, at, bt, blockIdx.x, blockIdx.y, TI, msh);
}

// the end of a child's Proces_Region in one launch: update_parent (the parent's m, rm over the
// footprint from the child's sums) and the child's m2phlb (its own sp, phlb from its m), which
// touch disjoint arrays -- the first nub workgroups the one, the rest the tiles of the other
__global__ void __launch_bounds__(256) k_update_m2phlb(ZR P, ZR C, int nf, int nub, const double *at,
                                                       const double *bt, int TI, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    extern __shared__ double msh[];
    if ((int)blockIdx.x < nub) {
        update_parent_at(P, C, nf, (long long)blockIdx.x * blockDim.x + threadIdx.x);
        return;
    }
    const int t = blockIdx.x - nub, nbx = (C.im + TI - 1) / TI;
    m2phlb_tile(C, P.dxyp, at, bt, t % nbx, t / nbx, TI, msh);
}

// ------------------------------------------------------------------ region 1's x halos
// pydyn/zoom/advect.py:283-288 (and check_cfl.py:113-117 for m alone): after dynamu, every row
// js..je of m and the tracer arrays gets its periodic halo q(0) = q(im), q(im+1) = q(1).
__global__ void __launch_bounds__(256) k_x_halo(ZR R, int js, int je, int nf, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const int nj = je - js + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nj * R.lm * nf) return;
    const int j = js + (int)(idx % nj), l = (int)((idx / nj) % R.lm) + 1, f = (int)(idx / ((long long)nj * R.lm));
    const HV &v = f == 0 ? R.m : R.q[(f - 1) / R.nt];
    const int t = f == 0 ? 1 : (f - 1) % R.nt + 1;
    v(0, j, l, t) = v(R.im, j, l, t);
    v(R.im + 1, j, l, t) = v(1, j, l, t);
}

// ------------------------------------------------------------------ outer interface faces
// pydyn/zoom/advect.py:121-157: a face slab (am at i = k, all j and l of the halo'd array; or bm
// at j = k, all i and l) saved and zeroed (restore = 0), or put back (restore = 1).
__global__ void __launch_bounds__(256) k_slab(HV f, int axis, int k, int n0, int n1, double *save, int restore,
                                              int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n0 * n1) return;
    const int a = idx % n0, l = idx / n0;  // a over the other horizontal axis (from its lower bound)
    double &v = axis == 0 ? f(k, f.l1 + a, f.l2 + l) : f(f.l0 + a, k, f.l2 + l);
    if (restore) {
        v = save[idx];
    } else {
        save[idx] = v;
        v = 0.0;
    }
}

// ------------------------------------------------------------------ whole arrays
__global__ void __launch_bounds__(256) k_copy_flat(const double *src, double *dst, long long n)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) dst[idx] = src[idx];
}

__global__ void __launch_bounds__(256) k_scale_flat(double *a, long long n, double factor)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) a[idx] = a[idx] * factor;
}

// the cyclic x halo faces of region 1's fluxes: am(im+1) = am(1) for every j, l
__global__ void __launch_bounds__(256) k_am_halo(ZR R, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const int n0 = R.jm + 2, n1 = R.lm + 2;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n0 * n1) return;
    const int j = idx % n0, l = idx / n0;
    R.am(R.im + 1, j, l) = R.am(1, j, l);
}

// ------------------------------------------------------------------ a child's x rows
// pydyn/zoom/advect.py:301-366 (and check_cfl.py:121-145 mass-only): one workgroup of XC_THREADS
// lanes (a cell a lane for rows of up to 128 cells) per (j, l) row of cells xref-1..im-xref+2 (n cells; the two ends are the parent's, held) and its
// n-1 faces am(xref-1..im-xref+1).  The row's sub-cycle search scales its faces by (n-1)/n in
// place per retry (kernels.py:169-199 with the ends held), the limiter runs on every cell of the
// row first, the sub-steps update the interior cells, and the faces are multiplied back.
// X_COMPRESS: compress_xedges before and uncompress_xedges after (pydyn/zoom/exchange.py:109-144) made
// on the row itself: the row's end cells read as compress leaves them (cell 0 the halo cell, cell 1
// the sum of cells 1..xref for m and rm, the east end alike), and cells 1..xref (im-xref+1..im)
// written as uncompress leaves them, the cell's value over xref -- the cells the two kernels write
// and nothing reads in between.  X_OUTER: outer_x_faces_zeroed (pydyn/zoom/advect.py:121-136): the
// row's first and last faces taken as zero and left in memory as they are (zeroed, then restored).
constexpr int X_COMPRESS = FUSE_COMPRESS, X_OUTER = FUSE_OUTER;
// Lanes a row: with many tracers each row's workgroup loops over its tracer range, loading the
// next tracer while it steps the current one; two waves a row keep twice the loads in flight per
// CU as one (the LDS a row needs, not the waves, bounds the rows a CU holds), which is what the
// loop's memory latency asks for.
constexpr int XC_THREADS = 128;

// cell c of a child's x row (Fortran i0 + c, i0 = xref - 1) as compress_xedges leaves it
__device__ __forceinline__ double xrow_cell(const HV &v, int c, int n, int i0, int r, int im, int j, int l, int t,
                                            bool compress, bool sum)
{
    if (compress) {
        if (c == 0) return v(0, j, l, t);
        if (c == n - 1) return v(im + 1, j, l, t);
        if (sum && c == 1) {
            double s = v(1, j, l, t);
            for (int k = 2; k <= r; k++) s = s + v(k, j, l, t);
            return s;
        }
        if (sum && c == n - 2) {
            double s = v(im - r + 1, j, l, t);
            for (int k = im - r + 2; k <= im; k++) s = s + v(k, j, l, t);
            return s;
        }
    }
    return v(i0 + c, j, l, t);
}

// cell c of a child's x row written back: as uncompress_xedges leaves it, or (uncompressed) as is;
// the ends (the parent's cells) only when held is false
__device__ __forceinline__ void xrow_store(const HV &v, int c, int n, int i0, int r, int im, int j, int l, int t,
                                           bool compress, bool ends, double x)
{
    if (compress) {
        if (c == 0 || c == n - 1) return;  // inside 1..xref / im-xref+1..im: uncompress writes them
        if (c == 1 || c == n - 2) {
            const double lo = x / (double)r;
            const int k0 = c == 1 ? 1 : im - r + 1;
            for (int k = k0; k < k0 + r; k++) v(k, j, l, t) = lo;
            return;
        }
    }
    if (!ends && (c == 0 || c == n - 1)) return;
    v(i0 + c, j, l, t) = x;
}

template <bool MASS_ONLY, int FUSE>
__global__ void __launch_bounds__(XC_THREADS) k_x_child(ZR C, int js, int je, int ls, int le, int limits, int nmax,
                                                int fail_bit, int *words, int skip_mask, int tchunk, int spec_k)
{
    // tchunk > 0: tracer ranges (grid y), several workgroups a row, each searching and moving the
    // row's mass on its own; none writes m or am, the caller runs the mass-only form afterwards
    // checked(spec_k): a search past nmax is a failed check of a checked step (spec_fail)
    // FUSE: X_COMPRESS, X_OUTER
    if (zskip(words, skip_mask)) return;
    extern __shared__ double lds[];
    const int nj = je - js + 1;
    const int b = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);  // the range slowest
    const int row = b % gridDim.x, bz = b / gridDim.x;
    const int j = js + row % nj, l = ls + row / nj;
    const int t0 = tchunk > 0 ? bz * tchunk + 1 : 1;  // Fortran tracer indices t0 .. t1
    const int t1 = tchunk > 0 && t0 - 1 + tchunk < C.nt ? t0 - 1 + tchunk : C.nt;
    const int lane = threadIdx.x;
    const int i0 = C.xr - 1;  // Fortran i of the row's cell 0
    const int n = C.im - 2 * C.xr + 4, nf = n - 1;
    constexpr bool cz = (FUSE & X_COMPRESS) != 0, oz = (FUSE & X_OUTER) != 0;
    const int xr = C.xr, im = C.im;
    double *M = lds, *M0 = M + n, *A = M0 + n, *R = A + nf, *F = R + 4 * n;

    for (int c = lane; c < n; c += XC_THREADS) M0[c] = M[c] = xrow_cell(C.m, c, n, i0, xr, im, j, l, 1, cz, true);
    for (int k = lane; k < nf; k += XC_THREADS) A[k] = oz && (k == 0 || k == nf - 1) ? 0.0 : C.am(i0 + k, j, l);
    __syncthreads();

    int nl = 1;
    for (;;) {
        bool ok = true;
        for (int k = lane; k < nf; k += XC_THREADS) {
            const double ak = A[k];
            const int u = ak >= 0.0 ? k : k + 1;
            const bool held = u == 0 || u == n - 1;
            const double alo = held ? 0.0 : A[u - 1], ahi = held ? 0.0 : A[u];
            double mx = M[u];
            for (int s = 0; s < nl; s++) {
                if (!(fabs(ak / mx) < 1.0)) {
                    ok = false;
                    break;
                }
                if (!held) mx = (mx + alo) - ahi;
            }
        }
        if (__syncthreads_and(ok)) break;
        nl += 1;
        __syncthreads();
        for (int k = lane; k < nf; k += XC_THREADS) A[k] = shrink(A[k], nl);
        __syncthreads();
        if (nl > nmax) {
            if (lane == 0) {
                if (checked(spec_k))
                    spec_fail(words, spec_k);
                else
                    atomicOr(words + W_FLAGS, fail_bit);
            }
            return;
        }
    }

    if (MASS_ONLY) {
        for (int s = 0; s < nl; s++) {
            for (int c = lane + 1; c < n - 1; c += XC_THREADS) M[c] = (M[c] + A[c - 1]) - A[c];
            __syncthreads();
        }
        for (int c = lane + 1; c < n - 1; c += XC_THREADS) xrow_store(C.m, c, n, i0, xr, im, j, l, 1, cz, false, M[c]);
    } else {
        // each tracer's moments of the lane's (at most two) cells are loaded during the sub-steps
        // of the tracer before
        constexpr int PRE = 2;
        const bool pre = n <= XC_THREADS * PRE;
        double qp[PRE][4];
        auto fetch = [&](int t) {
#pragma unroll
            for (int k = 0; k < PRE; k++) {
                const int c = lane + XC_THREADS * k;
                if (c < n)
#pragma unroll
                    for (int a = 0; a < 4; a++) qp[k][a] = xrow_cell(C.q[a], c, n, i0, xr, im, j, l, t, cz, a == 0);
            }
        };
        if (pre && t0 <= t1) fetch(t0);
        for (int t = t0; t <= t1; t++) {
            if (t > t0) {
                for (int c = lane; c < n; c += XC_THREADS) M[c] = M0[c];
            }
            for (int c = lane; c < n; c += XC_THREADS) {
                double rm, rx, ry, rz;
                if (pre) {
                    const int k = (c - lane) / XC_THREADS;
                    rm = k == 0 ? qp[0][0] : qp[1][0];
                    rx = k == 0 ? qp[0][1] : qp[1][1];
                    ry = k == 0 ? qp[0][2] : qp[1][2];
                    rz = k == 0 ? qp[0][3] : qp[1][3];
                } else {
                    rm = xrow_cell(C.q[0], c, n, i0, xr, im, j, l, t, cz, true);
                    rx = xrow_cell(C.q[1], c, n, i0, xr, im, j, l, t, cz, false);
                    ry = xrow_cell(C.q[2], c, n, i0, xr, im, j, l, t, cz, false);
                    rz = xrow_cell(C.q[3], c, n, i0, xr, im, j, l, t, cz, false);
                }
                R[c] = rm;
                R[n + c] = limits ? limit(rx, rm) : rx;
                R[2 * n + c] = ry;
                R[3 * n + c] = rz;
            }
            if (pre && t < t1) fetch(t + 1);
            __syncthreads();
            for (int s = 0; s < nl; s++) {
                for (int k = lane; k < nf; k += XC_THREADS) {
                    const double ak = A[k];
                    const int up = ak >= 0.0 ? k : k + 1;
                    const double alpha = ak * (1.0 / M[up]);
                    const FaceFlux ff = face_flux(ak, alpha, R[up], R[n + up], R[2 * n + up], R[3 * n + up]);
                    F[k] = ff.f;
                    F[nf + k] = ff.pf;
                    F[2 * nf + k] = ff.ft;
                    F[3 * nf + k] = ff.fw;
                }
                __syncthreads();
                for (int c = lane + 1; c < n - 1; c += XC_THREADS) {
                    const double m_old = M[c], a_lo = A[c - 1], a_hi = A[c];
                    const double m_new = (m_old + a_lo) - a_hi;
                    const double *f = F, *pf = F + nf, *fy = F + 2 * nf, *fz = F + 3 * nf;
                    const double rm_new = R[c] - (f[c] - f[c - 1]);
                    double rx = slope_update(R[n + c], pf[c - 1], pf[c], a_lo, a_hi, rm_new, f[c - 1], f[c], m_old, m_new);
                    if (limits) rx = limit(rx, rm_new);
                    R[c] = rm_new;
                    R[n + c] = rx;
                    R[2 * n + c] = R[2 * n + c] - (fy[c] - fy[c - 1]);
                    R[3 * n + c] = R[3 * n + c] - (fz[c] - fz[c - 1]);
                    M[c] = m_new;
                }
                __syncthreads();
            }
            for (int c = lane; c < n; c += XC_THREADS) {
#pragma unroll
                for (int a = 0; a < 4; a++) xrow_store(C.q[a], c, n, i0, xr, im, j, l, t, cz, true, R[a * n + c]);
            }
            __syncthreads();
        }
        if (tchunk > 0) return;
        for (int c = lane + 1; c < n - 1; c += XC_THREADS) xrow_store(C.m, c, n, i0, xr, im, j, l, 1, cz, false, M[c]);
    }
    for (int k = lane; k < nf; k += XC_THREADS)
        if (!(oz && (k == 0 || k == nf - 1))) C.am(i0 + k, j, l) = A[k] * (double)nl;  // "restore old am"
}

// ------------------------------------------------------------------ a child's y columns
// pydyn/zoom/advect.py:369-432 (check_cfl.py:171-201 mass-only): per layer l, the columns
// is..ie with cells yref-1..jm-yref+2 (n cells, the ends held) and faces bm(yref..jm-yref+1).
// The layer's search: nloop = 1, 2, ... until every face passes its trial; on a retry the whole
// halo'd layer of bm is scaled (the scope's faces only in the mass-only form).  One workgroup
// per layer; nloop[l] for the sub-step kernel.
template <bool MASS_ONLY>
__global__ void __launch_bounds__(1024) k_y_child_search(ZR C, int is, int ie, int ls, int nmax, int *nloop,
                                                         int fail_bit, int *words, int skip_mask)
{
    const int l = ls + blockIdx.x;
    if (zskip(words, skip_mask)) return;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int j0 = C.yr - 1, n = C.jm - 2 * C.yr + 4, nf = n - 1, ni = ie - is + 1;
    int nl = 1;
    for (;;) {
        bool ok = true;
        for (int idx = tid; idx < ni * nf; idx += nth) {
            const int i = is + idx % ni, k = idx / ni;
            const double bk = C.bm(i, j0 + 1 + k, l);
            const int u = bk >= 0.0 ? k : k + 1;
            const bool held = u == 0 || u == n - 1;
            const double blo = held ? 0.0 : C.bm(i, j0 + u, l), bhi = held ? 0.0 : C.bm(i, j0 + 1 + u, l);
            double mx = C.m(i, j0 + u, l);
            for (int s = 0; s < nl; s++) {
                if (!(fabs(bk / mx) < 1.0)) {
                    ok = false;
                    break;
                }
                if (!held) mx = (mx + blo) - bhi;
            }
        }
        if (__syncthreads_and(ok)) break;
        nl += 1;
        if (MASS_ONLY) {
            for (int idx = tid; idx < ni * nf; idx += nth) {
                double &b = C.bm(is + idx % ni, j0 + 1 + idx / ni, l);
                b = shrink(b, nl);
            }
        } else {
            const int n0 = C.im + 2, n1 = C.jm + 2;
            for (int idx = tid; idx < n0 * n1; idx += nth) {
                double &b = C.bm(C.bm.l0 + idx % n0, C.bm.l1 + idx / n0, l);
                b = shrink(b, nl);
            }
        }
        __threadfence_block();
        __syncthreads();
        if (nl > nmax) {
            if (tid == 0) {
                atomicOr(words + W_FLAGS, fail_bit);
                nloop[blockIdx.x] = nl;
            }
            return;
        }
    }
    if (tid == 0) nloop[blockIdx.x] = nl;
}

// Where a cell of a child's y column (Fortran j0 + c, j0 = yref - 1) comes from and goes to when
// compress_yedges runs before and uncompress_yedges after (pydyn/zoom/exchange.py:147-177): the
// halo cell moved next to the interface (row 0 or jm + 1), the interface's cells 1..yref
// (jm-yref+1..jm) summed into one for m and rm, and afterwards the interface cell's value over
// yref on each of those cells; the moved halo cells lie inside them and are not stored.
struct YCell {
    int row;      // the row read for rxm, rym, rzm (and m, rm when sum == 1)
    int sum0;     // the first row summed for m and rm ...
    int sum;      // ... and how many
    int store0;   // the first row written ...
    int stores;   // ... how many (0: none), each the value over `div`
    double div;
};

__device__ __forceinline__ YCell ycell(int c, int n, int j0, int r, int jm, bool compress)
{
    YCell y{j0 + c, j0 + c, 1, j0 + c, 1, 1.0};
    if (!compress) return y;
    if (c == 0 || c == n - 1) {
        y.row = y.sum0 = c == 0 ? 0 : jm + 1;
        y.stores = 0;
    } else if (c == 1 || c == n - 2) {
        y.sum0 = y.store0 = c == 1 ? 1 : jm - r + 1;
        y.sum = y.stores = r;
        y.div = (double)r;
    }
    return y;
}

__device__ __forceinline__ double ycol_load(const HV &v, const YCell &y, int i, int l, int t, bool sum)
{
    if (!sum) return v(i, y.row, l, t);
    double s = v(i, y.sum0, l, t);
    for (int k = 1; k < y.sum; k++) s = s + v(i, y.sum0 + k, l, t);
    return s;
}

__device__ __forceinline__ void ycol_store(const HV &v, const YCell &y, int i, int l, int t, double x)
{
    const double o = y.stores > 1 ? x / y.div : x;
    for (int k = 0; k < y.stores; k++) v(i, y.store0 + k, l, t) = o;
}

// The sub-steps of the layer's columns: one thread per cell of TI whole columns; each sub-step
// every thread reads its cell's neighbourhood, computes its two faces' fluxes and its update,
// then -- after the workgroup's barrier -- writes it.  The end cells are held; the limiter of
// the first sub-step is taken on the values as read (the reference limits rym on every cell of
// the scope, ends included, before the loop; the ends are written back limited).
// With tracer ranges (tchunk > 0, grid z) several workgroups take the same columns: each moves
// the masses through the sub-steps on its own (in LDS), none writes them, and the caller runs
// the mass-only form afterwards.
// nloop null: every layer at nloop 1 (a Check_CFL rehearsal at cap 1 passed, or a checked step:
// checked(spec_k), which first makes that rehearsal's test of every face of the columns at nloop 1).
// fuse (nloop null only): Y_COMPRESS -- compress_yedges before and uncompress_yedges after made on
// the columns themselves (a workgroup holds whole columns); Y_OUTER -- outer_y_faces_zeroed
// (pydyn/zoom/advect.py:139-157): the columns' first and last faces taken as zero, left in memory.
constexpr int Y_COMPRESS = FUSE_COMPRESS, Y_OUTER = FUSE_OUTER;

template <bool MASS_ONLY, int FUSE>
__global__ void __launch_bounds__(1024) k_y_child_steps(ZR C, int is, int ie, int ls, int TI, const int *nloop,
                                                        int limits, int *words, int skip_mask, int tchunk, int spec_k)
{
    if (zskip(words, skip_mask)) return;
    // neighbouring tiles, which share cache lines, on one XCD (as xcd_grid2), the tracer range slowest
    const int nbx = gridDim.x, nby = gridDim.y;
    const int b = xcd_tile((blockIdx.z * nby + blockIdx.y) * nbx + blockIdx.x, nbx * nby * gridDim.z);
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int t0 = tchunk > 0 ? bz * tchunk + 1 : 1;  // Fortran tracer indices t0 .. t1
    const int t1 = tchunk > 0 && t0 - 1 + tchunk < C.nt ? t0 - 1 + tchunk : C.nt;
    const int l = ls + by;
    const int j0 = C.yr - 1, n = C.jm - 2 * C.yr + 4;
    const int ti = threadIdx.x % TI, c = threadIdx.x / TI;
    const int i = is + bx * TI + ti;
    const bool mine = c < n && i <= ie;
    const int ii = mine ? i : is, cc = mine ? c : 0;
    const int jc = j0 + cc;  // Fortran j of the cell
    const bool inner = cc > 0 && cc < n - 1;
    const int nl = nloop ? nloop[by] : 1;
    constexpr bool cz = (FUSE & Y_COMPRESS) != 0, oz = (FUSE & Y_OUTER) != 0;
    const int yr = C.yr, jm = C.jm;
    // the column's masses as the step starts (compressed when cz)
    const YCell yc = ycell(cc, n, j0, yr, jm, cz);
    auto mass = [&](int k) { return ycol_load(C.m, ycell(k, n, j0, yr, jm, cz), ii, l, 1, true); };
    if (checked(spec_k)) {
        // the rehearsal's dynamvm at cap 1 (pydyn/zoom/check_cfl.py:171-201): the face above each
        // cell but the last, |b / m(upwind)| < 1 on the masses as they are before the step
        bool ok = true;
        if (mine && cc < n - 1) {
            const double bk = oz && (cc == 0 || cc == n - 2) ? 0.0 : C.bm(ii, jc + 1, l);
            ok = fabs(bk / mass(bk >= 0.0 ? cc : cc + 1)) < 1.0;
        }
        if (__any(!ok) && (threadIdx.x & 63) == 0) spec_fail(words, spec_k);
    }
    // faces below (bm(jc)) and above (bm(jc+1)) an interior cell
    const double b_lo = inner && !(oz && cc == 1) ? C.bm(ii, jc, l) : 0.0;
    const double b_hi = inner && !(oz && cc == n - 2) ? C.bm(ii, jc + 1, l) : 0.0;
    // each tracer's moments of the workgroup's cells pass through LDS ([moment][cell][ti]): every
    // cell is read from HBM once and its neighbours from LDS
    extern __shared__ double yst[];
    const int stride = n * TI, me = cc * TI + ti;
    const int u_lo = b_lo >= 0.0 ? cc - 1 : cc, u_hi = b_hi >= 0.0 ? cc : cc + 1;  // cells of the line
    const bool staged = tchunk > 0 || cz;  // the masses of the sub-step from LDS, not memory
    double m_cur = ycol_load(C.m, yc, ii, l, 1, true);
    for (int s = 0; s < nl; s++) {
        const double m_old = (MASS_ONLY || tchunk == 0) && !cz ? C.m(ii, jc, l) : m_cur;
        const double m_new = (m_old + b_lo) - b_hi;
        if (!MASS_ONLY) {
            const bool lim0 = limits && s == 0;
            // the Courant numbers of the sub-step, the same for every tracer
            double beta_lo, beta_hi;
            if (staged) {  // the line's masses at the sub-step's start, from LDS
                double *ym = yst + 4 * stride;
                if (mine) ym[me] = m_old;
                __syncthreads();
                beta_lo = inner ? b_lo / ym[u_lo * TI + ti] : 0.0;
                beta_hi = inner ? b_hi / ym[u_hi * TI + ti] : 0.0;
            } else {  // as written by the sub-step before
                beta_lo = inner ? b_lo / C.m(ii, j0 + u_lo, l) : 0.0;
                beta_hi = inner ? b_hi / C.m(ii, j0 + u_hi, l) : 0.0;
            }
            // each tracer's moments are loaded while the tracer before it is worked on
            double nx[4] = {0.0, 0.0, 0.0, 0.0};
            if (t0 <= t1) {
                for (int k = 0; k < 4; k++) nx[k] = ycol_load(C.q[k], yc, ii, l, t0, k == 0);
            }
            for (int t = t0; t <= t1; t++) {
                const double rm = nx[0], rx0 = nx[1], ry0 = nx[2], rz0 = nx[3];
                if (t < t1) {
                    for (int k = 0; k < 4; k++) nx[k] = ycol_load(C.q[k], yc, ii, l, t + 1, k == 0);
                }
                if (mine) {
                    yst[me] = rm;
                    yst[stride + me] = rx0;
                    yst[2 * stride + me] = ry0;
                    yst[3 * stride + me] = rz0;
                }
                __syncthreads();
                double rym = ry0;
                if (lim0) rym = limit(rym, rm);
                double out0 = rm, out1 = rx0, out2 = rym, out3 = rz0;
                if (inner) {
                    FaceFlux lo, hi;
                    {
                        const int u = u_lo * TI + ti;
                        const double rmu = yst[u], ryu = lim0 ? limit(yst[2 * stride + u], rmu) : yst[2 * stride + u];
                        lo = face_flux(b_lo, beta_lo, rmu, ryu, yst[stride + u], yst[3 * stride + u]);
                    }
                    {
                        const int u = u_hi * TI + ti;
                        const double rmu = yst[u], ryu = lim0 ? limit(yst[2 * stride + u], rmu) : yst[2 * stride + u];
                        hi = face_flux(b_hi, beta_hi, rmu, ryu, yst[stride + u], yst[3 * stride + u]);
                    }
                    const double inv_mnew = 1.0 / m_new;
                    const double flow_conv = b_lo - b_hi, flow_mean = 3.0 * (b_lo + b_hi), mass_x3 = 3.0 * m_old;
                    const double rm_new = rm + (lo.f - hi.f);
                    const double f_sum = lo.f + hi.f;
                    double ry = rym + ((((lo.pf - hi.pf) - flow_conv * rym) + flow_mean * rm_new) - mass_x3 * f_sum) * inv_mnew;
                    if (limits) ry = limit(ry, rm_new);
                    out0 = rm_new;
                    out2 = ry;
                    out1 = out1 + (lo.ft - hi.ft);
                    out3 = out3 + (lo.fw - hi.fw);
                }
                if (mine && (inner || lim0)) {  // no other thread reads the cell from memory
                    ycol_store(C.q[0], yc, ii, l, t, out0);
                    ycol_store(C.q[1], yc, ii, l, t, out1);
                    ycol_store(C.q[2], yc, ii, l, t, out2);
                    ycol_store(C.q[3], yc, ii, l, t, out3);
                }
                __syncthreads();  // the staged values are read before the next tracer's land
            }
            m_cur = m_new;
        } else {
            m_cur = m_new;
        }
        __syncthreads();
        if (mine && inner && (MASS_ONLY || tchunk == 0)) ycol_store(C.m, yc, ii, l, 1, m_new);
        __threadfence_block();
        __syncthreads();
    }
}

// the whole halo'd layer of bm multiplied back by its nloop ("restore old bm's")
__global__ void __launch_bounds__(256) k_y_child_restore(ZR C, int ls, const int *nloop, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const int l = ls + blockIdx.y, nl = nloop[blockIdx.y];
    if (nl == 1) return;  // times 1.0 is the identity
    const int n0 = C.im + 2, n1 = C.jm + 2;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n0 * n1) return;
    double &b = C.bm(C.bm.l0 + idx % n0, C.bm.l1 + idx / n0, l);
    b = b * (double)nl;
}


// ------------------------------------------------------------------ region 1's reduced grid
// pydyn/redgrid.py:57-92 on the rows jred[b] (1-based) with c = im / imred cells per cluster, around
// region 1's dynamu (zoom_x_reduced).  One 64-lane workgroup per (reduced row, level), the row
// staged in LDS, since the clusters are summed (and later split) in place.
//
// To the reduced grid: the row's uniform m (into mu) and am (into au) kept for the split back; am
// subsampled, am(i) = am(i c) for i = 1..imred, am(imred+1..im) = -1, am(0) = am(imred); m and
// every tracer array summed over each cluster from its first cell, the tracers beyond the reduced
// row zeroed (m is not).
__global__ void __launch_bounds__(64) k_uni2red(ZR R, HV mu, HV au, int tracers, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    extern __shared__ double row[];  // im + 2
    const int j = R.jred[blockIdx.x], l = blockIdx.y + 1, lane = threadIdx.x;
    const int n = R.imred[j - 1], c = R.im / n;
    for (int i = lane; i <= R.im + 1; i += 64) {
        const double v = R.am(i, j, l);
        row[i] = v;
        au(i, j, l) = v;
    }
    __syncthreads();
    for (int i = 1 + lane; i <= R.im; i += 64) R.am(i, j, l) = i <= n ? row[i * c] : -1.0;
    if (lane == 0) R.am(0, j, l) = row[n * c];
    __syncthreads();
    const int nf = tracers ? 1 + 4 * R.nt : 1;
    for (int f = 0; f < nf; f++) {
        const HV &v = f == 0 ? R.m : R.q[(f - 1) % 4];
        const int t = f == 0 ? 1 : (f - 1) / 4 + 1;
        for (int i = lane; i < R.im; i += 64) {
            const double x = v(i + 1, j, l, t);
            row[i] = x;
            if (f == 0) mu(i + 1, j, l) = x;
        }
        __syncthreads();
        for (int k = lane; k < n; k += 64) {
            double acc = row[k * c];
            for (int a = 1; a < c; a++) acc = acc + row[k * c + a];
            v(k + 1, j, l, t) = acc;
        }
        if (f > 0)  // the tracers beyond the reduced row are zeroed (m is not)
            for (int i = n + lane; i < R.im; i += 64) v(i + 1, j, l, t) = 0.0;
        __syncthreads();
    }
}

// Back to the uniform grid: the clusters split with the row's uniform masses advanced by its
// uniform faces, m_uni(i) = (mu(i) + au(i-1)) - au(i): q(i) = (m_uni(i) / m(cluster)) q(cluster) for
// every tracer array, m = m_uni; am put back as it was before the subsampling (the reference's copy
// of the faces); then the halo (pydyn/zoom/advect.py:237-277, the reference's quirk): with tracers
// q(0) = q(im) for m and every tracer array on a row with fewer cells than im; mass-only (Check_CFL's
// rehearsal) on the rows js..je its x advects m(0) = m(im), and m(im + 1) = m(1) on a row of im.
__global__ void __launch_bounds__(64) k_red2uni(ZR R, HV mu, HV au, int tracers, int js, int je, int *words,
                                                int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    extern __shared__ double sh[];
    double *red = sh, *mred = sh + R.im;
    const int j = R.jred[blockIdx.x], l = blockIdx.y + 1, lane = threadIdx.x;
    const int n = R.imred[j - 1], c = R.im / n;
    for (int k = lane; k < n; k += 64) mred[k] = R.m(k + 1, j, l);
    __syncthreads();
    const int nq = tracers ? 4 * R.nt : 0;
    for (int f = 0; f < nq; f++) {
        const HV &v = R.q[f % 4];
        const int t = f / 4 + 1;
        for (int k = lane; k < n; k += 64) red[k] = v(k + 1, j, l, t);
        __syncthreads();
        for (int i = lane; i < n * c; i += 64) {
            const int k = i / c;
            const double m_new = (mu(i + 1, j, l) + au(i, j, l)) - au(i + 1, j, l);
            v(i + 1, j, l, t) = (m_new / mred[k]) * red[k];
        }
        __syncthreads();
    }
    for (int i = lane; i < R.im; i += 64) R.m(i + 1, j, l) = (mu(i + 1, j, l) + au(i, j, l)) - au(i + 1, j, l);
    for (int i = lane; i <= R.im + 1; i += 64) R.am(i, j, l) = au(i, j, l);
    __syncthreads();
    if (tracers) {
        if (n == R.im) return;
        for (int f = lane; f < 1 + nq; f += 64) {
            const HV &v = f == 0 ? R.m : R.q[(f - 1) % 4];
            const int t = f == 0 ? 1 : (f - 1) / 4 + 1;
            v(0, j, l, t) = v(R.im, j, l, t);
        }
    } else if (lane == 0 && j >= js && j <= je) {
        R.m(0, j, l) = R.m(R.im, j, l);
        if (n == R.im) R.m(R.im + 1, j, l) = R.m(1, j, l);
    }
}

// ------------------------------------------------------------------ launchers
static int nfields(const ZR &R, int tracers) { return 1 + (tracers ? 4 * R.nt : 0); }

int zoom_put_x(tm5_ctx *ctx, const ZR &P, const ZR &C, int tracers)
{
    if (C.whole) return 0;
    const int nf = nfields(C, tracers);
    TM5_LAUNCH(ctx, k_put_x, zgrid((long long)C.jm * C.lm * nf), dim3(256), 0, ctx->stream, P, C, nf, ctx->words,
               zm(ctx, TM5_ERRORS | TM5_REHEARSAL * !tracers));
    return status(hipGetLastError());
}

int zoom_put_y(tm5_ctx *ctx, const ZR &P, const ZR &C, int tracers)
{
    const int nf = nfields(C, tracers);
    TM5_LAUNCH(ctx, k_put_y, zgrid((long long)C.im * C.lm * nf), dim3(256), 0, ctx->stream, P, C, nf, ctx->words,
               zm(ctx, TM5_ERRORS | TM5_REHEARSAL * !tracers));
    return status(hipGetLastError());
}

int zoom_compress(tm5_ctx *ctx, const ZR &C, int axis, int a0, int a1, int l0, int l1, int tracers, bool un)
{
    const int nf = tracers ? 1 + C.nt : 1;
    const long long n = (long long)(a1 - a0 + 1) * (l1 - l0 + 1) * nf;
    if (n <= 0) return 0;
    const int mask = zm(ctx, TM5_ERRORS | TM5_REHEARSAL * !tracers);
    if (un)
        TM5_LAUNCH(ctx, k_uncompress, zgrid(n), dim3(256), 0, ctx->stream, C, axis, a0, a1, l0, l1, nf, ctx->words, mask);
    else
        TM5_LAUNCH(ctx, k_compress, zgrid(n), dim3(256), 0, ctx->stream, C, axis, a0, a1, l0, l1, nf, ctx->words, mask);
    return status(hipGetLastError());
}

int zoom_mix_edges(tm5_ctx *ctx, const ZR &C, int xedges, int tracers)
{
    const int nf = nfields(C, tracers);
    const RingBlocks ring = ring_blocks(C, xedges);
    const long long n = (long long)ring.count() * C.lm * nf;
    if (n == 0) return 0;
    const size_t lds = (size_t)ring.count() * (C.xr * C.yr + 1) * sizeof(double);
    if (nf >= 32 && lds <= 65536)  // many fields: a workgroup per (level, field)
        TM5_LAUNCH(ctx, k_mix_edges_lds, dim3((unsigned)(C.lm * nf)), dim3(256), lds, ctx->stream, C, ring, ctx->words,
                   zm(ctx, TM5_ERRORS | TM5_REHEARSAL * !tracers));
    else
        TM5_LAUNCH(ctx, k_mix_edges, zgrid(n), dim3(256), 0, ctx->stream, C, ring, nf, ctx->words,
                   zm(ctx, TM5_ERRORS | TM5_REHEARSAL * !tracers));
    return status(hipGetLastError());
}

int zoom_update_parent(tm5_ctx *ctx, const ZR &P, const ZR &C, int tracers)
{
    const int nf = tracers ? 1 + C.nt : 1;
    const long long n = (long long)(C.im / C.xr) * (C.jm / C.yr) * C.lm * nf;
    TM5_LAUNCH(ctx, k_update_parent, zgrid(n), dim3(256), 0, ctx->stream, P, C, nf, ctx->words,
               zm(ctx, TM5_ERRORS | TM5_REHEARSAL * !tracers));
    return status(hipGetLastError());
}

// the tile of the m2phlb kernels (the context's m2phlb1 tile)
static int m2phlb_ti(const tm5_ctx *ctx, int lm)
{
    int TI = ctx->tile[TM5_TILE_M2PHLB];
    while (TI > 1 && (lm + 1) * TI > 8192) TI /= 2;
    return TI;
}

// update_parent (with tracers) and the child's m2phlb in one launch
int zoom_update_m2phlb(tm5_ctx *ctx, const ZR &P, const ZR &C, const double *at, const double *bt, int skip_mask)
{
    const int nf = 1 + C.nt;
    const long long nu = (long long)(C.im / C.xr) * (C.jm / C.yr) * C.lm * nf;
    const int TI = m2phlb_ti(ctx, C.lm), nub = (int)((nu + 255) / 256);
    const long long tiles = (long long)((C.im + TI - 1) / TI) * C.jm;
    TM5_LAUNCH(ctx, k_update_m2phlb, dim3((unsigned)(nub + tiles)), dim3(256), (size_t)(C.lm + 1) * TI * sizeof(double),
               ctx->stream, P, C, nf, nub, at, bt, TI, ctx->words, zm(ctx, skip_mask));
    return status(hipGetLastError());
}

int zoom_coarsen(tm5_ctx *ctx, const ZR &P, const ZR &C, int skip_mask)
{
    const long long imp = C.im / C.xr, jmp = C.jm / C.yr;
    const long long n = ((imp + 1) * jmp + imp * (jmp + 1) + 2 * imp * jmp) * C.lm;
    TM5_LAUNCH(ctx, k_coarsen, zgrid(n), dim3(256), 0, ctx->stream, P, C, ctx->words, zm(ctx, skip_mask));
    return status(hipGetLastError());
}

int zoom_coarsen_ambm(tm5_ctx *ctx, const ZR &P, const ZR &C, int skip_mask)
{
    const long long n = (long long)(C.jm + C.im) * C.lm;
    TM5_LAUNCH(ctx, k_coarsen_ambm, zgrid(n), dim3(256), 0, ctx->stream, P, C, ctx->words, zm(ctx, skip_mask));
    return status(hipGetLastError());
}

int zoom_m2phlb(tm5_ctx *ctx, const ZR &R, const double *pdxyp, const double *at, const double *bt, int skip_mask,
                int clock_n)
{
    const int TI = m2phlb_ti(ctx, R.lm);
    TM5_LAUNCH(ctx, k_m2phlb_zoom, dim3((unsigned)((R.im + TI - 1) / TI), (unsigned)R.jm), dim3(256),
               (size_t)(R.lm + 1) * TI * sizeof(double), ctx->stream, R, pdxyp, at, bt, TI, ctx->words,
               zm(ctx, skip_mask), clock_n);
    return status(hipGetLastError());
}

int zoom_x_halo(tm5_ctx *ctx, const ZR &R, int js, int je, int tracers, int skip_mask)
{
    if (je < js) return 0;
    const int nf = nfields(R, tracers);
    TM5_LAUNCH(ctx, k_x_halo, zgrid((long long)(je - js + 1) * R.lm * nf), dim3(256), 0, ctx->stream, R, js, je, nf,
               ctx->words, zm(ctx, skip_mask));
    return status(hipGetLastError());
}

int zoom_slab(tm5_ctx *ctx, const HV &f, int axis, int k, int n0, int n1, double *save, int restore, int skip_mask)
{
    TM5_LAUNCH(ctx, k_slab, zgrid((long long)n0 * n1), dim3(256), 0, ctx->stream, f, axis, k, n0, n1, save, restore,
               ctx->words, zm(ctx, skip_mask));
    return status(hipGetLastError());
}

int zoom_copy(tm5_ctx *ctx, const double *src, double *dst, long long n)
{
    TM5_LAUNCH(ctx, k_copy_flat, zgrid(n), dim3(256), 0, ctx->stream, src, dst, n);
    return status(hipGetLastError());
}

int zoom_scale(tm5_ctx *ctx, double *a, long long n, double factor)
{
    TM5_LAUNCH(ctx, k_scale_flat, zgrid(n), dim3(256), 0, ctx->stream, a, n, factor);
    return status(hipGetLastError());
}

int zoom_am_halo(tm5_ctx *ctx, const ZR &R, int skip_mask)
{
    TM5_LAUNCH(ctx, k_am_halo, zgrid((long long)(R.jm + 2) * (R.lm + 2)), dim3(256), 0, ctx->stream, R, ctx->words,
               zm(ctx, skip_mask));
    return status(hipGetLastError());
}

// spec_k: a checked step's x (the search capped at nmax, a failure spec_fail); FUSE: X_COMPRESS,
// X_OUTER (the compress / uncompress and the outer faces' zeroing made inside the kernel)
template <int FUSE>
static int x_child_f(tm5_ctx *ctx, const ZR &C, int js, int je, int ls, int le, int limits, int nmax, int mass_only,
                     int fail_bit, int skip_mask, int spec_k)
{
    const int n = C.im - 2 * C.xr + 4;
    if (n < 3 || n > 512 || je < js || le < ls) return TM5_EINVAL;  // LDS: 11 n doubles
    const size_t lds = (size_t)(2 * n + (n - 1) + 4 * n + 4 * (n - 1)) * sizeof(double);
    const dim3 grid((unsigned)((je - js + 1) * (le - ls + 1)));
    int ranges = 1, tchunk = C.nt;
    if (!mass_only && C.nt >= 32) {  // many tracers: tracer ranges over the grid, then the masses
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_x_child<false, FUSE>, XC_THREADS, lds) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        tchunk = tracer_chunk(grid.x, (long long)per_cu * (ctx->ncu > 0 ? ctx->ncu : 256), C.nt,
                              ctx->tile[TM5_TILE_RANGE_MIN]);
        ranges = (C.nt + tchunk - 1) / tchunk;
    }
    if (!mass_only)
        TM5_LAUNCH_NAMED(ctx, "k_x_child<false>", (k_x_child<false, FUSE>), dim3(grid.x, ranges), dim3(XC_THREADS), lds,
                         ctx->stream, C, js, je, ls, le, limits, nmax, fail_bit, ctx->words, zm(ctx, skip_mask),
                         ranges > 1 ? tchunk : 0, spec_k);
    if (mass_only || ranges > 1)
        TM5_LAUNCH_NAMED(ctx, "k_x_child<true>", (k_x_child<true, FUSE>), grid, dim3(XC_THREADS), lds, ctx->stream, C, js, je,
                         ls, le, 0, nmax, fail_bit, ctx->words, zm(ctx, skip_mask), 0, spec_k);
    return status(hipGetLastError());
}

int zoom_x_child(tm5_ctx *ctx, const ZR &C, int js, int je, int ls, int le, int limits, int nmax, int mass_only,
                 int fail_bit, int skip_mask, int spec_k, int fuse)
{
    switch (fuse) {
    case 0: return x_child_f<0>(ctx, C, js, je, ls, le, limits, nmax, mass_only, fail_bit, skip_mask, spec_k);
    case 1: return x_child_f<1>(ctx, C, js, je, ls, le, limits, nmax, mass_only, fail_bit, skip_mask, spec_k);
    case 2: return x_child_f<2>(ctx, C, js, je, ls, le, limits, nmax, mass_only, fail_bit, skip_mask, spec_k);
    case 3: return x_child_f<3>(ctx, C, js, je, ls, le, limits, nmax, mass_only, fail_bit, skip_mask, spec_k);
    default: return TM5_EINVAL;
    }
}

// known_one: every layer at nloop 1 without a search (Check_CFL's rehearsal at cap 1 passed on
// the same bits); checked(spec_k): a checked step, whose sub-step kernel makes that test itself;
// fuse (known_one only): Y_COMPRESS, Y_OUTER made inside the sub-step kernel
int zoom_y_child(tm5_ctx *ctx, const ZR &C, int is, int ie, int ls, int le, int limits, int nmax, int mass_only,
                 int fail_bit, int skip_mask, bool known_one, int spec_k, int fuse)
{
    if (checked(spec_k)) known_one = true;
    if (fuse && !known_one) return TM5_EINVAL;  // a search reads the faces and masses from memory
    const int n = C.jm - 2 * C.yr + 4;
    const int nlay = le - ls + 1;
    if (n < 3 || n > 1024 || ie < is || nlay < 1 || nlay > ctx->y_cap) return TM5_EINVAL;
    int *nloop = known_one ? nullptr : ctx->y_nloop;
    if (!known_one) {
        if (mass_only)
            TM5_LAUNCH(ctx, k_y_child_search<true>, dim3(nlay), dim3(1024), 0, ctx->stream, C, is, ie, ls, nmax, nloop,
                       fail_bit, ctx->words, zm(ctx, skip_mask));
        else
            TM5_LAUNCH(ctx, k_y_child_search<false>, dim3(nlay), dim3(1024), 0, ctx->stream, C, is, ie, ls, nmax,
                       nloop, fail_bit, ctx->words, zm(ctx, skip_mask));
    }
    const int cap = C.nt >= 8 ? ctx->tile[TM5_TILE_Y_MANY] : 16;
    const int TI = (1024 / n) < cap ? (1024 / n) : cap;
    const int ni = ie - is + 1;
    const dim3 grid((unsigned)((ni + TI - 1) / TI), (unsigned)nlay), block((unsigned)((TI * n + 63) / 64 * 64));
    // many tracers: tracer ranges over the grid (one workgroup of these a CU), then the masses
    const int tchunk = !mass_only && C.nt >= 32
                           ? tracer_chunk((long long)grid.x * grid.y, ctx->ncu > 0 ? ctx->ncu : 256, C.nt,
                                          ctx->tile[TM5_TILE_RANGE_MIN])
                           : C.nt;
    const int ranges = C.nt > 0 ? (C.nt + tchunk - 1) / tchunk : 1;
    auto steps = [&](auto fz) {
        constexpr int F = decltype(fz)::value;
        if (!mass_only)
            TM5_LAUNCH_NAMED(ctx, "k_y_child_steps<false>", (k_y_child_steps<false, F>), dim3(grid.x, grid.y, ranges),
                             block, (size_t)5 * n * TI * sizeof(double), ctx->stream, C, is, ie, ls, TI, nloop, limits,
                             ctx->words, zm(ctx, skip_mask), ranges > 1 ? tchunk : 0, spec_k);
        if (mass_only || ranges > 1)
            TM5_LAUNCH_NAMED(ctx, "k_y_child_steps<true>", (k_y_child_steps<true, F>), grid, block, 0, ctx->stream, C,
                             is, ie, ls, TI, nloop, 0, ctx->words, zm(ctx, skip_mask), 0,
                             mass_only ? spec_k : SPEC_NONE);
    };
    switch (fuse) {
    case 0: steps(std::integral_constant<int, 0>{}); break;
    case 1: steps(std::integral_constant<int, 1>{}); break;
    case 2: steps(std::integral_constant<int, 2>{}); break;
    case 3: steps(std::integral_constant<int, 3>{}); break;
    default: return TM5_EINVAL;
    }
    if (known_one) return status(hipGetLastError());  // times 1 is the identity: nothing to restore
    const int n0 = C.im + 2, n1 = C.jm + 2;
    TM5_LAUNCH(ctx, k_y_child_restore, dim3((unsigned)((n0 * n1 + 255) / 256), (unsigned)nlay), dim3(256), 0,
               ctx->stream, C, ls, nloop, ctx->words, zm(ctx, skip_mask));
    return status(hipGetLastError());
}

}  // namespace tm5

// ------------------------------------------------------------------ the C ABI (include/tm5hip.h)
using namespace tm5;

#define ZTRY(x)                   \
    do {                          \
        int rc_ = (x);            \
        if (rc_ != 0) return rc_; \
    } while (0)

// Region 1's dynamu on its reduced grid (pydyn/zoom/advect.py:253-297, pydyn/cfl.py:30-57): the
// uniform m and am kept, the clusters summed and the faces subsampled, every row advected with
// its own cell count, the halos, the clusters split with the uniformly advanced mass, am put back.
// spec_k: a checked step's (the search capped at max_nloop - 1, a failure spec_fail)
// Region 1's dynamu on its reduced grid (pydyn/zoom/advect.py:237-277): the reduced rows to their
// clusters, dynamu on every row (a reduced row with its imred cells; the periodic halo of rows
// js..je written by the x kernel, which puts am back as it read it), the reduced rows split back.
int zoom_x_reduced(tm5_ctx *c, const tm5_zregion &d, int js, int je, int limits, int max_nloop, int mass_only, int fail,
                   int skip, int spec_k)
{
    const ZR R = zr(d);
    if (!d.jred || !d.imred || !d.scratch || R.im > 4096) return TM5_EINVAL;
    const long long nm = (long long)(R.im + 4) * (R.jm + 4) * R.lm;
    HV mu = R.m, au = R.am;  // the reduced rows' uniform m and am, laid out as m and am
    mu.p = d.scratch;
    au.p = d.scratch + nm;
    const dim3 rows((unsigned)R.nred, (unsigned)R.lm);
    const int tracers = mass_only ? 0 : 1;
    TM5_LAUNCH(c, k_uni2red, rows, dim3(64), (size_t)(R.im + 2) * sizeof(double), c->stream, R, mu, au, tracers,
               c->words, skip);
    ZTRY(launch_x(c, R.im, R.jm, R.lm, sub3(R.m, 1, 1, 1), mass_only ? no_tracers() : subq(R, 1, 1, 1),
                  sub3(R.am, 0, 1, 1), mass_only ? 0 : limits, max_nloop - 1, fail, skip, R.imred, nullptr, nullptr,
                  spec_k, nullptr, js - 1, je - 1, nullptr, 1));
    TM5_LAUNCH(c, k_red2uni, rows, dim3(64), (size_t)2 * R.im * sizeof(double), c->stream, R, mu, au, tracers, js, je,
               c->words, skip);
    return status(hipGetLastError());
}

static bool zbad(const tm5_zregion *r) { return r == nullptr || r->m.data == nullptr || r->im < 1 || r->jm < 1 || r->lm < 1; }
static bool zglobal(const tm5_zregion *r) { return r->xcyc == 1 && r->touch_sp == 1 && r->touch_np == 1; }
static int zskipmask(const tm5_ctx *c, int rehearsal)
{
    return zm(c, rehearsal ? (TM5_ERRORS | TM5_REHEARSAL) : TM5_ERRORS);
}

extern "C" int tm5_zoom_put_xedges(tm5_ctx *c, const tm5_zregion *p, const tm5_zregion *ch, int tracers)
{
    if (!c || zbad(p) || zbad(ch)) return TM5_EINVAL;
    return zoom_put_x(c, zr(*p), zr(*ch), tracers);
}

extern "C" int tm5_zoom_put_yedges(tm5_ctx *c, const tm5_zregion *p, const tm5_zregion *ch, int tracers)
{
    if (!c || zbad(p) || zbad(ch)) return TM5_EINVAL;
    return zoom_put_y(c, zr(*p), zr(*ch), tracers);
}

static bool xedges_apply(const tm5_zregion *r) { return !(r->xref == 1 || r->whole); }
static bool yedges_apply(const tm5_zregion *r) { return !(r->yref == 1 || zglobal(r)); }

extern "C" int tm5_zoom_compress_xedges(tm5_ctx *c, const tm5_zregion *r, int js, int je, int ls, int le, int tracers)
{
    if (!c || zbad(r)) return TM5_EINVAL;
    return xedges_apply(r) ? zoom_compress(c, zr(*r), 0, js, je, ls, le, tracers, false) : 0;
}

extern "C" int tm5_zoom_uncompress_xedges(tm5_ctx *c, const tm5_zregion *r, int js, int je, int ls, int le, int tracers)
{
    if (!c || zbad(r)) return TM5_EINVAL;
    return xedges_apply(r) ? zoom_compress(c, zr(*r), 0, js, je, ls, le, tracers, true) : 0;
}

extern "C" int tm5_zoom_compress_yedges(tm5_ctx *c, const tm5_zregion *r, int is, int ie, int ls, int le, int tracers)
{
    if (!c || zbad(r)) return TM5_EINVAL;
    return yedges_apply(r) ? zoom_compress(c, zr(*r), 1, is, ie, ls, le, tracers, false) : 0;
}

extern "C" int tm5_zoom_uncompress_yedges(tm5_ctx *c, const tm5_zregion *r, int is, int ie, int ls, int le, int tracers)
{
    if (!c || zbad(r)) return TM5_EINVAL;
    return yedges_apply(r) ? zoom_compress(c, zr(*r), 1, is, ie, ls, le, tracers, true) : 0;
}

extern "C" int tm5_zoom_mix_edges(tm5_ctx *c, const tm5_zregion *r, int tracers)
{
    if (!c || zbad(r) || !r->edge) return TM5_EINVAL;
    if (r->zref != 1) return TM5_EINVAL;  // "mix_edges: zooming in z direction not supported!"
    const bool xedges = xedges_apply(r);
    if (!xedges && zglobal(r)) return 0;
    return zoom_mix_edges(c, zr(*r), xedges ? 1 : 0, tracers);
}

extern "C" int tm5_zoom_update_parent(tm5_ctx *c, const tm5_zregion *p, const tm5_zregion *ch, int tracers)
{
    if (!c || zbad(p) || zbad(ch) || ch->zref != 1) return TM5_EINVAL;
    return zoom_update_parent(c, zr(*p), zr(*ch), tracers);
}

extern "C" int tm5_zoom_coarsen_block(tm5_ctx *c, const tm5_zregion *p, const tm5_zregion *ch)
{
    if (!c || zbad(p) || zbad(ch) || ch->zref != 1) return TM5_EINVAL;
    return zoom_coarsen(c, zr(*p), zr(*ch), zm(c, TM5_ERRORS));
}

extern "C" int tm5_zoom_coarsen_ambm(tm5_ctx *c, const tm5_zregion *p, const tm5_zregion *ch)
{
    if (!c || zbad(p) || zbad(ch)) return TM5_EINVAL;
    return zoom_coarsen_ambm(c, zr(*p), zr(*ch), zm(c, TM5_ERRORS));
}

extern "C" int tm5_zoom_m2phlb(tm5_ctx *c, const tm5_zregion *r, const tm5_zregion *p, const double *at,
                               const double *bt)
{
    if (!c || zbad(r) || !r->dxyp || !at || !bt || (p && !p->dxyp)) return TM5_EINVAL;
    return zoom_m2phlb(c, zr(*r), p ? p->dxyp : nullptr, at, bt, zm(c, TM5_ERRORS), 0);
}

extern "C" int tm5_zoom_dynam0(tm5_ctx *c, const tm5_zregion *r, const double *bt, int ndyn, int rehearsal)
{
    if (!c || zbad(r) || !bt || r->tref_abs < 1) return TM5_EINVAL;
    const ZR R = zr(*r);
    const double dt = (double)ndyn / (2.0 * (double)r->tref_abs);
    const int skip = zskipmask(c, rehearsal);
    ZTRY(launch_dynam0(c, R.im, R.jm, R.lm, sub3(R.pu, 0, 1, 1), sub3(R.pv, 1, 1, 1), 1.0, bt, dt, sub3(R.am, 0, 1, 1),
                       sub3(R.bm, 1, 1, 1), sub3(R.cm, 1, 1, 0), skip, R.xcyc));
    if (R.xcyc == 1) ZTRY(zoom_am_halo(c, R, skip));
    return 0;
}

extern "C" int tm5_zoom_outer_faces(tm5_ctx *c, const tm5_zregion *r, int axis, int restore, double *save,
                                    int rehearsal)
{
    if (!c || zbad(r) || !save || axis < 0 || axis > 1) return TM5_EINVAL;
    const ZR R = zr(*r);
    const int skip = zskipmask(c, rehearsal);
    if (axis == 0) {
        const int n0 = R.jm + 2, n1 = R.lm + 2;
        ZTRY(zoom_slab(c, R.am, 0, R.xr - 1, n0, n1, save, restore, skip));
        return zoom_slab(c, R.am, 0, R.im - R.xr + 1, n0, n1, save + (long long)n0 * n1, restore, skip);
    }
    const int n0 = R.im + 2, n1 = R.lm + 2;
    if (R.touch_sp != 1) ZTRY(zoom_slab(c, R.bm, 1, R.yr, n0, n1, save, restore, skip));
    if (R.touch_np != 1) ZTRY(zoom_slab(c, R.bm, 1, R.jm - R.yr + 2, n0, n1, save + (long long)n0 * n1, restore, skip));
    return 0;
}

extern "C" int tm5_zoom_advect_x(tm5_ctx *c, const tm5_zregion *r, int js, int je, int limits, int max_nloop,
                                 int mass_only)
{
    if (!c || zbad(r) || max_nloop < 1) return TM5_EINVAL;
    const ZR R = zr(*r);
    const int skip = zskipmask(c, mass_only);
    if (zglobal(r)) {
        const int fail = mass_only ? TM5_REHEARSAL : TM5_FAIL_X;
        if (r->nred == 0) {
            ZTRY(launch_x(c, R.im, R.jm, R.lm, sub3(R.m, 1, 1, 1), mass_only ? no_tracers() : subq(R, 1, 1, 1),
                          sub3(R.am, 0, 1, 1), mass_only ? 0 : limits, max_nloop - 1, fail, skip));
            return zoom_x_halo(c, R, js, je, !mass_only, skip);
        }
        return zoom_x_reduced(c, *r, js, je, limits, max_nloop, mass_only, fail, skip, SPEC_NONE);
    }
    const int tr = mass_only ? 0 : 1;
    ZTRY(tm5_zoom_compress_xedges(c, r, js, je, 1, R.lm, tr));
    ZTRY(zoom_x_child(c, R, js, je, 1, R.lm, limits, max_nloop - 1, mass_only, mass_only ? TM5_REHEARSAL : TM5_FAIL_XC,
                      skip, SPEC_NONE, 0));
    if (mass_only) {  // dynamum's own uncompressm and mixm_edges (pydyn/zoom/check_cfl.py:143-144)
        ZTRY(tm5_zoom_uncompress_xedges(c, r, js, je, 1, R.lm, 0));
        ZTRY(tm5_zoom_mix_edges(c, r, 0));
    }
    return 0;
}

extern "C" int tm5_zoom_advect_y(tm5_ctx *c, const tm5_zregion *r, int is, int ie, int limits, int max_nloop,
                                 int mass_only)
{
    if (!c || zbad(r) || max_nloop < 1) return TM5_EINVAL;
    const ZR R = zr(*r);
    const int skip = zskipmask(c, mass_only);
    if (zglobal(r)) {
        const int fail = mass_only ? TM5_REHEARSAL : TM5_FAIL_Y;
        return launch_y(c, R.im, R.jm, R.lm, sub3(R.m, 1, 1, 1), mass_only ? no_tracers() : subq(R, 1, 1, 1),
                        sub3(R.bm, 1, 1, 1), mass_only ? 0 : limits, mass_only ? max_nloop : max_nloop - 1, fail, skip);
    }
    const int tr = mass_only ? 0 : 1;
    ZTRY(tm5_zoom_compress_yedges(c, r, is, ie, 1, R.lm, tr));
    ZTRY(zoom_y_child(c, R, is, ie, 1, R.lm, limits, mass_only ? max_nloop : max_nloop - 1, mass_only,
                      mass_only ? TM5_REHEARSAL : TM5_FAIL_YC, skip, false, SPEC_NONE, 0));
    if (mass_only) {
        ZTRY(tm5_zoom_uncompress_yedges(c, r, is, ie, 1, R.lm, 0));
        ZTRY(tm5_zoom_mix_edges(c, r, 0));
    }
    return 0;
}

extern "C" int tm5_zoom_advect_z(tm5_ctx *c, const tm5_zregion *r, int is, int ie, int js, int je, int limits,
                                 int iter_limit, int mass_only)
{
    if (!c || zbad(r) || iter_limit < 1) return TM5_EINVAL;
    const ZR R = zr(*r);
    const int fail = mass_only ? TM5_REHEARSAL : TM5_FAIL_Z;
    return launch_z_scoped(c, R.im, R.jm, R.lm, sub3(R.m, 1, 1, 1), mass_only ? no_tracers() : subq(R, 1, 1, 1),
                           sub3(R.cm, 1, 1, 0), is - 1, ie - 1, js - 1, je - 1, mass_only ? 0 : limits, iter_limit,
                           fail, zskipmask(c, mass_only));
}

// ------------------------------------------------------------------ sources and chemistry alone
// sources_region's do_add_2d (pydyn/zoom/physics.py:55-64): rm + x, rzm - x at (level, tracer)
// (1-based) on the cells the region owns; x (im, jm).
__global__ void __launch_bounds__(256) k_zoom_inject(ZR R, View3 x, int level, int tracer, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)R.im * R.jm) return;
    const int i = (int)(idx % R.im), j = (int)(idx / R.im);
    if (!R.owned[(long long)j * R.im + i]) return;
    const double v = x(i, j, 0);
    double &rm = R.q[0](i + 1, j + 1, level, tracer);
    double &rz = R.q[3](i + 1, j + 1, level, tracer);
    rm = rm + v;
    rz = rz - v;
}

// chemistry_region (pydyn/zoom/physics.py:67-78): rm and the slopes of tracers t0..t0+count-1
// (1-based) times the factor, decay (the scope's, (ier-isr+1, jer-jsr+1, lm)) on the owned cells
// and 1.0 elsewhere.
__global__ void __launch_bounds__(256) k_zoom_decay(ZR R, View3 decay, int t0, int count, int *words, int skip_mask)
{
    if (zskip(words, skip_mask)) return;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)R.im * R.jm * R.lm) return;
    const int i = (int)(idx % R.im), j = (int)((idx / R.im) % R.jm), l = (int)(idx / ((long long)R.im * R.jm));
    const int fi = i + 1, fj = j + 1;
    double f = 1.0;
    if (fi >= R.isr && fi <= R.ier && fj >= R.jsr && fj <= R.jer) f = decay(fi - R.isr, fj - R.jsr, l);
    if (!R.owned[(long long)j * R.im + i]) f = 1.0;
    for (int t = t0; t < t0 + count; t++)
        for (int a = 0; a < 4; a++) {
            double &q = R.q[a](fi, fj, l + 1, t);
            q = q * f;
        }
}

extern "C" int tm5_zoom_inject(tm5_ctx *c, const tm5_zregion *r, const tm5_view *x, int level, int tracer)
{
    if (!c || zbad(r) || !x || !x->data || !r->owned || level < 1 || level > r->lm || tracer < 1 || tracer > r->nt)
        return TM5_EINVAL;
    const ZR R = zr(*r);
    TM5_LAUNCH(c, k_zoom_inject, zgrid((long long)R.im * R.jm), dim3(256), 0, c->stream, R, view3(x), level, tracer,
               c->words, zm(c, TM5_ERRORS));
    return status(hipGetLastError());
}

extern "C" int tm5_zoom_decay(tm5_ctx *c, const tm5_zregion *r, const tm5_view *decay, int t0, int count)
{
    if (!c || zbad(r) || !decay || !decay->data || !r->owned || t0 < 1 || count < 0 || t0 + count - 1 > r->nt)
        return TM5_EINVAL;
    if (count == 0) return 0;
    const ZR R = zr(*r);
    TM5_LAUNCH(c, k_zoom_decay, zgrid((long long)R.im * R.jm * R.lm), dim3(256), 0, c->stream, R, view3(decay), t0,
               count, c->words, zm(c, TM5_ERRORS));
    return status(hipGetLastError());
}

// The column pass of a region; zp (non-null): with the step's next dynamw fused into it (the split
// form's zpost) on cm over the scope zp->i0..i1, j0..j1 (0-based), limits and checked step spec_k.
int tm5::zoom_physics(tm5_ctx *c, const tm5_zregion *r, const tm5_zphysics *p, const char *ops, const ZPost *zp)
{
    if (!c || zbad(r) || !p || !ops || !column_supported(p->lmax) || r->lm > 64 || p->lmax > r->lm) return TM5_EINVAL;
    ColumnArgs a{};
    if (!strcmp(ops, "v")) {
    } else if (!strcmp(ops, "csv")) {
        a.pre = 1;
    } else if (!strcmp(ops, "vsc")) {
        a.post = 1;
    } else if (!strcmp(ops, "vsccsv")) {
        a.twice = 1;
    } else if (!strcmp(ops, "vscMcsv")) {
        a.twice = 1;
        a.mid = 1;
    } else {
        return TM5_EINVAL;
    }
    const ZR R = zr(*r);
    a.im = R.im;
    a.jm = R.jm;
    a.lm = R.lm;
    a.m = sub3(R.m, 1, 1, 1);
    a.q = subq(R, 1, 1, 1);
    a.entu = view3(&p->entu);
    a.detu = view3(&p->detu);
    a.entd = view3(&p->entd);
    a.detd = view3(&p->detd);
    a.dkg = view3(&p->dkg);
    a.top = iview2(&p->cloud_top);
    a.lfs = iview2(&p->cloud_lfs);
    a.dt = p->dt_conv;
    a.owned = r->owned;
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
    a.dxyp = r->dxyp;
    a.at = p->at;
    a.bt = p->bt;
    a.sp = sub3(R.sp, 1, 1, 1);
    a.phlb = sub3(R.phlb, 1, 1, 1);
    if ((a.pre || a.post || a.twice) && (!p->emis.data || !p->loss.data)) return TM5_EINVAL;
    if (a.mid && (!p->at || !p->bt || !r->dxyp)) return TM5_EINVAL;
    if (zp && (zp->post || zp->pre)) {
        a.cm = sub3(R.cm, 1, 1, 0);
        a.zlimits = zp->limits;
        a.spec_k = zp->spec_k;
        a.zpost = zp->post;
        a.zi0 = zp->i0, a.zi1 = zp->i1, a.zj0 = zp->j0, a.zj1 = zp->j1;
        a.zpre = zp->pre;
        a.zpi0 = zp->pi0, a.zpi1 = zp->pi1, a.zpj0 = zp->pj0, a.zpj1 = zp->pj1;
        a.zhi0 = zp->hi0, a.zhi1 = zp->hi1, a.zhj0 = zp->hj0, a.zhj1 = zp->hj1;
    }
    return launch_column(c, a, p->lmax, zm(c, TM5_ERRORS));
}

extern "C" int tm5_zoom_physics(tm5_ctx *c, const tm5_zregion *r, const tm5_zphysics *p, const char *ops)
{
    return tm5::zoom_physics(c, r, p, ops, nullptr);
}

extern "C" int tm5_zoom_copy(tm5_ctx *c, const double *src, double *dst, int64_t n)
{
    if (!c || !src || !dst || n < 0) return TM5_EINVAL;
    return n ? zoom_copy(c, src, dst, n) : 0;
}

extern "C" int tm5_zoom_scale(tm5_ctx *c, double *a, int64_t n, double factor)
{
    if (!c || !a || n < 0) return TM5_EINVAL;
    return n ? zoom_scale(c, a, n, factor) : 0;
}
