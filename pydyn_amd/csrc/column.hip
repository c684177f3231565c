// Column physics in registers: convection (convec) and the fused middle of a dynamic step.
//
// Between the two 'v' operators of a step ("xyz v s c | m2phlb1 | c s v zyx") only the sources,
// the chemistry and m2phlb1 run; none of them changes the air mass, so both convection calls
// build the same matrix and invert it to the same bits.  k_column<N> builds conv1 once per
// column, inverts it, and applies one of the sequences "v", "csv", "vsc", "vsccsv" or
// "vscMcsv" (M = m2phlb1) to every tracer of the column in one pass; in a zoom tree only on the
// columns the region owns.
//
// Two columns per 64-lane wavefront: lanes 0-31 hold one column, lanes 32-63 the next (half
// h = lane / 32, lane li = lane % 32 of the half), N = lmax known at compile time so that the
// matrix lives in registers:
//   build      lane c owns matrix column c (the up/downdraft recursions, pydyn/kernels.py:350-396),
//              in registers, then written to the half's LDS matrix;
//   LU         lane r holds row r (read back from LDS); right-looking Crout without pivoting,
//              each element updated in increasing k as in the reference's left-looking loops
//              (pydyn/kernels.py:404-407);
//   inverse    lane c solves for the right-hand side e_c (forward, then back substitution in the
//              reference's order, pydyn/kernels.py:409-419);
//   apply      lane l holds row l of C: rm(l) = sum_k C(l,k) rm(k) from zero in increasing k,
//              slopes times C(l,l) (pydyn/kernels.py:448-457).
// A value uniform over a column (a pivot row, an L or U entry, the column's rm) is written to the
// half's LDS once and read back by every lane of the half: one LDS read serves both halves.
// Levels 32 and up (lm <= 64) sit in a second register set of the lane, where only the sources
// and the chemistry act (N < 32).
//
// The matrix of a column with cloud top T and level of free sinking F is dense in its leading
// DB x DB block (DB = max(T, F) + 1) and tridiagonal outside it; the entries outside are exact
// zeros (the reference's where(act, .., 0.0) and 0 / m).  LU fill-in stays inside that envelope,
// so the products outside it all have a +-0 factor and are skipped: subtracting +-0 leaves
// every non-zero value as it is and can only change the sign of a zero, and no zero's sign
// reaches a result (zeros only ever enter products, sums with non-zero terms, or numerators),
// so the outputs are the reference's bits for any non-singular matrix of finite entries.  A
// wave works to the larger envelope of its two columns: the other column's extra products have
// a +-0 factor too.
#include "tm5_lib.h"

namespace tm5 {

__device__ __forceinline__ bool skipped_k(const int *words, int mask)
{
    return (words[W_FLAGS] & mask) != 0;  // written by earlier kernels of the stream
}

// Lanes of one wavefront exchange values through LDS: the wavefront's LDS operations are
// performed in program order, the wait lets its writes land before the reads that follow, and
// the fences keep the compiler from moving LDS accesses across this point.
__device__ __forceinline__ void wsync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A uniform value the compiler must take as redefined here: comparisons with it are then made
// where they are used, as scalar compares, instead of being hoisted and kept as lane masks.
__device__ __forceinline__ int fresh(int v)
{
    asm volatile("" : "+s"(v));
    return v;
}

// lane k's value of each 32-lane half, to every lane of that half (ds_swizzle in bit-mask mode:
// and-mask 0, or-mask k; no LDS memory is touched)
template <int K> __device__ __forceinline__ double half_bcast(double v)
{
    static_assert(K >= 0 && K < 32, "a lane of the half");
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffLL), K << 5);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), K << 5);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// the larger of a value over the two halves of the wavefront (uniform)
__device__ __forceinline__ int wave_max2(int v)
{
    const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 32);
    return a > b ? a : b;
}

template <int N> struct ColumnLds {
    static constexpr int LD = (N % 2 == 0) ? N + 1 : N;  // odd: a row read by every lane spreads over the banks
    static constexpr int BUILD = 15 * 32;                  // the build's per-level vectors (share the matrix's space)
    static constexpr int ZSTAGE = 5 * 64 + 4 * 65;          // the fused dynamw's staging and face fluxes
    static constexpr int R1 = BUILD > ZSTAGE ? BUILD : ZSTAGE;
    static constexpr int R0 = N * LD > R1 ? N * LD : R1;
    // per half: matrix / build vectors, the column's m at every level, the broadcast slot of rm;
    // a multiple of 32 plus 8 doubles, so the two halves' broadcast reads fall in different banks
    static constexpr int S = ((R0 + 96 + 31) / 32) * 32 + 8;
};

// Step K of the LU and the steps after it (a template recursion: the pivot lane must be a
// compile-time constant of the swizzle).  Lane r holds row r in A.  Row K, final after step
// K-1, comes from lane K of the half: its diagonal, the next entry and the dense block in chunks
// of four columns (a chunk's columns past DB have a +-0 factor).  Rows r <= K take the update
// with a zero multiplier: r - (+-0) is r up to a zero's sign.
template <int N, int K> __device__ __forceinline__ void lu_steps(double (&A)[N], int li, int DB)
{
    if constexpr (K < N) {
        const int DBk = fresh(DB);
        double prow[N];
        prow[K] = half_bcast<K>(A[K]);
        if (K + 1 < N) prow[K + 1] = half_bcast<K>(A[K + 1]);
#pragma unroll
        for (int c0 = K + 2; c0 < N; c0 += 4) {
            if (c0 < DBk) {
#pragma unroll
                for (int c = c0; c < c0 + 4 && c < N; c++) prow[c] = half_bcast<K>(A[c]);
            }
        }
        const double q = A[K] / prow[K];
        A[K] = li > K ? q : A[K];
        const double lrk = li > K ? q : 0.0;
        if (K + 1 < N) A[K + 1] = A[K + 1] - lrk * prow[K + 1];
#pragma unroll
        for (int c0 = K + 2; c0 < N; c0 += 4) {
            if (c0 < DBk) {
#pragma unroll
                for (int c = c0; c < c0 + 4 && c < N; c++) A[c] = A[c] - lrk * prow[c];
            }
        }
        lu_steps<N, K + 1>(A, li, DB);
    }
}

// What the flagged sequence needs of a column besides C: the decay factors of the lane's two
// levels (fdlibm exp of minus the summed loss rates times dt) and the column's emission.
struct ColumnSeq {
    double decay, decay1, emis, x;
};

__device__ __forceinline__ ColumnSeq column_seq_setup(const ColumnArgs &a, int i, int j, int li)
{
    ColumnSeq s{1.0, 1.0, 0.0, 0.0};
    const int l1 = li + 32;
    if (a.pre || a.twice || a.post) {
        if (li < a.lm) {
            const double *L = &a.loss(i, j, li);
            const double total = ((L[0] + L[a.loss_sr]) + L[2 * a.loss_sr]) + L[3 * a.loss_sr];
            s.decay = fdlibm_exp(-(total * a.dt_chem));
        }
        if (l1 < a.lm) {
            const double *L = &a.loss(i, j, l1);
            const double total = ((L[0] + L[a.loss_sr]) + L[2 * a.loss_sr]) + L[3 * a.loss_sr];
            s.decay1 = fdlibm_exp(-(total * a.dt_chem));
        }
        s.emis = a.emis(i, j, 0);
        s.x = ((s.emis * a.factor) * a.ratio) * a.efrac;
    }
    return s;
}

// v on one tracer: rm <- row li of the operator (R) times the column's rm, summed from zero in
// increasing k over the levels < N, the slopes times the diagonal entry (pydyn/kernels.py:448-457);
// rv is the half's LDS vector of the column's rm.
template <int N>
__device__ __forceinline__ void column_v(const double (&R)[N], double diag, int li, double *rv, double &rm, double &rx,
                                         double &ry, double &rz)
{
    wsync();
    rv[li] = rm;
    wsync();
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) acc = acc + R[k] * rv[k];
    if (li < N) {
        rm = acc;
        rx = diag * rx;
        ry = diag * ry;
        rz = diag * rz;
    }
}

// Whether tracer t takes the emission at the lane's two cells, at levels lev (here) and lev1
// (here1), and the amount: do_add_2d's rm + x, rzm - x at the level (pydyn/zoom/physics.py:55-64);
// an ensemble member's emission field is its scale times the region's, scaled first.
__device__ __forceinline__ double column_emission(const ColumnArgs &a, const ColumnSeq &s, int lev, int lev1, int t,
                                                  bool &here, bool &here1, double &es)
{
    here = lev == a.level && t == a.tracer;
    here1 = lev1 == a.level && t == a.tracer;
    es = 1.0;
    if (!a.escale) return s.x;
    es = a.escale[t];
    here = lev == a.level && es != 0.0;
    here1 = lev1 == a.level && es != 0.0;
    return (((s.emis * es) * a.factor) * a.ratio) * a.efrac;
}

// a level no cell has (the second cell of a lane that holds one)
constexpr int NO_LEVEL = -(1 << 30);

// The decay of every moment of the lane's two cells (chemistry's c).
__device__ __forceinline__ void column_c(const ColumnSeq &s, double (&q)[8])
{
#pragma unroll
    for (int m = 0; m < 4; m++) q[m] = q[m] * s.decay;
#pragma unroll
    for (int m = 4; m < 8; m++) q[m] = q[m] * s.decay1;
}

// One tracer's flagged sequence around v: "c s" if pre, v, "s c c s v" if twice, "s c" if post,
// on two cells of the lane: level lev (q[0..3] = rm, rxm, rym, rzm; decay s.decay) and level
// lev1 (q[4..7]; s.decay1).  V: v acts on the first cell, at level li with R row li of C;
// without V both cells lie above the convective levels, where v changes nothing.
template <int N, bool V>
__device__ __forceinline__ void column_seq(const ColumnArgs &a, const ColumnSeq &s, const double (&R)[N], double diag,
                                           int li, int lev, int lev1, int t, double *rv, double (&q)[8])
{
    bool here, here1;
    double es;
    const double xt = column_emission(a, s, lev, lev1, t, here, here1, es);
    auto s_op = [&]() {
        if (here) {
            q[0] = q[0] + xt;
            q[3] = q[3] - xt;
        }
        if (here1) {
            q[4] = q[4] + xt;
            q[7] = q[7] - xt;
        }
    };
    if (a.pre) {
        column_c(s, q);
        s_op();
    }
    if (V) column_v<N>(R, diag, li, rv, q[0], q[1], q[2], q[3]);
    if (a.twice) {
        s_op();
        column_c(s, q);
        column_c(s, q);
        s_op();
        if (V) column_v<N>(R, diag, li, rv, q[0], q[1], q[2], q[3]);
    }
    if (a.post) {
        s_op();
        column_c(s, q);
    }
}

// zpre's columns: its rectangle less its hole
__device__ __forceinline__ bool zpre_has(const ColumnArgs &a, int i, int j)
{
    return i >= a.zpi0 && i <= a.zpi1 && j >= a.zpj0 && j <= a.zpj1 &&
           !(i >= a.zhi0 && i <= a.zhi1 && j >= a.zhj0 && j <= a.zhj1);
}

// column_seq with v on the lane's first cell, applied only where `on` (a column the region owns):
// elsewhere q is left as it is, but the lanes still pass the half's LDS exchanges
template <int N>
__device__ __forceinline__ void column_seq_if(bool on, const ColumnArgs &a, const ColumnSeq &s, const double (&R)[N],
                                              double diag, int li, int lev1, int t, double *rv, double (&q)[8])
{
    double w[8];
#pragma unroll
    for (int m = 0; m < 8; m++) w[m] = q[m];
    column_seq<N, true>(a, s, R, diag, li, li, lev1, t, rv, w);
#pragma unroll
    for (int m = 0; m < 8; m++) q[m] = on ? w[m] : q[m];
}

template <int N, bool ADJ>
__global__ void __launch_bounds__(256, 3) k_column(ColumnArgs a, int *words, int skip_mask)
{
    if (skipped_k(words, skip_mask)) return;
    constexpr int LD = ColumnLds<N>::LD, S = ColumnLds<N>::S;
    static_assert(N < 32, "a half-wave holds the convective levels");
    extern __shared__ double lds[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int h = lane >> 5, li = lane & 31, l1 = li + 32;
    const int ncol = a.im * a.jm;
    int col = (xcd_tile(blockIdx.x, gridDim.x) * 4 + wave) * 2 + h;
    const bool active = col < ncol;
    if (!active) col = ncol - 1;
    const int i = col % a.im, j = col / a.im;
    const int lm = a.lm;
    double *H = lds + (size_t)(wave * 2 + h) * S;
    double *Mt = H;  // N x LD: conv1, then L\U, then C (column c written by lane c)
    double *eu = H, *du = H + 32, *ed = H + 64, *dd = H + 96, *dk = H + 128, *eum = H + 160, *edm = H + 192;
    double *amu = H + 224, *amd = H + 256, *cd = H + 288, *cdl = H + 320;
    double *ind = H + 352, *kpd = H + 384, *inu = H + 416, *kpu = H + 448;  // the recursions' inflows, keeps
    double *mc = H + ColumnLds<N>::R0;  // m at all lm levels (m2phlb1)
    double *rv = mc + 64;               // the column's rm by level, for v

    // zfused: the first dynamw (n_advz 1, c = cm / 1) moves m before convection and m2phlb1 see it
    const bool zf = !ADJ && a.zfused;
    // the dynamw after the sequence: zfused's second, or zpost's on the z scope's columns (not in
    // the split form's factor pass, whose apply takes it)
    const bool zq = !ADJ && !a.cbuf && (a.zfused || a.zpost);
    const bool zin = a.zfused || (i >= a.zi0 && i <= a.zi1 && j >= a.zj0 && j <= a.zj1);
    // zpre: the matrix (and m2phlb1) from the mass after the dynamw the pass takes on this column
    // (in the split form's factor pass, whose apply moves the tracers)
    const bool zpc = !ADJ && a.zpre && zpre_has(a, i, j);
    // Every input the column reads from HBM is loaded here, together, so that their latencies
    // overlap (the exchanges through LDS below would hold each later group of loads back): the
    // masses, the faces of the lane's two cells, the convection inputs, the cloud levels, and for
    // a sequence on tracers the loss rates and the emission (decay factors made at once).
    const double m0_l = li < lm ? a.m(i, j, li) : 0.0, m0_h = l1 < lm ? a.m(i, j, l1) : 0.0;
    double zc_lo = 0.0, zc_hi = 0.0, zc_lo1 = 0.0, zc_hi1 = 0.0;  // cm below and above each cell
    if (zq || zpc) {
        if (li < lm) {
            zc_lo = a.cm(i, j, li);
            zc_hi = a.cm(i, j, li + 1);
        }
        if (l1 < lm) {
            zc_lo1 = a.cm(i, j, l1);
            zc_hi1 = a.cm(i, j, l1 + 1);
        }
    }
    double e_u = 0.0, d_u = 0.0, e_d = 0.0, d_d = 0.0, dk_l = 0.0;
    if (li < N) {
        e_u = a.entu(i, j, li);
        d_u = a.detu(i, j, li);
        e_d = a.entd(i, j, li);
        d_d = a.detd(i, j, li);
        dk_l = a.dkg(i, j, li);
    }
    const int ctop = a.top(i, j), clfs = a.lfs(i, j);
    // a column the region does not own is computed like the others but keeps its tracers; the
    // middle m2phlb1 is every column's
    const bool own = a.owned == nullptr || a.owned[(long long)j * a.im + i] != 0;
    const ColumnSeq cs = a.cbuf ? ColumnSeq{1.0, 1.0, 0.0, 0.0} : column_seq_setup(a, i, j, li);
    double m_l = m0_l, m_h = m0_h;
    if (zpc) {
        if (li < lm) m_l = (m0_l + zc_lo / 1.0) - zc_hi / 1.0;
        if (l1 < lm) m_h = (m0_h + zc_lo1 / 1.0) - zc_hi1 / 1.0;
    }
    if (zf) {
        if (li < lm) m_l = (m0_l + zc_lo / 1.0) - zc_hi / 1.0;
        if (l1 < lm) m_h = (m0_h + zc_lo1 / 1.0) - zc_hi1 / 1.0;
        if (checked(a.spec_k)) {
            // a checked step: Check_CFL's dynamwm at cap 1 on both dynamw's masses (m before and
            // after the first), every interior face k tested by lane k and k + 32 of its column
            H[li] = m0_l;
            H[l1] = m0_h;
            mc[li] = m_l;
            mc[l1] = m_h;
            wsync();
            bool ok = true, bad = false;
#pragma unroll
            for (int h2 = 0; h2 < 2; h2++) {
                const int k = h2 ? l1 : li;
                if (k >= 1 && k < lm) {
                    const double ck = h2 ? zc_lo1 : zc_lo;  // cm(k)
                    const int u = ck >= 0.0 ? k - 1 : k;
                    ok = ok && fabs(ck / H[u]) < 1.0 && fabs(ck / mc[u]) < 1.0;
                }
            }
            if (li == 0) bad = a.cm(i, j, 0) != 0.0 || a.cm(i, j, lm) != 0.0;
            const bool any_bad = __any(bad), any_fail = __any(!ok);  // every lane votes
            if (lane == 0) {
                if (any_bad) atomicOr(words + W_FLAGS, TM5_BAD_Z);
                if (any_fail) spec_fail(words, a.spec_k);
            }
            wsync();  // H is read: the build vectors take its space
        }
    } else if (zq && checked(a.spec_k)) {
        // zpost in a checked step: Check_CFL's dynamwm at cap 1 on the masses as they stand, every
        // column of the region (a column outside the z scope is only tested)
        mc[li] = m_l;
        mc[l1] = m_h;
        wsync();
        bool ok = true, bad = false;
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++) {
            const int k = h2 ? l1 : li;
            if (active && k >= 1 && k < lm) {
                const double ck = h2 ? zc_lo1 : zc_lo;  // cm(k)
                ok = ok && fabs(ck / mc[ck >= 0.0 ? k - 1 : k]) < 1.0;
            }
        }
        if (active && li == 0) bad = a.cm(i, j, 0) != 0.0 || a.cm(i, j, lm) != 0.0;
        const bool any_bad = __any(bad), any_fail = __any(!ok);  // every lane votes
        if (lane == 0) {
            if (any_bad) atomicOr(words + W_FLAGS, TM5_BAD_Z);
            if (any_fail) spec_fail(words, a.spec_k);
        }
    }
    mc[li] = m_l;
    mc[l1] = m_h;
    amu[li] = amd[li] = 0.0;
    if (li < N) {
        eu[li] = e_u;
        du[li] = d_u;
        ed[li] = e_d;
        dd[li] = d_d;
        dk[li] = dk_l;
        eum[li] = e_u / m_l;  // the fu(k, k-1) and fd(k-1, k-1) quotients, every level at once
        edm[li] = -e_d / m_l;
    }
    wsync();

    // ---- build conv1: lane c owns column c (pydyn/kernels.py:366-395)
    // levels above the cloud top (updraft) or the level of free sinking (downdraft) contribute
    // exact zeros (the reference's where(act, .., 0.0)); a wave runs to the larger of its two
    // columns' levels, the other column's extra levels giving the same zeros
    const int kup = ctop < N ? ctop : N, kdn = clfs < N ? clfs : N;
    const int db = (kup > kdn ? kup : kdn) + 1 < N ? (kup > kdn ? kup : kdn) + 1 : N;
    const int KUP = wave_max2(kup), KDN = wave_max2(kdn), DB = wave_max2(db);
    // Each draft's recursion in three passes: the column's mass flux level by level (the same
    // in every lane of the half: adds and compares only), then every level's keep factor in the
    // level's own lane (one division for all levels), then the lane's matrix column, run down
    // the levels with those factors.  Every value is the same expression on the same operands.
    double colv[N + 1];  // rows 0..N of this lane's column of the flux matrix; rows 0 and N stay zero
#pragma unroll
    for (int k = 0; k <= N; k++) colv[k] = 0.0;
    {  // downdraft levels k = kdn..2: inflow into level k-1 and its mass flux
        double prev = 0.0;
#pragma unroll
        for (int k = N; k >= 2; k--) {
            if (k <= KDN) {
                const bool on = clfs >= k;
                const double inflow = prev - ed[k - 1];
                const double aa = inflow + dd[k - 1];
                const double amd_k = (on && aa < 0.0) ? aa : 0.0;
                if (li == 0) {
                    ind[k - 1] = inflow;
                    amd[k - 1] = amd_k;
                }
                prev = amd_k;
            }
        }
    }
    {  // updraft levels k = 1..kup
        double prev = 0.0;
#pragma unroll
        for (int k = 1; k <= N; k++) {
            if (k <= KUP) {
                const bool on = ctop >= k;
                const double inflow = prev + eu[k - 1];
                const double aa = inflow - du[k - 1];
                const double amu_k = (on && aa > 0.0) ? aa : 0.0;
                if (li == 0) {
                    inu[k - 1] = inflow;
                    amu[k] = amu_k;
                }
                prev = amu_k;
            }
        }
    }
    wsync();
    // lane j: the downdraft's keep at level k = j + 1 (ed, dd of row j), the updraft's at k = j + 1
    // (eu, du of row j)
    if (li >= 1 && li + 1 <= KDN) {
        const double inflow = ind[li];
        const double aa = inflow + d_d;
        kpd[li] = aa < 0.0 ? max0(1.0 + d_d / inflow) : 0.0;
    }
    if (li + 1 <= KUP) {
        const double inflow = inu[li];
        const double aa = inflow - d_u;
        kpu[li] = aa > 0.0 ? max0(1.0 - d_u / inflow) : 0.0;
    }
    wsync();
    {  // downdraft rows k-1 for k = kdn..2
        double run = 0.0;
#pragma unroll
        for (int k = N; k >= 2; k--) {
            if (k <= KDN) {
                const bool on = clfs >= k;
                const double keep = kpd[k - 1];
                double v = 0.0;
                if (li >= k)
                    v = on ? run * keep : 0.0;
                else if (li == k - 1)
                    v = on ? edm[k - 1] * keep : 0.0;
                colv[k - 1] = v;
                run = v;
            }
        }
    }
    {  // updraft rows k = 1..kup, added in front: row k is fu + fd
        double run = 0.0;
#pragma unroll
        for (int k = 1; k <= N; k++) {
            double v = 0.0;
            if (k <= KUP) {
                const bool on = ctop >= k;
                const double keep = kpu[k - 1];
                if (li < k - 1)
                    v = on ? run * keep : 0.0;
                else if (li == k - 1)
                    v = on ? eum[k - 1] * keep : 0.0;
                run = v;
            }
            if (k < N) colv[k] = v + colv[k];
        }
    }
    wsync();
    // the four corrections of interface k (pydyn/kernels.py:388-391), every k at once in lane k:
    // lane k takes -cu(k) then -cdk(k) on row k; lane k-1 takes -cd(k) then +cdl(k)
    double q0 = 0.0, q3 = 0.0;
    if (li >= 1 && li < N) {
        q0 = amu[li] / mc[li];
        cd[li] = amd[li] / mc[li - 1];
        cdl[li] = dk[li - 1] / mc[li - 1];
        q3 = dk[li - 1] / mc[li];
    }
    wsync();
    {
        const double cdn = li + 1 < N ? cd[li + 1] : 0.0, cln = li + 1 < N ? cdl[li + 1] : 0.0;
#pragma unroll
        for (int k = 1; k < N; k++) {
            if (li == k) colv[k] = (colv[k] - q0) - q3;
            if (li + 1 == k) colv[k] = (colv[k] - cdn) + cln;
        }
    }
    wsync();  // the build vectors are read: the matrix takes their space
    if (li < N) {
#pragma unroll
        for (int r = 0; r < N; r++) {
            double v = (-a.dt) * (colv[r] - colv[r + 1]);
            if (r == li) v = v + 1.0;
            Mt[r * LD + li] = v;
        }
    }
    wsync();
    // conv1 row r into lane r's registers
    double A[N];
    {
        const int r = li < N ? li : N - 1;
#pragma unroll
        for (int c = 0; c < N; c++) A[c] = Mt[r * LD + c];
    }
    if constexpr (!ADJ) {
        if (a.build_only) {  // the LAPACK kind: conv1's rows for k_lapack_inverse, nothing else
            if (active && li < N) {
                double *Cr = a.cbuf + ((long long)col * N + li) * N;
#pragma unroll
                for (int c = 0; c < N; c++) Cr[c] = A[c];
            }
            return;
        }
    }

    // ---- LU, lanes over rows.  Entry (r, c) is non-zero only inside the dense block (r, c < DB)
    // or on the three diagonals, and LU fill-in stays inside that envelope; the products skipped
    // outside it have a factor that is exactly +-0.
    lu_steps<N, 0>(A, li, DB);
    // the factors to LDS, row r by lane r: what the substitutions read (the three diagonals, and
    // for rows inside the dense block its columns with the last chunk's padding)
    if (li < N) {
#pragma unroll
        for (int c = 0; c < N; c++)
            if ((c + 1 >= li && c <= li + 1) || (li < DB && c < DB + 3)) Mt[li * LD + c] = A[c];
    }
    wsync();

    // ---- inverse: lane c solves L U x = e_c
    double y[N];
#pragma unroll
    for (int r = 0; r < N; r++) y[r] = (li == r) ? 1.0 : 0.0;
#pragma unroll
    for (int r = 1; r < N; r++) {
        if (r < fresh(DB)) {
            double lrow[N];
#pragma unroll
            for (int s = 0; s < r - 1; s++) lrow[s] = Mt[r * LD + s];
#pragma unroll
            for (int s = 0; s < r - 1; s++) y[r] = y[r] - lrow[s] * y[s];
        }
        y[r] = y[r] - Mt[r * LD + r - 1] * y[r - 1];
    }
#pragma unroll
    for (int r = N - 1; r >= 0; r--) {
        const int DBr = fresh(DB);
        double urow[N];  // row r of U inside the envelope
        urow[r] = Mt[r * LD + r];
        if (r + 1 < N) urow[r + 1] = Mt[r * LD + r + 1];
#pragma unroll
        for (int s0 = r + 2; s0 < N; s0 += 4) {
            if (s0 < DBr) {
#pragma unroll
                for (int s = s0; s < s0 + 4 && s < N; s++) urow[s] = Mt[r * LD + s];
            }
        }
        double acc = y[r];
        if (r + 1 < N) acc = acc - urow[r + 1] * y[r + 1];
#pragma unroll
        for (int s0 = r + 2; s0 < N; s0 += 4) {
            if (s0 < DBr) {
#pragma unroll
                for (int s = s0; s < s0 + 4 && s < N; s++) acc = acc - urow[s] * y[s];
            }
        }
        y[r] = acc / urow[r];
    }
    // C through LDS: lane c writes column c; lane l reads row l
    wsync();
    if (li < N) {
#pragma unroll
        for (int r = 0; r < N; r++) Mt[r * LD + li] = y[r];
    }
    wsync();
    double crow[ADJ ? 1 : N];
    if (!ADJ) {
        const int r = li < N ? li : N - 1;
#pragma unroll
        for (int k = 0; k < N; k++) crow[k] = Mt[r * LD + k];
    }
    const double diag = li < N ? Mt[li * LD + li] : 0.0;
    if constexpr (!ADJ) {
        if (a.cbuf && active && li < N) {
            double *Cr = a.cbuf + ((long long)col * N + li) * N;
#pragma unroll
            for (int k = 0; k < N; k++) Cr[k] = crow[k];
        }
    }

    // zfused: dynamw's sub-step through LDS staging Z (the matrix's space, free once C is read):
    // Z[0, 64) the masses at the first sub-step's start, then rm, rxm, rym, rzm by level; the
    // second sub-step starts from mc.  Each lane takes the two faces of each of its cells.
    static_assert(ColumnLds<N>::R0 >= ColumnLds<N>::ZSTAGE, "the staging fits the matrix's space");
    double *Z = H;
    double zm0 = 0.0, zm0h = 0.0, zm2 = 0.0, zm2h = 0.0;
    if (zq || a.zpre) {
        if (li < lm) {
            zm0 = m0_l;
            zm2 = (m_l + zc_lo / 1.0) - zc_hi / 1.0;
        }
        if (l1 < lm) {
            zm0h = m0_h;
            zm2h = (m_h + zc_lo1 / 1.0) - zc_hi1 / 1.0;
        }
    }
    const bool zlim = a.zlimits != 0;
    // face k of the staged sub-step (pydyn/kernels.py:324-346, as k_z_steps' z_face at n_advz 1)
    auto z_face = [&](int k, double cf, const double *Mz) {
        FaceFlux ff{0.0, 0.0, 0.0, 0.0};
        if (k <= 0 || k >= lm) return ff;  // the closed ends
        const int u = cf >= 0.0 ? k - 1 : k;
        const double gamma = cf / Mz[u];
        const double rmu = Z[64 + u];
        const double rzu = zlim ? limit(Z[256 + u], rmu) : Z[256 + u];
        return face_flux(cf, gamma, rmu, rzu, Z[128 + u], Z[192 + u]);
    };
    // the sub-step's face fluxes, face k's f, pf, ft, fw at F[k], F[ZF + k], F[2 ZF + k], F[3 ZF + k]:
    // each interior face is made once, by the lane of the cell above it, and read by both cells
    constexpr int ZF = 65;
    double *F = Z + 5 * 64;
    auto z_cell = [&](int l, double cm_lo, double cm_hi, double m_old, double m_new, double &rm, double &rx,
                      double &ry, double &rz) {
        const double c_lo = cm_lo / 1.0, c_hi = cm_hi / 1.0;
        FaceFlux lo{0.0, 0.0, 0.0, 0.0}, hi{0.0, 0.0, 0.0, 0.0};  // the closed ends carry nothing
        if (l >= 1) lo = FaceFlux{F[l], F[ZF + l], F[2 * ZF + l], F[3 * ZF + l]};
        if (l + 1 < lm) hi = FaceFlux{F[l + 1], F[ZF + l + 1], F[2 * ZF + l + 1], F[3 * ZF + l + 1]};
        const double rzm = zlim ? limit(rz, rm) : rz;
        const double rm_new = (rm + lo.f) - hi.f;  // advance: rm + f(l-1) - f(l)
        double rzn = slope_update(rzm, lo.pf, hi.pf, c_lo, c_hi, rm_new, lo.f, hi.f, m_old, m_new);
        if (zlim) rzn = limit(rzn, rm_new);
        rx = (rx + lo.ft) - hi.ft;
        ry = (ry + lo.fw) - hi.fw;
        rm = rm_new;
        rz = rzn;
    };

    // dynamw at n_advz 1 on one tracer: m0 -> m1 (first), m1 -> m2 (second)
    auto z_op = [&](bool second, bool moves, double (&q)[8]) {
        wsync();  // earlier reads of Z and of the matrix precede these writes
        if (!second) {
            if (li < lm) Z[li] = zm0;
            if (l1 < lm) Z[l1] = zm0h;
        }
        if (li < lm) {
#pragma unroll
            for (int m = 0; m < 4; m++) Z[64 * (m + 1) + li] = q[m];
        }
        if (l1 < lm) {
#pragma unroll
            for (int m = 0; m < 4; m++) Z[64 * (m + 1) + l1] = q[4 + m];
        }
        wsync();
        const double *Mz = second ? mc : Z;
        auto put = [&](int k, double cm_k) {  // face k, whose cm is the lower face's of cell k
            const FaceFlux ff = z_face(k, cm_k / 1.0, Mz);
            F[k] = ff.f;
            F[ZF + k] = ff.pf;
            F[2 * ZF + k] = ff.ft;
            F[3 * ZF + k] = ff.fw;
        };
        if (li >= 1 && li < lm) put(li, zc_lo);
        if (l1 < lm) put(l1, zc_lo1);
        wsync();
        if (!moves) return;  // a column outside the dynamw's scope is not moved
        if (li < lm) z_cell(li, zc_lo, zc_hi, second ? mc[li] : zm0, second ? zm2 : mc[li], q[0], q[1], q[2], q[3]);
        if (l1 < lm)
            z_cell(l1, zc_lo1, zc_hi1, second ? mc[l1] : zm0h, second ? zm2h : mc[l1], q[4], q[5], q[6], q[7]);
    };

    // ---- apply: the flagged sequence around v for each tracer of the column, levels li and l1
    // (the factor pass of the split form stops here: k_column_apply takes the tracers)
    const int ntl = a.cbuf ? 0 : a.q.nt;
    double lam_e = 0.0;  // ADJ: the emission's lambda from this lane's level
    for (int t = 0; t < ntl; t++) {
        const long long o = a.q.off(i, j, li < lm ? li : 0, t), o1 = a.q.off(i, j, l1 < lm ? l1 : 0, t);
        double q[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (li < lm) {
#pragma unroll
            for (int m = 0; m < 4; m++) q[m] = a.q.q[m][o];
        }
        if (l1 < lm) {
#pragma unroll
            for (int m = 0; m < 4; m++) q[4 + m] = a.q.q[m][o1];
        }
        if constexpr (ADJ) {  // the transposes in reverse: v^T; s^T adds (rm - rzm) times dx/demis at the level
            bool here, here1;
            double es;
            column_emission(a, cs, li, l1, t, here, here1, es);
            const double kt = ((es * a.factor) * a.ratio) * a.efrac;
            auto s_ad = [&]() {
                if (here) lam_e = lam_e + kt * (q[0] - q[3]);
                if (here1) lam_e = lam_e + kt * (q[4] - q[7]);
            };
            // v^T: lane k holds column k of C
            auto v_op = [&]() { column_v<N>(y, diag, li, rv, q[0], q[1], q[2], q[3]); };
            if (a.post) {
                column_c(cs, q);
                s_ad();
            }
            if (a.twice) {
                v_op();
                s_ad();
                column_c(cs, q);
                column_c(cs, q);
                s_ad();
            }
            v_op();
            if (a.pre) {
                s_ad();
                column_c(cs, q);
            }
        } else {
            if (zf || a.zpre) z_op(false, zf || zpc, q);
            if ((zq || a.zpre) && a.owned)  // a column the region does not own keeps its tracers through the sequence
                column_seq_if<N>(own, a, cs, crow, diag, li, l1, t, rv, q);
            else
                column_seq<N, true>(a, cs, crow, diag, li, li, l1, t, rv, q);
            if (zq) z_op(true, zin, q);
        }
        if (active && (own || (zq && zin) || zpc)) {
            if (li < lm) {
#pragma unroll
                for (int m = 0; m < 4; m++) a.q.q[m][o] = q[m];
            }
            if (l1 < lm) {
#pragma unroll
                for (int m = 0; m < 4; m++) a.q.q[m][o1] = q[4 + m];
            }
        }
    }
    if (ADJ) {
        if (active && own && (li == a.level || l1 == a.level)) a.lam_emis(i, j, 0) = a.lam_emis(i, j, 0) + lam_e;
        return;
    }
    if (zq && zin && active) {  // the second dynamw's mass
        if (li < lm) a.m(i, j, li) = zm2;
        if (l1 < lm) a.m(i, j, l1) = zm2h;
    } else if (zpc && active && !a.cbuf) {  // zpre's alone (the split form's: k_z_mass_post)
        if (li < lm) a.m(i, j, li) = m_l;
        if (l1 < lm) a.m(i, j, l1) = m_h;
    }
    // m2phlb1 between the two chemistry calls (pydyn/zoom/massflow.py:282-290, 24-32): sp from
    // the column's mass summed upward, phlb = at + bt sp
    if (a.mid && active) {
        double p = mc[0];
        for (int l = 1; l < lm; l++) p = p + mc[l];
        p = p * 9.80665 / a.dxyp[j];
        if (li == 0) a.sp(i, j, 0) = p;
        for (int l = li; l <= lm; l += 32) a.phlb(i, j, l) = a.at[l] + a.bt[l] * p;
    }
}

// The apply of the split form, for many tracers: k_column's sequence on every tracer, from the
// rows of C its factor pass stored in cbuf.  A workgroup takes blockDim.x / 32 (8, 16 or 32: the
// context's TM5_TILE_APPLY) neighbouring columns of a row (two columns a wave, lanes over levels
// as in k_column) and a range of tracers (grid z),
// so each level's cache line of a tracer is read and written by one workgroup at about the same
// time: its waves meet at a barrier after every tracer (a rendezvous only; nothing is exchanged
// between waves), and each wave loads the tracers ahead of the one it works on.  Levels 32 and
// up (only sources and chemistry act there) follow for the range's tracers at once, cell by cell.
constexpr int APPLY_COLS_MAX = 32;
template <int N>
__global__ void __launch_bounds__(APPLY_COLS_MAX * 32)
    k_column_apply(ColumnArgs a, const double *cbuf, int tchunk, int *words, int skip_mask)
{
    if (skipped_k(words, skip_mask)) return;
    constexpr int AHEAD = 2;                         // tracers loaded ahead
    __shared__ double rvs[APPLY_COLS_MAX * 32];      // each column's rm by level, for v
    __shared__ double dec_hi[APPLY_COLS_MAX * 32];   // each column's decay factors at levels 32 and up
    const int cols = blockDim.x >> 5;
    // tiles renumbered XCD-contiguously (xcd_tile), the tracer range slowest
    const int nbx = gridDim.x, nby = gridDim.y;
    const int b = xcd_tile((blockIdx.z * nby + blockIdx.y) * nbx + blockIdx.x, nbx * nby * gridDim.z);
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int h = lane >> 5, li = lane & 31, lm = a.lm;
    const int t0 = bz * tchunk, nt = a.q.nt - t0 < tchunk ? a.q.nt - t0 : tchunk;  // tracers t0 .. t0 + nt - 1
    // the row's tiles start on 64-byte boundaries of the tracer arrays (every level and tracer of
    // the row has the row's offset modulo 8 doubles when those strides are multiples of 8), so
    // neighbouring tiles do not share the sectors at their edges
    const int j = by;
    const int e = (a.q.sl % 8 == 0 && a.q.st % 8 == 0) ? (int)(((uintptr_t)(a.q.q[0] + a.q.off(0, j, 0, 0)) >> 3) & 7) : 0;
    int i = bx * cols - e + wave * 2 + h;
    const bool active = i >= 0 && i < a.im;
    if (!active) i = i < 0 ? 0 : a.im - 1;
    const bool store = active && (a.owned == nullptr || a.owned[(long long)j * a.im + i] != 0);
    double *rv = rvs + (wave * 2 + h) * 32;
    const double *Cr = cbuf + ((long long)j * a.im + i) * (N * N) + (long long)(li < N ? li : N - 1) * N;
    double crow[N];
#pragma unroll
    for (int k = 0; k < N; k++) crow[k] = Cr[k];
    const double diag = li < N ? Cr[li] : 0.0;
    const ColumnSeq cs = column_seq_setup(a, i, j, li);
    const bool in = li < lm;
    const long long o = a.q.off(i, j, in ? li : 0, t0), st = a.q.st;
    double ahead[AHEAD][4];
#pragma unroll
    for (int d = 0; d < AHEAD; d++) {
#pragma unroll
        for (int m = 0; m < 4; m++) ahead[d][m] = in && d < nt ? a.q.q[m][o + d * st] : 0.0;
    }
    for (int t = 0; t < nt; t++) {
        double q[8] = {ahead[0][0], ahead[0][1], ahead[0][2], ahead[0][3], 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d + 1 < AHEAD; d++) {
#pragma unroll
            for (int m = 0; m < 4; m++) ahead[d][m] = ahead[d + 1][m];
        }
        const int tn = t + AHEAD;
#pragma unroll
        for (int m = 0; m < 4; m++) ahead[AHEAD - 1][m] = in && tn < nt ? a.q.q[m][o + tn * st] : 0.0;
        column_seq<N, true>(a, cs, crow, diag, li, li, NO_LEVEL, t0 + t, rv, q);
        if (store && in) {
#pragma unroll
            for (int m = 0; m < 4; m++) a.q.q[m][o + t * st] = q[m];
        }
        __builtin_amdgcn_s_barrier();
    }
    // levels 32 .. lm - 1 of every tracer: cell (t, 32 + c) of the column by lane li + 32 n
    const int nhi = lm - 32;
    if (nhi <= 0) return;
    double *dh = dec_hi + (wave * 2 + h) * 32;
    if (li < nhi) dh[li] = cs.decay1;
    wsync();
    const int cells = nhi * nt;
    constexpr int B = 4;  // cells per lane in flight
    for (int c0 = 0; c0 < cells; c0 += 32 * B) {
        double v[B][8];
        long long oc[B];
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int c = c0 + u * 32 + li;
            const int t = c < cells ? c / nhi : 0, lev = c < cells ? 32 + c % nhi : 32;
            oc[u] = a.q.off(i, j, lev, t0 + t);
#pragma unroll
            for (int m = 0; m < 4; m++) v[u][m] = c < cells ? a.q.q[m][oc[u]] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int c = c0 + u * 32 + li;
            if (c >= cells) continue;
            const int t = c / nhi, lev = 32 + c % nhi;
            const ColumnSeq ch{dh[lev - 32], 1.0, cs.emis, cs.x};
#pragma unroll
            for (int m = 4; m < 8; m++) v[u][m] = 0.0;
            column_seq<N, false>(a, ch, crow, diag, li, lev, NO_LEVEL, t0 + t, rv, v[u]);
            if (store) {
#pragma unroll
                for (int m = 0; m < 4; m++) a.q.q[m][oc[u]] = v[u][m];
            }
        }
    }
}

// The apply with the step's dynamw around the sequence fused: every tracer of the tile's columns
// takes k_column_apply's sequence (on the columns the region owns) and, at n_advz 1 (c = cm / 1,
// pydyn/kernels.py:324-346, as k_z_steps' sub-step), zpre's dynamw before it on the columns it
// names and zpost's after it on the columns of its scope.  Each lane holds the tracer's cells at
// levels li and li + 32, so the column's every level is in its half's registers.  A face's
// Courant number and upwind cell depend on the masses only and are found once for each dynamw;
// per tracer the moments pass through the half's LDS, lane k makes face k's fluxes (and face
// k + 32's), and both cells of each face read them back.  No workgroup writes m (the tracer
// ranges of a column all read it): k_z_mass_post moves it afterwards.  A checked step's zpost is
// tested here, every column's faces at cap 1, by the first tracer range (zpre's was tested where
// the rest of its scope moved).
constexpr int APPLY_Z_COLS = 8;
constexpr int APPLY_Z_F = 65;                               // a flux plane: faces 0 .. 64
constexpr int APPLY_Z_LDS = 5 * 64 + 4 * APPLY_Z_F + 32 + 3;  // masses, moments, fluxes, rm for v
template <int N>
__global__ void __launch_bounds__(APPLY_Z_COLS * 32)
    k_column_apply_z(ColumnArgs a, const double *cbuf, int tchunk, int *words, int skip_mask)
{
    if (skipped_k(words, skip_mask)) return;
    extern __shared__ double zsh[];
    const int cols = blockDim.x >> 5;
    const int nbx = gridDim.x, nby = gridDim.y;
    const int b = xcd_tile((blockIdx.z * nby + blockIdx.y) * nbx + blockIdx.x, nbx * nby * gridDim.z);
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int h = lane >> 5, li = lane & 31, l1 = li + 32, lm = a.lm;
    const int t0 = bz * tchunk, nt = a.q.nt - t0 < tchunk ? a.q.nt - t0 : tchunk;
    const int j = by;
    const int e = (a.q.sl % 8 == 0 && a.q.st % 8 == 0) ? (int)(((uintptr_t)(a.q.q[0] + a.q.off(0, j, 0, 0)) >> 3) & 7) : 0;
    int i = bx * cols - e + wave * 2 + h;
    const bool active = i >= 0 && i < a.im;
    if (!active) i = i < 0 ? 0 : a.im - 1;
    const bool own = active && (a.owned == nullptr || a.owned[(long long)j * a.im + i] != 0);
    const bool zin = active && a.zpost && i >= a.zi0 && i <= a.zi1 && j >= a.zj0 && j <= a.zj1;
    const bool zpc = active && a.zpre && zpre_has(a, i, j);
    double *Z = zsh + (wave * 2 + h) * APPLY_Z_LDS;  // Z[0, 64) masses, then rm, rxm, rym, rzm by level
    double *F = Z + 5 * 64;                          // f, pf, ft, fw by face
    double *rv = F + 4 * APPLY_Z_F;
    const double *Cr = cbuf + ((long long)j * a.im + i) * (N * N) + (long long)(li < N ? li : N - 1) * N;
    double crow[N];
#pragma unroll
    for (int k = 0; k < N; k++) crow[k] = Cr[k];
    const double diag = li < N ? Cr[li] : 0.0;
    const ColumnSeq cs = column_seq_setup(a, i, j, li);
    const bool in0 = li < lm, in1 = l1 < lm;
    // the masses as they stand, after zpre's dynamw (mp, on its columns) and after zpost's (m2),
    // and the faces below and above each of the lane's cells
    const double m0 = in0 ? a.m(i, j, li) : 0.0, m0h = in1 ? a.m(i, j, l1) : 0.0;
    const double zc_lo = in0 ? a.cm(i, j, li) : 0.0, zc_hi = in0 ? a.cm(i, j, li + 1) : 0.0;
    const double zc_lo1 = in1 ? a.cm(i, j, l1) : 0.0, zc_hi1 = in1 ? a.cm(i, j, l1 + 1) : 0.0;
    const double mp = (m0 + zc_lo / 1.0) - zc_hi / 1.0, mph = (m0h + zc_lo1 / 1.0) - zc_hi1 / 1.0;
    const double mb = zpc ? mp : m0, mbh = zpc ? mph : m0h;  // before zpost's dynamw
    const double m2 = (mb + zc_lo / 1.0) - zc_hi / 1.0, m2h = (mbh + zc_lo1 / 1.0) - zc_hi1 / 1.0;
    const bool zlim = a.zlimits != 0;
    // face k = li (interior: 1 <= k < lm) and face k1 = l1 (k1 < lm): c = cm / 1, upwind cell and,
    // from the masses staged in Z, the Courant number (the same for every tracer)
    const bool fa = li >= 1 && in0, fb = in1;
    const double ca = zc_lo / 1.0, cb = zc_lo1 / 1.0;
    const int ua = ca >= 0.0 ? li - 1 : li, ub = cb >= 0.0 ? l1 - 1 : l1;
    double gpa = 0.0, gpb = 0.0;  // zpre's
    if (a.zpre) {
        if (in0) Z[li] = m0;
        if (in1) Z[l1] = m0h;
        wsync();
        gpa = fa ? ca / Z[ua] : 0.0;
        gpb = fb ? cb / Z[ub] : 0.0;
        wsync();
    }
    if (in0) Z[li] = mb;
    if (in1) Z[l1] = mbh;
    wsync();
    if (a.zpost && checked(a.spec_k) && bz == 0) {
        // Check_CFL's dynamwm at cap 1: every interior face k by lane k and k + 32 of its column
        bool ok = true, bad = false;
        if (active) {
#pragma unroll
            for (int h2 = 0; h2 < 2; h2++) {
                const int k = h2 ? l1 : li;
                if (k >= 1 && k < lm) {
                    const double ck = h2 ? zc_lo1 : zc_lo;  // cm(k)
                    ok = ok && fabs(ck / Z[ck >= 0.0 ? k - 1 : k]) < 1.0;
                }
            }
            if (li == 0) bad = a.cm(i, j, 0) != 0.0 || a.cm(i, j, lm) != 0.0;
        }
        const bool any_bad = __any(bad), any_fail = __any(!ok);  // every lane votes
        if (lane == 0) {
            if (any_bad) atomicOr(words + W_FLAGS, TM5_BAD_Z);
            if (any_fail) spec_fail(words, a.spec_k);
        }
    }
    const double ga = fa ? ca / Z[ua] : 0.0, gb = fb ? cb / Z[ub] : 0.0;  // zpost's
    auto put_face = [&](int k, int u, double cf, double g) {
        const double rmu = Z[64 + u];
        const double rzu = zlim ? limit(Z[256 + u], rmu) : Z[256 + u];
        const FaceFlux ff = face_flux(cf, g, rmu, rzu, Z[128 + u], Z[192 + u]);
        F[k] = ff.f;
        F[APPLY_Z_F + k] = ff.pf;
        F[2 * APPLY_Z_F + k] = ff.ft;
        F[3 * APPLY_Z_F + k] = ff.fw;
    };
    auto z_cell = [&](int l, double cm_lo, double cm_hi, double m_old, double m_new, double &rm, double &rx,
                      double &ry, double &rz) {
        const double c_lo = cm_lo / 1.0, c_hi = cm_hi / 1.0;
        FaceFlux lo{0.0, 0.0, 0.0, 0.0}, hi{0.0, 0.0, 0.0, 0.0};  // the closed ends carry nothing
        if (l >= 1) lo = FaceFlux{F[l], F[APPLY_Z_F + l], F[2 * APPLY_Z_F + l], F[3 * APPLY_Z_F + l]};
        if (l + 1 < lm) hi = FaceFlux{F[l + 1], F[APPLY_Z_F + l + 1], F[2 * APPLY_Z_F + l + 1], F[3 * APPLY_Z_F + l + 1]};
        const double rzm = zlim ? limit(rz, rm) : rz;
        const double rm_new = (rm + lo.f) - hi.f;
        double rzn = slope_update(rzm, lo.pf, hi.pf, c_lo, c_hi, rm_new, lo.f, hi.f, m_old, m_new);
        if (zlim) rzn = limit(rzn, rm_new);
        rx = (rx + lo.ft) - hi.ft;
        ry = (ry + lo.fw) - hi.fw;
        rm = rm_new;
        rz = rzn;
    };
    // one dynamw on the tracer's cells: the moments staged, the faces made, the cells moved where
    // `moves` (the half's lanes all pass the exchanges)
    auto z_op = [&](double g0, double g1, bool moves, double mo, double mn, double moh, double mnh, double (&q)[8]) {
        wsync();  // the previous staged values and fluxes are read
        if (in0) {
#pragma unroll
            for (int m = 0; m < 4; m++) Z[64 * (m + 1) + li] = q[m];
        }
        if (in1) {
#pragma unroll
            for (int m = 0; m < 4; m++) Z[64 * (m + 1) + l1] = q[4 + m];
        }
        wsync();
        if (fa) put_face(li, ua, ca, g0);
        if (fb) put_face(l1, ub, cb, g1);
        wsync();
        if (moves) {
            if (in0) z_cell(li, zc_lo, zc_hi, mo, mn, q[0], q[1], q[2], q[3]);
            if (in1) z_cell(l1, zc_lo1, zc_hi1, moh, mnh, q[4], q[5], q[6], q[7]);
        }
    };
    const long long o = a.q.off(i, j, in0 ? li : 0, t0), o1 = a.q.off(i, j, in1 ? l1 : 0, t0), st = a.q.st;
    for (int t = 0; t < nt; t++) {
        double q[8];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            q[m] = in0 ? a.q.q[m][o + t * st] : 0.0;
            q[4 + m] = in1 ? a.q.q[m][o1 + t * st] : 0.0;
        }
        if (a.zpre) z_op(gpa, gpb, zpc, m0, mp, m0h, mph, q);
        // the sequence (a column the region does not own keeps its tracers through it; its lanes
        // still take part in the half's exchanges)
        if (a.owned)
            column_seq_if<N>(own, a, cs, crow, diag, li, in1 ? l1 : NO_LEVEL, t0 + t, rv, q);
        else
            column_seq<N, true>(a, cs, crow, diag, li, li, in1 ? l1 : NO_LEVEL, t0 + t, rv, q);
        if (a.zpost) z_op(ga, gb, zin, mb, m2, mbh, m2h, q);
        if (own || zin || zpc) {
            if (in0) {
#pragma unroll
                for (int m = 0; m < 4; m++) a.q.q[m][o + t * st] = q[m];
            }
            if (in1) {
#pragma unroll
                for (int m = 0; m < 4; m++) a.q.q[m][o1 + t * st] = q[4 + m];
            }
        }
        __builtin_amdgcn_s_barrier();
    }
}

// The fused dynamw's masses: m <- (m + cm(l) / 1) - cm(l + 1) / 1 for zpre's (on its columns) and
// then zpost's (on its scope) -- k_z_steps' mass at n_advz 1 -- one thread a cell
__global__ void __launch_bounds__(256) k_z_mass_post(ColumnArgs a, int *words, int skip_mask)
{
    if (skipped_k(words, skip_mask)) return;
    const long long n = (long long)a.im * a.jm * a.lm, c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % a.im), j = (int)((c / a.im) % a.jm), l = (int)(c / ((long long)a.im * a.jm));
    const bool pre = a.zpre && zpre_has(a, i, j);
    const bool post = a.zpost && i >= a.zi0 && i <= a.zi1 && j >= a.zj0 && j <= a.zj1;
    if (!pre && !post) return;
    double mm = a.m(i, j, l);
    if (pre) mm = (mm + a.cm(i, j, l) / 1.0) - a.cm(i, j, l + 1) / 1.0;
    if (post) mm = (mm + a.cm(i, j, l) / 1.0) - a.cm(i, j, l + 1) / 1.0;
    a.m(i, j, l) = mm;
}

// the split form from this many tracers on (a factor pass, then k_column_apply)
constexpr int SPLIT_NT = 8;

int tracer_chunk(long long tiles, long long slots, int nt, int min_chunk)
{
    const int most = nt / min_chunk > 1 ? nt / min_chunk : 1;
    int ranges = 1;
    double best = -1.0;
    for (int r = 1; r <= most; r++) {
        const int chunk = (nt + r - 1) / r, used = (nt + chunk - 1) / chunk;
        const double w = (double)tiles * used / (double)slots, eff = w / ceil(w);
        if (eff > best) {
            best = eff;
            ranges = used;
        }
        if (eff >= 0.9) break;
    }
    return (nt + ranges - 1) / ranges;
}

// ext (non-null): the split form's matrices go to ext (ext_n doubles) instead of the context's
// buffer, which is then neither allocated nor grown; too small an ext is TM5_ESCRATCH
template <int N>
static int launch_column_n(tm5_ctx *ctx, const ColumnArgs &args, int skip_mask, double *ext, size_t ext_n)
{
    if (args.zfused && (args.adjoint || args.owned || !args.cm.p)) return TM5_EINVAL;  // a region's step only
    if (args.zpost && (args.zfused || args.adjoint || args.lapack || !args.cm.p || args.zi0 < 0 || args.zj0 < 0 ||
                       args.zi1 >= args.im || args.zj1 >= args.jm || args.zi1 < args.zi0 || args.zj1 < args.zj0))
        return TM5_EINVAL;  // a forward step's, on a scope inside the region
    if (args.zpre && (args.zfused || args.adjoint || args.lapack || !args.cm.p || args.zpi0 < 0 || args.zpj0 < 0 ||
                      args.zpi1 >= args.im || args.zpj1 >= args.jm))
        return TM5_EINVAL;  // a forward step's, on a rectangle inside the region
    const size_t lds = (size_t)8 * ColumnLds<N>::S * sizeof(double);  // four waves, two columns each
    const int ncol = args.im * args.jm;
    const dim3 grid((ncol + 7) / 8), block(256);
    static const char *name[] = {"k_column<0>", "k_column<1>", "k_column<2>", "k_column<3>", "k_column<4>",
                                 "k_column<5>", "k_column<6>", "k_column<7>", "k_column<8>", "k_column<9>",
                                 "k_column<10>", "k_column<11>", "k_column<12>", "k_column<13>", "k_column<14>",
                                 "k_column<15>", "k_column<16>", "k_column<17>", "k_column<18>", "k_column<19>",
                                 "k_column<20>", "k_column<21>", "k_column<22>", "k_column<23>", "k_column<24>",
                                 "k_column<25>", "k_column<26>", "k_column<27>", "k_column<28>", "k_column<29>",
                                 "k_column<30>", "k_column<31>"};
    if (args.adjoint) {
        TM5_LAUNCH_NAMED(ctx, "k_column_adjoint", (k_column<N, true>), grid, block, lds, ctx->stream, args, ctx->words,
                         skip_mask);
        return status(hipGetLastError());
    }
    if (args.lapack && (args.zfused || args.mid || args.pre || args.twice || args.post)) return TM5_EINVAL;
    if (!args.lapack && (args.zfused || args.q.nt < SPLIT_NT)) {
        TM5_LAUNCH_NAMED(ctx, name[N], (k_column<N, false>), grid, block, lds, ctx->stream, args, ctx->words,
                         skip_mask);
        return status(hipGetLastError());
    }
    // the split form: the matrices of every column into the context's buffer, then the tracers
    const size_t need = (size_t)ncol * N * N;
    if (ext && ext_n < need) return TM5_ESCRATCH;
    if (!ext && ctx->cbuf_n < need) {
        hipError_t e = hipSuccess;
        if (ctx->cbuf) {
            e = hipStreamSynchronize(ctx->stream);  // earlier launches may still read it
            if (e == hipSuccess) e = hipFree(ctx->cbuf);
            if (e != hipSuccess) return status(e);
            ctx->cbuf = nullptr;
            ctx->cbuf_n = 0;
        }
        e = hipMalloc((void **)&ctx->cbuf, need * sizeof(double));
        if (e != hipSuccess) return status(e);
        ctx->cbuf_n = need;
    }
    double *buf = ext ? ext : ctx->cbuf;
    ColumnArgs f = args;
    f.cbuf = buf;
    f.build_only = args.lapack;
    TM5_LAUNCH_NAMED(ctx, "k_column_factor", (k_column<N, false>), grid, block, lds, ctx->stream, f, ctx->words,
                     skip_mask);
    int rc = status(hipGetLastError());
    if (rc) return rc;
    if (args.lapack) {  // conv1 -> C in place, DGETRF2 + DGETRI per column
        rc = launch_lapack_inverse(ctx, ncol, N, buf, buf, nullptr, skip_mask);
        if (rc) return rc;
    }
    // tracer ranges enough for the tiles to fill the device's workgroup slots (16 waves a CU)
    const int nt = args.q.nt;
    if (nt < 1) return 0;  // no tracers: the matrices only
    const int cols = ctx->tile[TM5_TILE_APPLY];  // 8, 16 or 32 (tm5_ctx_set_tile)
    const int tiles = (args.im + 7 + cols - 1) / cols * args.jm;
    const int tchunk = tracer_chunk(tiles, 32 / cols * (ctx->ncu > 0 ? ctx->ncu : 256), nt,
                                    ctx->tile[TM5_TILE_RANGE_MIN]);
    const dim3 agrid((args.im + 7 + cols - 1) / cols, args.jm, (nt + tchunk - 1) / tchunk);
    const dim3 ablock(cols * 32);
    if (args.zpost || args.zpre) {  // the step's dynamw around the sequence too, then their masses
        const int zcols = cols < APPLY_Z_COLS ? cols : APPLY_Z_COLS;
        const int ztiles = (args.im + 7 + zcols - 1) / zcols * args.jm;
        const int zchunk = tracer_chunk(ztiles, 32 / zcols * (ctx->ncu > 0 ? ctx->ncu : 256), nt,
                                        ctx->tile[TM5_TILE_RANGE_MIN]);
        const dim3 zgrid((args.im + 7 + zcols - 1) / zcols, args.jm, (nt + zchunk - 1) / zchunk);
        TM5_LAUNCH_NAMED(ctx, "k_column_apply_z", k_column_apply_z<N>, zgrid, dim3(zcols * 32),
                         (size_t)zcols * APPLY_Z_LDS * sizeof(double), ctx->stream, args, buf, zchunk, ctx->words,
                         skip_mask);
        int rz = status(hipGetLastError());
        if (rz) return rz;
        const long long ncell = (long long)args.im * args.jm * args.lm;
        TM5_LAUNCH(ctx, k_z_mass_post, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, ctx->stream, args,
                   ctx->words, skip_mask);
        return status(hipGetLastError());
    }
    TM5_LAUNCH_NAMED(ctx, "k_column_apply", k_column_apply<N>, agrid, ablock, 0, ctx->stream, args, buf, tchunk,
                     ctx->words, skip_mask);
    return status(hipGetLastError());
}

// the split form's working set, in doubles, for lmax N: N x N a column
size_t column_matrix_doubles(int lmax) { return (size_t)lmax * (size_t)lmax; }

// the split form from this many tracers on; below it the one-kernel form, which needs no buffer
int column_split_nt() { return SPLIT_NT; }

// the lmax values compiled in register form (two columns a wave, four waves a workgroup); others
// go through the LDS kernel (convec.hip)
bool column_supported(int n)
{
    switch (n) {
    case 4: case 5: case 8: case 10: case 15: case 20: case 25: case 26:
        return true;
    default:
        return false;
    }
}

int launch_column(tm5_ctx *ctx, const ColumnArgs &args, int lmax, int skip_mask, double *ext, size_t ext_n)
{
    if (args.lm > 64 || lmax > args.lm) return TM5_EINVAL;
    switch (lmax) {
    case 4: return launch_column_n<4>(ctx, args, skip_mask, ext, ext_n);
    case 5: return launch_column_n<5>(ctx, args, skip_mask, ext, ext_n);
    case 8: return launch_column_n<8>(ctx, args, skip_mask, ext, ext_n);
    case 10: return launch_column_n<10>(ctx, args, skip_mask, ext, ext_n);
    case 15: return launch_column_n<15>(ctx, args, skip_mask, ext, ext_n);
    case 20: return launch_column_n<20>(ctx, args, skip_mask, ext, ext_n);
    case 25: return launch_column_n<25>(ctx, args, skip_mask, ext, ext_n);
    case 26: return launch_column_n<26>(ctx, args, skip_mask, ext, ext_n);
    default: return TM5_EINVAL;
    }
}

}  // namespace tm5
