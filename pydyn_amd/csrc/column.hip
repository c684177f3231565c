// Column physics in registers: convection (convec) and the fused middle of a dynamic step.
//
// Between the two 'v' operators of a step ("xyz v s c | m2phlb1 | c s v zyx") only the sources,
// the chemistry and m2phlb1 run; none of them changes the air mass, so both convection calls
// build the same matrix and invert it to the same bits.  k_column<N> builds conv1 once per
// column, inverts it, and applies one of the sequences "v", "csv", "vsc", "vsccsv" or
// "vscMcsv" (M = m2phlb1) to every tracer of the column in one pass; in a zoom tree only on the
// columns the region owns.
//
// One column per 64-lane wavefront, N = lmax known at compile time so that the matrix lives in
// registers:
//   build      lane c owns matrix column c (the up/downdraft recursions, pydyn/kernels.py:350-396),
//              through LDS;
//   LU         lane r holds row r; right-looking Crout without pivoting, each element updated in
//              increasing k as in the reference's left-looking loops (pydyn/kernels.py:404-407);
//   inverse    lane c solves for the right-hand side e_c (forward, then back substitution in the
//              reference's order, pydyn/kernels.py:409-419);
//   apply      lane l holds row l of C: rm(l) = sum_k C(l,k) rm(k) from zero in increasing k,
//              slopes times C(l,l) (pydyn/kernels.py:448-457).
// Uniform values (a pivot, an L or U entry) come from the owning lane by readlane.
//
// The matrix of a column with cloud top T and level of free sinking F is dense in its leading
// DB x DB block (DB = max(T, F) + 1) and tridiagonal outside it; the entries outside are exact
// zeros (the reference's where(act, .., 0.0) and 0 / m).  LU fill-in stays inside that envelope,
// so the products outside it all have a +-0 factor and are skipped: subtracting +-0 leaves
// every non-zero value as it is and can only change the sign of a zero, and no zero's sign
// reaches a result (zeros only ever enter products, sums with non-zero terms, or numerators),
// so the outputs are the reference's bits for any non-singular matrix of finite entries.
#include "tm5_lib.h"

namespace tm5 {

__device__ __forceinline__ bool skipped_k(const int *words, int mask)
{
    return (words[W_FLAGS] & mask) != 0;  // written by earlier kernels of the stream
}

__device__ __forceinline__ double bcast(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}


constexpr int VEC = 64 + 13 * 32;  // per-level vectors of a wave's LDS

// m2phlb1 of one column (pydyn/zoom/massflow.py:24-32): every lane holds m of its level in m_l
__device__ __forceinline__ void column_m2phlb1(const ColumnArgs &a, int i, int j, double m_l, const double *mc)
{
    double p = bcast(m_l, 0);
    for (int l = 1; l < a.lm; l++) p = p + mc[l];
    p = p * 9.80665 / a.dxyp[j];
    const int lane = threadIdx.x & 63;
    if (lane == 0) a.sp(i, j, 0) = p;
    for (int l = lane; l <= a.lm; l += 64) a.phlb(i, j, l) = a.at[l] + a.bt[l] * p;
}

// ADJ: the transposed sequence on lambdas (a.q), C^T from the columns of C each lane holds, the
// emission's lambda accumulated into a.lam_emis.
template <int N, bool ADJ>
__global__ void __launch_bounds__(256) k_column(ColumnArgs a, int *words, int skip_mask)
{
    if (skipped_k(words, skip_mask)) return;
    constexpr int LD = (N % 2 == 0) ? N + 1 : N;
    extern __shared__ double lds[];
    // the column is uniform over the wavefront: said so, the branches on it are scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int ncol = a.im * a.jm;
    int col = xcd_tile(blockIdx.x, gridDim.x) * 4 + wave;
    const bool active = col < ncol;
    if (!active) col = ncol - 1;
    const int i = col % a.im, j = col / a.im;
    const int lm = a.lm;
    // per wave: fu / conv1 / C^T (N+1)*LD, fd (N+1)*LD, vectors
    static_assert(N < 32, "the per-level vectors hold 32 levels");
    double *U = lds + (size_t)wave * (2 * (N + 1) * LD + VEC);
    double *D = U + (N + 1) * LD;
    double *vec = D + (N + 1) * LD;
    double *mc = vec;  // all lm levels (m2phlb1)
    double *eu = vec + 64, *du = eu + 32, *ed = du + 32, *dd = ed + 32, *dk = dd + 32;
    double *amu = dk + 32, *amd = amu + 32, *eum = amd + 32, *edm = eum + 32;
    double *cu = edm + 32, *cd = cu + 32, *cdl = cd + 32, *cdk = cdl + 32;

    const double m_l = lane < lm ? a.m(i, j, lane) : 0.0;
    mc[lane] = m_l;
    // a column the region does not own is computed like the others (every wave reaches every
    // barrier) but keeps its tracers; the middle m2phlb1 is every column's
    const bool own = a.owned == nullptr || a.owned[(long long)j * a.im + i] != 0;
    for (int x = lane; x < (N + 1) * LD; x += 64) U[x] = D[x] = 0.0;
    if (lane < 32) amu[lane] = amd[lane] = 0.0;
    __syncthreads();
    if (lane < N) {
        const double e_u = a.entu(i, j, lane), e_d = a.entd(i, j, lane);
        eu[lane] = e_u;
        du[lane] = a.detu(i, j, lane);
        ed[lane] = e_d;
        dd[lane] = a.detd(i, j, lane);
        dk[lane] = a.dkg(i, j, lane);
        eum[lane] = e_u / m_l;   // the fu(k, k-1) and fd(k-1, k-1) quotients, every level at once
        edm[lane] = -e_d / m_l;
    }
    __syncthreads();

    // ---- build conv1: lane c owns column c (pydyn/kernels.py:366-395)
    const int ctop = __builtin_amdgcn_readfirstlane(a.top(i, j)), clfs = __builtin_amdgcn_readfirstlane(a.lfs(i, j));
    // levels above the cloud top (updraft) or the level of free sinking (downdraft) contribute
    // exact zeros (the reference's where(act, .., 0.0)): their iterations are skipped
    const int kup = ctop < N ? ctop : N, kdn = clfs < N ? clfs : N;
    // the dense leading block of conv1; outside it the matrix is tridiagonal
    const int DB = __builtin_amdgcn_readfirstlane((kup > kdn ? kup : kdn) + 1 < N ? (kup > kdn ? kup : kdn) + 1 : N);
    {
        double run = 0.0, prev = 0.0;
        for (int k = 1; k <= kup; k++) {
            const bool on = ctop >= k;
            const double inflow = prev + eu[k - 1];
            const double aa = inflow - du[k - 1];
            const bool rising = aa > 0.0;
            const double keep = rising ? max0(1.0 - du[k - 1] / inflow) : 0.0;
            const double amu_k = (on && rising) ? aa : 0.0;
            if (lane < N) {
                double v = 0.0;
                if (lane < k - 1)
                    v = on ? run * keep : 0.0;
                else if (lane == k - 1)
                    v = on ? eum[k - 1] * keep : 0.0;
                U[k * LD + lane] = v;
                run = v;
            }
            if (lane == 0) amu[k] = amu_k;
            prev = amu_k;
        }
    }
    {
        double run = 0.0, prev = 0.0;
        for (int k = kdn; k >= 2; k--) {
            const bool on = clfs >= k;
            const double inflow = prev - ed[k - 1];
            const double aa = inflow + dd[k - 1];
            const bool sinking = aa < 0.0;
            const double keep = sinking ? max0(1.0 + dd[k - 1] / inflow) : 0.0;
            const double amd_k = (on && sinking) ? aa : 0.0;
            if (lane < N) {
                double v = 0.0;
                if (lane >= k)
                    v = on ? run * keep : 0.0;
                else if (lane == k - 1)
                    v = on ? edm[k - 1] * keep : 0.0;
                D[(k - 1) * LD + lane] = v;
                run = v;
            }
            if (lane == 0) amd[k - 1] = amd_k;
            prev = amd_k;
        }
    }
    __syncthreads();
    // the four corrections of interface k (pydyn/kernels.py:388-391), every k at once in lane k
    {
        double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
        if (lane >= 1 && lane < N) {
            q0 = amu[lane] / mc[lane];
            q1 = amd[lane] / mc[lane - 1];
            q2 = dk[lane - 1] / mc[lane - 1];
            q3 = dk[lane - 1] / mc[lane];
        }
        if (lane >= 1 && lane < N) {
            cu[lane] = q0;
            cd[lane] = q1;
            cdl[lane] = q2;
            cdk[lane] = q3;
        }
        __syncthreads();
    }
    if (lane < N) {
        for (int k = 1; k < N; k++) {
            double v = U[k * LD + lane] + D[k * LD + lane];
            if (lane == k) v = v - cu[k];
            if (lane == k - 1) v = v - cd[k];
            if (lane == k - 1) v = v + cdl[k];
            if (lane == k) v = v - cdk[k];
            D[k * LD + lane] = v;
        }
    }
    __syncthreads();
    // conv1 row r into lane r's registers
    double A[N];
#pragma unroll
    for (int c = 0; c < N; c++) {
        const int r = lane < N ? lane : N - 1;
        double v = (-a.dt) * (D[r * LD + c] - D[(r + 1) * LD + c]);
        if (r == c) v = v + 1.0;
        A[c] = v;
    }

    // ---- LU, lanes over rows.  Entry (r, c) is non-zero only inside the dense block (r, c < DB)
    // or on the three diagonals, and LU fill-in stays inside that envelope; the products skipped
    // outside it have a factor that is exactly +-0.
    // Rows r <= k take the update with a zero multiplier: r - (+-0) is r up to a zero's sign.
#pragma unroll
    for (int k = 0; k < N; k++) {
        const double piv = bcast(A[k], k);
        const double q = A[k] / piv;
        A[k] = lane > k ? q : A[k];
        const double lrk = lane > k ? q : 0.0;
        if (k + 1 < N) A[k + 1] = A[k + 1] - lrk * bcast(A[k + 1], k);
#pragma unroll
        for (int c = k + 2; c < N; c++) {
            if (c < DB) A[c] = A[c] - lrk * bcast(A[c], k);
        }
    }

    // ---- inverse: lane c solves L U x = e_c
    double y[N];
#pragma unroll
    for (int r = 0; r < N; r++) y[r] = (lane == r) ? 1.0 : 0.0;
#pragma unroll
    for (int r = 1; r < N; r++) {
        if (r < DB) {
#pragma unroll
            for (int s = 0; s < r - 1; s++) y[r] = y[r] - bcast(A[s], r) * y[s];
        }
        y[r] = y[r] - bcast(A[r - 1], r) * y[r - 1];
    }
#pragma unroll
    for (int r = N - 1; r >= 0; r--) {
        double acc = y[r];
        if (r + 1 < N) acc = acc - bcast(A[r + 1], r) * y[r + 1];
#pragma unroll
        for (int s = r + 2; s < N; s++) {
            if (s < DB) acc = acc - bcast(A[s], r) * y[s];
        }
        y[r] = acc / bcast(A[r], r);
    }
    // C^T through LDS: lane c wrote column c; lane l reads row l
    __syncthreads();
    if (lane < N) {
#pragma unroll
        for (int r = 0; r < N; r++) U[r * LD + lane] = y[r];
    }
    __syncthreads();
    double crow[ADJ ? 1 : N];
    if (!ADJ) {
        const int r = lane < N ? lane : N - 1;
#pragma unroll
        for (int k = 0; k < N; k++) crow[k] = U[r * LD + k];
    }
    const double diag = lane < N ? U[lane * LD + lane] : 0.0;

    // ---- apply: the flagged sequence around v for each tracer of the column
    const bool sc = a.pre || a.twice || a.post;
    double decay = 1.0, x = 0.0, emis_ij = 0.0;
    if (sc) {
        if (lane < lm) {
            const double *L = &a.loss(i, j, lane);
            const double total = ((L[0] + L[a.loss_sr]) + L[2 * a.loss_sr]) + L[3 * a.loss_sr];
            decay = fdlibm_exp(-(total * a.dt_chem));
        }
        emis_ij = a.emis(i, j, 0);
        x = ((emis_ij * a.factor) * a.ratio) * a.efrac;
    }
    double lam_e = 0.0;  // ADJ: this lane's share of the emission's lambda
    for (int t = 0; t < a.q.nt; t++) {
        const long long o = a.q.off(i, j, lane < lm ? lane : 0, t);
        double rm = 0.0, rx = 0.0, ry = 0.0, rz = 0.0;
        if (lane < lm) {
            rm = a.q.q[0][o];
            rx = a.q.q[1][o];
            ry = a.q.q[2][o];
            rz = a.q.q[3][o];
        }
        // do_add_2d: rm + x, rzm - x at the level (pydyn/zoom/physics.py:55-64); an ensemble member's
        // emission field is its scale times the region's, scaled first
        bool here = lane == a.level && t == a.tracer;
        double xt = x;
        if (a.escale) {
            const double es = a.escale[t];
            here = lane == a.level && es != 0.0;
            xt = (((emis_ij * es) * a.factor) * a.ratio) * a.efrac;
        }
        auto s_op = [&]() {
            if (here) {
                rm = rm + xt;
                rz = rz - xt;
            }
        };
        auto c_op = [&]() {  // the decay of every moment
            rm = rm * decay;
            rx = rx * decay;
            ry = ry * decay;
            rz = rz * decay;
        };
        auto v_op = [&]() {  // rm <- C rm over levels < N (lane k: C^T, its column), slopes times diag(C)
            double acc = 0.0;
            if (ADJ) {
#pragma unroll
                for (int k = 0; k < N; k++) acc = acc + y[k] * bcast(rm, k);
            } else {
#pragma unroll
                for (int k = 0; k < N; k++) acc = acc + crow[ADJ ? 0 : k] * bcast(rm, k);
            }
            if (lane < N) {
                rm = acc;
                rx = diag * rx;
                ry = diag * ry;
                rz = diag * rz;
            }
        };
        if (ADJ) {  // the transposes in reverse: v^T; s^T adds (rm - rzm) times dx/demis at the level
            const double es = a.escale ? a.escale[t] : 1.0;
            const double kt = ((es * a.factor) * a.ratio) * a.efrac;
            auto s_ad = [&]() {
                if (here) lam_e = lam_e + kt * (rm - rz);
            };
            if (a.post) {
                c_op();
                s_ad();
            }
            if (a.twice) {
                v_op();
                s_ad();
                c_op();
                c_op();
                s_ad();
            }
            v_op();
            if (a.pre) {
                s_ad();
                c_op();
            }
        } else {
        if (a.pre) {
            c_op();
            s_op();
        }
        v_op();
        if (a.twice) {
            s_op();
            c_op();
            c_op();
            s_op();
            v_op();
        }
        if (a.post) {
            s_op();
            c_op();
        }
        }
        if (active && own && lane < lm) {
            a.q.q[0][o] = rm;
            a.q.q[1][o] = rx;
            a.q.q[2][o] = ry;
            a.q.q[3][o] = rz;
        }
    }
    if (ADJ) {
        if (active && own && lane == a.level) a.lam_emis(i, j, 0) = a.lam_emis(i, j, 0) + lam_e;
        return;
    }
    // m2phlb1 between the two chemistry calls (pydyn/zoom/massflow.py:282-290)
    if (a.mid && active) column_m2phlb1(a, i, j, m_l, mc);
}

template <int N>
static int launch_column_n(tm5_ctx *ctx, const ColumnArgs &args, int skip_mask)
{
    constexpr int LD = (N % 2 == 0) ? N + 1 : N;
    const size_t per_wave = (size_t)(2 * (N + 1) * LD + VEC) * sizeof(double);
    const int ncol = args.im * args.jm;
    static const char *name[] = {"k_column<0>", "k_column<1>", "k_column<2>", "k_column<3>", "k_column<4>",
                                 "k_column<5>", "k_column<6>", "k_column<7>", "k_column<8>", "k_column<9>",
                                 "k_column<10>", "k_column<11>", "k_column<12>", "k_column<13>", "k_column<14>",
                                 "k_column<15>", "k_column<16>", "k_column<17>", "k_column<18>", "k_column<19>",
                                 "k_column<20>", "k_column<21>", "k_column<22>", "k_column<23>", "k_column<24>",
                                 "k_column<25>", "k_column<26>", "k_column<27>", "k_column<28>", "k_column<29>",
                                 "k_column<30>", "k_column<31>"};
    if (args.adjoint) {
        TM5_LAUNCH_NAMED(ctx, "k_column_adjoint", (k_column<N, true>), dim3((ncol + 3) / 4), dim3(256), 4 * per_wave,
                         ctx->stream, args, ctx->words, skip_mask);
        return status(hipGetLastError());
    }
    TM5_LAUNCH_NAMED(ctx, name[N], (k_column<N, false>), dim3((ncol + 3) / 4), dim3(256), 4 * per_wave, ctx->stream, args, ctx->words,
                       skip_mask);
    return status(hipGetLastError());
}

// the lmax values compiled in register form (a wave's LDS stays under 16 KB, four waves a
// workgroup); others go through the LDS kernel (convec.hip)
bool column_supported(int n)
{
    switch (n) {
    case 4: case 5: case 8: case 10: case 15: case 20: case 25: case 26:
        return true;
    default:
        return false;
    }
}

int launch_column(tm5_ctx *ctx, const ColumnArgs &args, int lmax, int skip_mask)
{
    if (args.lm > 64 || lmax > args.lm) return TM5_EINVAL;
    switch (lmax) {
    case 4: return launch_column_n<4>(ctx, args, skip_mask);
    case 5: return launch_column_n<5>(ctx, args, skip_mask);
    case 8: return launch_column_n<8>(ctx, args, skip_mask);
    case 10: return launch_column_n<10>(ctx, args, skip_mask);
    case 15: return launch_column_n<15>(ctx, args, skip_mask);
    case 20: return launch_column_n<20>(ctx, args, skip_mask);
    case 25: return launch_column_n<25>(ctx, args, skip_mask);
    case 26: return launch_column_n<26>(ctx, args, skip_mask);
    default: return TM5_EINVAL;
    }
}

}  // namespace tm5
