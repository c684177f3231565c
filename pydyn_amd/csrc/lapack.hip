// The LAPACK-kind convection inverse: reference LAPACK's DGETRF (the recursive DGETRF2) and DGETRI
// with lwork = n (its unblocked sweep) on each column's matrix, as the reference calls them per
// column (pydyn/lapack.py:43-66) with the reference BLAS loop orders it transcribes for its
// device (pydyn/gpu/csrc/convec.cuh:17-206: DTRSM skips a zero right-hand side entry, DTRMV a
// zero x entry, DGEMM and DGEMV skip nothing).
//
// Two matrices per 64-lane wavefront, one per 32-lane half (N <= 26 < 32 rows), each matrix
// column-major in its half's LDS: every LDS instruction then serves two matrices, and a half
// never reads the other half's matrix, so the halves may take different branches (pivots, zero
// skips) without meeting.  The recursion of DGETRF2 depends only on the shape, so it is unrolled
// at compile time (lu2).  Every entry receives its updates in the reference's order: the loops
// run across lanes only where the reference's loop iterations touch different entries -- rows in
// DGEMM / DGEMV / DSCAL, columns in DTRSM (whose columns are independent) -- and every loop that
// carries an entry's sequence of updates stays inside one lane.  Interchanges (DLASWP's rows,
// DGETRI's columns) move no data until the result is stored.
#include "tm5_lib.h"

namespace tm5 {

namespace {

constexpr double SFMIN = 2.2250738585072014e-308;  // DLAMCH('S')

// the wave's LDS writes land before any lane's reads that follow (a matrix never spans waves)
__device__ __forceinline__ void lsync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x, computed where it stands: a select that takes it stays a select (no branch around its
// arithmetic), so the loads and products of a loop can be issued ahead of their uses.  The asm is
// empty and emits nothing.
__device__ __forceinline__ double pin(double x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// LANES lanes per matrix (lane: 0..LANES-1 within the matrix's half); MPG matrices a workgroup
// (two waves: 22 KB of LDS at N = 26)
constexpr int LANES = 32, MPG = 4;

// Row interchanges move no data.  perm[r] is the physical row (in the LDS matrix, leading
// dimension LD) of logical row r; DLASWP's interchange of rows k and p exchanges perm[k] and
// perm[p].  The reference interchanges rows of a block of columns only just before its
// arithmetic next reads them (DGETRF2 swaps a12 after the left half is factored and the left
// block after the right half), and no arithmetic reads a block between a pivot's choice and its
// interchange there, so every operation sees the operands it sees in the reference.
//
// DGETRF2 on the logical submatrix of rows R0..R0+M-1 and columns C0..C0+NN-1.  Returns info
// (0, or the 1-based row of the first exactly zero pivot in the whole matrix), the same on every
// lane of the half.
template <int LD, int M, int NN, int R0, int C0>
__device__ __forceinline__ int lu2(double *a, int *perm, int lane)
{
    if constexpr (M == 0 || NN == 0) {
        return 0;
    } else if constexpr (M == 1) {
        return a[perm[R0] + LD * C0] == 0.0 ? R0 + 1 : 0;
    } else if constexpr (NN == 1) {
        // the column's physical rows and values: every lane reads them all (broadcasts), and its own
        int ph[M];
        double v[M];
#pragma unroll
        for (int i = 0; i < M; i++) ph[i] = perm[R0 + i];
        const int own = perm[R0 + (lane < M ? lane : M - 1)];  // loads without a branch
#pragma unroll
        for (int i = 0; i < M; i++) v[i] = a[ph[i] + LD * C0];
        const double vown = a[own + LD * C0];
        // IDAMAX: the first entry of largest magnitude (a select chain: the trees a reduction would
        // take measured slower, their extra instructions outweighing the shorter chain)
        int p = 0, php = ph[0];
        double big = fabs(v[0]), piv = v[0];
#pragma unroll
        for (int i = 1; i < M; i++) {
            const bool more = fabs(v[i]) > big;
            big = more ? fabs(v[i]) : big;
            piv = more ? v[i] : piv;
            php = more ? ph[i] : php;
            p = more ? i : p;
        }
        lsync();  // every lane has read perm before it changes
        if (piv == 0.0) return R0 + 1;
        if (lane == 0 && p != 0) {  // logical rows R0 and R0 + p exchange their physical rows
            perm[R0] = php;
            perm[R0 + p] = ph[0];
        }
        // logical row R0 + lane after the interchange: row R0's old values if lane == p
        if (lane >= 1 && lane < M) {
            const int row = lane == p ? ph[0] : own;
            const double x = lane == p ? v[0] : vown;
            if (fabs(piv) >= SFMIN) {
                const double r = 1.0 / piv;
                a[row + LD * C0] = r * x;
            } else {
                a[row + LD * C0] = x / piv;
            }
        }
        lsync();
        return 0;
    } else {
        constexpr int MN = M < NN ? M : NN, N1 = MN / 2, N2 = NN - N1;
        int info = lu2<LD, M, N1, R0, C0>(a, perm, lane);
        int ph[N1];  // the physical rows of logical rows R0..R0+N1-1 (final for this node)
#pragma unroll
        for (int i = 0; i < N1; i++) ph[i] = perm[R0 + i];
        // DTRSM L L N U (N1 x N2): lane j solves column C0+N1+j with the unit lower L of the left
        // block's top rows, the column in registers; a zero b(k) skips its updates (a select: the
        // same values)
        if (lane < N2) {
            double *col = a + LD * (C0 + N1 + lane);
            double bv[N1];
#pragma unroll
            for (int i = 0; i < N1; i++) bv[i] = col[ph[i]];
#pragma unroll
            for (int k = 0; k < N1; k++) {
                const double bk = bv[k];
#pragma unroll
                for (int i = k + 1; i < N1; i++) {
                    const double u = pin(bv[i] - bk * a[ph[i] + LD * (C0 + k)]);
                    bv[i] = bk == 0.0 ? bv[i] : u;
                }
            }
#pragma unroll
            for (int i = 0; i < N1; i++) col[ph[i]] = bv[i];
        }
        lsync();
        // DGEMM N N, alpha -1: a22 -= a21 a12, each entry's terms in l order, entries across lanes
        // (an entry reads no other entry of a22, so the passes are independent: the lanes past the
        // last entry compute it again and store nothing)
        constexpr int ME = M - N1, TOT = ME * N2;
#pragma unroll
        for (int it = 0; it < (TOT + LANES - 1) / LANES; it++) {
            const int e0 = lane + LANES * it, e = e0 < TOT ? e0 : TOT - 1;
            const int i = e % ME, j = e / ME;
            const int pi = perm[R0 + N1 + i];
            double *cj = a + LD * (C0 + N1 + j);
            double c = cj[pi];
#pragma unroll
            for (int l = 0; l < N1; l++) {
                const double t = -1.0 * cj[ph[l]];
                c = c + t * a[pi + LD * (C0 + l)];
            }
            if (e0 < TOT) cj[pi] = c;
        }
        lsync();
        const int i2 = lu2<LD, M - N1, N2, R0 + N1, C0 + N1>(a, perm, lane);
        if (info == 0 && i2 > 0) info = i2;
        return info;
    }
}

// DGETRI (n = N, lwork = n) on the factors in a (leading dimension N, logical row order): DTRTI2
// and the sweep inv(A) L = inv(U) from the last column; its column interchanges are the caller's
// (they move data only).  Returns info.
template <int N>
__device__ __forceinline__ int inv_from_lu(double *a, int lane)  // inlined: a stays an LDS address
{
    int zero = 0;  // the first exactly zero diagonal entry (1-based), as DTRTRI's check
    double dinv[N];  // DTRTI2's 1 / a(j, j), all issued at once (no column before j changes a(j, j))
#pragma unroll
    for (int k = N - 1; k >= 0; k--) {
        const double d = a[k + N * k];
        zero = d == 0.0 ? k + 1 : zero;
        dinv[k] = d;
    }
    if (zero != 0) return zero;
#pragma unroll
    for (int k = 0; k < N; k++) dinv[k] = 1.0 / dinv[k];
    // DTRTI2 (upper, non-unit): column j, lane i its entry i < j, in registers.  The loops are
    // unrolled and a zero x(k) skips its step by a select, so every load of a column is issued
    // before the first is used.
#pragma unroll
    for (int j = 0; j < N; j++) {
        double *x = a + N * j;
        // the column's loads first: x(0:j) as it is (steps before k touch entries below k) and the
        // lane's row of the columns before it (a(k, k) on lanes >= k, unused past k)
        double xs[N > 1 ? N - 1 : 1], al[N > 1 ? N - 1 : 1];
#pragma unroll
        for (int k = 0; k < j; k++) {
            xs[k] = x[k];
            al[k] = a[(lane < k ? lane : k) + N * k];
        }
        const double inv = dinv[j];
        const double ajj = -inv;
        const double x0 = x[lane < j ? lane : 0];
        double xi = lane < j ? x0 : 0.0;
        // DTRMV U N N on x(0:j); a zero x(k) skips step k
#pragma unroll
        for (int k = 0; k < j; k++) {
            const double xk = xs[k], alk = al[k];
            const double sum = pin(xi + xk * alk), scl = pin(xi * alk);
            const bool below = lane < k && xk != 0.0, diag = lane == k && xk != 0.0;
            xi = below ? sum : (diag ? scl : xi);
        }
        if (lane < j) xi = ajj * xi;  // DSCAL
        lsync();  // every lane has read x before it is written
        if (lane < j) x[lane] = xi;
        if (lane == 0) x[j] = inv;
        lsync();
    }
    // the sweep: column j from the columns after it (final already), entries across lanes.  DGETRI
    // copies L's column j (below the diagonal) to work and zeroes it there, then adds
    // -work(k) a(:, k) over k > j; every lane reads the copied entries straight from the column.
#pragma unroll
    for (int j = N - 1; j >= 0; j--) {
        double *cj = a + N * j;
        if (lane < N) {
            const double own = cj[lane];
            double c = lane > j ? 0.0 : own;
#pragma unroll
            for (int k = j + 1; k < N; k++) {  // DGEMV N, alpha -1, beta one
                const double t = -1.0 * cj[k];
                c = c + t * a[lane + N * k];
            }
            lsync();  // every lane has read column j before it is written
            cj[lane] = c;
        }
        lsync();
    }
    return 0;
}

}  // namespace

// ncol row-major N x N matrices A to their inverses Y (A == Y allowed): MPG matrices a workgroup
// of MPG / 2 waves, one matrix a 32-lane half.  A zero pivot sets *info to 1 (info null:
// TM5_SINGULAR in the flag word); that matrix's Y is what the factorization left.
template <int N>
__global__ void __launch_bounds__(MPG * LANES) k_lapack_inverse(long long ncol, const double *A, double *Y,
                                                                 int *info_out, int *words, int skip_mask)
{
    static_assert(N <= LANES, "a matrix's rows are the lanes of a half wave");
    if ((words[W_FLAGS] & skip_mask) != 0) return;
    __shared__ double lds[MPG][N * N];
    __shared__ int perms[MPG][N];
    const int slot = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    const long long q = (long long)blockIdx.x * MPG + slot;
    if (q >= ncol) return;  // the whole half: no workgroup barrier follows
    double *a = lds[slot];
    int *perm = perms[slot];
    const double *src = A + q * N * N;
    if (lane < N) {
        for (int c = 0; c < N; c++) a[lane + N * c] = src[lane * N + c];
        perm[lane] = lane;
    }
    lsync();
    int info = lu2<N, N, N, 0, 0>(a, perm, lane);
    {  // the factors into logical row order: lane c gathers column c
        int ph[N];
        double v[N];
#pragma unroll
        for (int r = 0; r < N; r++) ph[r] = perm[r];
        if (lane < N) {
#pragma unroll
            for (int r = 0; r < N; r++) v[r] = a[ph[r] + N * lane];
        }
        lsync();
        if (lane < N) {
#pragma unroll
            for (int r = 0; r < N; r++) a[r + N * lane] = v[r];
        }
        lsync();
    }
    if (info == 0) info = inv_from_lu<N>(a, lane);
    if (info != 0 && lane == 0) {
        if (info_out)
            atomicOr(info_out, 1);
        else
            atomicOr(words + W_FLAGS, TM5_SINGULAR);
    }
    double *dst = Y + q * N * N;
    if (lane < N) {
        // DGETRI's column interchanges (j from n-1 down, columns j and ipiv(j)) move column i of
        // the sweep's result to column perm[i]; a failed inverse leaves the factors in place
#pragma unroll
        for (int c = 0; c < N; c++) dst[lane * N + (info == 0 ? perm[c] : c)] = a[lane + N * c];
    }
}

int launch_lapack_inverse(tm5_ctx *ctx, long long ncol, int n, const double *A, double *Y, int *info, int skip_mask)
{
    if (ncol < 0 || !A || !Y) return TM5_EINVAL;
    if (ncol == 0) return 0;
    const dim3 grid((unsigned)((ncol + MPG - 1) / MPG)), block(MPG * LANES);
    switch (n) {
#define TM5_LAPACK_CASE(K)                                                                                        \
    case K:                                                                                                       \
        TM5_LAUNCH_NAMED(ctx, "k_lapack_inverse", k_lapack_inverse<K>, grid, block, 0, ctx->stream, ncol, \
                         A, Y, info, ctx->words, skip_mask);                                                      \
        break;
        TM5_LAPACK_CASE(4)
        TM5_LAPACK_CASE(5)
        TM5_LAPACK_CASE(8)
        TM5_LAPACK_CASE(10)
        TM5_LAPACK_CASE(15)
        TM5_LAPACK_CASE(20)
        TM5_LAPACK_CASE(25)
        TM5_LAPACK_CASE(26)
#undef TM5_LAPACK_CASE
    default:
        return TM5_EINVAL;
    }
    return status(hipGetLastError());
}

}  // namespace tm5
