// The LAPACK-kind convection inverse: reference LAPACK's DGETRF (the recursive DGETRF2) and DGETRI
// with lwork = n (its unblocked sweep) on each column's matrix, as the reference calls them per
// column (pydyn/lapack.py:43-66) with the reference BLAS loop orders it transcribes for its
// device (pydyn/gpu/csrc/convec.cuh:17-206: DTRSM skips a zero right-hand side entry, DTRMV a
// zero x entry, DGEMM and DGEMV skip nothing).
//
// One 64-lane wavefront per matrix, the matrix column-major in the wave's LDS.  The recursion of
// DGETRF2 depends only on the shape, so it is unrolled at compile time (lu2<M, N>).  Every entry
// receives its updates in the reference's order: the loops run across lanes only where the
// reference's loop iterations touch different entries -- rows in DGEMM / DGEMV / DSCAL / the
// column interchanges, columns in DLASWP and DTRSM (whose columns are independent) -- and every
// loop that carries an entry's sequence of updates stays inside one lane.
#include "tm5_lib.h"

namespace tm5 {

namespace {

constexpr double SFMIN = 2.2250738585072014e-308;  // DLAMCH('S')

// the wave's LDS writes land before any lane's reads that follow (one wave per matrix)
__device__ __forceinline__ void lsync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DLASWP(ncols, a, lda, k1, k2, ipiv, 1), 1-based k1..k2: lane c swaps within column c
__device__ __forceinline__ void lane_swaps(int ncols, double *a, int lda, int k1, int k2, const int *ipiv, int lane)
{
    for (int k = k1; k <= k2; k++) {
        const int p = ipiv[k - 1];
        if (p != k && lane < ncols) {
            double *col = a + (long long)lda * lane;
            const double t = col[k - 1];
            col[k - 1] = col[p - 1];
            col[p - 1] = t;
        }
    }
    lsync();
}

// DGETRF2 on the M x NN matrix at a (leading dimension lda); pivots 1-based into ipiv.  Returns
// info (0, or the 1-based index of the first exactly zero pivot), the same on every lane.
template <int M, int NN>
__device__ int lu2(double *a, int lda, int *ipiv, int lane)
{
    if constexpr (M == 0 || NN == 0) {
        return 0;
    } else if constexpr (M == 1) {
        if (lane == 0) ipiv[0] = 1;
        lsync();
        return a[0] == 0.0 ? 1 : 0;
    } else if constexpr (NN == 1) {
        // IDAMAX on every lane (the first entry of largest magnitude): the same answer everywhere
        int p = 0;
        double big = fabs(a[0]);
        for (int i = 1; i < M; i++) {
            const double v = fabs(a[i]);
            if (v > big) {
                big = v;
                p = i;
            }
        }
        const double piv = a[p];
        lsync();  // every lane has read the column before it changes
        if (lane == 0) ipiv[0] = p + 1;
        if (piv == 0.0) {
            lsync();
            return 1;
        }
        if (lane == 0 && p != 0) {
            const double t = a[0];
            a[0] = a[p];
            a[p] = t;
        }
        lsync();
        const double a0 = a[0];
        if (lane >= 1 && lane < M) {
            if (fabs(a0) >= SFMIN) {
                const double r = 1.0 / a0;
                a[lane] = r * a[lane];
            } else {
                a[lane] = a[lane] / a0;
            }
        }
        lsync();
        return 0;
    } else {
        constexpr int MN = M < NN ? M : NN, N1 = MN / 2, N2 = NN - N1;
        int info = lu2<M, N1>(a, lda, ipiv, lane);
        double *a12 = a + (long long)lda * N1, *a21 = a + N1, *a22 = a + N1 + (long long)lda * N1;
        lane_swaps(N2, a12, lda, 1, N1, ipiv, lane);
        // DTRSM L L N U (N1 x N2): lane j solves column j of a12 with the unit lower a(0:N1, 0:N1)
        if (lane < N2) {
            double *b = a12 + (long long)lda * lane;
            for (int k = 0; k < N1; k++) {
                const double bk = b[k];
                if (bk == 0.0) continue;
                for (int i = k + 1; i < N1; i++) b[i] = b[i] - bk * a[i + (long long)lda * k];
            }
        }
        lsync();
        // DGEMM N N, alpha -1: a22 -= a21 a12, each entry's terms in l order, entries across lanes
        for (int e = lane; e < (M - N1) * N2; e += 64) {
            const int i = e % (M - N1), j = e / (M - N1);
            double c = a22[i + (long long)lda * j];
            for (int l = 0; l < N1; l++) {
                const double t = -1.0 * a12[l + (long long)lda * j];
                c = c + t * a21[i + (long long)lda * l];
            }
            a22[i + (long long)lda * j] = c;
        }
        lsync();
        const int i2 = lu2<M - N1, N2>(a22, lda, ipiv + N1, lane);
        if (info == 0 && i2 > 0) info = i2 + N1;
        if (lane >= N1 && lane < MN) ipiv[lane] = ipiv[lane] + N1;
        lsync();
        lane_swaps(N1, a, lda, N1 + 1, MN, ipiv, lane);
        return info;
    }
}

// DGETRI (n = N, lwork = n) on the factors in a (leading dimension N): DTRTI2, the sweep
// inv(A) L = inv(U) from the last column, the column interchanges.  Returns info.
template <int N>
__device__ int inv_from_lu(double *a, const int *ipiv, double *work, int lane)
{
    for (int k = 0; k < N; k++)
        if (a[k + N * k] == 0.0) return k + 1;
    // DTRTI2 (upper, non-unit): column j, lane i its entry i < j, in registers
    for (int j = 0; j < N; j++) {
        double *x = a + N * j;
        const double inv = 1.0 / x[j];
        const double ajj = -inv;
        double xi = lane < j ? x[lane] : 0.0;
        // DTRMV U N N on x(0:j): step k reads x[k] as it still is (steps before k touch entries below k)
        for (int k = 0; k < j; k++) {
            const double xk = x[k];
            if (xk == 0.0) continue;
            if (lane < k) xi = xi + xk * a[lane + N * k];
            if (lane == k) xi = xi * a[k + N * k];
        }
        if (lane < j) xi = ajj * xi;  // DSCAL
        lsync();  // every lane has read x before it is written
        if (lane < j) x[lane] = xi;
        if (lane == 0) x[j] = inv;
        lsync();
    }
    // the sweep: column j from the columns after it (final already), entries across lanes
    for (int j = N - 1; j >= 0; j--) {
        double *cj = a + N * j;
        if (lane > j && lane < N) {
            work[lane] = cj[lane];
            cj[lane] = 0.0;
        }
        lsync();
        if (lane < N) {
            double c = cj[lane];
            for (int k = j + 1; k < N; k++) {  // DGEMV N, alpha -1, beta one
                const double t = -1.0 * work[k];
                c = c + t * a[lane + N * k];
            }
            cj[lane] = c;
        }
        lsync();
    }
    // the column interchanges, last first: lane i swaps within row i
    for (int j = N - 2; j >= 0; j--) {
        const int p = ipiv[j] - 1;
        if (p != j && lane < N) {
            const double t = a[lane + N * j];
            a[lane + N * j] = a[lane + N * p];
            a[lane + N * p] = t;
        }
    }
    lsync();
    return 0;
}

}  // namespace

// ncol row-major N x N matrices A to their inverses Y (A == Y allowed): four waves a workgroup,
// one matrix a wave.  A zero pivot sets *info to 1 (info null: TM5_SINGULAR in the flag word);
// that matrix's Y is what the factorization left.
template <int N>
__global__ void __launch_bounds__(256) k_lapack_inverse(long long ncol, const double *A, double *Y, int *info_out,
                                                        int *words, int skip_mask)
{
    if ((words[W_FLAGS] & skip_mask) != 0) return;
    __shared__ double lds[4][N * N + N];
    __shared__ int piv[4][N];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long q = (long long)blockIdx.x * 4 + wave;
    if (q >= ncol) return;  // the whole wave: no workgroup barrier follows
    double *a = lds[wave], *work = a + N * N;
    int *ipiv = piv[wave];
    const double *src = A + q * N * N;
    if (lane < N)
        for (int c = 0; c < N; c++) a[lane + N * c] = src[lane * N + c];
    lsync();
    int info = lu2<N, N>(a, N, ipiv, lane);
    if (info == 0) info = inv_from_lu<N>(a, ipiv, work, lane);
    if (info != 0 && lane == 0) {
        if (info_out)
            atomicOr(info_out, 1);
        else
            atomicOr(words + W_FLAGS, TM5_SINGULAR);
    }
    lsync();
    double *dst = Y + q * N * N;
    if (lane < N)
        for (int c = 0; c < N; c++) dst[lane * N + c] = a[lane + N * c];
}

int launch_lapack_inverse(tm5_ctx *ctx, long long ncol, int n, const double *A, double *Y, int *info, int skip_mask)
{
    if (ncol < 0 || !A || !Y) return TM5_EINVAL;
    if (ncol == 0) return 0;
    const dim3 grid((unsigned)((ncol + 3) / 4)), block(256);
    switch (n) {
#define TM5_LAPACK_CASE(K)                                                                                        \
    case K:                                                                                                       \
        TM5_LAUNCH_NAMED(ctx, "k_lapack_inverse", k_lapack_inverse<K>, grid, block, 0, ctx->stream, ncol, A, Y,  \
                         info, ctx->words, skip_mask);                                                                  \
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
