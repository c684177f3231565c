/* The oracle's LAPACK-kind inverse: reference LAPACK's DGETRF (recursive DGETRF2) and DGETRI
   called with lwork = n (its unblocked sweep), as the reference runs them per column
   (pydyn/lapack.py:43-66, LapackInverse) and transcribes them for its device
   (pydyn/gpu/csrc/convec.cuh:17-206: reference BLAS loop orders; DTRSM skips a zero right-hand
   side entry, DTRMV a zero x entry, DGEMM and DGEMV skip nothing).  Restated from netlib's
   published DGETRF2 / DGETRI / DTRTI2 and the reference BLAS, in this file's own code.

   Test infrastructure only.  Matrices are column-major n x n with leading dimension lda;
   pivots 1-based, as LAPACK's. */
#include <math.h>

#define A_(a, lda, i, j) ((a)[(i) + (long)(lda) * (j)])

static const double SFMIN = 2.2250738585072014e-308; /* DLAMCH('S') */

/* IDAMAX, 0-based: the first entry of largest magnitude */
static int amax_index(int n, const double *x)
{
    int best = 0;
    double big = fabs(x[0]);
    for (int i = 1; i < n; i++) {
        const double v = fabs(x[i]);
        if (v > big) {
            big = v;
            best = i;
        }
    }
    return best;
}

/* DLASWP(ncols, a, lda, k1, k2, ipiv, 1): rows k1..k2 (1-based) swapped with their pivots */
static void swap_rows(int ncols, double *a, int lda, int k1, int k2, const int *ipiv)
{
    for (int k = k1; k <= k2; k++) {
        const int p = ipiv[k - 1];
        if (p == k) continue;
        for (int c = 0; c < ncols; c++) {
            const double t = A_(a, lda, k - 1, c);
            A_(a, lda, k - 1, c) = A_(a, lda, p - 1, c);
            A_(a, lda, p - 1, c) = t;
        }
    }
}

/* B := inv(L) B, L unit lower m x m (DTRSM L L N U, alpha one), column by column */
static void lower_solve(int m, int ncols, const double *l, int lda, double *b, int ldb)
{
    for (int c = 0; c < ncols; c++)
        for (int k = 0; k < m; k++) {
            const double bk = A_(b, ldb, k, c);
            if (bk == 0.0) continue;
            for (int i = k + 1; i < m; i++) A_(b, ldb, i, c) = A_(b, ldb, i, c) - bk * A_(l, lda, i, k);
        }
}

/* C := C - A B (DGEMM N N, alpha -1, beta one): per column of C, the terms in k order */
static void sub_product(int m, int ncols, int kk, const double *a, int lda, const double *b, int ldb, double *c,
                        int ldc)
{
    for (int j = 0; j < ncols; j++)
        for (int k = 0; k < kk; k++) {
            const double t = -1.0 * A_(b, ldb, k, j);
            for (int i = 0; i < m; i++) A_(c, ldc, i, j) = A_(c, ldc, i, j) + t * A_(a, lda, i, k);
        }
}

/* DGETRF2: info 0, or the 1-based index of the first exactly zero pivot */
int oracle_dgetrf2(int m, int n, double *a, int lda, int *ipiv)
{
    if (m == 0 || n == 0) return 0;
    if (m == 1) {
        ipiv[0] = 1;
        return a[0] == 0.0 ? 1 : 0;
    }
    if (n == 1) {
        const int p = amax_index(m, a);
        ipiv[0] = p + 1;
        if (a[p] == 0.0) return 1;
        if (p != 0) {
            const double t = a[0];
            a[0] = a[p];
            a[p] = t;
        }
        if (fabs(a[0]) >= SFMIN) {
            const double r = 1.0 / a[0];
            for (int i = 1; i < m; i++) a[i] = r * a[i];
        } else {
            for (int i = 1; i < m; i++) a[i] = a[i] / a[0];
        }
        return 0;
    }
    const int mn = m < n ? m : n, n1 = mn / 2, n2 = n - n1;
    int info = oracle_dgetrf2(m, n1, a, lda, ipiv);
    double *a12 = &A_(a, lda, 0, n1), *a21 = &A_(a, lda, n1, 0), *a22 = &A_(a, lda, n1, n1);
    swap_rows(n2, a12, lda, 1, n1, ipiv);
    lower_solve(n1, n2, a, lda, a12, lda);
    sub_product(m - n1, n2, n1, a21, lda, a12, lda, a22, lda);
    const int i2 = oracle_dgetrf2(m - n1, n2, a22, lda, ipiv + n1);
    if (info == 0 && i2 > 0) info = i2 + n1;
    for (int k = n1; k < mn; k++) ipiv[k] += n1;
    swap_rows(n1, a, lda, n1 + 1, mn, ipiv);
    return info;
}

/* DGETRI with lwork = n: DTRTI2 on U, then inv(A) L = inv(U) column by column from the last,
   then the column interchanges in reverse; work holds n doubles.  info as DGETRI's. */
int oracle_dgetri(int n, double *a, int lda, const int *ipiv, double *work)
{
    if (n == 0) return 0;
    for (int k = 0; k < n; k++)
        if (A_(a, lda, k, k) == 0.0) return k + 1;
    for (int j = 0; j < n; j++) { /* DTRTI2 (upper, non-unit) */
        A_(a, lda, j, j) = 1.0 / A_(a, lda, j, j);
        const double ajj = -A_(a, lda, j, j);
        double *x = &A_(a, lda, 0, j);
        for (int k = 0; k < j; k++) { /* DTRMV U N N: x := T x with T = a(0:j-1, 0:j-1) */
            const double xk = x[k];
            if (xk == 0.0) continue;
            for (int i = 0; i < k; i++) x[i] = x[i] + xk * A_(a, lda, i, k);
            x[k] = x[k] * A_(a, lda, k, k);
        }
        for (int i = 0; i < j; i++) x[i] = ajj * x[i]; /* DSCAL */
    }
    for (int j = n - 1; j >= 0; j--) {
        for (int i = j + 1; i < n; i++) {
            work[i] = A_(a, lda, i, j);
            A_(a, lda, i, j) = 0.0;
        }
        for (int k = j + 1; k < n; k++) { /* DGEMV N, alpha -1, beta one */
            const double t = -1.0 * work[k];
            for (int i = 0; i < n; i++) A_(a, lda, i, j) = A_(a, lda, i, j) + t * A_(a, lda, i, k);
        }
    }
    for (int j = n - 2; j >= 0; j--) {
        const int p = ipiv[j] - 1;
        if (p == j) continue;
        for (int i = 0; i < n; i++) {
            const double t = A_(a, lda, i, j);
            A_(a, lda, i, j) = A_(a, lda, i, p);
            A_(a, lda, i, p) = t;
        }
    }
    return 0;
}

/* ncol row-major n x n matrices A to their inverses Y (row-major); returns the number of
   columns whose factorization or inverse reported a zero pivot (their Y as left) */
int oracle_inverse_rows(long ncol, int n, const double *A, double *Y)
{
    double a[64 * 64], work[64];
    int ipiv[64], bad = 0;
    if (n < 1 || n > 64) return -1;
    for (long q = 0; q < ncol; q++) {
        for (int r = 0; r < n; r++)
            for (int c = 0; c < n; c++) A_(a, n, r, c) = A[q * n * n + r * n + c];
        int e = oracle_dgetrf2(n, n, a, n, ipiv);
        if (e == 0) e = oracle_dgetri(n, a, n, ipiv, work);
        if (e != 0) bad++;
        for (int r = 0; r < n; r++)
            for (int c = 0; c < n; c++) Y[q * n * n + r * n + c] = A_(a, n, r, c);
    }
    return bad;
}
