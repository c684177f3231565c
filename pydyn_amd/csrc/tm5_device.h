// Device-side arithmetic shared by the TM5 transport kernels.
//
// Every expression keeps the association order of the reference (pydyn's NumPy operators,
// pydyn/kernels.py and pydyn/ops.py) so that results are bit-identical to it.  The library is
// compiled with -ffp-contract=off: no multiply-add is fused, every operation rounds once.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tm5 {

// ------------------------------------------------------------------ NumPy semantics
// np.clip(s, -|rm|, |rm|) with array bounds (pydyn/kernels.py:65-67): max by `>` then min by
// `<`, a NaN in the running value kept.  With rm = +-0 every finite s gives +0.
__device__ __forceinline__ double limit(double s, double rm)
{
    const double hi = fabs(rm), lo = -hi;
    const double t = __builtin_isnan(s) ? s : (s > lo ? s : lo);
    return __builtin_isnan(t) ? t : (t < hi ? t : hi);
}

// np.maximum(0.0, x): x on a tie (so -0.0 stays -0.0) and for NaN.
__device__ __forceinline__ double max0(double x) { return 0.0 > x ? 0.0 : x; }

// ------------------------------------------------------------------ the slopes scheme
struct FaceFlux {
    double f, pf, ft, fw;  // tracer mass, its first moment, the two carried cross slopes
};

// pydyn/kernels.py:75-84 for one face: flux, its Courant number c, the upwind cell's tracer
// mass rm_u, slope along the flow s_u and the two cross slopes t_u, w_u.
__device__ __forceinline__ FaceFlux face_flux(double flux, double c, double rm_u, double s_u, double t_u,
                                              double w_u)
{
    FaceFlux q;
    q.f = flux >= 0.0 ? c * (rm_u + (1.0 - c) * s_u) : c * (rm_u - (1.0 + c) * s_u);
    q.pf = flux * ((c * c) * s_u - 3.0 * q.f);
    q.ft = c * t_u;
    q.fw = c * w_u;
    return q;
}

// pydyn/kernels.py:87-92 for one cell with lower face (lo) and upper face (hi).
__device__ __forceinline__ double slope_update(double s, double pf_lo, double pf_hi, double a_lo, double a_hi,
                                               double rm_new, double f_lo, double f_hi, double m_old, double m_new)
{
    const double inner = ((pf_lo - pf_hi) - (a_lo - a_hi) * s) + 3.0 * ((a_lo + a_hi) * rm_new - (f_lo + f_hi) * m_old);
    return s + inner / m_new;
}

// pydyn/ops.py:304-307: dynamu/dynamv's retry scaling of the current flux.
// Workgroups go to the eight XCDs round-robin (workgroup b to XCD b % 8), each XCD with its
// own L2.  Renumbered by this bijection of [0, nb), each XCD takes a contiguous range of tiles,
// so neighbouring tiles, which share cache lines, meet in one L2 instead of fetching the lines
// once per XCD.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    constexpr int X = 8;
    const int per = nb / X, rem = nb % X, x = b % X, k = b / X;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

// The workgroup's (x, y) in a 2-d grid renumbered like xcd_tile: neighbouring tiles, which share
// cache lines at their edges, run on the same XCD and so share its L2.
__device__ __forceinline__ void xcd_grid2(int &bx, int &by)
{
    const int nbx = gridDim.x;
    const int t = xcd_tile(blockIdx.y * nbx + blockIdx.x, nbx * gridDim.y);
    bx = t % nbx;
    by = t / nbx;
}

__device__ __forceinline__ double shrink(double f, int n) { return f * (double)(n - 1) / (double)n; }

// ------------------------------------------------------------------ the fdlibm exp
// exp(x) by the method of Sun's fdlibm e_exp.c: x = k ln2 + r, a degree-5 Remez polynomial in
// r^2 for r (e^r + 1) / (e^r - 1), e^r = 1 + 2r / (R - r), scaled by 2^k.  The same operations
// in the same order as the oracle's C copy, so host and device agree bit for bit.
__device__ inline double fdlibm_exp(double x)
{
    const double one = 1.0, huge = 1.0e300, twom1000 = 9.33263618503218878990e-302;
    const double o_threshold = 7.09782712893383973096e+02;
    const double u_threshold = -7.45133219101941108420e+02;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double p1 = 1.66666666666666019037e-01, p2 = -2.77777777770155933842e-03;
    const double p3 = 6.61375632143793436117e-05, p4 = -1.65339022054652515390e-06;
    const double p5 = 4.13813679705723846039e-08;

    const uint64_t bits = (uint64_t)__double_as_longlong(x);
    uint32_t hx = (uint32_t)(bits >> 32);
    const int neg = (int)(hx >> 31);
    hx &= 0x7fffffffu;
    double hi = 0.0, lo = 0.0;
    int k = 0;
    if (hx >= 0x40862E42u) {
        if (hx >= 0x7ff00000u) {
            if (((hx & 0xfffffu) | (uint32_t)bits) != 0) return x + x;
            return neg ? 0.0 : x;
        }
        if (x > o_threshold) return huge * huge;
        if (x < u_threshold) return twom1000 * twom1000;
    }
    if (hx > 0x3fd62e42u) {
        if (hx < 0x3FF0A2B2u) {
            if (neg) {
                hi = x + ln2_hi;
                lo = -ln2_lo;
                k = -1;
            } else {
                hi = x - ln2_hi;
                lo = ln2_lo;
                k = 1;
            }
        } else {
            k = (int)(inv_ln2 * x + (neg ? -0.5 : 0.5));
            const double t = (double)k;
            hi = x - t * ln2_hi;
            lo = t * ln2_lo;
        }
        x = hi - lo;
    } else if (hx < 0x3e300000u) {
        if (huge + x > one) return one + x;
    }
    const double t = x * x;
    const double c = x - t * (p1 + t * (p2 + t * (p3 + t * (p4 + t * p5))));
    if (k == 0) return one - ((x * c) / (c - 2.0) - x);
    const double y = one - ((lo - (x * c) / (2.0 - c)) - hi);
    const uint64_t yb = (uint64_t)__double_as_longlong(y);
    if (k >= -1021) return __longlong_as_double((long long)(yb + ((uint64_t)(int64_t)k << 52)));
    return __longlong_as_double((long long)(yb + ((uint64_t)(int64_t)(k + 1000) << 52))) * twom1000;
}

}  // namespace tm5
