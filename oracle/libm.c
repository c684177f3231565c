/* exp, log and pow for the oracle (test infrastructure), by the methods of Sun's fdlibm e_exp.c,
   e_log.c and e_pow.c (freely distributable, Sun Microsystems 1993):
     exp  reduce x = k ln2 + r with |r| <= ln2/2 (ln2 split into a 32-bit-exact high part and a
          low correction), approximate r (e^r + 1) / (e^r - 1) by a degree-5 Remez polynomial in
          r^2, reconstruct e^r = 1 + 2r / (R(r) - r) and scale by 2^k;
     log  x = 2^k (1 + f) with sqrt(2)/2 < 1 + f < sqrt(2); s = f / (2 + f), log(1 + f) =
          f - s (f - R(s)) with R a degree-14 Remez polynomial in s (even powers), or
          f - (hfsq - s (hfsq + R)) with hfsq = f^2 / 2 away from 1; plus k ln2 in two parts;
     pow  log2(|x|) to extra precision as t1 + t2 (|x| = 2^n (1 + f) reduced against 1 or 1.5,
          (x - b)/(x + b) and its odd Remez series, 2/(3 ln2) in two parts), y split into a
          head with 32 low bits zero and a tail, y log2|x| = p_h + p_l, then 2^(p_h + p_l) as
          2^n e^(r) with the exp method above; the special cases of C99 Annex F.
   Written as plain C with no FMA contraction (-ffp-contract=off) so that it rounds like the
   device's copy (pydyn_amd/csrc/tm5_device.h).  Constants are given by their bit patterns.

   The library also exports exp/log/pow/cos/sin under their C names: loaded into the reference
   with its own `use_libm(path)`, it makes the reference compute with these exp, log and pow
   (cos and sin are forwarded to the C library).  Oracle code calls oracle_*_map. */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

static double from_bits(uint64_t b)
{
    double x;
    memcpy(&x, &b, 8);
    return x;
}

static uint64_t to_bits(double x)
{
    uint64_t b;
    memcpy(&b, &x, 8);
    return b;
}

static double fd_exp(double x)
{
    const double one = 1.0, huge = 1.0e300, twom1000 = 9.33263618503218878990e-302;
    const double o_threshold = 7.09782712893383973096e+02;
    const double u_threshold = -7.45133219101941108420e+02;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double p1 = 1.66666666666666019037e-01, p2 = -2.77777777770155933842e-03;
    const double p3 = 6.61375632143793436117e-05, p4 = -1.65339022054652515390e-06;
    const double p5 = 4.13813679705723846039e-08;

    uint64_t bits = to_bits(x);
    uint32_t hx = (uint32_t)(bits >> 32);
    int neg = (int)(hx >> 31);
    hx &= 0x7fffffffu;
    double hi = 0.0, lo = 0.0;
    int k = 0;

    if (hx >= 0x40862E42u) { /* |x| >= 709.78...: overflow, underflow, inf, NaN */
        if (hx >= 0x7ff00000u) {
            if (((hx & 0xfffffu) | (uint32_t)bits) != 0) return x + x; /* NaN */
            return neg ? 0.0 : x;                                        /* exp(+-inf) */
        }
        if (x > o_threshold) return huge * huge;
        if (x < u_threshold) return twom1000 * twom1000;
    }
    if (hx > 0x3fd62e42u) { /* |x| > ln2/2 */
        if (hx < 0x3FF0A2B2u) { /* |x| < 3 ln2 / 2: k = +-1 */
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
            double t = (double)k;
            hi = x - t * ln2_hi;
            lo = t * ln2_lo;
        }
        x = hi - lo;
    } else if (hx < 0x3e300000u) { /* |x| < 2^-28 */
        if (huge + x > one) return one + x;
    }

    double t = x * x;
    double c = x - t * (p1 + t * (p2 + t * (p3 + t * (p4 + t * p5))));
    if (k == 0) return one - ((x * c) / (c - 2.0) - x);
    double y = one - ((lo - (x * c) / (2.0 - c)) - hi);
    if (k >= -1021) return from_bits(to_bits(y) + ((uint64_t)(int64_t)k << 52));
    return from_bits(to_bits(y) + ((uint64_t)(int64_t)(k + 1000) << 52)) * twom1000;
}

static int32_t hi_word(double x) { return (int32_t)(to_bits(x) >> 32); }
static uint32_t lo_word(double x) { return (uint32_t)to_bits(x); }
static double with_hi(double x, int32_t hi)
{
    return from_bits(((uint64_t)(uint32_t)hi << 32) | (to_bits(x) & 0xffffffffu));
}
static double lo_cleared(double x) { return from_bits(to_bits(x) & 0xffffffff00000000ull); }

static double fd_log(double x)
{
    const double ln2_hi = from_bits(0x3fe62e42fee00000ull), ln2_lo = from_bits(0x3dea39ef35793c76ull);
    const double two54 = from_bits(0x4350000000000000ull);
    const double Lg1 = from_bits(0x3fe5555555555593ull), Lg2 = from_bits(0x3fd999999997fa04ull);
    const double Lg3 = from_bits(0x3fd2492494229359ull), Lg4 = from_bits(0x3fcc71c51d8e78afull);
    const double Lg5 = from_bits(0x3fc7466496cb03deull), Lg6 = from_bits(0x3fc39a09d078c69full);
    const double Lg7 = from_bits(0x3fc2f112df3e5244ull);
    int32_t hx = hi_word(x);
    const uint32_t lx = lo_word(x);
    int k = 0;
    if (hx < 0x00100000) {                                         /* x < 2^-1022 */
        if (((hx & 0x7fffffff) | (int32_t)lx) == 0) return -HUGE_VAL; /* log(+-0) */
        if (hx < 0) return NAN;                                      /* log(-x) */
        k -= 54;
        x *= two54; /* subnormal: scale up */
        hx = hi_word(x);
    }
    if (hx >= 0x7ff00000) return x + x;
    k += (hx >> 20) - 1023;
    hx &= 0x000fffff;
    int i = (hx + 0x95f64) & 0x100000;
    x = with_hi(x, hx | (i ^ 0x3ff00000)); /* normalise x or x/2 */
    k += i >> 20;
    const double f = x - 1.0;
    if ((0x000fffff & (2 + hx)) < 3) { /* |f| < 2^-20 */
        if (f == 0.0) {
            if (k == 0) return 0.0;
            const double dk = (double)k;
            return dk * ln2_hi + dk * ln2_lo;
        }
        const double R = f * f * (0.5 - 0.33333333333333333 * f);
        if (k == 0) return f - R;
        const double dk = (double)k;
        return dk * ln2_hi - ((R - dk * ln2_lo) - f);
    }
    const double s = f / (2.0 + f);
    const double dk = (double)k;
    const double z = s * s;
    i = hx - 0x6147a;
    const double w = z * z;
    const int j = 0x6b851 - hx;
    const double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    const double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    i |= j;
    const double R = t2 + t1;
    if (i > 0) {
        const double hfsq = 0.5 * f * f;
        if (k == 0) return f - (hfsq - s * (hfsq + R));
        return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
    }
    if (k == 0) return f - s * (f - R);
    return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

static double fd_pow(double x, double y)
{
    const double bp[2] = {1.0, 1.5};
    const double dp_h[2] = {0.0, from_bits(0x3fe2b80340000000ull)};
    const double dp_l[2] = {0.0, from_bits(0x3e4cfdeb43cfd006ull)};
    const double two53 = from_bits(0x4340000000000000ull), huge = 1.0e300, tiny = 1.0e-300;
    const double L1 = from_bits(0x3fe3333333333303ull), L2 = from_bits(0x3fdb6db6db6fabffull);
    const double L3 = from_bits(0x3fd55555518f264dull), L4 = from_bits(0x3fd17460a91d4101ull);
    const double L5 = from_bits(0x3fcd864a93c9db65ull), L6 = from_bits(0x3fca7e284a454eefull);
    const double P1 = from_bits(0x3fc555555555553eull), P2 = from_bits(0xbf66c16c16bebd93ull);
    const double P3 = from_bits(0x3f11566aaf25de2cull), P4 = from_bits(0xbebbbd41c5d26bf1ull);
    const double P5 = from_bits(0x3e66376972bea4d0ull);
    const double lg2 = from_bits(0x3fe62e42fefa39efull), lg2_h = from_bits(0x3fe62e4300000000ull);
    const double lg2_l = from_bits(0xbe205c610ca86c39ull);
    const double ovt = 8.0085662595372944372e-17; /* -(1024 - log2(ovfl + .5 ulp)) */
    const double cp = from_bits(0x3feec709dc3a03fdull), cp_h = from_bits(0x3feec709e0000000ull);
    const double cp_l = from_bits(0xbe3e2fe0145b01f5ull);
    const double ivln2 = from_bits(0x3ff71547652b82feull), ivln2_h = from_bits(0x3ff7154760000000ull);
    const double ivln2_l = from_bits(0x3e54ae0bf85ddf44ull);

    const int32_t hx = hi_word(x), hy = hi_word(y);
    const uint32_t lx = lo_word(x), ly = lo_word(y);
    int32_t ix = hx & 0x7fffffff;
    const int32_t iy = hy & 0x7fffffff;

    if ((iy | (int32_t)ly) == 0) return 1.0; /* x^0 = 1, also for NaN x */
    if (x == 1.0) return 1.0;                /* 1^y = 1, also for NaN y (C99) */
    if (ix > 0x7ff00000 || (ix == 0x7ff00000 && lx != 0) || iy > 0x7ff00000 || (iy == 0x7ff00000 && ly != 0))
        return x + y; /* NaN */
    /* yisint: 0 y not an integer, 1 an odd integer, 2 an even integer (x < 0 only) */
    int yisint = 0;
    if (hx < 0) {
        if (iy >= 0x43400000)
            yisint = 2;
        else if (iy >= 0x3ff00000) {
            const int k = (iy >> 20) - 0x3ff;
            if (k > 20) {
                const uint32_t j = ly >> (52 - k);
                if ((j << (52 - k)) == ly) yisint = 2 - (int)(j & 1);
            } else if (ly == 0) {
                const int32_t j = iy >> (20 - k);
                if ((j << (20 - k)) == iy) yisint = 2 - (j & 1);
            }
        }
    }
    if (ly == 0) {
        if (iy == 0x7ff00000) { /* y = +-inf */
            if (((ix - 0x3ff00000) | (int32_t)lx) == 0) return 1.0; /* (-1)^+-inf = 1 (C99) */
            if (ix >= 0x3ff00000) return hy >= 0 ? y : 0.0;        /* |x| > 1 */
            return hy < 0 ? -y : 0.0;                                /* |x| < 1 */
        }
        if (iy == 0x3ff00000) return hy < 0 ? 1.0 / x : x; /* y = +-1 */
        if (hy == 0x40000000) return x * x;                /* y = 2 */
        if (hy == 0x3fe00000 && hx >= 0) return sqrt(x);   /* y = 0.5, x >= +0 */
    }
    double ax = fabs(x);
    if (lx == 0 && (ix == 0x7ff00000 || ix == 0 || ix == 0x3ff00000)) { /* x = +-0, +-inf, +-1 */
        double z = ax;
        if (hy < 0) z = 1.0 / z;
        if (hx < 0) {
            if (((ix - 0x3ff00000) | yisint) == 0)
                z = NAN; /* (-1)^non-integer */
            else if (yisint == 1)
                z = -z;
        }
        return z;
    }
    int n = (int)((uint32_t)hx >> 31) ^ 1; /* 1 for x >= 0, 0 for x < 0 */
    if ((n | yisint) == 0) return NAN;     /* (x < 0)^non-integer */
    double s = 1.0;
    if ((n | (yisint - 1)) == 0) s = -1.0; /* (x < 0)^odd */

    double t1, t2;
    if (iy > 0x41e00000) { /* |y| > 2^31 */
        if (iy > 0x43f00000) { /* |y| > 2^64: over- or underflow */
            if (ix <= 0x3fefffff) return hy < 0 ? huge * huge : tiny * tiny;
            if (ix >= 0x3ff00000) return hy > 0 ? huge * huge : tiny * tiny;
        }
        if (ix < 0x3fefffff) return hy < 0 ? s * huge * huge : s * tiny * tiny;
        if (ix > 0x3ff00000) return hy > 0 ? s * huge * huge : s * tiny * tiny;
        /* |1 - x| <= 2^-20: log(x) by x - x^2/2 + x^3/3 - x^4/4 */
        const double t = ax - 1.0;
        const double w = (t * t) * (0.5 - t * (0.3333333333333333333333 - t * 0.25));
        const double u = ivln2_h * t;
        const double v = t * ivln2_l - w * ivln2;
        t1 = lo_cleared(u + v);
        t2 = v - (t1 - u);
    } else {
        n = 0;
        if (ix < 0x00100000) { /* subnormal x */
            ax *= two53;
            n -= 53;
            ix = hi_word(ax);
        }
        n += (ix >> 20) - 0x3ff;
        const int32_t j = ix & 0x000fffff;
        int k;
        ix = j | 0x3ff00000; /* normalise */
        if (j <= 0x3988E)
            k = 0; /* |x| < sqrt(3/2) */
        else if (j < 0xBB67A)
            k = 1; /* |x| < sqrt(3) */
        else {
            k = 0;
            n += 1;
            ix -= 0x00100000;
        }
        ax = with_hi(ax, ix);
        /* ss = s_h + s_l = (x - 1)/(x + 1) or (x - 1.5)/(x + 1.5) */
        double u = ax - bp[k];
        double v = 1.0 / (ax + bp[k]);
        const double ss = u * v;
        const double s_h = lo_cleared(ss);
        double t_h = from_bits((uint64_t)(uint32_t)(((ix >> 1) | 0x20000000) + 0x00080000 + (k << 18)) << 32);
        double t_l = ax - (t_h - bp[k]);
        const double s_l = v * ((u - s_h * t_h) - s_h * t_l);
        /* log(ax) */
        double s2 = ss * ss;
        double r = s2 * s2 * (L1 + s2 * (L2 + s2 * (L3 + s2 * (L4 + s2 * (L5 + s2 * L6)))));
        r += s_l * (s_h + ss);
        s2 = s_h * s_h;
        t_h = lo_cleared(3.0 + s2 + r);
        t_l = r - ((t_h - 3.0) - s2);
        u = s_h * t_h;
        v = s_l * t_h + t_l * ss;
        const double p_h = lo_cleared(u + v);
        const double p_l = v - (p_h - u);
        const double z_h = cp_h * p_h;
        const double z_l = cp_l * p_h + p_l * cp + dp_l[k];
        /* log2(ax) = n + dp_h + z_h + z_l */
        const double t = (double)n;
        t1 = lo_cleared(((z_h + z_l) + dp_h[k]) + t);
        t2 = z_l - (((t1 - t) - dp_h[k]) - z_h);
    }
    /* (y1 + y2) (t1 + t2) with y1 y's head */
    const double y1 = lo_cleared(y);
    const double p_l = (y - y1) * t1 + y * t2;
    double p_h = y1 * t1;
    double z = p_l + p_h;
    int32_t j = hi_word(z);
    int32_t i = (int32_t)lo_word(z);
    if (j >= 0x40900000) { /* z >= 1024 */
        if (((j - 0x40900000) | i) != 0) return s * huge * huge;
        if (p_l + ovt > z - p_h) return s * huge * huge;
    } else if ((j & 0x7fffffff) >= 0x4090cc00) { /* z <= -1075 */
        if (((j - (int32_t)0xc090cc00) | i) != 0) return s * tiny * tiny;
        if (p_l <= z - p_h) return s * tiny * tiny;
    }
    /* 2^(p_h + p_l) */
    i = j & 0x7fffffff;
    int k = (i >> 20) - 0x3ff;
    n = 0;
    if (i > 0x3fe00000) { /* |z| > 0.5: n = [z + 0.5] */
        n = j + (0x00100000 >> (k + 1));
        k = ((n & 0x7fffffff) >> 20) - 0x3ff;
        const double t = from_bits((uint64_t)(uint32_t)(n & ~(0x000fffff >> k)) << 32);
        n = ((n & 0x000fffff) | 0x00100000) >> (20 - k);
        if (j < 0) n = -n;
        p_h -= t;
    }
    const double t = lo_cleared(p_l + p_h);
    const double u = t * lg2_h;
    const double v = (p_l - (t - p_h)) * lg2 + t * lg2_l;
    z = u + v;
    const double w = v - (z - u);
    const double tt = z * z;
    const double tt1 = z - tt * (P1 + tt * (P2 + tt * (P3 + tt * (P4 + tt * P5))));
    const double r = (z * tt1) / (tt1 - 2.0) - (w + z * w);
    z = 1.0 - (r - z);
    j = hi_word(z);
    j += (int32_t)((uint32_t)n << 20);
    if ((j >> 20) <= 0)
        z = ldexp(z, n); /* subnormal result */
    else
        z = with_hi(z, j);
    return s * z;
}

void oracle_exp_map(const double *x, double *y, size_t n)
{
    for (size_t i = 0; i < n; i++) y[i] = fd_exp(x[i]);
}

void oracle_log_map(const double *x, double *y, size_t n)
{
    for (size_t i = 0; i < n; i++) y[i] = fd_log(x[i]);
}

void oracle_pow_map(const double *x, const double *y, double *out, size_t n)
{
    for (size_t i = 0; i < n; i++) out[i] = fd_pow(x[i], y[i]);
}

/* ------------------------------------------------------------------ the C names */
double exp(double x) { return fd_exp(x); }

typedef double (*fn1)(double);

static void *next_symbol(const char *name) { return dlsym(RTLD_NEXT, name); }

double log(double x) { return fd_log(x); }
double pow(double x, double y) { return fd_pow(x, y); }
double cos(double x) { return ((fn1)next_symbol("cos"))(x); }
double sin(double x) { return ((fn1)next_symbol("sin"))(x); }
