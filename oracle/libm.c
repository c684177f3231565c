/* exp(x) for the oracle, by the method of Sun's fdlibm e_exp.c (freely distributable, Sun
   Microsystems 1993): reduce x = k ln2 + r with |r| <= ln2/2 (ln2 split into a 32-bit-exact
   high part and a low correction), approximate r (e^r + 1) / (e^r - 1) by a degree-5 Remez
   polynomial in r^2, reconstruct e^r = 1 + 2r / (R(r) - r) and scale by 2^k.  Written as plain
   C with no FMA contraction (-ffp-contract=off) so that it rounds like the device's copy.

   The library also exports exp/log/pow/cos/sin under their C names: loaded into the reference
   with its own `use_libm(path)`, it makes the reference compute its decay factors with this
   exp (the other four are forwarded to the C library: the transport step does not use them).
   Oracle code calls oracle_exp_map. */
#define _GNU_SOURCE
#include <dlfcn.h>
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

void oracle_exp_map(const double *x, double *y, size_t n)
{
    for (size_t i = 0; i < n; i++) y[i] = fd_exp(x[i]);
}

/* ------------------------------------------------------------------ the C names */
double exp(double x) { return fd_exp(x); }

typedef double (*fn1)(double);
typedef double (*fn2)(double, double);

static void *next_symbol(const char *name) { return dlsym(RTLD_NEXT, name); }

double log(double x) { return ((fn1)next_symbol("log"))(x); }
double cos(double x) { return ((fn1)next_symbol("cos"))(x); }
double sin(double x) { return ((fn1)next_symbol("sin"))(x); }
double pow(double x, double y) { return ((fn2)next_symbol("pow"))(x, y); }
