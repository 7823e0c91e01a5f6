/*
 * wamp_math.h -- pow and tanh of the FP64 step, written with IEEE basic
 * operations only (add, subtract, multiply, divide, fused multiply-add) and
 * exact bit manipulations, so the same source gives the same bits wherever
 * it is compiled without contraction (-ffp-contract=off): in the gfx950
 * kernels and in host C or C++.
 *
 * The step uses pow once per cell (the entropy p / rho^gamma of the THINC
 * indicator, cell_entropy SC:308-314) and tanh in THINC (thinc_ul / thinc_ur
 * SC:327-349).  Math libraries round these functions differently in the last
 * place, and the shock-interface flow amplifies a last-place difference
 * within a few hundred steps; with these two functions the FP64 step is a
 * fixed sequence of correctly rounded operations.
 *
 * Divisions go through wm_div (below); the series are evaluated by Horner's
 * rule with explicit fused multiply-adds (correctly rounded everywhere).
 *
 * Methods (both from the mathematics, no table):
 *  - log in double-double: x = 2^k m, m in [1/sqrt2, sqrt2); ln m =
 *    2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.172, s carried as a
 *    double-double and the odd series s^3/3 + ... + s^25/25 in double;
 *    k ln 2 with ln 2 split into two doubles.
 *  - exp of a double-double z: z = n ln2 + r, |r| <= ln2/2 (Cody-Waite
 *    reduction with a 32-bit head of ln 2, so n * head is exact), e^r by its
 *    Taylor series to r^13/13! with 1 + r kept exact, then scaling by 2^n.
 *  - pow(x, y) = exp(y ln x) with y ln x formed in double-double: within one
 *    ulp (about 0.6 ulp in the worst case of the error budget) for x > 0.
 *  - tanh(x) = (e^{2|x|} - 1) / (e^{2|x|} + 1), numerator and denominator
 *    as double-doubles from the double-double exponential (small arguments
 *    lose nothing to cancellation) and the quotient with one correction
 *    step; x for |x| < 2^-28, +-1 for |x| > 22.  Within one ulp.
 * tests/test_math.py checks both against 60-digit decimal arithmetic.
 */
#ifndef WAMP_MATH_H
#define WAMP_MATH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIP__) || defined(__HIPCC__)
#define WM_FN __host__ __device__ static inline
#else
#define WM_FN static inline
#endif

#define WM_LN2_HI 0x1.62e42fefa39efp-1   /* RN(ln 2) */
#define WM_LN2_LO 0x1.abc9e3b39803fp-56  /* RN(ln 2 - WM_LN2_HI) */
#define WM_L2H 0x1.62e42fee00000p-1      /* ln 2 to 32 significant bits */
#define WM_L2L 0x1.a39ef35793c76p-33     /* RN(ln 2 - WM_L2H) */
#define WM_INV_LN2 0x1.71547652b82fep+0  /* RN(1 / ln 2) */
#define WM_SQRT2 0x1.6a09e667f3bcdp+0    /* RN(sqrt 2) */

WM_FN uint64_t wm_bits(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
WM_FN double wm_double(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }
/* 2^n for -1022 <= n <= 1023 */
WM_FN double wm_pow2(int n) { return wm_double((uint64_t)(n + 1023) << 52); }

/* a / b, correctly rounded.  In device code: the hardware reciprocal, two
 * Newton steps and one correction (Markstein), which is the IEEE quotient
 * bit for bit for finite b with 2^-1000 <= |b| <= 2^1000, a = 0 or |a| >=
 * 2^-960 and |a / b| in [2^-1000, 2^700] (a zero quotient may carry the
 * other sign); every division below stays in that range (tests/test_gpu_arith.py
 * checks the sequence against IEEE division).  Host code divides. */
#if defined(__HIP_DEVICE_COMPILE__)
WM_FN double wm_div(double a, double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    const double q = a * y;
    return fma(fma(-b, q, a), y, q);
}
#else
WM_FN double wm_div(double a, double b) { return a / b; }
#endif

/* s + e = a + b exactly (Knuth) */
WM_FN double wm_two_sum(double a, double b, double *e)
{
    double s = a + b;
    double bb = s - a;
    *e = (a - (s - bb)) + (b - bb);
    return s;
}
/* s + e = a + b exactly, |a| >= |b| (Dekker) */
WM_FN double wm_fast_two_sum(double a, double b, double *e)
{
    double s = a + b;
    *e = b - (s - a);
    return s;
}

/* hi + lo ~ ln x for finite x > 0 (absolute error below 1e-18 + 2^-60 |ln x|) */
WM_FN double wm_log_dd(double x, double *lo)
{
    int k = 0;
    if (x < 0x1p-1022) {           /* subnormal: scale into the normal range */
        x = x * 0x1p54;
        k = -54;
    }
    uint64_t b = wm_bits(x);
    k += (int)((b >> 52) & 0x7ff) - 1023;
    double m = wm_double((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  /* [1, 2) */
    if (m > WM_SQRT2) {
        m = m * 0.5;
        k += 1;
    }
    const double num = m - 1.0;    /* exact (Sterbenz) */
    double dl;
    const double dh = wm_two_sum(m, 1.0, &dl);
    const double sh = wm_div(num, dh);
    const double r = fma(-sh, dh, num) - sh * dl;
    const double sl = wm_div(r, dh);   /* sh + sl = num / (dh + dl) */
    const double s2 = sh * sh;
    double t = fma(1.0 / 25.0, s2, 1.0 / 23.0);
    t = fma(t, s2, 1.0 / 21.0);
    t = fma(t, s2, 1.0 / 19.0);
    t = fma(t, s2, 1.0 / 17.0);
    t = fma(t, s2, 1.0 / 15.0);
    t = fma(t, s2, 1.0 / 13.0);
    t = fma(t, s2, 1.0 / 11.0);
    t = fma(t, s2, 1.0 / 9.0);
    t = fma(t, s2, 1.0 / 7.0);
    t = fma(t, s2, 1.0 / 5.0);
    t = fma(t, s2, 1.0 / 3.0);
    const double tail = sh * (s2 * t);            /* s^3/3 + s^5/5 + ... */
    double ml;
    const double mh = wm_fast_two_sum(2.0 * sh, 2.0 * (sl + tail), &ml);  /* ln m */
    const double kd = (double)k;
    const double ah = kd * WM_LN2_HI;
    const double ae = fma(kd, WM_LN2_HI, -ah) + kd * WM_LN2_LO;  /* k ln 2 = ah + ae */
    double e;
    const double h = wm_two_sum(ah, mh, &e);
    *lo = e + (ae + ml);
    return h;
}

/* e^(zh + zl) = 2^n (hi + lo), |zl| <= ulp(zh), |zh| < 700 */
WM_FN double wm_exp_dd(double zh, double zl, int *n_out, double *lo)
{
    const double shifter = 0x1.8p52;
    const double nd = (zh * WM_INV_LN2 + shifter) - shifter;   /* round to nearest */
    const double r1 = zh - nd * WM_L2H;                        /* exact */
    double re;
    const double rh = wm_two_sum(r1, -(nd * WM_L2L), &re);
    const double rl = re + zl;
    double q = fma(1.0 / 6227020800.0, rh, 1.0 / 479001600.0);  /* 1/13!, 1/12! */
    q = fma(q, rh, 1.0 / 39916800.0);
    q = fma(q, rh, 1.0 / 3628800.0);
    q = fma(q, rh, 1.0 / 362880.0);
    q = fma(q, rh, 1.0 / 40320.0);
    q = fma(q, rh, 1.0 / 5040.0);
    q = fma(q, rh, 1.0 / 720.0);
    q = fma(q, rh, 1.0 / 120.0);
    q = fma(q, rh, 1.0 / 24.0);
    q = fma(q, rh, 1.0 / 6.0);
    q = fma(q, rh, 0.5);
    q = q * (rh * rh);                                         /* e^r - 1 - r */
    double b;
    const double a = wm_two_sum(1.0, rh, &b);
    const double c = b + (q + rl * ((1.0 + rh) + q));     /* + rl e^rh */
    *n_out = (int)nd;
    return wm_fast_two_sum(a, c, lo);
}

/* x^y for x > 0 (0 for x = 0 and y > 0, NaN for x < 0 or NaN, inf for x = inf) */
WM_FN double wm_pow(double x, double y)
{
    if (x != x) return x;
    if (x <= 0.0) return x == 0.0 ? 0.0 : (x - x) / (x - x);
    if (x == INFINITY) return x;
    double ll;
    const double lh = wm_log_dd(x, &ll);
    const double zh0 = y * lh;
    const double ze = fma(y, lh, -zh0) + y * ll;
    double zl;
    const double zh = wm_fast_two_sum(zh0, ze, &zl);
    if (zh > 709.0) return INFINITY;
    if (zh < -708.0) return 0.0;
    int n;
    double el;
    const double eh = wm_exp_dd(zh, zl, &n, &el);
    (void)el;
    return eh * wm_pow2(n);
}

WM_FN double wm_tanh(double x)
{
    const double a = fabs(x);
    if (a != a) return x;
    if (a > 22.0) return x < 0.0 ? -1.0 : 1.0;
    if (a < 0x1p-28) return x;
    int n;
    double el;
    double eh = wm_exp_dd(2.0 * a, 0.0, &n, &el);
    const double sc = wm_pow2(n);
    eh = eh * sc;                  /* exact scaling */
    el = el * sc;
    double ne, de;
    const double nh = wm_two_sum(eh, -1.0, &ne);   /* e^2a - 1 = nh + nl */
    const double nl = ne + el;
    const double dh = wm_two_sum(eh, 1.0, &de);    /* e^2a + 1 = dh + dl */
    const double dl = de + el;
    const double t0 = wm_div(nh, dh);              /* one correction step of the quotient */
    const double r = (fma(-t0, dh, nh) + nl) - t0 * dl;
    const double t = t0 + wm_div(r, dh);
    return x < 0.0 ? -t : t;
}

#endif
