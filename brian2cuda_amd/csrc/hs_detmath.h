/*
 * hs_detmath.h -- the numerical contract of the hot path: elementary functions
 * that give the SAME BITS on the x86-64 host (gcc/clang, SSE2, -ffp-contract=off)
 * and on gfx950 (hipcc -ffp-contract=off).
 *
 * Why: Brian2's cpp_standalone evaluates model expressions with the host libm
 * (exp, log, sqrt, cos).  A GPU libm rounds differently in the last ulp, and in a
 * spiking network one ulp becomes a different spike train within a few hundred
 * milliseconds.  Every transcendental on the per-step path is therefore evaluated
 * here with IEEE-754 operations whose results are fixed by the standard
 * (+ - * /, sqrt, floor, frexp, ldexp, int<->float conversion), in a fixed order and
 * with no fused multiply-add.  The expression text is the contract; the host
 * oracle and the device kernels include this one header.
 *
 * Methods (all from the mathematics, no vendor code):
 *   exp  : Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split into
 *          a 32-bit head (k*head exact for |k| < 2^21) and a tail; exp(r) by its
 *          Taylor series to r^13 (truncation < 4e-18 relative); 2^k by ldexp.
 *   log  : x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s=(m-1)/(m+1),
 *          |s| <= 0.1716, odd series to s^21 (truncation < 2e-17).
 *   sqrt : the IEEE operation (correctly rounded by the standard).
 *   sin/cos(2 pi u): quadrant split of u (exact: u*4 is a power-of-two scaling),
 *          reflection to [0, pi/4], Taylor series to a^17 / a^18.
 * Accuracy versus a correctly rounded libm is checked by the CPU tests
 * (a few ulp); bit equality host vs device is checked by the GPU tests.
 */
#ifndef HS_DETMATH_H
#define HS_DETMATH_H

#ifndef __HIPCC_RTC__
#include <math.h>
#endif

#ifdef __HIP__
#define HS_HD __host__ __device__ static inline __attribute__((always_inline))
#else
#define HS_HD static inline
#endif

/* ln2 = HS_LN2_HI + HS_LN2_LO; HS_LN2_HI has 32 significant bits. */
#define HS_LN2_HI 0.6931471803691238
#define HS_LN2_LO 1.9082149292705877e-10
#define HS_LOG2E 1.4426950408889634
#define HS_SQRT_HALF 0.7071067811865476
#define HS_PI_2 1.5707963267948966
/* single precision split: head has 12 significant bits (k*head exact, |k|<2^12) */
#define HS_LN2F_HI 0.693115234375f
#define HS_LN2F_LO 3.194618329871446e-05f
#define HS_LOG2EF 1.44269504f

HS_HD double hs_exp(double x)
{
    /* keep 2^k in the normal range; NaN falls through the comparisons */
    if (x > 709.0) x = 709.0;
    if (x < -708.0) x = -708.0;
    double kd = floor(x * HS_LOG2E + 0.5);
    double r = (x - kd * HS_LN2_HI) - kd * HS_LN2_LO;
    double p = 1.6059043836821613e-10;           /* 1/13! */
    p = 2.08767569878681e-09 + r * p;            /* 1/12! */
    p = 2.505210838544172e-08 + r * p;           /* 1/11! */
    p = 2.755731922398589e-07 + r * p;           /* 1/10! */
    p = 2.7557319223985893e-06 + r * p;          /* 1/9!  */
    p = 2.48015873015873e-05 + r * p;            /* 1/8!  */
    p = 0.0001984126984126984 + r * p;           /* 1/7!  */
    p = 0.001388888888888889 + r * p;            /* 1/6!  */
    p = 0.008333333333333333 + r * p;            /* 1/5!  */
    p = 0.041666666666666664 + r * p;            /* 1/4!  */
    p = 0.16666666666666666 + r * p;             /* 1/3!  */
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    return ldexp(p, (int)kd);
}

HS_HD float hs_expf(float x)
{
    if (x > 88.0f) x = 88.0f;
    if (x < -87.0f) x = -87.0f;
    float kf = floorf(x * HS_LOG2EF + 0.5f);
    float r = (x - kf * HS_LN2F_HI) - kf * HS_LN2F_LO;
    float p = 0.00019841270113829523f;           /* 1/7! */
    p = 0.0013888889225199819f + r * p;
    p = 0.008333333767950535f + r * p;
    p = 0.0416666679084301f + r * p;
    p = 0.1666666716337204f + r * p;
    p = 0.5f + r * p;
    p = 1.0f + r * p;
    p = 1.0f + r * p;
    return ldexpf(p, (int)kf);
}

/* natural log for finite x > 0 (callers guarantee it: uniforms in (0, 1]) */
HS_HD double hs_log(double x)
{
    int e;
    double m = frexp(x, &e); /* m in [0.5, 1) */
    if (m < HS_SQRT_HALF) {
        m = m * 2.0;
        e = e - 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double q = 0.047619047619047616;             /* 1/21 */
    q = 0.05263157894736842 + s2 * q;            /* 1/19 */
    q = 0.058823529411764705 + s2 * q;           /* 1/17 */
    q = 0.06666666666666667 + s2 * q;            /* 1/15 */
    q = 0.07692307692307693 + s2 * q;            /* 1/13 */
    q = 0.09090909090909091 + s2 * q;            /* 1/11 */
    q = 0.1111111111111111 + s2 * q;             /* 1/9  */
    q = 0.14285714285714285 + s2 * q;            /* 1/7  */
    q = 0.2 + s2 * q;                            /* 1/5  */
    q = 0.3333333333333333 + s2 * q;             /* 1/3  */
    double lm = 2.0 * s + (2.0 * s) * (s2 * q);
    double ed = (double)e;
    return ed * HS_LN2_HI + (ed * HS_LN2_LO + lm);
}

/* square root: IEEE-754 requires it correctly rounded, like + - * /, so the
   hardware/library operation is already deterministic (x86 sqrtsd; the gfx950
   f64 expansion rsq + Newton + fix-up is correctly rounded) */
HS_HD double hs_sqrt(double x)
{
    return sqrt(x);
}

/* sin(a), cos(a) for a in [0, pi/4] by Taylor series */
HS_HD double hs_sin_red(double a)
{
    double a2 = a * a;
    double p = 2.8114572543455206e-15;           /* 1/17! */
    p = -7.647163731819816e-13 + a2 * p;         /* -1/15! */
    p = 1.6059043836821613e-10 + a2 * p;         /* 1/13! */
    p = -2.505210838544172e-08 + a2 * p;         /* -1/11! */
    p = 2.7557319223985893e-06 + a2 * p;         /* 1/9! */
    p = -0.0001984126984126984 + a2 * p;         /* -1/7! */
    p = 0.008333333333333333 + a2 * p;           /* 1/5! */
    p = -0.16666666666666666 + a2 * p;           /* -1/3! */
    return a + a * (a2 * p);
}

HS_HD double hs_cos_red(double a)
{
    double a2 = a * a;
    double p = 1.5619206968586225e-16;           /* 1/18! */
    p = -4.779477332387385e-14 + a2 * p;         /* -1/16! */
    p = 1.1470745597729725e-11 + a2 * p;         /* 1/14! */
    p = -2.08767569878681e-09 + a2 * p;          /* -1/12! */
    p = 2.755731922398589e-07 + a2 * p;          /* 1/10! */
    p = -2.48015873015873e-05 + a2 * p;          /* -1/8! */
    p = 0.001388888888888889 + a2 * p;           /* 1/6! */
    p = -0.041666666666666664 + a2 * p;          /* -1/4! */
    p = 0.5 + a2 * p;                            /* 1/2 - a^2/4! + ... */
    return 1.0 - a2 * p;
}

/* cos(2 pi u) for u in [0, 1) */
HS_HD double hs_cos2pi(double u)
{
    double t = u * 4.0;
    double qd = floor(t);
    double f = t - qd;  /* exact */
    int q = (int)qd & 3;
    double s, c;
    if (f <= 0.5) {
        double a = f * HS_PI_2;
        s = hs_sin_red(a);
        c = hs_cos_red(a);
    } else {
        double a = (1.0 - f) * HS_PI_2;
        s = hs_cos_red(a);
        c = hs_sin_red(a);
    }
    /* cos(q pi/2 + th): q=0 c, q=1 -s, q=2 -c, q=3 s */
    if (q == 0) return c;
    if (q == 1) return -s;
    if (q == 2) return -c;
    return s;
}

/* Box-Muller: one standard normal from u1 in (0,1], u2 in [0,1) */
HS_HD double hs_box_muller(double u1, double u2)
{
    return hs_sqrt(-2.0 * hs_log(u1)) * hs_cos2pi(u2);
}

/* ---- elementary functions of generated code (double) ---------------------
 * sin, cos, tan: x = k pi/2 + r, k the integer nearest to x 2/pi and r from a
 * three-part split of pi/2 (33 + 33 + 53 bits: k*hi and k*mid are exact for
 * |k| < 2^20, |x| < 1.6e6; beyond that r loses about log2(|k| / 2^20) bits),
 * then the series of hs_sin_red / hs_cos_red on |r| <= pi/4 and the quadrant
 * k mod 4.  Infinite or NaN x gives NaN.
 * sinh, tanh, expm1, exprel: Taylor series for |x| < 0.5 (no cancellation near
 * 0; truncation below 2^-60 relative), exp beyond; cosh from exp.
 * log1p: log(u) * x / (u - 1) with u = 1 + x (exact where u - 1 is).
 * arctan: odd; x > 1 -> pi/2 - arctan(1/x); x > tan(pi/12) ->
 * pi/6 + arctan((sqrt3 x - 1) / (x + sqrt3)); then the series to t^27 on
 * |t| <= 0.268 (truncation below 2^-59).  arcsin x = arctan(x / sqrt((1-x)(1+x))),
 * arccos x = 2 arctan(sqrt((1-x) / (1+x))); |x| > 1 gives NaN.
 */
#define HS_PIO2_HI 1.5707963267341256     /* pi/2 to 33 significant bits */
#define HS_PIO2_MID 6.077100506303966e-11 /* the next 33 bits */
#define HS_PIO2_LO 2.0222662487959506e-21 /* the rest */
#define HS_2_PI 0.6366197723675814
#define HS_PI 3.141592653589793
#define HS_PI_6 0.5235987755982989
#define HS_SQRT3 1.7320508075688772
#define HS_TAN_PI_12 0.2679491924311227

HS_HD void hs_sincos(double x, double *s, double *c)
{
    if (!(x - x == 0.0)) { /* inf or NaN */
        *s = x - x;
        *c = x - x;
        return;
    }
    const double kd = floor(x * HS_2_PI + 0.5);
    const double r = ((x - kd * HS_PIO2_HI) - kd * HS_PIO2_MID) - kd * HS_PIO2_LO;
    const double sr = hs_sin_red(r), cr = hs_cos_red(r);
    const int q = (int)(kd - 4.0 * floor(kd * 0.25)); /* k mod 4 */
    if (q == 0) {
        *s = sr;
        *c = cr;
    } else if (q == 1) {
        *s = cr;
        *c = -sr;
    } else if (q == 2) {
        *s = -sr;
        *c = -cr;
    } else {
        *s = -cr;
        *c = sr;
    }
}

HS_HD double hs_sin(double x)
{
    double s, c;
    hs_sincos(x, &s, &c);
    return s;
}

HS_HD double hs_cos(double x)
{
    double s, c;
    hs_sincos(x, &s, &c);
    return c;
}

HS_HD double hs_tan(double x)
{
    double s, c;
    hs_sincos(x, &s, &c);
    return s / c;
}

/* (e^x - 1) / x for |x| < 0.5: 1 + x/2! + x^2/3! + ... + x^16/17! */
HS_HD double hs_exprel_series(double x)
{
    double q = 2.8114572543455206e-15;
    q = 4.779477332387385e-14 + x * q;
    q = 7.647163731819816e-13 + x * q;
    q = 1.1470745597729725e-11 + x * q;
    q = 1.6059043836821613e-10 + x * q;
    q = 2.08767569878681e-09 + x * q;
    q = 2.505210838544172e-08 + x * q;
    q = 2.755731922398589e-07 + x * q;
    q = 2.7557319223985893e-06 + x * q;
    q = 2.48015873015873e-05 + x * q;
    q = 0.0001984126984126984 + x * q;
    q = 0.001388888888888889 + x * q;
    q = 0.008333333333333333 + x * q;
    q = 0.041666666666666664 + x * q;
    q = 0.16666666666666666 + x * q;
    q = 0.5 + x * q;
    return 1.0 + x * q;
}

HS_HD double hs_expm1(double x)
{
    if (x > -0.5 && x < 0.5) return x * hs_exprel_series(x);
    return hs_exp(x) - 1.0;
}

HS_HD double hs_exprel(double x)
{
    if (x > -0.5 && x < 0.5) return hs_exprel_series(x);
    return hs_expm1(x) / x;
}

HS_HD double hs_log1p(double x)
{
    const double u = 1.0 + x;
    if (u == 1.0) return x;
    return hs_log(u) * (x / (u - 1.0));
}

HS_HD double hs_sinh(double x)
{
    if (x > -0.5 && x < 0.5) { /* x (1 + x^2/3! + ... + x^16/17!) */
        const double x2 = x * x;
        double q = 2.8114572543455206e-15;
        q = 7.647163731819816e-13 + x2 * q;
        q = 1.6059043836821613e-10 + x2 * q;
        q = 2.505210838544172e-08 + x2 * q;
        q = 2.7557319223985893e-06 + x2 * q;
        q = 0.0001984126984126984 + x2 * q;
        q = 0.008333333333333333 + x2 * q;
        q = 0.16666666666666666 + x2 * q;
        return x + x * (x2 * q);
    }
    const double e = hs_exp(x < 0.0 ? -x : x);
    const double r = 0.5 * (e - 1.0 / e);
    return x < 0.0 ? -r : r;
}

HS_HD double hs_cosh(double x)
{
    const double e = hs_exp(x < 0.0 ? -x : x);
    return 0.5 * (e + 1.0 / e);
}

HS_HD double hs_tanh(double x)
{
    if (x > -0.5 && x < 0.5) {
        const double em = hs_expm1(2.0 * x);
        return em / (em + 2.0);
    }
    const double e = hs_exp(x < 0.0 ? -2.0 * x : 2.0 * x);
    const double r = 1.0 - 2.0 / (e + 1.0);
    return x < 0.0 ? -r : r;
}

HS_HD double hs_atan(double x)
{
    if (x != x) return x;
    double t = x < 0.0 ? -x : x, base = 0.0;
    const int inv = t > 1.0;
    if (inv) t = 1.0 / t;
    if (t > HS_TAN_PI_12) {
        t = (HS_SQRT3 * t - 1.0) / (t + HS_SQRT3);
        base = HS_PI_6;
    }
    const double t2 = t * t;
    double p = -0.037037037037037035;
    p = 0.04 + t2 * p;
    p = -0.043478260869565216 + t2 * p;
    p = 0.047619047619047616 + t2 * p;
    p = -0.05263157894736842 + t2 * p;
    p = 0.058823529411764705 + t2 * p;
    p = -0.06666666666666667 + t2 * p;
    p = 0.07692307692307693 + t2 * p;
    p = -0.09090909090909091 + t2 * p;
    p = 0.1111111111111111 + t2 * p;
    p = -0.14285714285714285 + t2 * p;
    p = 0.2 + t2 * p;
    p = -0.3333333333333333 + t2 * p;
    double r = base + (t + t * (t2 * p));
    if (inv) r = HS_PI_2 - r;
    return x < 0.0 ? -r : r;
}

HS_HD double hs_asin(double x)
{
    if (!(x >= -1.0 && x <= 1.0)) return (x - x) / (x - x); /* NaN */
    if (x == 1.0) return HS_PI_2;
    if (x == -1.0) return -HS_PI_2;
    return hs_atan(x / hs_sqrt((1.0 - x) * (1.0 + x)));
}

HS_HD double hs_acos(double x)
{
    if (!(x >= -1.0 && x <= 1.0)) return (x - x) / (x - x); /* NaN */
    if (x == -1.0) return HS_PI;
    return 2.0 * hs_atan(hs_sqrt((1.0 - x) / (1.0 + x)));
}

/* ---- single precision (the noise of the FP32 form) ------------------------
 * Same methods with float arithmetic throughout: the series stop where the
 * truncation falls below 2^-26 relative.  IEEE float + - * / sqrt are correctly
 * rounded on both sides (x86 SSE; gfx950 with correctly rounded float divide
 * and sqrt, clang's default for HIP).
 */
HS_HD float hs_logf(float x)
{
    int e;
    float m = frexpf(x, &e); /* m in [0.5, 1) */
    if (m < 0.70710677f) {
        m = m * 2.0f;
        e = e - 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);          /* |s| <= 0.1716 */
    float s2 = s * s;
    float q = 0.11111111f;                        /* 1/9; next term 2 s^11/11 < 2e-9 |lm| */
    q = 0.14285715f + s2 * q;                     /* 1/7 */
    q = 0.2f + s2 * q;                            /* 1/5 */
    q = 0.33333334f + s2 * q;                     /* 1/3 */
    float lm = 2.0f * s + (2.0f * s) * (s2 * q);
    float ed = (float)e;                          /* |e| < 2^8: ed * HS_LN2F_HI exact */
    return ed * HS_LN2F_HI + (ed * HS_LN2F_LO + lm);
}

HS_HD float hs_sqrtf(float x)
{
    return sqrtf(x);
}

/* sin(a), cos(a) for a in [0, pi/4]: to a^9 (trunc. 1.7e-9), to a^10 (1.1e-10) */
HS_HD float hs_sin_redf(float a)
{
    float a2 = a * a;
    float p = 2.7557319e-06f;                     /* 1/9! */
    p = -0.0001984127f + a2 * p;                  /* -1/7! */
    p = 0.008333334f + a2 * p;                    /* 1/5! */
    p = -0.16666667f + a2 * p;                    /* -1/3! */
    return a + a * (a2 * p);
}

HS_HD float hs_cos_redf(float a)
{
    float a2 = a * a;
    float p = 2.7557319e-07f;                     /* 1/10! */
    p = -2.4801588e-05f + a2 * p;                 /* -1/8! */
    p = 0.0013888889f + a2 * p;                   /* 1/6! */
    p = -0.041666668f + a2 * p;                   /* -1/4! */
    p = 0.5f + a2 * p;
    return 1.0f - a2 * p;
}

/* cos(2 pi u) for u in [0, 1) */
HS_HD float hs_cos2pif(float u)
{
    float t = u * 4.0f;
    float qf = floorf(t);
    float f = t - qf; /* exact */
    int q = (int)qf & 3;
    float s, c;
    if (f <= 0.5f) {
        float a = f * 1.5707964f;
        s = hs_sin_redf(a);
        c = hs_cos_redf(a);
    } else {
        float a = (1.0f - f) * 1.5707964f;
        s = hs_cos_redf(a);
        c = hs_sin_redf(a);
    }
    if (q == 0) return c;
    if (q == 1) return -s;
    if (q == 2) return -c;
    return s;
}

HS_HD float hs_box_mullerf(float u1, float u2)
{
    return hs_sqrtf(-2.0f * hs_logf(u1)) * hs_cos2pif(u2);
}

#endif /* HS_DETMATH_H */
