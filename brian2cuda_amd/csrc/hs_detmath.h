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
