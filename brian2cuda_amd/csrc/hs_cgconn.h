/*
 * hs_cgconn.h -- the operators of connect conditions and probabilities given as
 * expressions (Synapses.connect(condition=..., p=...)) that C does not spell
 * the way Brian2's expressions mean them: floor division and the remainder
 * with the sign of the divisor (Python's // and %), and a power with a
 * non-literal exponent.  Shared by the kernels compiled at run time and the
 * host oracle, so both evaluate an expression with the same operations.
 * Integer operands are long long (i, j, N_pre, N_post and integer literals),
 * real ones double.
 */
#ifndef HS_CGCONN_H
#define HS_CGCONN_H

#include "hs_detmath.h"

HS_HD long long hs_cg_floordiv(long long a, long long b)
{
    long long q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}
HS_HD double hs_cg_floordiv(double a, double b) { return floor(a / b); }
HS_HD double hs_cg_floordiv(long long a, double b) { return floor((double)a / b); }
HS_HD double hs_cg_floordiv(double a, long long b) { return floor(a / (double)b); }

HS_HD long long hs_cg_pymod(long long a, long long b)
{
    long long r = a % b;
    if (r != 0 && ((r < 0) != (b < 0))) r += b;
    return r;
}
HS_HD double hs_cg_pymod(double a, double b)
{
    double r = fmod(a, b);
    if (r != 0.0 && ((r < 0.0) != (b < 0.0))) r += b;
    return r;
}
HS_HD double hs_cg_pymod(long long a, double b) { return hs_cg_pymod((double)a, b); }
HS_HD double hs_cg_pymod(double a, long long b) { return hs_cg_pymod(a, (double)b); }

/* a**b for a runtime exponent: exp(b log a) with the library's exp and log;
   0**b and negative bases with integral b as the real power defines them */
HS_HD double hs_cg_pow(double a, double b)
{
    if (a > 0.0) return hs_exp(b * hs_log(a));
    if (a == 0.0) return b > 0.0 ? 0.0 : (b == 0.0 ? 1.0 : __builtin_inf());
    const double n = floor(b);
    if (n != b) return __builtin_nan("");
    const double m = hs_exp(b * hs_log(-a));
    return fmod(n, 2.0) != 0.0 ? -m : m;
}

#endif /* HS_CGCONN_H */
