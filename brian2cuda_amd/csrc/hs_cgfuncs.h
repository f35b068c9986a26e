/*
 * hs_cgfuncs.h -- the trigonometric, hyperbolic and related functions that
 * generated code (code models, synaptic code, connect expressions) may call,
 * under the names the generated text uses (hs_cg_sin ... hs_cg_exprel, mapped
 * from Brian2's sin ... exprel by macros after the engine sources).  Each
 * evaluates hs_detmath.h's double function: the device and the host oracle
 * compute the same bits, and an FP32 argument gets the double result rounded
 * once to float.
 */
#ifndef HS_CGFUNCS_H
#define HS_CGFUNCS_H

#include "hs_detmath.h"
#include "hs_philox.h"

#define HS_CG_FUNC1(name, fn)                                                 \
    HS_HD double name(double x) { return fn(x); }                             \
    HS_HD float name(float x) { return (float)fn((double)x); }                \
    HS_HD double name(int x) { return fn((double)x); }                        \
    HS_HD double name(long long x) { return fn((double)x); }

HS_CG_FUNC1(hs_cg_sin, hs_sin)
HS_CG_FUNC1(hs_cg_cos, hs_cos)
HS_CG_FUNC1(hs_cg_tan, hs_tan)
HS_CG_FUNC1(hs_cg_sinh, hs_sinh)
HS_CG_FUNC1(hs_cg_cosh, hs_cosh)
HS_CG_FUNC1(hs_cg_tanh, hs_tanh)
HS_CG_FUNC1(hs_cg_arcsin, hs_asin)
HS_CG_FUNC1(hs_cg_arccos, hs_acos)
HS_CG_FUNC1(hs_cg_arctan, hs_atan)
HS_CG_FUNC1(hs_cg_expm1, hs_expm1)
HS_CG_FUNC1(hs_cg_log1p, hs_log1p)
HS_CG_FUNC1(hs_cg_exprel, hs_exprel)

#undef HS_CG_FUNC1

/* the stream of rand() (normal = 0) or randn() (normal = 1) call site k < 16,
   or of noise variable k (normal = 2, k >= 1), of a code model whose group
   draws from stream (below 2^24): the group's stream in the low 24 bits,
   1 + k + 16 normal above, apart from its xi (stream itself) and from every
   other call site */
HS_HD unsigned int hs_cg_rnd_stream(unsigned int stream, int k, int normal)
{
    return (stream & 0x00FFFFFFu) | ((unsigned int)(1 + k + 16 * normal) << 24);
}

/* the stream of rand() (normal = 0) or randn() (normal = 1) call site k < 16 of
   pathway pid's synaptic code: the pathway in the low 24 bits, 64 + k + 16 normal
   above (apart from every code-model stream, whose top byte is below 64) */
HS_HD unsigned int hs_cg_syn_stream(unsigned int pid, int k, int normal)
{
    return (pid & 0x00FFFFFFu) | ((unsigned int)(64 + k + 16 * normal) << 24);
}

/* the step word of a synapse draw: target j in the high 32 bits, the step's low
   32 bits below (so the Philox counter is (i, step_lo, j, stream); runs shorter
   than 2^32 steps never repeat a counter) */
HS_HD long long hs_cg_syn_ctr(long long step, int j)
{
    return (long long)(((unsigned long long)(unsigned int)j << 32) | (unsigned long long)(unsigned int)step);
}

/* the stream of rand() (normal = 0) or randn() (normal = 1) call site k < 16 of
   a connect condition or probability: the connect call's stream in the low 24
   bits, 96 + k + 16 normal above */
HS_HD unsigned int hs_cg_conn_stream(unsigned int stream, int k, int normal)
{
    return (stream & 0x00FFFFFFu) | ((unsigned int)(96 + k + 16 * normal) << 24);
}

/* the step word of a connect-code draw: j in the high 32 bits over the tag
   0x43524E44 ("CRND") plus 256 x the synapse copy (connect(n=...); 0 in
   conditions and probabilities, which see the pair once) */
HS_HD long long hs_cg_conn_ctr(long long j, int copy)
{
    return (long long)(((unsigned long long)(unsigned int)j << 32) |
                       (0x43524E44ull + ((unsigned long long)(unsigned int)copy << 8)));
}

#endif /* HS_CGFUNCS_H */
