/*
 * hs_philox.h -- counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
 * "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11) and the draws the hot
 * path takes from it.
 *
 * Replaces the cuRAND host generator + pre-generated buffers of the reference
 * (brian2cuda/templates/rand.cu:46-561, brian2cuda/brianlib/curand_buffer.h:20-160)
 * and the per-thread curand states (rand.cu:25-42): a draw is a pure function of
 * (seed, stream, index, step), so there is no buffer to refill, no state to
 * initialise and no dependence on launch geometry (both precisions draw from the
 * same Philox words).
 *
 * Counter layout of one draw:  ctr = (index, step_lo, step_hi, stream),
 *                              key = (seed_lo, seed_hi).
 * One Philox call gives 4 words -> two 53-bit uniforms -> one Box-Muller normal
 * (the FP32 form: two 24-bit uniforms and float arithmetic).
 */
#ifndef HS_PHILOX_H
#define HS_PHILOX_H

#ifndef __HIPCC_RTC__
#include <stdint.h>
#endif
#include "hs_detmath.h"

#define HS_PHILOX_M0 0xD2511F53u
#define HS_PHILOX_M1 0xCD9E8D57u
#define HS_PHILOX_W0 0x9E3779B9u
#define HS_PHILOX_W1 0xBB67AE85u

typedef struct {
    uint32_t v[4];
} hs_u32x4;

HS_HD hs_u32x4 hs_philox4x32_10(hs_u32x4 c, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)HS_PHILOX_M0 * (uint64_t)c.v[0];
        uint64_t p1 = (uint64_t)HS_PHILOX_M1 * (uint64_t)c.v[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        hs_u32x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += HS_PHILOX_W0;
        k1 += HS_PHILOX_W1;
    }
    return c;
}

/* 2^-53 */
#define HS_TWO_M53 1.1102230246251565e-16

/* uniform in [0, 1) with 53 random bits from two words */
HS_HD double hs_u53(uint32_t a, uint32_t b)
{
    uint64_t k = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return (double)k * HS_TWO_M53;
}

/* standard normal for (seed, stream, index, step) -- identical bits on host and device */
HS_HD double hs_randn(uint64_t seed, uint32_t stream, uint32_t index, int64_t step)
{
    hs_u32x4 c;
    c.v[0] = index;
    c.v[1] = (uint32_t)((uint64_t)step);
    c.v[2] = (uint32_t)((uint64_t)step >> 32);
    c.v[3] = stream;
    hs_u32x4 w = hs_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double u1 = hs_u53(w.v[0], w.v[1]) + HS_TWO_M53; /* (0, 1] */
    double u2 = hs_u53(w.v[2], w.v[3]);              /* [0, 1) */
    return hs_box_muller(u1, u2);
}

/* the FP32 form's standard normal: the same Philox words, 24-bit uniforms
   u1 = (w0>>8 + 1) 2^-24 in (0, 1], u2 = (w2>>8) 2^-24 in [0, 1), and the
   Box-Muller transform in float arithmetic */
HS_HD float hs_randn_f32(uint64_t seed, uint32_t stream, uint32_t index, int64_t step)
{
    hs_u32x4 c;
    c.v[0] = index;
    c.v[1] = (uint32_t)((uint64_t)step);
    c.v[2] = (uint32_t)((uint64_t)step >> 32);
    c.v[3] = stream;
    hs_u32x4 w = hs_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    float u1 = (float)((w.v[0] >> 8) + 1u) * 5.9604645e-08f; /* 2^-24 */
    float u2 = (float)(w.v[2] >> 8) * 5.9604645e-08f;
    return hs_box_mullerf(u1, u2);
}

/* uniform in [0,1) for (seed, stream, index, step) */
HS_HD double hs_rand(uint64_t seed, uint32_t stream, uint32_t index, int64_t step)
{
    hs_u32x4 c;
    c.v[0] = index;
    c.v[1] = (uint32_t)((uint64_t)step);
    c.v[2] = (uint32_t)((uint64_t)step >> 32);
    c.v[3] = stream;
    hs_u32x4 w = hs_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return hs_u53(w.v[0], w.v[1]);
}

#endif /* HS_PHILOX_H */
