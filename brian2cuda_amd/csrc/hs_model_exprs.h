/*
 * hs_model_exprs.h -- double and float instances of the model expressions
 * (hs_model_body.inc) plus the host-side derivation of their per-group
 * constants.  Included by the HIP kernels and by the host oracle.
 */
#ifndef HS_MODEL_EXPRS_H
#define HS_MODEL_EXPRS_H

#include "hs_detmath.h"

/* the exact instances: IEEE division, no contraction (the numerical contract) */
#define HS_DIV(a, b) ((a) / (b))
#define HS_FAST_CONTRACT

#define R double
#define HS_SFX(name) name##_f64
#define HS_EXP hs_exp
#include "hs_model_body.inc"
#undef R
#undef HS_SFX
#undef HS_EXP

#define R float
#define HS_SFX(name) name##_f32
#define HS_EXP hs_expf
#include "hs_model_body.inc"
#undef R
#undef HS_SFX
#undef HS_EXP
#undef HS_DIV
#undef HS_FAST_CONTRACT

/*
 * Brian2's `timestep` function (brian2cuda/cuda_generator.py:997-1006):
 * (int64_t)((t + 1e-3*dt)/dt).  Used for the refractory period in steps.
 */
HS_HD long long hs_timestep(double t, double dt)
{
    return (long long)((t + 1e-3 * dt) / dt);
}

/*
 * Per-group constants, computed once on the host in double with the
 * deterministic exp so that every caller derives the same bits.
 * Returns the number of constants written to k (<= 16), or -1 for a bad model
 * or parameters outside the closed form (LIF_CUBA: taum, taue, taui > 0 and
 * taue, taui != taum, whose Ax divide by taux - taum).
 * p holds every parameter of the model, optional ones included (hs_model_params).
 *   LIF_CUBA : params El,Vt,Vr,taum,taue,taui,ref,free -> k = Em,Ee,Ei,Ae,Ai,El,free
 *   LIF_NOISE: params tau,mu,sigma,theta,Vr,ref,stream -> k = a,b,mu
 *   POISSON  : params rate [,stream] -> k = dt
 *   HH_COBA  : params Cm,gl,El,EK,ENa,gNa,gK,VT,Ee,Ei,taue,taui,vth,ref
 *              -> k = Cm,gl,El,EK,ENa,gNa,gK,VT,Ee,Ei,dec_e,dec_i,dt
 */
static inline int hs_model_constants(int model, const double *p, double dt, double *k)
{
    if (model == 1) { /* HS_MODEL_LIF_CUBA */
        double taum = p[3], taue = p[4], taui = p[5];
        if (!(taum > 0.0 && taue > 0.0 && taui > 0.0) || taue == taum || taui == taum) return -1;
        double Em = hs_exp(-dt / taum), Ee = hs_exp(-dt / taue), Ei = hs_exp(-dt / taui);
        k[0] = Em;
        k[1] = Ee;
        k[2] = Ei;
        k[3] = taue / (taue - taum) * (Ee - Em);
        k[4] = taui / (taui - taum) * (Ei - Em);
        k[5] = p[0];
        k[6] = p[7];
        return 7;
    }
    if (model == 2) { /* HS_MODEL_LIF_NOISE */
        double tau = p[0], mu = p[1], sigma = p[2];
        k[0] = dt / tau;
        k[1] = hs_sqrt(dt) * (sigma * hs_sqrt(tau) / tau);
        k[2] = mu;
        return 3;
    }
    if (model == 3) { /* HS_MODEL_HH_COBA */
        for (int i = 0; i < 10; ++i) k[i] = p[i];
        k[10] = hs_exp(-dt / p[10]);
        k[11] = hs_exp(-dt / p[11]);
        k[12] = dt;
        return 13;
    }
    if (model == 4) /* HS_MODEL_SPIKEGEN: no state update */
        return 0;
    if (model == 5) { /* HS_MODEL_POISSON: the firing probability is rates * k[0] */
        if (!(p[0] >= 0.0)) return -1;
        k[0] = dt;
        return 1;
    }
    return -1;
}

/*
 * A model's parameter vector with its optional trailing parameters filled in
 * (zero = the default): out[16] from np given values.  Returns the model's full
 * count, or -1 if np is not a count the model accepts.
 */
static inline int hs_model_params(int model, const double *p, int np, double *out)
{
    int lo, hi;
    if (model == 1) { lo = 7; hi = 8; }       /* LIF_CUBA [, free] */
    else if (model == 2) { lo = hi = 7; }     /* LIF_NOISE */
    else if (model == 3) { lo = hi = 14; }    /* HH_COBA */
    else if (model == 4) { lo = hi = 1; }     /* SPIKEGEN */
    else if (model == 5) { lo = 1; hi = 2; }  /* POISSON [, stream] */
    else return -1;
    if (np < lo || np > hi || (np > 0 && !p)) return -1;
    for (int i = 0; i < 16; ++i) out[i] = i < np ? p[i] : 0.0;
    return hi;
}

#endif /* HS_MODEL_EXPRS_H */
