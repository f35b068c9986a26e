/*
 * cpu_ref_mc.c -- THE ORACLE for multicompartment cells (test infrastructure
 * only; see cpu_ref.c for the rules).
 *
 * A batch of independent SpatialNeuron trees, each advanced exactly as the
 * reference's spatial state updater does it, serially per tree:
 *   prep     spatialstateupdate.cu:18-93 (itself cpp_standalone's
 *            spatialneuron_prepare.cpp): inverse axial resistances, the
 *            tridiagonal bands _ab_star0/1/2, boundary terms _invr0/_invrn and
 *            the homogeneous right-hand sides _b_plus/_b_minus
 *   step     1. channel currents linearised at the start of the step
 *               (_gtot_all, _I0_all; :108-113) and the gates advanced
 *            2. Thomas elimination per section with three right-hand sides
 *               (_tridiagsolve_kernel, :128-192)
 *            3. the coupling system over the section end points, built and
 *               solved in the reference's loop order (_coupling_kernel, :202-307)
 *            4. v = v_star + B[parent] u_minus + B[i+1] u_plus (_combine_kernel,
 *               :316-351)
 * The channel expressions are the shared contract (hs_model_body.inc,
 * hs_mc_channels), evaluated in double or float; the solve is double.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../brian2cuda_amd/csrc/hs_model_exprs.h"
#include "cpu_ref.h"

typedef struct {
    int nc, ns;            /* compartments, sections */
    int c0, s0;            /* offsets into the batch arrays */
    int *first, *end;      /* per section: [first, end) local compartments */
    int *mparent;          /* _morph_parent_i: node of the section's start */
    int *nchild;           /* per node */
    int *child;            /* per node, row length rowlen */
    int *idxchild;         /* per section: index among its parent node's children */
    int rowlen;
    double *invr, *ab0, *ab1, *ab2, *bplus, *bminus, *invr0, *invrn;
    double *vstar, *uplus, *uminus, *c;
    double *Pdiag, *Pparent, *Pchildren, *B;
} ref_tree;

struct ref_mc {
    int prec;
    double dt, Ri, vth;
    int ntrees;
    int64_t ncomp;
    double chan[6];
    double *v, *m, *n, *h, *I, *area, *Cm, *gtot, *I0;
    int32_t *spikes;
    int64_t step;
    ref_tree *t;
};

static void prep_tree(ref_mc *M, ref_tree *T, const double *rl1, const double *rl2)
{
    const int N = T->nc;
    const double *area = M->area + T->c0, *Cm = M->Cm + T->c0;
    const double Ri = M->Ri, dt = M->dt;
    /* spatialstateupdate.cu:39-83 */
    for (int i = 1; i < N; i++) T->invr[i] = 1.0 / (Ri * (1 / rl2[i - 1] + 1 / rl1[i]));
    for (int s = 0; s < T->ns; s++) T->invr[T->first[s]] = 0;
    for (int i = 0; i < N; i++) T->ab1[i] = (-(Cm[i] / dt) - T->invr[i] / area[i]);
    for (int i = 1; i < N; i++) {
        T->ab0[i] = T->invr[i] / area[i - 1];
        T->ab2[i - 1] = T->invr[i] / area[i];
        T->ab1[i - 1] -= T->invr[i] / area[i - 1];
    }
    for (int s = 0; s < T->ns; s++) {
        const int first = T->first[s], last = T->end[s] - 1;
        const double invr0 = rl1[first] / Ri, invrn = rl2[last] / Ri;
        T->invr0[s] = invr0;
        T->invrn[s] = invrn;
        T->ab1[first] -= (invr0 / area[first]);
        T->ab1[last] -= (invrn / area[last]);
        T->bplus[last] = -(invrn / area[last]);
        T->bminus[first] = -(invr0 / area[first]);
    }
}

ref_mc *ref_mc_create(int precision, double dt, int ntrees, const int32_t *tree_comp, const int32_t *tree_sec,
                      const int32_t *sec_first, const int32_t *sec_n, const int32_t *sec_parent, const double *area,
                      const double *rl1, const double *rl2, const double *Cm, double Ri, const double *chan,
                      double vth)
{
    if ((precision != 64 && precision != 32) || ntrees <= 0) return NULL;
    ref_mc *M = (ref_mc *)calloc(1, sizeof(ref_mc));
    M->prec = precision;
    M->dt = dt;
    M->Ri = Ri;
    M->vth = vth;
    M->ntrees = ntrees;
    M->ncomp = tree_comp[ntrees];
    memcpy(M->chan, chan, sizeof(M->chan));
    size_t nc = (size_t)M->ncomp;
    M->v = calloc(nc, sizeof(double));
    M->m = calloc(nc, sizeof(double));
    M->n = calloc(nc, sizeof(double));
    M->h = calloc(nc, sizeof(double));
    M->I = calloc(nc, sizeof(double));
    M->area = malloc(nc * sizeof(double));
    M->Cm = malloc(nc * sizeof(double));
    M->gtot = calloc(nc, sizeof(double));
    M->I0 = calloc(nc, sizeof(double));
    memcpy(M->area, area, nc * sizeof(double));
    memcpy(M->Cm, Cm, nc * sizeof(double));
    M->spikes = calloc((size_t)ntrees, sizeof(int32_t));
    M->t = calloc((size_t)ntrees, sizeof(ref_tree));
    for (int k = 0; k < ntrees; k++) {
        ref_tree *T = &M->t[k];
        T->c0 = tree_comp[k];
        T->nc = tree_comp[k + 1] - tree_comp[k];
        T->s0 = tree_sec[k];
        T->ns = tree_sec[k + 1] - tree_sec[k];
        const int ns = T->ns, N = T->nc;
        T->first = malloc(ns * sizeof(int));
        T->end = malloc(ns * sizeof(int));
        T->mparent = malloc(ns * sizeof(int));
        T->idxchild = calloc(ns, sizeof(int));
        T->nchild = calloc(ns + 1, sizeof(int));
        T->rowlen = 1;
        for (int s = 0; s < ns; s++) {
            T->first[s] = sec_first[T->s0 + s];
            T->end[s] = sec_first[T->s0 + s] + sec_n[T->s0 + s];
            const int p = sec_parent[T->s0 + s];
            T->mparent[s] = p < 0 ? 0 : p + 1;
            T->idxchild[s] = T->nchild[T->mparent[s]]++;
        }
        for (int i = 0; i <= ns; i++)
            if (T->nchild[i] > T->rowlen) T->rowlen = T->nchild[i];
        T->child = calloc((size_t)(ns + 1) * T->rowlen, sizeof(int));
        for (int s = 0; s < ns; s++) T->child[T->mparent[s] * T->rowlen + T->idxchild[s]] = s + 1;
        T->invr = calloc(N, sizeof(double));
        T->ab0 = calloc(N, sizeof(double));
        T->ab1 = calloc(N, sizeof(double));
        T->ab2 = calloc(N, sizeof(double));
        T->bplus = calloc(N, sizeof(double));
        T->bminus = calloc(N, sizeof(double));
        T->invr0 = calloc(ns, sizeof(double));
        T->invrn = calloc(ns, sizeof(double));
        T->vstar = calloc(N, sizeof(double));
        T->uplus = calloc(N, sizeof(double));
        T->uminus = calloc(N, sizeof(double));
        T->c = calloc(N, sizeof(double));
        T->Pdiag = calloc(ns + 1, sizeof(double));
        T->Pparent = calloc(ns, sizeof(double));
        T->Pchildren = calloc((size_t)(ns + 1) * T->rowlen, sizeof(double));
        T->B = calloc(ns + 1, sizeof(double));
        prep_tree(M, T, rl1 + T->c0, rl2 + T->c0);
    }
    return M;
}

static double *mc_var(ref_mc *M, int var)
{
    switch (var) {
    case 0: return M->v;
    case 1: return M->m;
    case 2: return M->n;
    case 3: return M->h;
    case 4: return M->I;
    default: return NULL;
    }
}

int ref_mc_set(ref_mc *M, int var, const double *values)
{
    double *a = mc_var(M, var);
    if (!a) return -1;
    for (int64_t i = 0; i < M->ncomp; i++) {
        a[i] = values[i];
        if (M->prec == 32 && var >= 1 && var <= 3) a[i] = (double)(float)values[i];
    }
    return 0;
}

int ref_mc_get(ref_mc *M, int var, double *out)
{
    double *a = mc_var(M, var);
    if (!a) return -1;
    memcpy(out, a, sizeof(double) * (size_t)M->ncomp);
    return 0;
}

static void tree_step(ref_mc *M, ref_tree *T, int tree)
{
    const int N = T->nc, ns = T->ns;
    double *v = M->v + T->c0;
    const double *Cm = M->Cm + T->c0, *area = M->area + T->c0, *Ipt = M->I + T->c0;
    const double dt = M->dt;
    const double v_soma_old = v[0];
    /* 1. channels: gtot, I0 at the start of the step, gates advanced */
    for (int j = 0; j < N; j++) {
        const int64_t gj = T->c0 + j;
        double gtot, I0;
        if (M->prec == 64) {
            double k[7] = {M->chan[0], M->chan[1], M->chan[2], M->chan[3], M->chan[4], M->chan[5], dt};
            double g[3] = {M->m[gj], M->n[gj], M->h[gj]};
            hs_mc_channels_f64(v[j], g, k, &gtot, &I0);
            M->m[gj] = g[0]; M->n[gj] = g[1]; M->h[gj] = g[2];
        } else {
            float k[7] = {(float)M->chan[0], (float)M->chan[1], (float)M->chan[2], (float)M->chan[3],
                          (float)M->chan[4], (float)M->chan[5], (float)dt};
            float g[3] = {(float)M->m[gj], (float)M->n[gj], (float)M->h[gj]};
            float gt, i0;
            hs_mc_channels_f32((float)v[j], g, k, &gt, &i0);
            M->m[gj] = g[0]; M->n[gj] = g[1]; M->h[gj] = g[2];
            gtot = gt;
            I0 = i0;
        }
        M->gtot[gj] = gtot;
        M->I0[gj] = I0 + Ipt[j] / area[j]; /* point current (amp) enters as current density */
    }
    const double *gt = M->gtot + T->c0, *i0 = M->I0 + T->c0;
    /* 2. Thomas per section, three right-hand sides (spatialstateupdate.cu:151-191) */
    for (int s = 0; s < ns; s++) {
        const int js = T->first[s], je = T->end[s];
        double ai, bi, m;
        for (int j = js; j < je; j++) {
            T->vstar[j] = -(Cm[j] / dt * v[j]) - i0[j];
            T->uplus[j] = T->bplus[j];
            T->uminus[j] = T->bminus[j];
            bi = T->ab1[j] - gt[j];
            if (j < N - 1) T->c[j] = T->ab0[j + 1];
            if (j > 0) {
                ai = T->ab2[j - 1];
                m = 1.0 / (bi - ai * T->c[j - 1]);
                T->c[j] = T->c[j] * m;
                T->vstar[j] = (T->vstar[j] - ai * T->vstar[j - 1]) * m;
                T->uplus[j] = (T->uplus[j] - ai * T->uplus[j - 1]) * m;
                T->uminus[j] = (T->uminus[j] - ai * T->uminus[j - 1]) * m;
            } else {
                T->c[0] = T->c[0] / bi;
                T->vstar[0] = T->vstar[0] / bi;
                T->uplus[0] = T->uplus[0] / bi;
                T->uminus[0] = T->uminus[0] / bi;
            }
        }
        for (int j = je - 2; j >= js; j--) {
            T->vstar[j] = T->vstar[j] - T->c[j] * T->vstar[j + 1];
            T->uplus[j] = T->uplus[j] - T->c[j] * T->uplus[j + 1];
            T->uminus[j] = T->uminus[j] - T->c[j] * T->uminus[j + 1];
        }
    }
    /* 3. coupling system (spatialstateupdate.cu:224-305) */
    const int nB = ns + 1, rl = T->rowlen;
    for (int i = 0; i < nB - 1; i++) {
        const int ip = T->mparent[i], ic = T->idxchild[i];
        const int first = T->first[i], last = T->end[i] - 1;
        const double invr0 = T->invr0[i], invrn = T->invrn[i];
        if (i == 0) {
            T->Pdiag[0] = T->uminus[first] - 1;
            T->Pchildren[0 * rl + 0] = T->uplus[first];
            T->B[0] = -T->vstar[first];
        } else {
            T->Pdiag[ip] += (1 - T->uminus[first]) * invr0;
            T->Pchildren[ip * rl + ic] = -T->uplus[first] * invr0;
            T->B[ip] += T->vstar[first] * invr0;
        }
        T->Pdiag[i + 1] = (1 - T->uplus[last]) * invrn;
        T->Pparent[i] = -T->uminus[last] * invrn;
        T->B[i + 1] = T->vstar[last] * invrn;
    }
    for (int i = nB - 1; i >= 0; i--) {
        const int nch = T->nchild[i];
        for (int k = 0; k < nch; k++) {
            const int j = T->child[i * rl + k];
            const double subfac = T->Pchildren[i * rl + k] / T->Pdiag[j];
            T->Pdiag[i] = T->Pdiag[i] - subfac * T->Pparent[j - 1];
            T->B[i] = T->B[i] - subfac * T->B[j];
        }
    }
    T->B[0] = T->B[0] / T->Pdiag[0];
    for (int i = 1; i < nB; i++) {
        const int j = T->mparent[i - 1];
        T->B[i] = T->B[i] - T->Pparent[i - 1] * T->B[j];
        T->B[i] = T->B[i] / T->Pdiag[i];
    }
    /* 4. combine (spatialstateupdate.cu:339-349) */
    for (int s = 0; s < ns; s++) {
        const int ip = T->mparent[s];
        for (int j = T->first[s]; j < T->end[s]; j++)
            v[j] = T->vstar[j] + T->B[ip] * T->uminus[j] + T->B[s + 1] * T->uplus[j];
    }
    if (v[0] > M->vth && !(v_soma_old > M->vth)) M->spikes[tree]++;
}

int ref_mc_run(ref_mc *M, int64_t nsteps)
{
    for (int64_t s = 0; s < nsteps; s++) {
        for (int k = 0; k < M->ntrees; k++) tree_step(M, &M->t[k], k);
        M->step++;
    }
    return 0;
}

int ref_mc_spike_counts(ref_mc *M, int32_t *out)
{
    memcpy(out, M->spikes, sizeof(int32_t) * M->ntrees);
    return 0;
}

void ref_mc_destroy(ref_mc *M)
{
    if (!M) return;
    for (int k = 0; k < M->ntrees; k++) {
        ref_tree *T = &M->t[k];
        free(T->first); free(T->end); free(T->mparent); free(T->idxchild); free(T->nchild); free(T->child);
        free(T->invr); free(T->ab0); free(T->ab1); free(T->ab2); free(T->bplus); free(T->bminus);
        free(T->invr0); free(T->invrn); free(T->vstar); free(T->uplus); free(T->uminus); free(T->c);
        free(T->Pdiag); free(T->Pparent); free(T->Pchildren); free(T->B);
    }
    free(M->t); free(M->v); free(M->m); free(M->n); free(M->h); free(M->I); free(M->area); free(M->Cm);
    free(M->gtot); free(M->I0); free(M->spikes);
    free(M);
}
