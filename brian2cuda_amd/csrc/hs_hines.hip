/*
 * hs_hines.hip -- multicompartment cells on gfx950: a batch of independent
 * SpatialNeuron trees, one wave64 workgroup per tree, the tree's state held on
 * chip (registers + LDS) for all the steps of a launch.
 *
 * Per step, in the reference's arithmetic (spatialstateupdate.cu), parallel
 * where the reference's serial loops have independent iterations:
 *   1. lanes over compartments: HH channels linearised at the start of the
 *      step (_gtot, _I0) and gates advanced (exponential Euler); the main
 *      diagonal and right-hand side of each compartment go to LDS
 *   2. lanes over sections: Thomas elimination + back substitution with three
 *      right-hand sides (v_star, u_plus, u_minus) walking the section in LDS
 *   3. the coupling system over section end points: built node-parallel (each
 *      node adds its children's terms in the reference's order), eliminated
 *      level by level from the leaves, substituted level by level from the root
 *      -- every node sees exactly the operations of the serial loops
 *   4. lanes over compartments: v = v_star + B[parent] u_minus + B[i+1] u_plus
 * The solve is FP64; the channel arithmetic is FP64 or FP32.
 */
#include <stdarg.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "../../include/hs.h"
#include "hs_device.h"

#define HS_HINES_MAXC 1024
#define HS_HINES_MAXS 255
#define HS_HINES_RL 4 /* children per node */
#define HS_HINES_TU 16 /* sections up to this length run the unrolled Thomas sweep */
#ifndef HS_HINES_PF
#define HS_HINES_PF 4 /* its prefetch distance for the off-diagonal bands */
#endif

template <typename R>
struct HinesDesc {
    const int32_t *tree_comp, *tree_sec;
    /* per compartment */
    double *v;
    R *m, *n, *h;
    const double *ab0, *ab1, *ab2, *cdt, *iext;
    const int32_t *csec; /* tree-local section of each compartment */
    /* per section */
    const int32_t *first, *end, *mparent;
    const double *invr0, *invrn, *bplus, *bminus;
    /* per node (tree-local node k of tree t at tree_sec[t] + t + k) */
    const int32_t *nchild, *child, *order, *lvl_start; /* order: nodes by depth; lvl_start per tree at node offset */
    const int32_t *nlevels;                            /* per tree */
    int32_t *spikes;                                   /* per tree */
    R chan[7];
    double vth;
    int cmax, smax;
};

template <typename R>
__device__ __forceinline__ void hs_mc_chan(R v, R *g, const R *k, R *gt, R *i0);
template <>
__device__ __forceinline__ void hs_mc_chan<double>(double v, double *g, const double *k, double *gt, double *i0)
{
    hs_mc_channels_f64(v, g, k, gt, i0);
}
template <>
__device__ __forceinline__ void hs_mc_chan<float>(float v, float *g, const float *k, float *gt, float *i0)
{
    hs_mc_channels_f32(v, g, k, gt, i0);
}

/* up to 8 compartments per lane: two waves per SIMD (2048 trees then fit the
   chip at once: 256 CUs x 4 SIMDs x 2); the register budget this sets covers
   the unrolled Thomas sweep without spills.  16 per lane: one wave per SIMD and
   the plain sweep (the unrolled one would spill) */
template <typename R, int CPL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CPL <= 8 ? 2 : 1))) void hs_k_hines(HinesDesc<R> d, int nsteps)
{
    extern __shared__ double lds[];
    const int tree = blockIdx.x, lane = threadIdx.x;
    const int c0 = d.tree_comp[tree], nc = d.tree_comp[tree + 1] - c0;
    const int s0 = d.tree_sec[tree], ns = d.tree_sec[tree + 1] - s0;
    const int n0 = s0 + tree; /* node offset */
    const int CM = d.cmax, SM = d.smax;
    double *A = lds;            /* main diagonal, then c (superdiagonal after elimination) */
    double *Vs = A + CM;        /* right-hand side, then v_star */
    double *Up = Vs + CM;
    double *Um = Up + CM;
    double *Pd = Um + CM;       /* SM + 1 */
    double *Pp = Pd + SM + 1;   /* SM */
    double *Pc = Pp + SM;       /* (SM + 1) * RL */
    double *Bn = Pc + (SM + 1) * HS_HINES_RL;
    double *i0s = Bn + SM + 1;  /* SM: invr0 */
    double *ins = i0s + SM;     /* SM: invrn */
    double *bps = ins + SM;     /* SM: b_plus at the last compartment */
    double *bms = bps + SM;     /* SM: b_minus at the first compartment */
    int *fst = (int *)(bms + SM);
    int *lst = fst + SM;        /* end (exclusive) */
    int *mpr = lst + SM;
    int *nch = mpr + SM;        /* SM + 1 */
    int *chl = nch + SM + 1;    /* (SM + 1) * RL */
    int *ord = chl + (SM + 1) * HS_HINES_RL; /* SM + 1 */
    int *lvs = ord + SM + 1;    /* SM + 2 */

    /* the off-diagonal bands are constant: read from global memory (L1/L2
       resident after the first step) to keep the workgroup's LDS small enough
       for every tree of a large batch to be resident at once */
    const double *__restrict__ gab0 = d.ab0 + c0;
    const double *__restrict__ gab2 = d.ab2 + c0;
    for (int s = lane; s < ns; s += HS_WAVE) {
        fst[s] = d.first[s0 + s];
        lst[s] = d.end[s0 + s];
        mpr[s] = d.mparent[s0 + s];
        i0s[s] = d.invr0[s0 + s];
        ins[s] = d.invrn[s0 + s];
        bps[s] = d.bplus[s0 + s];
        bms[s] = d.bminus[s0 + s];
    }
    for (int k = lane; k <= ns; k += HS_WAVE) {
        nch[k] = d.nchild[n0 + k];
        for (int r = 0; r < HS_HINES_RL; ++r) chl[k * HS_HINES_RL + r] = d.child[(n0 + k) * HS_HINES_RL + r];
        ord[k] = d.order[n0 + k];
    }
    const int nlev = d.nlevels[tree];
    for (int k = lane; k <= nlev; k += HS_WAVE) lvs[k] = d.lvl_start[n0 + tree + k];

    double v[CPL], ab1[CPL], cdt[CPL], iext[CPL];
    R gm[CPL], gn[CPL], gh[CPL];
    int sec[CPL];
#pragma unroll
    for (int r = 0; r < CPL; ++r) {
        const int j = lane + HS_WAVE * r;
        const bool ok = j < nc;
        v[r] = ok ? d.v[c0 + j] : 0.0;
        ab1[r] = ok ? d.ab1[c0 + j] : 0.0;
        cdt[r] = ok ? d.cdt[c0 + j] : 0.0;
        iext[r] = ok ? d.iext[c0 + j] : 0.0;
        gm[r] = ok ? d.m[c0 + j] : (R)0;
        gn[r] = ok ? d.n[c0 + j] : (R)0;
        gh[r] = ok ? d.h[c0 + j] : (R)0;
        sec[r] = ok ? d.csec[c0 + j] : 0;
    }
    R kc[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) kc[q] = d.chan[q];
    int spikes = 0;
    __syncthreads();

    for (int step = 0; step < nsteps; ++step) {
        const double v_soma_old = v[0]; /* lane 0 owns compartment 0 */
        /* 1. channels (spatialstateupdate.cu:108-113) and the rows of the system */
#pragma unroll
        for (int r = 0; r < CPL; ++r) {
            const int j = lane + HS_WAVE * r;
            if (j < nc) {
                R g3[3] = {gm[r], gn[r], gh[r]};
                R gt, i0;
                hs_mc_chan<R>((R)v[r], g3, kc, &gt, &i0);
                gm[r] = g3[0];
                gn[r] = g3[1];
                gh[r] = g3[2];
                const double I0 = (double)i0 + iext[r];
                A[j] = ab1[r] - (double)gt;
                Vs[j] = -(cdt[r] * v[r]) - I0;
            }
        }
        __syncthreads();
        /* 2. Thomas per section (spatialstateupdate.cu:151-191) */
        for (int s = lane; s < ns; s += HS_WAVE) {
            const int js = fst[s], je = lst[s], last = je - 1;
            double cp = 0.0, vp = 0.0, upp = 0.0, ump = 0.0;
            if (CPL <= 8 && je - js <= HS_HINES_TU) {
                /* a short section: the two off-diagonal bands (global memory,
                   an L2 round trip each) are loaded HS_HINES_PF compartments
                   ahead, not one; the diagonal and right-hand side come from LDS
                   as below; same operations, same order */
                double Ga[HS_HINES_TU], Gc[HS_HINES_TU];
#pragma unroll
                for (int q = 0; q < HS_HINES_TU; ++q) Ga[q] = Gc[q] = 0.0;
#pragma unroll
                for (int q = 0; q < HS_HINES_PF; ++q) {
                    const int j = js + q;
                    if (j < je) {
                        Gc[q] = (j < nc - 1) ? gab0[j + 1] : 0.0;
                        Ga[q] = j > 0 ? gab2[j - 1] : 0.0;
                    }
                }
                const double bp = bps[s], bm = bms[s];
#pragma unroll
                for (int q = 0; q < HS_HINES_TU; ++q) {
                    const int j = js + q;
                    if (q + HS_HINES_PF < HS_HINES_TU && j + HS_HINES_PF < je) {
                        const int jn = j + HS_HINES_PF;
                        Gc[q + HS_HINES_PF] = (jn < nc - 1) ? gab0[jn + 1] : 0.0;
                        Ga[q + HS_HINES_PF] = gab2[jn - 1];
                    }
                    if (j < je) {
                        const double b = A[j], ai = Ga[q];
                        double cj = Gc[q], vs = Vs[j];
                        double up = (j == last) ? bp : 0.0;
                        double um = (q == 0) ? bm : 0.0;
                        if (j > 0) {
                            if (q == 0) {
                                const double m = 1.0 / b;
                                cj = cj * m;
                                vs = vs * m;
                                up = up * m;
                                um = um * m;
                            } else {
                                const double m = 1.0 / (b - ai * cp);
                                cj = cj * m;
                                vs = (vs - ai * vp) * m;
                                up = (up - ai * upp) * m;
                                um = (um - ai * ump) * m;
                            }
                        } else {
                            cj = cj / b;
                            vs = vs / b;
                            up = up / b;
                            um = um / b;
                        }
                        A[j] = cj;
                        Vs[j] = vs;
                        Up[j] = up;
                        Um[j] = um;
                        cp = cj;
                        vp = vs;
                        upp = up;
                        ump = um;
                    }
                }
#pragma unroll
                for (int q = HS_HINES_TU - 2; q >= 0; --q) {
                    const int j = js + q;
                    if (j <= je - 2) {
                        vp = Vs[j] - A[j] * vp;
                        upp = Up[j] - A[j] * upp;
                        ump = Um[j] - A[j] * ump;
                        Vs[j] = vp;
                        Up[j] = upp;
                        Um[j] = ump;
                    }
                }
                continue;
            }
            /* the inputs of compartment j+1 are loaded while j is computed (they
               are not written before their own iteration) */
            double b_n = A[js], vs_n = Vs[js];
            double c_n = (js < nc - 1) ? gab0[js + 1] : 0.0;
            double a_n = js > 0 ? gab2[js - 1] : 0.0;
            for (int j = js; j < je; ++j) {
                const double b = b_n, ai = a_n;
                double vs = vs_n;
                double cj = c_n;
                if (j + 1 < je) {
                    b_n = A[j + 1];
                    vs_n = Vs[j + 1];
                    c_n = (j + 1 < nc - 1) ? gab0[j + 2] : 0.0;
                    a_n = gab2[j];
                }
                double up = (j == last) ? bps[s] : 0.0;
                double um = (j == js) ? bms[s] : 0.0;
                if (j > 0) {
                    if (j == js) {
                        /* section start: the subdiagonal is exactly 0 (its axial
                           conductance is cut), so (x - 0*prev)*m == x*m */
                        const double m = 1.0 / b;
                        cj = cj * m;
                        vs = vs * m;
                        up = up * m;
                        um = um * m;
                    } else {
                        const double m = 1.0 / (b - ai * cp);
                        cj = cj * m;
                        vs = (vs - ai * vp) * m;
                        up = (up - ai * upp) * m;
                        um = (um - ai * ump) * m;
                    }
                } else {
                    cj = cj / b;
                    vs = vs / b;
                    up = up / b;
                    um = um / b;
                }
                A[j] = cj;
                Vs[j] = vs;
                Up[j] = up;
                Um[j] = um;
                cp = cj;
                vp = vs;
                upp = up;
                ump = um;
            }
            if (je - 2 >= js) {
                double c_b = A[je - 2], v_b = Vs[je - 2], u_b = Up[je - 2], m_b = Um[je - 2];
                for (int j = je - 2; j >= js; --j) {
                    const double c = c_b, v0 = v_b, u0 = u_b, m0 = m_b;
                    if (j - 1 >= js) {
                        c_b = A[j - 1];
                        v_b = Vs[j - 1];
                        u_b = Up[j - 1];
                        m_b = Um[j - 1];
                    }
                    vp = v0 - c * vp;
                    upp = u0 - c * upp;
                    ump = m0 - c * ump;
                    Vs[j] = vp;
                    Up[j] = upp;
                    Um[j] = ump;
                }
            }
        }
        __syncthreads();
        /* 3a. build the coupling system (spatialstateupdate.cu:224-265): node k's
           row gets the term of its own section, then its children's in order */
        for (int k = lane; k <= ns; k += HS_WAVE) {
            double pd, b;
            if (k == 0) {
                const int f = fst[0];
                pd = Um[f] - 1;
                Pc[0] = Up[f];
                b = -Vs[f];
            } else {
                const int i = k - 1, l = lst[i] - 1;
                pd = (1 - Up[l]) * ins[i];
                Pp[i] = -Um[l] * ins[i];
                b = Vs[l] * ins[i];
            }
            for (int q = 0; q < nch[k]; ++q) {
                const int c = chl[k * HS_HINES_RL + q]; /* child node = child section + 1 */
                const int cs = c - 1;
                if (cs == 0) continue; /* the root section's start is node 0 itself */
                const int f = fst[cs];
                pd += (1 - Um[f]) * i0s[cs];
                Pc[k * HS_HINES_RL + q] = -Up[f] * i0s[cs];
                b += Vs[f] * i0s[cs];
            }
            Pd[k] = pd;
            Bn[k] = b;
        }
        __syncthreads();
        /* 3b. eliminate the children of every node, deepest level first (:268-287) */
        for (int L = nlev - 1; L >= 0; --L) {
            for (int q = lvs[L] + lane; q < lvs[L + 1]; q += HS_WAVE) {
                const int i = ord[q];
                const int nci = nch[i];
                double pd = Pd[i], b = Bn[i];
                /* every child's operands are loaded (and its factor divided out)
                   before the row update, which then runs in child order */
                double sf[HS_HINES_RL], pp[HS_HINES_RL], bj[HS_HINES_RL];
#pragma unroll
                for (int c = 0; c < HS_HINES_RL; ++c) {
                    if (c < nci) {
                        const int j = chl[i * HS_HINES_RL + c];
                        sf[c] = Pc[i * HS_HINES_RL + c] / Pd[j];
                        pp[c] = Pp[j - 1];
                        bj[c] = Bn[j];
                    }
                }
#pragma unroll
                for (int c = 0; c < HS_HINES_RL; ++c) {
                    if (c < nci) {
                        pd = pd - sf[c] * pp[c];
                        b = b - sf[c] * bj[c];
                    }
                }
                Pd[i] = pd;
                Bn[i] = b;
            }
            __syncthreads();
        }
        /* 3c. substitute from the root (:290-296) */
        if (lane == 0) Bn[0] = Bn[0] / Pd[0];
        __syncthreads();
        for (int L = 1; L < nlev; ++L) {
            for (int q = lvs[L] + lane; q < lvs[L + 1]; q += HS_WAVE) {
                const int i = ord[q];
                const int jp = mpr[i - 1];
                double b = Bn[i] - Pp[i - 1] * Bn[jp];
                Bn[i] = b / Pd[i];
            }
            __syncthreads();
        }
        /* 4. combine (spatialstateupdate.cu:339-349) */
#pragma unroll
        for (int r = 0; r < CPL; ++r) {
            const int j = lane + HS_WAVE * r;
            if (j < nc) {
                const int s = sec[r];
                v[r] = Vs[j] + Bn[mpr[s]] * Um[j] + Bn[s + 1] * Up[j];
            }
        }
        if (lane == 0 && v[0] > d.vth && !(v_soma_old > d.vth)) ++spikes;
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < CPL; ++r) {
        const int j = lane + HS_WAVE * r;
        if (j < nc) {
            d.v[c0 + j] = v[r];
            d.m[c0 + j] = gm[r];
            d.n[c0 + j] = gn[r];
            d.h[c0 + j] = gh[r];
        }
    }
    if (lane == 0) d.spikes[tree] += spikes;
}

/* ------------------------------------------------------------------------ */
/* host side                                                                 */
/* ------------------------------------------------------------------------ */
#define HS_HIP(call)                                                                                        \
    do {                                                                                                    \
        hipError_t _e = (call);                                                                             \
        if (_e != hipSuccess)                                                                               \
            return hs_rt_fail(HS_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                              __LINE__);                                                                    \
    } while (0)

struct hs_hines {
    int prec;
    double dt;
    int ntrees, ncomp, nsec, nnode, cmax, smax;
    std::vector<void *> allocs;
    HinesDesc<double> d64;
    HinesDesc<float> d32;
    long long steps_per_launch = 1000;
    int64_t steps = 0;
    std::vector<double> area, I; /* host copies: the point current is stored as I / area */
    double last_wall = 0, last_gpu_ms = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

static int hs_h_alloc(hs_hines *H, size_t bytes, void **p)
{
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return hs_rt_fail(HS_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    }
    H->allocs.push_back(*p);
    return HS_OK;
}

template <typename T>
static int hs_h_upload(hs_hines *H, const std::vector<T> &host, const T **dev)
{
    void *p = nullptr;
    int r = hs_h_alloc(H, sizeof(T) * host.size(), &p);
    if (r != HS_OK) return r;
    if (!host.empty())
        HS_HIP(hipMemcpyAsync(p, host.data(), sizeof(T) * host.size(), hipMemcpyHostToDevice, hs_rt_stream()));
    *dev = (const T *)p;
    return HS_OK;
}

/* the pointer and size fields shared by the double and float descriptors */
template <typename R>
static void hs_h_copy_common(HinesDesc<R> &o, const HinesDesc<double> &d)
{
    o.tree_comp = d.tree_comp;
    o.tree_sec = d.tree_sec;
    o.v = d.v;
    o.ab0 = d.ab0;
    o.ab1 = d.ab1;
    o.ab2 = d.ab2;
    o.cdt = d.cdt;
    o.iext = d.iext;
    o.csec = d.csec;
    o.first = d.first;
    o.end = d.end;
    o.mparent = d.mparent;
    o.invr0 = d.invr0;
    o.invrn = d.invrn;
    o.bplus = d.bplus;
    o.bminus = d.bminus;
    o.nchild = d.nchild;
    o.child = d.child;
    o.order = d.order;
    o.lvl_start = d.lvl_start;
    o.nlevels = d.nlevels;
    o.spikes = d.spikes;
    o.vth = d.vth;
    o.cmax = d.cmax;
    o.smax = d.smax;
}

extern "C" void hs_hines_destroy(hs_hines *H)
{
    if (!H) return;
    (void)hipStreamSynchronize(hs_rt_stream());
    for (void *p : H->allocs) (void)hipFree(p);
    if (H->e0) (void)hipEventDestroy(H->e0);
    if (H->e1) (void)hipEventDestroy(H->e1);
    delete H;
}

extern "C" int hs_hines_create(int precision, double dt, int n_trees, const int32_t *tree_comp,
                               const int32_t *tree_sec, const int32_t *sec_first, const int32_t *sec_n,
                               const int32_t *sec_parent, const double *area, const double *rl1, const double *rl2,
                               const double *Cm, double Ri, const double *chan, double v_spike, hs_hines **out)
{
    int r = hs_rt_need_init();
    if (r != HS_OK) return r;
    if (!out || n_trees <= 0 || (precision != HS_F64 && precision != HS_F32) || !(dt > 0) || !tree_comp ||
        !tree_sec || !sec_first || !sec_n || !sec_parent || !area || !rl1 || !rl2 || !Cm || !chan)
        return hs_rt_fail(HS_ERR_ARG, "hs_hines_create: bad arguments");
    const int NC = tree_comp[n_trees], NS = tree_sec[n_trees];
    /* validate the morphology and derive the tables of the reference's prep */
    std::vector<double> ab0(NC, 0.0), ab1(NC, 0.0), ab2(NC, 0.0), cdt(NC), invr(NC, 0.0);
    std::vector<double> invr0(NS), invrn(NS), bplus(NS), bminus(NS);
    std::vector<int32_t> first(NS), endc(NS), mparent(NS), csec(NC, 0);
    std::vector<int32_t> nchild(NS + n_trees, 0), child((size_t)(NS + n_trees) * HS_HINES_RL, 0);
    std::vector<int32_t> order(NS + n_trees), lvl(NS + 2 * n_trees, 0), nlev(n_trees);
    int cmax = 1, smax = 1;
    for (int t = 0; t < n_trees; ++t) {
        const int c0 = tree_comp[t], N = tree_comp[t + 1] - c0;
        const int s0 = tree_sec[t], ns = tree_sec[t + 1] - s0;
        const int n0 = s0 + t;
        if (N <= 0 || ns <= 0 || N > HS_HINES_MAXC || ns > HS_HINES_MAXS)
            return hs_rt_fail(HS_ERR_UNSUPPORTED, "tree %d: %d compartments / %d sections (limits %d / %d)", t, N,
                              ns, HS_HINES_MAXC, HS_HINES_MAXS);
        cmax = std::max(cmax, N);
        smax = std::max(smax, ns);
        int expect = 0;
        std::vector<int> depth(ns + 1, 0);
        for (int s = 0; s < ns; ++s) {
            const int p = sec_parent[s0 + s];
            if (sec_first[s0 + s] != expect || sec_n[s0 + s] <= 0 || (s == 0) != (p < 0) || p >= s)
                return hs_rt_fail(HS_ERR_ARG, "tree %d section %d: sections must be depth-first, contiguous", t, s);
            expect += sec_n[s0 + s];
            first[s0 + s] = sec_first[s0 + s];
            endc[s0 + s] = sec_first[s0 + s] + sec_n[s0 + s];
            const int mp = p < 0 ? 0 : p + 1;
            mparent[s0 + s] = mp;
            if (nchild[n0 + mp] >= HS_HINES_RL)
                return hs_rt_fail(HS_ERR_UNSUPPORTED, "tree %d: more than %d children at a branch point", t,
                                  HS_HINES_RL);
            child[(size_t)(n0 + mp) * HS_HINES_RL + nchild[n0 + mp]++] = s + 1;
            depth[s + 1] = depth[mp] + 1;
            for (int j = first[s0 + s]; j < endc[s0 + s]; ++j) csec[c0 + j] = s;
        }
        if (expect != N) return hs_rt_fail(HS_ERR_ARG, "tree %d: sections cover %d of %d compartments", t, expect, N);
        /* nodes by depth */
        int maxd = 0;
        for (int k = 0; k <= ns; ++k) maxd = std::max(maxd, depth[k]);
        int q = 0;
        for (int L = 0; L <= maxd; ++L) {
            lvl[n0 + t + L] = q;
            for (int k = 0; k <= ns; ++k)
                if (depth[k] == L) order[n0 + q++] = k;
        }
        lvl[n0 + t + maxd + 1] = q;
        nlev[t] = maxd + 1;
        /* spatialstateupdate.cu:39-83, tree-local indices, same operation order */
        const double *ar = area + c0, *r1 = rl1 + c0, *r2 = rl2 + c0, *cm = Cm + c0;
        double *iv = invr.data() + c0, *a0 = ab0.data() + c0, *a1 = ab1.data() + c0, *a2 = ab2.data() + c0;
        for (int i = 1; i < N; i++) iv[i] = 1.0 / (Ri * (1 / r2[i - 1] + 1 / r1[i]));
        for (int s = 0; s < ns; s++) iv[first[s0 + s]] = 0;
        for (int i = 0; i < N; i++) a1[i] = (-(cm[i] / dt) - iv[i] / ar[i]);
        for (int i = 1; i < N; i++) {
            a0[i] = iv[i] / ar[i - 1];
            a2[i - 1] = iv[i] / ar[i];
            a1[i - 1] -= iv[i] / ar[i - 1];
        }
        for (int s = 0; s < ns; s++) {
            const int f = first[s0 + s], l = endc[s0 + s] - 1;
            const double i0 = r1[f] / Ri, in = r2[l] / Ri;
            invr0[s0 + s] = i0;
            invrn[s0 + s] = in;
            a1[f] -= (i0 / ar[f]);
            a1[l] -= (in / ar[l]);
            bplus[s0 + s] = -(in / ar[l]);
            bminus[s0 + s] = -(i0 / ar[f]);
        }
        for (int i = 0; i < N; i++) cdt[c0 + i] = cm[i] / dt;
    }
    hs_hines *H = new hs_hines();
    H->prec = precision;
    H->dt = dt;
    H->ntrees = n_trees;
    H->ncomp = NC;
    H->nsec = NS;
    H->nnode = NS + n_trees;
    H->cmax = cmax;
    H->smax = smax;
    HinesDesc<double> d;
    memset(&d, 0, sizeof(d));
    int rr = HS_OK;
#define UP(vec, field)                                    \
    if (rr == HS_OK) rr = hs_h_upload(H, vec, &d.field);
    std::vector<int32_t> tc(tree_comp, tree_comp + n_trees + 1), ts(tree_sec, tree_sec + n_trees + 1);
    UP(tc, tree_comp) UP(ts, tree_sec) UP(ab0, ab0) UP(ab1, ab1) UP(ab2, ab2) UP(cdt, cdt) UP(csec, csec)
    UP(first, first) UP(endc, end) UP(mparent, mparent) UP(invr0, invr0) UP(invrn, invrn) UP(bplus, bplus)
    UP(bminus, bminus) UP(nchild, nchild) UP(child, child) UP(order, order) UP(lvl, lvl_start) UP(nlev, nlevels)
#undef UP
    std::vector<double> zeros(NC, 0.0);
    if (rr == HS_OK) rr = hs_h_upload(H, zeros, &d.iext);
    void *pv = nullptr, *pm = nullptr, *pn = nullptr, *ph = nullptr, *ps = nullptr;
    const size_t rs = precision == HS_F64 ? sizeof(double) : sizeof(float);
    if (rr == HS_OK) rr = hs_h_alloc(H, sizeof(double) * NC, &pv);
    if (rr == HS_OK) rr = hs_h_alloc(H, rs * NC, &pm);
    if (rr == HS_OK) rr = hs_h_alloc(H, rs * NC, &pn);
    if (rr == HS_OK) rr = hs_h_alloc(H, rs * NC, &ph);
    if (rr == HS_OK) rr = hs_h_alloc(H, sizeof(int32_t) * n_trees, &ps);
    if (rr != HS_OK) {
        hs_hines_destroy(H);
        return rr;
    }
    hipStream_t st = hs_rt_stream();
    if (hipMemsetAsync(pv, 0, sizeof(double) * NC, st) != hipSuccess ||
        hipMemsetAsync(pm, 0, rs * NC, st) != hipSuccess || hipMemsetAsync(pn, 0, rs * NC, st) != hipSuccess ||
        hipMemsetAsync(ph, 0, rs * NC, st) != hipSuccess ||
        hipMemsetAsync(ps, 0, sizeof(int32_t) * n_trees, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        hs_hines_destroy(H);
        return hs_rt_fail(HS_ERR_HIP, "hs_hines_create: initialising device arrays failed");
    }
    d.v = (double *)pv;
    d.spikes = (int32_t *)ps;
    d.vth = v_spike;
    d.cmax = cmax;
    d.smax = smax;
    H->d64 = d;
    hs_h_copy_common(H->d32, d);
    H->area.assign(area, area + NC);
    H->I.assign(NC, 0.0);
    H->d64.m = (double *)pm;
    H->d64.n = (double *)pn;
    H->d64.h = (double *)ph;
    H->d32.m = (float *)pm;
    H->d32.n = (float *)pn;
    H->d32.h = (float *)ph;
    for (int q = 0; q < 6; ++q) {
        H->d64.chan[q] = chan[q];
        H->d32.chan[q] = (float)chan[q];
    }
    H->d64.chan[6] = dt;
    H->d32.chan[6] = (float)dt;
    if (hipEventCreate(&H->e0) != hipSuccess || hipEventCreate(&H->e1) != hipSuccess) {
        hs_hines_destroy(H);
        return hs_rt_fail(HS_ERR_HIP, "hipEventCreate failed");
    }
    *out = H;
    return HS_OK;
}

static int hs_h_var(hs_hines *H, int var, void **p, bool *gate)
{
    *gate = var >= 1 && var <= 3;
    switch (var) {
    case 0: *p = H->d64.v; return HS_OK;
    case 1: *p = H->d64.m; return HS_OK;
    case 2: *p = H->d64.n; return HS_OK;
    case 3: *p = H->d64.h; return HS_OK;
    case 4: *p = (void *)H->d64.iext; return HS_OK;
    default: return hs_rt_fail(HS_ERR_ARG, "bad variable %d", var);
    }
}

extern "C" int hs_hines_set_state(hs_hines *H, int var, const double *values)
{
    if (!H || !values) return hs_rt_fail(HS_ERR_ARG, "null argument");
    void *p;
    bool gate;
    int r = hs_h_var(H, var, &p, &gate);
    if (r != HS_OK) return r;
    hipStream_t st = hs_rt_stream();
    if (var == 4) {
        /* a point current enters the membrane equation as the density I / area */
        std::vector<double> ia(H->ncomp);
        for (int i = 0; i < H->ncomp; ++i) {
            H->I[i] = values[i];
            ia[i] = values[i] / H->area[i];
        }
        HS_HIP(hipMemcpyAsync(p, ia.data(), sizeof(double) * H->ncomp, hipMemcpyHostToDevice, st));
        HS_HIP(hipStreamSynchronize(st));
        return HS_OK;
    }
    if (gate && H->prec == HS_F32) {
        std::vector<float> f(H->ncomp);
        for (int i = 0; i < H->ncomp; ++i) f[i] = (float)values[i];
        HS_HIP(hipMemcpyAsync(p, f.data(), sizeof(float) * H->ncomp, hipMemcpyHostToDevice, st));
        HS_HIP(hipStreamSynchronize(st));
    } else {
        HS_HIP(hipMemcpyAsync(p, values, sizeof(double) * H->ncomp, hipMemcpyHostToDevice, st));
        HS_HIP(hipStreamSynchronize(st));
    }
    return HS_OK;
}

extern "C" int hs_hines_get_state(hs_hines *H, int var, double *values)
{
    if (!H || !values) return hs_rt_fail(HS_ERR_ARG, "null argument");
    void *p;
    bool gate;
    int r = hs_h_var(H, var, &p, &gate);
    if (r != HS_OK) return r;
    hipStream_t st = hs_rt_stream();
    if (var == 4) {
        memcpy(values, H->I.data(), sizeof(double) * H->ncomp);
        return HS_OK;
    }
    if (gate && H->prec == HS_F32) {
        std::vector<float> f(H->ncomp);
        HS_HIP(hipMemcpyAsync(f.data(), p, sizeof(float) * H->ncomp, hipMemcpyDeviceToHost, st));
        HS_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < H->ncomp; ++i) values[i] = f[i];
    } else {
        HS_HIP(hipMemcpyAsync(values, p, sizeof(double) * H->ncomp, hipMemcpyDeviceToHost, st));
        HS_HIP(hipStreamSynchronize(st));
    }
    return HS_OK;
}

extern "C" int hs_hines_set_option(hs_hines *H, int key, long long value)
{
    if (!H || key != 0 || value < 1 || value > (1 << 24)) return hs_rt_fail(HS_ERR_ARG, "bad option");
    H->steps_per_launch = value;
    return HS_OK;
}

template <typename R>
static hipError_t hs_h_launch(const HinesDesc<R> &d, int ntrees, int cmax, int k, size_t lds, hipStream_t s)
{
    const int cpl = (cmax + HS_WAVE - 1) / HS_WAVE;
    if (cpl <= 2) hipLaunchKernelGGL((hs_k_hines<R, 2>), dim3(ntrees), dim3(HS_WAVE), lds, s, d, k);
    else if (cpl <= 4) hipLaunchKernelGGL((hs_k_hines<R, 4>), dim3(ntrees), dim3(HS_WAVE), lds, s, d, k);
    else if (cpl <= 8) hipLaunchKernelGGL((hs_k_hines<R, 8>), dim3(ntrees), dim3(HS_WAVE), lds, s, d, k);
    else hipLaunchKernelGGL((hs_k_hines<R, 16>), dim3(ntrees), dim3(HS_WAVE), lds, s, d, k);
    return hipGetLastError();
}

extern "C" int hs_hines_run(hs_hines *H, int64_t nsteps)
{
    if (!H || nsteps < 0) return hs_rt_fail(HS_ERR_ARG, "bad arguments");
    const int CM = H->cmax, SM = H->smax;
    const size_t lds = sizeof(double) * (4 * (size_t)CM + (SM + 1) + SM + (SM + 1) * HS_HINES_RL + (SM + 1) + 4 * SM) +
                       sizeof(int) * (3 * (size_t)SM + (SM + 1) + (SM + 1) * HS_HINES_RL + (SM + 1) + (SM + 2));
    if (lds > 160 * 1024) return hs_rt_fail(HS_ERR_UNSUPPORTED, "tree too large for LDS (%zu bytes)", lds);
    hipStream_t st = hs_rt_stream();
    auto t0 = std::chrono::steady_clock::now();
    HS_HIP(hipEventRecord(H->e0, st));
    for (int64_t done = 0; done < nsteps;) {
        const int k = (int)std::min<int64_t>(H->steps_per_launch, nsteps - done);
        if (H->prec == HS_F64) HS_HIP(hs_h_launch<double>(H->d64, H->ntrees, CM, k, lds, st));
        else HS_HIP(hs_h_launch<float>(H->d32, H->ntrees, CM, k, lds, st));
        done += k;
    }
    HS_HIP(hipEventRecord(H->e1, st));
    HS_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    float ms = 0;
    HS_HIP(hipEventElapsedTime(&ms, H->e0, H->e1));
    H->last_gpu_ms = ms;
    H->last_wall = std::chrono::duration<double>(t1 - t0).count();
    H->steps += nsteps;
    return HS_OK;
}

extern "C" int hs_hines_spike_counts(hs_hines *H, int32_t *counts)
{
    if (!H || !counts) return hs_rt_fail(HS_ERR_ARG, "null argument");
    hipStream_t st = hs_rt_stream();
    HS_HIP(hipMemcpyAsync(counts, H->d64.spikes, sizeof(int32_t) * H->ntrees, hipMemcpyDeviceToHost, st));
    HS_HIP(hipStreamSynchronize(st));
    return HS_OK;
}

extern "C" int hs_hines_stats(hs_hines *H, double *wall_s, double *gpu_ms, int64_t *steps)
{
    if (!H) return hs_rt_fail(HS_ERR_ARG, "null argument");
    if (wall_s) *wall_s = H->last_wall;
    if (gpu_ms) *gpu_ms = H->last_gpu_ms;
    if (steps) *steps = H->steps;
    return HS_OK;
}
