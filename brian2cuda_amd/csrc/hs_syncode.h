/*
 * hs_syncode.h -- the delivery kernel of pathways whose on_pre is code
 * (hs_net_pathway_code), compiled only into the kernels built at run time
 * (hs_codegen.hip), whose source defines HS_SCG_PATHS (the count) and, after
 * this header, one HsSynCg<R, c> per code pathway and precision:
 *   NS, NPOST, NPRE     synapse variables, target variables added to, source
 *                       variables read
 *   on_pre(sv, pre, eff, dt, t, i, j)  the statements for one synapse: sv its
 *                       variables (in and out), pre the source neuron's
 *                       variables, eff the additions to the target's
 *                       variables (zero on entry)
 *
 * Replaces the effect kernel of such a pathway (synapses.cu:12-164 with the
 * translated on_pre code, cuda_generator.py:463-527 using atomics for the
 * additive target writes).  Step T: for each delay slot d (block-strided), the
 * spikes of step T - dmin - d, their bundles, one thread per synapse; a synapse
 * sits in one bundle and a neuron spikes at most once a step, so no two threads
 * touch the same synapse variables; the target additions are hardware FP
 * atomics (several synapses share a target), into the state (plain form) or
 * the accumulator of step T (fused form, one step per launch).
 */
#ifndef HS_SYNCODE_H
#define HS_SYNCODE_H
#ifdef HS_SCG_PATHS

template <typename R, int C> struct HsSynCg;

template <typename R, int C>
__device__ __forceinline__ void hs_syn_code_path(const DPath<R> &p, const DGroup<R> *groups, long long T, int fused,
                                                 unsigned long long *ev)
{
    constexpr int NS = HsSynCg<R, C>::NS, NPOST = HsSynCg<R, C>::NPOST, NPRE = HsSynCg<R, C>::NPRE;
    const DGroup<R> &gs = groups[p.src];
    const DGroup<R> &gt = groups[p.tgt];
    R *tgt[NPOST > 0 ? NPOST : 1];
#pragma unroll
    for (int k = 0; k < NPOST; ++k)
        tgt[k] = fused ? gt.acc[p.post_var[k]] + hs_mod(T, gt.nacc) * (long long)gt.n : gt.var[p.post_var[k]];
    const double t = (double)T * p.dt;
    for (int d = blockIdx.x; d < p.D; d += gridDim.x) {
        const long long ts = T - p.dmin - d; /* emission step of the bundles due now */
        if (ts < 0) continue;
        const long long slot = hs_mod(ts, p.Q);
        for (int sl = 0; sl < p.nsub; ++sl) {
            const int cnt = hs_ld(p.hist_cnt + (slot * p.nsub + sl) * p.cstride);
            const int32_t *lst = p.hist + slot * p.nsub * p.subcap + (long long)sl * p.subcap;
            for (int k = 0; k < cnt; ++k) {
                const int i = hs_ld(lst + k);
                const long long row = (long long)i * p.D + d;
                const int b1 = hs_ld(p.bstart + row + 1);
                R pre[NPRE > 0 ? NPRE : 1];
#pragma unroll
                for (int v = 0; v < NPRE; ++v) pre[v] = hs_ld(gs.var[p.pre_var[v]] + i);
                for (int q = hs_ld(p.bstart + row) + (int)threadIdx.x; q < b1; q += blockDim.x) {
                    const int j = hs_ld(p.post + q);
                    R sv[NS > 0 ? NS : 1];
                    R eff[NPOST > 0 ? NPOST : 1];
#pragma unroll
                    for (int v = 0; v < NS; ++v) sv[v] = hs_ld(p.sv[v] + q);
#pragma unroll
                    for (int v = 0; v < NPOST; ++v) eff[v] = (R)0;
                    HsSynCg<R, C>::on_pre(sv, pre, eff, (R)p.dt, t, i, j);
#pragma unroll
                    for (int v = 0; v < NS; ++v) hs_st(p.sv[v] + q, sv[v]);
#pragma unroll
                    for (int v = 0; v < NPOST; ++v) hs_atomic_add(tgt[v] + j, eff[v]);
                    ++*ev;
                }
            }
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256) void hs_k_syn_code(const DPath<R> *__restrict__ paths, int np,
                                                     const DGroup<R> *__restrict__ groups, HsCtl *ctl, int koff,
                                                     int fused)
{
    const long long T = ctl->base + koff;
    unsigned long long ev = 0;
    for (int pi = 0; pi < np; ++pi) {
        const DPath<R> &p = paths[pi];
        if (p.dmode != HS_DELIVER_CODE) continue;
        if (p.cg == 0) hs_syn_code_path<R, 0>(p, groups, T, fused, &ev);
#if HS_SCG_PATHS > 1
        else if (p.cg == 1) hs_syn_code_path<R, 1>(p, groups, T, fused, &ev);
#endif
#if HS_SCG_PATHS > 2
        else if (p.cg == 2) hs_syn_code_path<R, 2>(p, groups, T, fused, &ev);
#endif
#if HS_SCG_PATHS > 3
        else if (p.cg == 3) hs_syn_code_path<R, 3>(p, groups, T, fused, &ev);
#endif
    }
    ev = hs_wave_sum(ev);
    hs_flush_events(ev, ctl, (long long)blockIdx.x * (blockDim.x / HS_WAVE) + threadIdx.x / HS_WAVE);
}

#endif /* HS_SCG_PATHS */
#endif /* HS_SYNCODE_H */
