/*
 * hs_syncode.h -- the delivery kernel of pathways whose on_pre is code
 * (hs_net_pathway_code), compiled only into the kernels built at run time
 * (hs_codegen.hip), whose source defines HS_SCG_PATHS (the count) and, after
 * this header, one HsSynCg<R, c> per code pathway and precision:
 *   NS, NPOST, NPRE     synapse variables, target variables added to, source
 *                       variables read
 *   on_pre(sv, pre, eff, dt, t, i, j, lastupdate)  the statements for one
 *                       synapse: sv its variables (in and out), pre the source
 *                       neuron's variables, eff the additions to the target's
 *                       variables (zero on entry), lastupdate the time of its
 *                       last on_pre or on_post (seconds)
 *   on_post(sv, pre, post, dt, t, i, j, lastupdate)  for each synapse of a
 *                       spiking target neuron (HAS_POST), post its variables
 *                       (read)
 *
 * Replaces the effect kernel of such a pathway (synapses.cu:12-164 with the
 * translated on_pre code, cuda_generator.py:463-527 using atomics for the
 * additive target writes).  Step T: for each delay slot d (block-strided), the
 * spikes of step T - dmin - d, their bundles, one thread per synapse; a synapse
 * sits in one bundle and a neuron spikes at most once a step, so no two threads
 * touch the same synapse variables; the target additions are hardware FP
 * atomics (several synapses share a target), into the state (plain form) or
 * the accumulator of step T (fused form, one step per launch).  on_post runs
 * after on_pre (Brian2's pre pathway, order -1, before the post pathway, +1),
 * per spike of the target group through the target-indexed synapse table, as
 * the post pathway's kernel does (synapses.cu:201-322).
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
                    HsSynCg<R, C>::on_pre(sv, pre, eff, (R)p.dt, t, i, j, (double)hs_ld(p.lastup + q) * p.dt, p.seed,
                                          p.pid, T);
#pragma unroll
                    for (int v = 0; v < NS; ++v) hs_st(p.sv[v] + q, sv[v]);
                    hs_st(p.lastup + q, (int32_t)T);
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

/* on_post of pathway p at step T: for each spike of the target group at T
   (block-strided), that neuron's synapses (qsyn[qstart[j] .. qstart[j+1]),
   thread-strided) */
template <typename R, int C>
__device__ __forceinline__ void hs_syn_post_path(const DPath<R> &p, const DGroup<R> *groups, long long T)
{
    if constexpr (HsSynCg<R, C>::HAS_POST) {
        constexpr int NS = HsSynCg<R, C>::NS, NPRE = HsSynCg<R, C>::NPRE, NPR = HsSynCg<R, C>::NPR;
        const DGroup<R> &gs = groups[p.src];
        const DGroup<R> &gt = groups[p.tgt];
        const double t = (double)T * p.dt;
        const long long slot = hs_mod(T, p.tQ);
        for (int sl = 0; sl < p.tnsub; ++sl) {
            const int cnt = hs_ld(p.thist_cnt + (slot * p.tnsub + sl) * HS_CNT_STRIDE);
            const int32_t *lst = p.thist + slot * p.tnsub * p.tsubcap + (long long)sl * p.tsubcap;
            for (int k = blockIdx.x; k < cnt; k += gridDim.x) {
                const int j = hs_ld(lst + k);
                R post[NPR > 0 ? NPR : 1];
#pragma unroll
                for (int v = 0; v < NPR; ++v) post[v] = hs_ld(gt.var[p.post_read_var[v]] + j);
                const int q1 = hs_ld(p.qstart + j + 1);
                for (int e = hs_ld(p.qstart + j) + (int)threadIdx.x; e < q1; e += blockDim.x) {
                    const int q = hs_ld(p.qsyn + e);
                    const int i = hs_ld(p.synpre + q);
                    R sv[NS > 0 ? NS : 1];
                    R pre[NPRE > 0 ? NPRE : 1];
#pragma unroll
                    for (int v = 0; v < NS; ++v) sv[v] = hs_ld(p.sv[v] + q);
#pragma unroll
                    for (int v = 0; v < NPRE; ++v) pre[v] = hs_ld(gs.var[p.pre_var[v]] + i);
                    HsSynCg<R, C>::on_post(sv, pre, post, (R)p.dt, t, i, j, (double)hs_ld(p.lastup + q) * p.dt,
                                           p.seed, p.pid, T);
#pragma unroll
                    for (int v = 0; v < NS; ++v) hs_st(p.sv[v] + q, sv[v]);
                    hs_st(p.lastup + q, (int32_t)T);
                }
            }
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256) void hs_k_syn_post(const DPath<R> *__restrict__ paths, int np,
                                                     const DGroup<R> *__restrict__ groups, HsCtl *ctl, int koff)
{
    const long long T = ctl->base + koff;
    for (int pi = 0; pi < np; ++pi) {
        const DPath<R> &p = paths[pi];
        if (p.dmode != HS_DELIVER_CODE || !p.qstart) continue;
        if (p.cg == 0) hs_syn_post_path<R, 0>(p, groups, T);
#if HS_SCG_PATHS > 1
        else if (p.cg == 1) hs_syn_post_path<R, 1>(p, groups, T);
#endif
#if HS_SCG_PATHS > 2
        else if (p.cg == 2) hs_syn_post_path<R, 2>(p, groups, T);
#endif
#if HS_SCG_PATHS > 3
        else if (p.cg == 3) hs_syn_post_path<R, 3>(p, groups, T);
#endif
    }
}

#endif /* HS_SCG_PATHS */
#endif /* HS_SYNCODE_H */
