/*
 * hs_plastic.hip -- spike-timing-dependent plasticity on scatter pathways: the
 * on_pre and on_post code of an STDP Synapses object, run after the step's
 * update and delivery launches (plain form: effects on the state; fused form
 * with one-step windows: effects into the accumulator of the step).
 *
 * Replaces, for these synapses, the effect kernel of the pre pathway
 * (brian2cuda/templates/synapses.cu:12-164 with synaptic_effects "synapse"
 * plus atomics on the target) and the post pathway's kernel over the synapses of
 * each spiking postsynaptic neuron (same template, launched for the post
 * pathway; synapses.cu:201-322 chooses full parallelism because the post code
 * only writes synapse variables).
 *
 * on_pre (step T): for each delay slot d, the spikes of step T - dmin - d; block
 * b takes slots b, b + grid, ...; the threads of the block take the synapses of
 * each bundle.  A synapse sits in one bundle and a neuron spikes at most once a
 * step, so no two threads touch the same synapse; the target effect is a
 * hardware FP atomic (several synapses share a target).
 * on_post (step T): the target group's spikes of step T, block-strided; the
 * threads take that neuron's synapses (qsyn).  Launched after on_pre: Brian2
 * runs the pre pathway (order -1) before the post pathway (order +1) within the
 * synapses slot.
 * The trace and weight arithmetic is the shared expression set
 * (hs_model_body.inc: hs_stdp_decay, hs_stdp_on_pre, hs_stdp_on_post).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device.h"
#include "hs_internal.h"
#include "hs_model_exprs.h"

template <typename R> struct HsStdp;
template <> struct HsStdp<double> {
    static __device__ __forceinline__ double decay(long long n, double dt, double tau)
    {
        return hs_stdp_decay_f64(n, dt, tau);
    }
    static __device__ __forceinline__ double pre(double *w, double *a, double *b, double fa, double fb, const double *sk)
    {
        return hs_stdp_on_pre_f64(w, a, b, fa, fb, sk);
    }
    static __device__ __forceinline__ void post(double *w, double *a, double *b, double fa, double fb, const double *sk)
    {
        hs_stdp_on_post_f64(w, a, b, fa, fb, sk);
    }
};
template <> struct HsStdp<float> {
    static __device__ __forceinline__ float decay(long long n, double dt, double tau)
    {
        return hs_stdp_decay_f32(n, dt, tau);
    }
    static __device__ __forceinline__ float pre(float *w, float *a, float *b, float fa, float fb, const float *sk)
    {
        return hs_stdp_on_pre_f32(w, a, b, fa, fb, sk);
    }
    static __device__ __forceinline__ void post(float *w, float *a, float *b, float fa, float fb, const float *sk)
    {
        hs_stdp_on_post_f32(w, a, b, fa, fb, sk);
    }
};

template <typename R>
__global__ __launch_bounds__(256) void hs_k_stdp_pre(const DPath<R> *paths, int np, HsCtl *ctl, int koff)
{
    const long long T = ctl->base + koff;
    unsigned long long ev = 0;
    for (int pi = 0; pi < np; ++pi) {
        const DPath<R> &p = paths[pi];
        if (p.dmode != HS_DELIVER_PLASTIC) continue;
        R sk[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) sk[j] = p.sk[j];
        /* plain form: the target state; fused form: the accumulator of step T,
           added to the state at the start of step T + 1 */
        R *tgt = p.acc ? p.acc + hs_mod(T, p.nacc) * p.n_tgt : p.target;
        for (int d = blockIdx.x; d < p.D; d += gridDim.x) {
            const long long ts = T - p.dmin - d; /* emission step of the bundles due now */
            if (ts < 0) continue;
            const long long slot = hs_mod(ts, p.Q);
            for (int sl = 0; sl < p.nsub; ++sl) {
                const int cnt = p.hist_cnt[(slot * p.nsub + sl) * p.cstride];
                const int32_t *lst = p.hist + slot * p.nsub * p.subcap + (long long)sl * p.subcap;
                for (int k = 0; k < cnt; ++k) {
                    const long long row = (long long)lst[k] * p.D + d;
                    const int b1 = p.bstart[row + 1];
                    for (int q = p.bstart[row] + (int)threadIdx.x; q < b1; q += blockDim.x) {
                        const long long n = T - (long long)p.lastup[q];
                        const R fa = HsStdp<R>::decay(n, p.dt, p.taupre);
                        const R fb = HsStdp<R>::decay(n, p.dt, p.taupost);
                        R w = p.wm[q], a = p.apre[q], b = p.apost[q];
                        const R e = HsStdp<R>::pre(&w, &a, &b, fa, fb, sk);
                        hs_atomic_add(tgt + p.post[q], e);
                        p.wm[q] = w;
                        p.apre[q] = a;
                        p.apost[q] = b;
                        p.lastup[q] = (int32_t)T;
                        ev++;
                    }
                }
            }
        }
    }
    ev = hs_wave_sum(ev);
    hs_flush_events(ev, ctl, (long long)blockIdx.x * (blockDim.x / HS_WAVE) + threadIdx.x / HS_WAVE);
}

template <typename R>
__global__ __launch_bounds__(256) void hs_k_stdp_post(const DPath<R> *paths, int np, HsCtl *ctl, int koff)
{
    const long long T = ctl->base + koff;
    for (int pi = 0; pi < np; ++pi) {
        const DPath<R> &p = paths[pi];
        if (p.dmode != HS_DELIVER_PLASTIC) continue;
        R sk[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) sk[j] = p.sk[j];
        const long long slot = hs_mod(T, p.tQ);
        for (int sl = 0; sl < p.tnsub; ++sl) {
            const int cnt = p.thist_cnt[(slot * p.tnsub + sl) * HS_CNT_STRIDE];
            const int32_t *lst = p.thist + slot * p.tnsub * p.tsubcap + (long long)sl * p.tsubcap;
            for (int k = blockIdx.x; k < cnt; k += gridDim.x) {
                const int j = lst[k];
                const int q1 = p.qstart[j + 1];
                for (int e = p.qstart[j] + (int)threadIdx.x; e < q1; e += blockDim.x) {
                    const int q = p.qsyn[e];
                    const long long n = T - (long long)p.lastup[q];
                    const R fa = HsStdp<R>::decay(n, p.dt, p.taupre);
                    const R fb = HsStdp<R>::decay(n, p.dt, p.taupost);
                    R w = p.wm[q], a = p.apre[q], b = p.apost[q];
                    HsStdp<R>::post(&w, &a, &b, fa, fb, sk);
                    p.wm[q] = w;
                    p.apre[q] = a;
                    p.apost[q] = b;
                    p.lastup[q] = (int32_t)T;
                }
            }
        }
    }
}

template <typename R>
hipError_t hs_launch_stdp(const DPath<R> *paths, int np, HsCtl *ctl, int koff, int grid, hipStream_t s)
{
    hs_k_stdp_pre<R><<<grid, 256, 0, s>>>(paths, np, ctl, koff);
    hs_k_stdp_post<R><<<grid, 256, 0, s>>>(paths, np, ctl, koff);
    return hipGetLastError();
}

template hipError_t hs_launch_stdp<double>(const DPath<double> *, int, HsCtl *, int, int, hipStream_t);
template hipError_t hs_launch_stdp<float>(const DPath<float> *, int, HsCtl *, int, int, hipStream_t);

/* StateMonitor of a synapse variable (Brian2's default when='start'): at the
   start of step T, row T - t0 of buf takes src[idx[k]] for the n recorded
   synapses (positions in bundle order) */
template <typename R>
__global__ __launch_bounds__(256) void hs_k_synmon(const R *__restrict__ src, const int32_t *__restrict__ idx, int n,
                                                    R *__restrict__ buf, long long rows, long long t0, int every,
                                                    const HsCtl *ctl, int koff)
{
    const long long dt0 = ctl->base + koff - t0;
    if (dt0 < 0 || dt0 % every != 0) return;
    const long long row = dt0 / every;
    if (row >= rows) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        buf[row * n + k] = src[idx[k]];
}

hipError_t hs_launch_synmon(int prec, const void *src, const int32_t *idx, int n, void *buf, long long rows,
                            long long t0, int every, const HsCtl *ctl, int koff, hipStream_t s)
{
    if (n <= 0 || every < 1) return hipSuccess;
    const int grid = (n + 255) / 256 > 1024 ? 1024 : (n + 255) / 256;
    if (prec == 64)
        hipLaunchKernelGGL(hs_k_synmon<double>, dim3(grid), dim3(256), 0, s, (const double *)src, idx, n,
                           (double *)buf, rows, t0, every, ctl, koff);
    else
        hipLaunchKernelGGL(hs_k_synmon<float>, dim3(grid), dim3(256), 0, s, (const float *)src, idx, n,
                           (float *)buf, rows, t0, every, ctl, koff);
    return hipGetLastError();
}
