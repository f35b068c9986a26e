/*
 * hs_ops.hip -- the individually testable kernels of the C ABI: threshold +
 * spikespace compaction, Philox draws, one state-update step, one pathway's
 * delivery, and device evaluation of the deterministic math.
 */
#include "hs_device.h"

/* ------------------------------------------------------------------------ */
/* standalone kernels                                                        */
/* ------------------------------------------------------------------------ */
template <typename R>
__global__ __launch_bounds__(1024) void hs_k_threshold(const R *v, int n, R vth, int32_t *last, long long step,
                                                       int ref_steps, int32_t *spikes, int32_t *count)
{
    __shared__ HsCompactLds L;
    for (int tile = blockIdx.x; tile * (int)blockDim.x < n; tile += gridDim.x) {
        const int i = tile * blockDim.x + threadIdx.x;
        bool spk = false;
        if (i < n) {
            const int nr = (step - (long long)last[i]) >= (long long)ref_steps;
            spk = (v[i] > vth) && nr;
            if (spk) last[i] = (int32_t)step;
        }
        hs_block_compact(spk, i, spikes, count, &L, nullptr, nullptr, nullptr, 0, step, nullptr);
    }
}

/* sorted form: pass 1 counts per tile, pass 2 scans the counts, pass 3 writes */
template <typename R>
__global__ __launch_bounds__(256) void hs_k_threshold_count(const R *v, int n, R vth, const int32_t *last,
                                                            long long step, int ref_steps, int32_t *tile_cnt)
{
    __shared__ int wc[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    bool spk = false;
    if (i < n) spk = (v[i] > vth) && ((step - (long long)last[i]) >= (long long)ref_steps);
    unsigned long long m = __ballot(spk);
    if (hs_lane() == 0) wc[threadIdx.x / HS_WAVE] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ __launch_bounds__(64) void hs_k_scan_tiles(int32_t *tile_cnt, int ntiles, int32_t *count)
{
    int carry = 0;
    for (int c = 0; c < ntiles; c += HS_WAVE) {
        int idx = c + hs_lane();
        int x = idx < ntiles ? tile_cnt[idx] : 0;
        int tot;
        int ex = hs_wave_excl_scan(x, &tot);
        if (idx < ntiles) tile_cnt[idx] = carry + ex;
        carry += tot;
    }
    if (hs_lane() == 0) *count = carry;
}

template <typename R>
__global__ __launch_bounds__(256) void hs_k_threshold_write(const R *v, int n, R vth, int32_t *last, long long step,
                                                            int ref_steps, const int32_t *tile_off, int32_t *spikes)
{
    __shared__ int wc[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    bool spk = false;
    if (i < n) spk = (v[i] > vth) && ((step - (long long)last[i]) >= (long long)ref_steps);
    unsigned long long m = __ballot(spk);
    const int wid = threadIdx.x / HS_WAVE;
    if (hs_lane() == 0) wc[wid] = __popcll(m);
    __syncthreads();
    int off = tile_off[blockIdx.x];
    for (int w = 0; w < wid; ++w) off += wc[w];
    if (spk) {
        spikes[off + __popcll(m & hs_lanemask_lt())] = i;
        last[i] = (int32_t)step;
    }
}

hipError_t hs_launch_threshold(int prec, const void *v, int n, double vth, int32_t *last, long long step,
                               int ref_steps, int sorted, int32_t *spikes, int32_t *count, int32_t *scratch,
                               hipStream_t s)
{
    hipError_t e;
    if (!sorted) {
        e = hipMemsetAsync(count, 0, sizeof(int32_t), s);
        if (e != hipSuccess) return e;
        int grid = (n + 255) / 256;
        if (grid < 1) grid = 1;
        if (prec == HS_F64)
            hipLaunchKernelGGL((hs_k_threshold<double>), dim3(grid), dim3(256), 0, s, (const double *)v, n, vth,
                               last, step, ref_steps, spikes, count);
        else
            hipLaunchKernelGGL((hs_k_threshold<float>), dim3(grid), dim3(256), 0, s, (const float *)v, n,
                               (float)vth, last, step, ref_steps, spikes, count);
        return hipGetLastError();
    }
    int ntiles = (n + 255) / 256;
    if (ntiles < 1) ntiles = 1;
    if (prec == HS_F64)
        hipLaunchKernelGGL((hs_k_threshold_count<double>), dim3(ntiles), dim3(256), 0, s, (const double *)v, n, vth,
                           last, step, ref_steps, scratch);
    else
        hipLaunchKernelGGL((hs_k_threshold_count<float>), dim3(ntiles), dim3(256), 0, s, (const float *)v, n,
                           (float)vth, last, step, ref_steps, scratch);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(hs_k_scan_tiles, dim3(1), dim3(64), 0, s, scratch, ntiles, count);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (prec == HS_F64)
        hipLaunchKernelGGL((hs_k_threshold_write<double>), dim3(ntiles), dim3(256), 0, s, (const double *)v, n, vth,
                           last, step, ref_steps, scratch, spikes);
    else
        hipLaunchKernelGGL((hs_k_threshold_write<float>), dim3(ntiles), dim3(256), 0, s, (const float *)v, n,
                           (float)vth, last, step, ref_steps, scratch, spikes);
    return hipGetLastError();
}

__global__ void hs_k_randn(unsigned long long seed, unsigned int stream, long long step, int n, int prec, void *out,
                           int uniform)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (uniform) ((double *)out)[i] = hs_rand(seed, stream, (uint32_t)i, step);
        else if (prec == HS_F64) ((double *)out)[i] = hs_randn(seed, stream, (uint32_t)i, step);
        else ((float *)out)[i] = hs_randn_f32(seed, stream, (uint32_t)i, step);
    }
}

hipError_t hs_launch_randn(unsigned long long seed, unsigned int stream, long long step, int n, int prec, void *out,
                           int uniform, hipStream_t s)
{
    int grid = (n + 255) / 256;
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(hs_k_randn, dim3(grid), dim3(256), 0, s, seed, stream, step, n, prec, out, uniform);
    return hipGetLastError();
}

template <typename R, int MODEL>
__global__ void hs_k_stateupdate(int n, R *v0, R *v1, R *v2, R *v3, R *v4, R *v5, const int32_t *last, const R *kd,
                                 unsigned long long seed, unsigned int stream, int ref_steps, long long step)
{
    constexpr int NV = HsNV<MODEL>::value;
    R *vars[6] = {v0, v1, v2, v3, v4, v5};
    R k[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) k[j] = kd[j];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        R s[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) s[v] = vars[v][i];
        const int nr = (step - (long long)last[i]) >= (long long)ref_steps;
        (void)hs_neuron_step<R, MODEL>(s, k, nr, (R)0, hs_noise<R, MODEL>(seed, stream, i, step));
#pragma unroll
        for (int v = 0; v < NV; ++v) vars[v][i] = s[v];
    }
}

template <typename R>
static hipError_t hs_launch_su(int model, int n, void *const *vars, const int32_t *last, const R *kd,
                               unsigned long long seed, unsigned int stream, int ref_steps, long long step,
                               hipStream_t s)
{
    R *v[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int nv = model == HS_MODEL_LIF_CUBA ? 3 : (model == HS_MODEL_LIF_NOISE ? 1 : 6);
    for (int i = 0; i < nv; ++i) v[i] = (R *)vars[i];
    int grid = (n + 255) / 256;
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    if (model == HS_MODEL_LIF_CUBA)
        hipLaunchKernelGGL((hs_k_stateupdate<R, HS_MODEL_LIF_CUBA>), dim3(grid), dim3(256), 0, s, n, v[0], v[1], v[2],
                           v[3], v[4], v[5], last, kd, seed, stream, ref_steps, step);
    else if (model == HS_MODEL_LIF_NOISE)
        hipLaunchKernelGGL((hs_k_stateupdate<R, HS_MODEL_LIF_NOISE>), dim3(grid), dim3(256), 0, s, n, v[0], v[1], v[2],
                           v[3], v[4], v[5], last, kd, seed, stream, ref_steps, step);
    else
        hipLaunchKernelGGL((hs_k_stateupdate<R, HS_MODEL_HH_COBA>), dim3(grid), dim3(256), 0, s, n, v[0], v[1], v[2],
                           v[3], v[4], v[5], last, kd, seed, stream, ref_steps, step);
    return hipGetLastError();
}

/* kd: device array of 16 constants already converted to the precision */
hipError_t hs_launch_stateupdate(int model, int prec, int n, void *const *vars, const int32_t *last, const double *kd,
                                 int nk, unsigned long long seed, unsigned int stream, int ref_steps, long long step,
                                 hipStream_t s)
{
    (void)nk;
    if (prec == HS_F64) return hs_launch_su<double>(model, n, vars, last, kd, seed, stream, ref_steps, step, s);
    return hs_launch_su<float>(model, n, vars, last, (const float *)(const void *)kd, seed, stream, ref_steps, step, s);
}

__global__ void hs_k_math(int fn, int n, const double *x, double *y)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a = x[i], r;
        if (fn == 0) r = hs_exp(a);
        else if (fn == 1) r = hs_log(a);
        else if (fn == 2) r = hs_sqrt(a);
        else if (fn == 3) r = hs_cos2pi(a);
        else if (fn == 4) r = (double)hs_expf((float)a);
        else if (fn == 5) r = (double)hs_logf((float)a);
        else if (fn == 6) r = (double)hs_sqrtf((float)a);
        else r = (double)hs_cos2pif((float)a);
        y[i] = r;
    }
}

hipError_t hs_launch_math(int fn, int n, const double *x, double *y, hipStream_t s)
{
    int grid = (n + 255) / 256;
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(hs_k_math, dim3(grid), dim3(256), 0, s, fn, n, x, y);
    return hipGetLastError();
}

template <typename R>
__global__ __launch_bounds__(256) void hs_k_deliver_one(DPath<R> p, HsCtl *ctl, unsigned long long *events)
{
    __shared__ HsDeliverLds L;
    const int nwb = blockDim.x / HS_WAVE;
    const long long rank = (long long)blockIdx.x * nwb + threadIdx.x / HS_WAVE;
    unsigned long long ev = 0;
    hs_deliver_path<R, false>(p, ctl->base, 1, &L, rank, (long long)gridDim.x * nwb, &ev);
    hs_flush_events(ev, events);
}

template <typename R>
hipError_t hs_launch_deliver_one_ctl(const DPath<R> &p, HsCtl *ctl, unsigned long long *events, int blocks,
                                     hipStream_t s)
{
    hipLaunchKernelGGL((hs_k_deliver_one<R>), dim3(blocks), dim3(256), 0, s, p, ctl, events);
    return hipGetLastError();
}

template hipError_t hs_launch_deliver_one_ctl<double>(const DPath<double> &, HsCtl *, unsigned long long *, int,
                                                      hipStream_t);
template hipError_t hs_launch_deliver_one_ctl<float>(const DPath<float> &, HsCtl *, unsigned long long *, int,
                                                     hipStream_t);

/* effect application of a flat event list: target[idx[k]] += w for k < n with
   hardware atomics (no spike list, no bundle lookup: the throughput of random
   atomics alone).  An index outside [0, n_target) is skipped and counted in bad. */
template <typename R>
__global__ __launch_bounds__(256) void hs_k_scatter_add(const int32_t *__restrict__ idx, long long n, int n_target,
                                                         R w, R *__restrict__ target, unsigned long long *bad)
{
    unsigned long long nb = 0;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const int t = idx[k];
        if ((unsigned int)t < (unsigned int)n_target) unsafeAtomicAdd(&target[t], w);
        else ++nb;
    }
    nb = hs_wave_sum(nb);
    if (hs_lane() == 0 && nb) atomicAdd(bad, nb);
}

hipError_t hs_launch_scatter_add(int prec, const int32_t *idx, long long n, int n_target, double w, void *target,
                                 unsigned long long *bad, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const long long g = (n + 255) / 256;
    const int grid = (int)(g > 65536 ? 65536 : g);
    if (prec == 64)
        hipLaunchKernelGGL(hs_k_scatter_add<double>, dim3(grid), dim3(256), 0, s, idx, n, n_target, w,
                           (double *)target, bad);
    else
        hipLaunchKernelGGL(hs_k_scatter_add<float>, dim3(grid), dim3(256), 0, s, idx, n, n_target, (float)w,
                           (float *)target, bad);
    return hipGetLastError();
}
