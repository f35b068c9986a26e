/*
 * hs_kernels.hip -- gfx950 kernels of the per-timestep hot path.
 *
 *   update  : deferred reset + StateMonitor sample + state update + threshold +
 *             wave64 ballot / prefix-count spike compaction into the spike ring
 *             (threshold.cu:19-45, reset.cu:3-27, common_group.cu:108-160,
 *             statemonitor.cu:74-100, spikemonitor.cu:165-196 in one pass)
 *   deliver : synaptic effects of every (spike, delay) bundle due this step,
 *             bundle descriptors flattened per wavefront so that lanes walk
 *             consecutive synapses (synapses.cu:12-164, spikequeue.h:432-516)
 *   fused   : update(t) and deliver(t) in one launch (effects go to the
 *             accumulator of step t, read by update(t+1)); valid when every
 *             delay is >= 1 step, so the two halves touch disjoint data
 *   finalize: applies what is pending at the end of a run
 * All kernels are wave64; block sizes are multiples of 64 up to 1024.
 */
#include "hs_internal.h"
#include "hs_model_exprs.h"
#include "hs_philox.h"

#define HS_WAVE 64

/* ------------------------------------------------------------------------ */
/* wave helpers                                                             */
/* ------------------------------------------------------------------------ */
__device__ __forceinline__ int hs_lane() { return threadIdx.x & (HS_WAVE - 1); }

__device__ __forceinline__ unsigned long long hs_lanemask_lt()
{
    return (1ull << hs_lane()) - 1ull;
}

/* exclusive prefix sum across the 64 lanes; *total gets the wave sum */
__device__ __forceinline__ int hs_wave_excl_scan(int x, int *total)
{
    int inc = x;
#pragma unroll
    for (int o = 1; o < HS_WAVE; o <<= 1) {
        int y = __shfl_up(inc, o, HS_WAVE);
        if (hs_lane() >= o) inc += y;
    }
    *total = __shfl(inc, HS_WAVE - 1, HS_WAVE);
    return inc - x;
}

template <typename R> struct HsOps;
template <> struct HsOps<double> {
    static __device__ __forceinline__ void cuba(double *v, double *ge, double *gi, int nr, const double *k)
    {
        hs_lif_cuba_update_f64(v, ge, gi, nr, k);
    }
    static __device__ __forceinline__ double noise(double V, double xi, const double *k)
    {
        return hs_lif_noise_update_f64(V, xi, k);
    }
    static __device__ __forceinline__ void hh(double *s, const double *k) { hs_hh_coba_update_f64(s, k); }
};
template <> struct HsOps<float> {
    static __device__ __forceinline__ void cuba(float *v, float *ge, float *gi, int nr, const float *k)
    {
        hs_lif_cuba_update_f32(v, ge, gi, nr, k);
    }
    static __device__ __forceinline__ float noise(float V, float xi, const float *k)
    {
        return hs_lif_noise_update_f32(V, xi, k);
    }
    static __device__ __forceinline__ void hh(float *s, const float *k) { hs_hh_coba_update_f32(s, k); }
};

template <int MODEL> struct HsNV;
template <> struct HsNV<HS_MODEL_LIF_CUBA> { static constexpr int value = HS_LIF_CUBA_NVARS; };
template <> struct HsNV<HS_MODEL_LIF_NOISE> { static constexpr int value = HS_LIF_NOISE_NVARS; };
template <> struct HsNV<HS_MODEL_HH_COBA> { static constexpr int value = HS_HH_NVARS; };

/* one neuron, one step: returns the threshold condition */
template <typename R, int MODEL>
__device__ __forceinline__ bool hs_neuron_step(R *s, const R *k, int not_ref, R vth, unsigned long long seed,
                                               unsigned int stream, int i, long long t)
{
    if (MODEL == HS_MODEL_LIF_CUBA) {
        HsOps<R>::cuba(&s[0], &s[1], &s[2], not_ref, k);
    } else if (MODEL == HS_MODEL_LIF_NOISE) {
        R xi = (R)hs_randn(seed, stream, (uint32_t)i, (int64_t)t);
        s[0] = HsOps<R>::noise(s[0], xi, k);
    } else {
        HsOps<R>::hh(s, k);
    }
    return (s[0] > vth) && not_ref;
}

/* ------------------------------------------------------------------------ */
/* block-level compaction: ballot per wave, prefix over waves, one atomic     */
/* per block on the spike count; indices ascending within the block          */
/* ------------------------------------------------------------------------ */
struct HsCompactLds {
    int wave_off[16];
    int total;
    int base;
    unsigned long long mbase;
};

__device__ __forceinline__ void hs_block_compact(bool pred, int idx, int32_t *list, int32_t *cnt, HsCompactLds *L,
                                                 int32_t *mon_i, int32_t *mon_t, unsigned long long *mon_n,
                                                 long long mon_cap, long long t)
{
    const int wid = threadIdx.x / HS_WAVE;
    const int nw = blockDim.x / HS_WAVE;
    unsigned long long m = __ballot(pred);
    if (hs_lane() == 0) L->wave_off[wid] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < nw; ++w) {
            int c = L->wave_off[w];
            L->wave_off[w] = tot;
            tot += c;
        }
        L->total = tot;
        L->base = tot ? atomicAdd(cnt, tot) : 0;
        L->mbase = (tot && mon_n) ? atomicAdd(mon_n, (unsigned long long)tot) : 0ull;
    }
    __syncthreads();
    if (pred) {
        int pos = L->wave_off[wid] + __popcll(m & hs_lanemask_lt());
        list[L->base + pos] = idx;
        if (mon_n) {
            unsigned long long r = L->mbase + (unsigned long long)pos;
            if ((long long)r < mon_cap) {
                mon_i[r] = idx;
                mon_t[r] = (int32_t)t;
            }
        }
    }
    __syncthreads();
}

/* ------------------------------------------------------------------------ */
/* update phase: one tile of blockDim neurons of one group                   */
/* ------------------------------------------------------------------------ */
template <typename R, int MODEL, bool FUSED>
__device__ void hs_update_tile(const DGroup<R> &g, int tile, long long t, long long no_reset_before,
                               HsCompactLds *L)
{
    constexpr int NV = HsNV<MODEL>::value;
    const int i = tile * blockDim.x + threadIdx.x;
    const bool valid = i < g.n;
    const int slot = (int)(t % g.Q);
    if (tile == 0 && threadIdx.x == 0) g.hist_cnt[(t + 1) % g.Q] = 0; /* ring slot of step t+1 */
    bool spk = false;
    if (valid) {
        const int32_t last = g.last[i];
        R s[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) s[v] = g.var[v][i];
        if (FUSED) {
            const int pb = (int)((t - 1) & 1);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                R *a = g.acc[pb][v];
                if (a) {
                    s[v] = s[v] + a[i];
                    a[i] = (R)0;
                }
            }
        }
        if (g.has_reset && (long long)last == t - 1 && t - 1 >= no_reset_before) s[0] = g.vr;
        for (int m = 0; m < g.nst; ++m) {
            if (i >= g.st_start[m] && i < g.st_stop[m]) {
                long long row = t - g.st_t0[m];
                if (row >= 0 && row < g.st_rows[m]) {
                    R val = s[0];
#pragma unroll
                    for (int v = 1; v < NV; ++v)
                        if (v == g.st_var[m]) val = s[v];
                    g.st_buf[m][row * (long long)(g.st_stop[m] - g.st_start[m]) + (i - g.st_start[m])] = val;
                }
            }
        }
        const int not_ref = (t - (long long)last) >= (long long)g.ref_steps;
        R k[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) k[j] = g.k[j];
        spk = hs_neuron_step<R, MODEL>(s, k, not_ref, g.vth, g.seed, g.stream, i, t);
        if (spk) g.last[i] = (int32_t)t;
#pragma unroll
        for (int v = 0; v < NV; ++v) g.var[v][i] = s[v];
    }
    hs_block_compact(spk, i, g.hist + (long long)slot * g.n, g.hist_cnt + slot, L, g.smon_i, g.smon_t,
                     g.smon_n, g.smon_cap, t);
}

template <typename R, bool FUSED>
__device__ void hs_update_phase(const DGroup<R> *groups, int ng, int total_tiles, long long t,
                                long long no_reset_before, HsCompactLds *L)
{
    for (int tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
        int gi = 0;
        while (gi + 1 < ng && tile >= groups[gi + 1].tile_begin) ++gi;
        const DGroup<R> &g = groups[gi];
        const int lt = tile - g.tile_begin;
        if (g.model == HS_MODEL_LIF_CUBA) hs_update_tile<R, HS_MODEL_LIF_CUBA, FUSED>(g, lt, t, no_reset_before, L);
        else if (g.model == HS_MODEL_LIF_NOISE) hs_update_tile<R, HS_MODEL_LIF_NOISE, FUSED>(g, lt, t, no_reset_before, L);
        else hs_update_tile<R, HS_MODEL_HH_COBA, FUSED>(g, lt, t, no_reset_before, L);
    }
}

/* ------------------------------------------------------------------------ */
/* delivery phase                                                            */
/* ------------------------------------------------------------------------ */
struct HsDeliverLds {
    int pref[HS_MAX_DELAY_SLOTS + 1];
};

/*
 * Every wave takes 64 (delay, spike) work items at a time; lane j resolves its
 * bundle (pre, d) to [s0, s0+len).  A wave scan of len flattens the 64 bundles
 * into one synapse index space which the lanes walk 64 at a time: consecutive
 * lanes read consecutive synapses of a bundle (coalesced post/weight reads),
 * each lane finds its bundle by a 6-step binary search over the scanned
 * offsets (cross-lane reads, no LDS traffic).  Effects are hardware atomics
 * (global_atomic_add_f64/_f32) into the target array.
 */
template <typename R, bool ACC>
__device__ void hs_deliver_path(const DPath<R> &p, long long t, HsDeliverLds *L, unsigned long long *ev_lane)
{
    const int D = p.D;
    /* prefix over the spike counts of the D due steps (wave 0) */
    if (threadIdx.x < HS_WAVE) {
        int carry = 0;
        for (int c = 0; c < D; c += HS_WAVE) {
            int di = c + hs_lane();
            int cnt = 0;
            if (di < D) {
                long long st = t - (long long)(p.dmin + di);
                long long slot = st % p.Q;
                if (slot < 0) slot += p.Q;
                cnt = p.hist_cnt[slot];
            }
            int tot;
            int ex = hs_wave_excl_scan(cnt, &tot);
            if (di < D) L->pref[di] = carry + ex;
            carry += tot;
        }
        if (hs_lane() == 0) L->pref[D] = carry;
    }
    __syncthreads();
    const int W = L->pref[D];
    R *target = ACC ? p.acc[t & 1] : p.target;
    const int nwb = blockDim.x / HS_WAVE;
    const long long wave_gid = (long long)blockIdx.x * nwb + threadIdx.x / HS_WAVE;
    const long long n_waves = (long long)gridDim.x * nwb;
    for (long long c = wave_gid; c * HS_WAVE < W; c += n_waves) {
        const int w = (int)(c * HS_WAVE) + hs_lane();
        int len = 0, s0 = 0, pre = 0;
        if (w < W) {
            int lo = 0, hi = D; /* pref[lo] <= w < pref[hi] */
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (L->pref[mid] <= w) lo = mid;
                else hi = mid;
            }
            const int kk = w - L->pref[lo];
            long long slot = (t - (long long)(p.dmin + lo)) % p.Q;
            if (slot < 0) slot += p.Q;
            pre = p.hist[slot * p.n_src + kk];
            const long long b = (long long)pre * D + lo;
            s0 = p.bstart[b];
            len = p.bstart[b + 1] - s0;
        }
        int tot;
        const int off = hs_wave_excl_scan(len, &tot);
        for (int j0 = 0; j0 < tot; j0 += HS_WAVE) {
            const int j = j0 + hs_lane();
            int lo = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                int cand = lo + step;
                int oc = __shfl(off, cand, HS_WAVE);
                if (oc <= j) lo = cand;
            }
            const int sb = __shfl(s0, lo, HS_WAVE);
            const int ob = __shfl(off, lo, HS_WAVE);
            const int pb = __shfl(pre, lo, HS_WAVE);
            if (j < tot) {
                const int syn = sb + (j - ob);
                const int tg = p.post[syn];
                R wv;
                if (p.wmode == HS_W_SCALAR) wv = p.wscalar;
                else if (p.wmode == HS_W_SYNAPSE) wv = p.w[syn];
                else wv = p.w[pb];
                unsafeAtomicAdd(&target[tg], wv);
            }
        }
        *ev_lane += (unsigned long long)tot;
    }
    __syncthreads();
}

__device__ __forceinline__ void hs_flush_events(unsigned long long ev, unsigned long long *events)
{
    /* lane 0 of every wave holds the wave's count (all lanes hold the same) */
    if (hs_lane() == 0 && ev) atomicAdd(events, ev);
}

/* ------------------------------------------------------------------------ */
/* engine kernels                                                            */
/* ------------------------------------------------------------------------ */
template <typename R, bool FUSED>
__global__ __launch_bounds__(1024) void hs_k_update(const DGroup<R> *groups, int ng, int total_tiles, HsCtl *ctl,
                                                    int koff)
{
    __shared__ HsCompactLds L;
    const long long t = ctl->base + koff;
    hs_update_phase<R, FUSED>(groups, ng, total_tiles, t, ctl->no_reset_before, &L);
}

template <typename R, bool ACC>
__global__ __launch_bounds__(1024) void hs_k_deliver(const DPath<R> *paths, int np, HsCtl *ctl, int koff)
{
    __shared__ HsDeliverLds L;
    const long long t = ctl->base + koff;
    unsigned long long ev = 0;
    for (int pi = 0; pi < np; ++pi) hs_deliver_path<R, ACC>(paths[pi], t, &L, &ev);
    hs_flush_events(ev, &ctl->events);
}

template <typename R>
__global__ __launch_bounds__(1024) void hs_k_fused(const DGroup<R> *groups, int ng, int total_tiles,
                                                   const DPath<R> *paths, int np, HsCtl *ctl, int koff)
{
    __shared__ HsCompactLds Lc;
    __shared__ HsDeliverLds Ld;
    const long long t = ctl->base + koff;
    hs_update_phase<R, true>(groups, ng, total_tiles, t, ctl->no_reset_before, &Lc);
    unsigned long long ev = 0;
    for (int pi = 0; pi < np; ++pi) hs_deliver_path<R, true>(paths[pi], t, &Ld, &ev);
    hs_flush_events(ev, &ctl->events);
}

/* end of a run at step t_end = ctl->base: apply the last step's accumulated
   effects (fused) and the resets of its spikers */
template <typename R, bool FUSED>
__global__ __launch_bounds__(1024) void hs_k_finalize(const DGroup<R> *groups, int ng, int total_tiles, HsCtl *ctl)
{
    const long long t = ctl->base;
    for (int tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
        int gi = 0;
        while (gi + 1 < ng && tile >= groups[gi + 1].tile_begin) ++gi;
        const DGroup<R> &g = groups[gi];
        const int i = (tile - g.tile_begin) * blockDim.x + threadIdx.x;
        if (i >= g.n) continue;
        if (FUSED) {
            const int pb = (int)((t - 1) & 1);
            for (int v = 0; v < g.nvars; ++v) {
                R *a = g.acc[pb][v];
                if (a) {
                    g.var[v][i] = g.var[v][i] + a[i];
                    a[i] = (R)0;
                }
            }
        }
        if (g.has_reset && (long long)g.last[i] == t - 1 && t - 1 >= ctl->no_reset_before) g.var[0][i] = g.vr;
    }
}

__global__ void hs_k_tick(HsCtl *ctl, long long k) { ctl->base += k; }
__global__ void hs_k_mark_reset(HsCtl *ctl) { ctl->no_reset_before = ctl->base; }

/* ------------------------------------------------------------------------ */
/* launchers                                                                 */
/* ------------------------------------------------------------------------ */
template <typename R>
hipError_t hs_launch_update(const DGroup<R> *groups, int ng, int total_tiles, HsCtl *ctl, int koff, int fused,
                            int block, int grid, hipStream_t s)
{
    if (fused) hipLaunchKernelGGL((hs_k_update<R, true>), dim3(grid), dim3(block), 0, s, groups, ng, total_tiles, ctl, koff);
    else hipLaunchKernelGGL((hs_k_update<R, false>), dim3(grid), dim3(block), 0, s, groups, ng, total_tiles, ctl, koff);
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_deliver(const DPath<R> *paths, int np, HsCtl *ctl, int koff, int fused, int block, int grid,
                             hipStream_t s)
{
    if (fused) hipLaunchKernelGGL((hs_k_deliver<R, true>), dim3(grid), dim3(block), 0, s, paths, np, ctl, koff);
    else hipLaunchKernelGGL((hs_k_deliver<R, false>), dim3(grid), dim3(block), 0, s, paths, np, ctl, koff);
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_fused(const DGroup<R> *groups, int ng, int total_tiles, const DPath<R> *paths, int np,
                           HsCtl *ctl, int koff, int block, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((hs_k_fused<R>), dim3(grid), dim3(block), 0, s, groups, ng, total_tiles, paths, np, ctl, koff);
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_finalize(const DGroup<R> *groups, int ng, int total_tiles, HsCtl *ctl, int fused, int block,
                              hipStream_t s)
{
    int grid = total_tiles < 4096 ? total_tiles : 4096;
    if (grid < 1) grid = 1;
    if (fused) hipLaunchKernelGGL((hs_k_finalize<R, true>), dim3(grid), dim3(block), 0, s, groups, ng, total_tiles, ctl);
    else hipLaunchKernelGGL((hs_k_finalize<R, false>), dim3(grid), dim3(block), 0, s, groups, ng, total_tiles, ctl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(hs_k_mark_reset, dim3(1), dim3(1), 0, s, ctl);
    return hipGetLastError();
}

hipError_t hs_launch_tick(HsCtl *ctl, long long k, hipStream_t s)
{
    hipLaunchKernelGGL(hs_k_tick, dim3(1), dim3(1), 0, s, ctl, k);
    return hipGetLastError();
}

template hipError_t hs_launch_update<double>(const DGroup<double> *, int, int, HsCtl *, int, int, int, int, hipStream_t);
template hipError_t hs_launch_update<float>(const DGroup<float> *, int, int, HsCtl *, int, int, int, int, hipStream_t);
template hipError_t hs_launch_deliver<double>(const DPath<double> *, int, HsCtl *, int, int, int, int, hipStream_t);
template hipError_t hs_launch_deliver<float>(const DPath<float> *, int, HsCtl *, int, int, int, int, hipStream_t);
template hipError_t hs_launch_fused<double>(const DGroup<double> *, int, int, const DPath<double> *, int, HsCtl *, int, int, int, hipStream_t);
template hipError_t hs_launch_fused<float>(const DGroup<float> *, int, int, const DPath<float> *, int, HsCtl *, int, int, int, hipStream_t);
template hipError_t hs_launch_finalize<double>(const DGroup<double> *, int, int, HsCtl *, int, int, hipStream_t);
template hipError_t hs_launch_finalize<float>(const DGroup<float> *, int, int, HsCtl *, int, int, hipStream_t);

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
        hs_block_compact(spk, i, spikes, count, &L, nullptr, nullptr, nullptr, 0, step);
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
        double z = uniform ? hs_rand(seed, stream, (uint32_t)i, step) : hs_randn(seed, stream, (uint32_t)i, step);
        if (prec == HS_F64) ((double *)out)[i] = z;
        else ((float *)out)[i] = (float)z;
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
        (void)hs_neuron_step<R, MODEL>(s, k, nr, (R)0, seed, stream, i, step);
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
        else r = (double)hs_expf((float)a);
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

template <typename R, bool ACC>
__global__ __launch_bounds__(256) void hs_k_deliver_one(DPath<R> p, HsCtl *ctl, unsigned long long *events)
{
    __shared__ HsDeliverLds L;
    unsigned long long ev = 0;
    hs_deliver_path<R, ACC>(p, ctl->base, &L, &ev);
    hs_flush_events(ev, events);
}

template <typename R>
hipError_t hs_launch_deliver_one_ctl(const DPath<R> &p, HsCtl *ctl, int acc, unsigned long long *events, int blocks,
                                     hipStream_t s)
{
    if (acc) hipLaunchKernelGGL((hs_k_deliver_one<R, true>), dim3(blocks), dim3(256), 0, s, p, ctl, events);
    else hipLaunchKernelGGL((hs_k_deliver_one<R, false>), dim3(blocks), dim3(256), 0, s, p, ctl, events);
    return hipGetLastError();
}

template hipError_t hs_launch_deliver_one_ctl<double>(const DPath<double> &, HsCtl *, int, unsigned long long *, int, hipStream_t);
template hipError_t hs_launch_deliver_one_ctl<float>(const DPath<float> &, HsCtl *, int, unsigned long long *, int, hipStream_t);
