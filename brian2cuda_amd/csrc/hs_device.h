/*
 * hs_device.h -- device building blocks shared by the engine kernels and the
 * standalone kernels: wave64 helpers, the per-neuron step, block compaction of
 * the spike list, and the delivery of (spike, delay) bundles.
 */
#ifndef HS_DEVICE_H
#define HS_DEVICE_H

#include "hs_internal.h"
#include "hs_model_exprs.h"
#include "hs_philox.h"

#define HS_WAVE 64

__device__ __forceinline__ int hs_lane() { return threadIdx.x & (HS_WAVE - 1); }

/*
 * Accesses to HBM arrays through pointers held in the DGroup / DPath
 * descriptors.  Such pointers are generic to the compiler, which then emits
 * FLAT instructions: those count against the LDS/scalar counter as well, so
 * every wait for a descriptor field also waits for the data loads in flight
 * and the state loads of a neuron serialise.  Casting to the global address
 * space gives global_load / global_store / global_atomic instead.
 */
#define HS_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ T hs_ld(const T *p)
{
    return *(const HS_GLOBAL T *)p;
}

template <typename T>
__device__ __forceinline__ void hs_st(T *p, T v)
{
    *(HS_GLOBAL T *)p = v;
}

/* relaxed device-scope atomic add whose result is not used (hardware FP
   atomics with -munsafe-fp-atomics) */
template <typename T>
__device__ __forceinline__ void hs_atomic_add(T *p, T v)
{
    (void)__hip_atomic_fetch_add((HS_GLOBAL T *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* relaxed device-scope atomic add returning the old value */
template <typename T>
__device__ __forceinline__ T hs_atomic_fetch_add(T *p, T v)
{
    return __hip_atomic_fetch_add((HS_GLOBAL T *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* x unchanged, but opaque to the optimiser (no instruction is emitted): a
   choice among values stays a choice among registers instead of becoming a
   load from a computed address, which would move the array to LDS */
template <typename T>
__device__ __forceinline__ T hs_opaque(T x)
{
    asm volatile("" : "+v"(x));
    return x;
}

__device__ __forceinline__ unsigned long long hs_lanemask_lt() { return (1ull << hs_lane()) - 1ull; }

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

/* a mod m in [0, m) for m > 0.  The engine sizes every ring (spike lists,
   accumulators, noise) to a power of two, so a step's slot is a mask instead
   of a signed 64-bit remainder (some 60 instructions per call); the standalone
   operators take rings of any size */
__device__ __forceinline__ long long hs_mod(long long a, long long m)
{
    if ((m & (m - 1)) == 0) return a & (m - 1);
    long long r = a % m;
    return r < 0 ? r + m : r;
}

/* x mod m for x in (-m, 2m): ring offsets relative to a base already reduced */
__device__ __forceinline__ int hs_wrap(int x, int m)
{
    if (x < 0) x += m;
    else if (x >= m) x -= m;
    return x;
}

/*
 * The FP32 fast form of the COBA Hodgkin-Huxley update (option HS_OPT_FAST_MATH):
 * the same expression text with the device exponential __expf, division as a
 * product with the hardware reciprocal (v_rcp_f32, 1 ulp) and FMA contraction.
 * Not bit-reproducible on the host: its error is measured against the FP64
 * oracle instead.
 */
#pragma push_macro("HS_HD")
#undef HS_HD
#define HS_HD __device__ static inline __attribute__((always_inline))
#define R float
#define HS_SFX(name) name##_f32fast
#define HS_EXP __expf
#define HS_DIV(a, b) ((a) * __builtin_amdgcn_rcpf(b))
#define HS_FAST_CONTRACT _Pragma("clang fp contract(fast)")
#include "hs_model_body.inc"
#undef R
#undef HS_SFX
#undef HS_EXP
#undef HS_DIV
#undef HS_FAST_CONTRACT
#pragma pop_macro("HS_HD")

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
    /* no fast form in FP64: the exact update */
    static __device__ __forceinline__ void hh_fast(double *s, const double *k) { hs_hh_coba_update_f64(s, k); }
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
    static __device__ __forceinline__ void hh_fast(float *s, const float *k) { hs_hh_coba_update_f32fast(s, k); }
};

/*
 * Code models (HS_MODEL_CODE + c): compiled only into the kernels that are
 * built at run time for a network holding such groups (hs_codegen.hip), whose
 * source defines HS_CG_MODELS (the count) and, after this header, one
 * HsCg<R, c> per model and precision:
 *   NV, NOISY                   variables, one normal draw per neuron-step
 *   step(s, not_ref, xi, dt, t, i)  update, then the threshold condition
 *   reset(s, dt, t, i)          the reset of a neuron that spiked at time t
 */
#ifdef HS_CG_MODELS
template <typename R, int C> struct HsCg;
template <int MODEL> struct HsNV {
    static constexpr int value = HsCg<double, MODEL - HS_MODEL_CODE>::NV;
};
template <int MODEL> __device__ constexpr bool hs_cg_noisy()
{
    if constexpr (HS_IS_CODE_MODEL(MODEL)) return HsCg<double, MODEL - HS_MODEL_CODE>::NOISY != 0;
    else return false;
}
#else
template <int MODEL> struct HsNV;
template <int MODEL> __device__ constexpr bool hs_cg_noisy() { return false; }
#endif
template <> struct HsNV<HS_MODEL_LIF_CUBA> { static constexpr int value = HS_LIF_CUBA_NVARS; };
template <> struct HsNV<HS_MODEL_LIF_NOISE> { static constexpr int value = HS_LIF_NOISE_NVARS; };
template <> struct HsNV<HS_MODEL_HH_COBA> { static constexpr int value = HS_HH_NVARS; };
template <> struct HsNV<HS_MODEL_HH_FAST> { static constexpr int value = HS_HH_NVARS; };
template <> struct HsNV<HS_MODEL_SPIKEGEN> { static constexpr int value = HS_SPIKEGEN_NVARS; };
template <> struct HsNV<HS_MODEL_POISSON> { static constexpr int value = HS_POISSON_NVARS; };

/* SpikeGeneratorGroup: does neuron i fire at step t (binary search in the step's ascending list) */
template <typename R>
__device__ __forceinline__ bool hs_spikegen_fires(const DGroup<R> &g, long long t, int i)
{
    const long long s = g.gen_period > 0 ? t % g.gen_period : t;
    if (s < 0 || s >= g.gen_steps) return false;
    int lo = hs_ld(g.gen_off + s), hi = hs_ld(g.gen_off + s + 1);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int v = hs_ld(g.gen_idx + mid);
        if (v == i) return true;
        if (v < i) lo = mid + 1;
        else hi = mid;
    }
    return false;
}

/* the model's noise draw for neuron i at step t (0 for noiseless models) */
template <typename R, int MODEL>
__device__ __forceinline__ R hs_noise(unsigned long long seed, unsigned int stream, int i, long long t)
{
    if (MODEL == HS_MODEL_LIF_NOISE || hs_cg_noisy<MODEL>()) {
        if (sizeof(R) == sizeof(float)) return (R)hs_randn_f32(seed, stream, (uint32_t)i, (int64_t)t);
        return (R)hs_randn(seed, stream, (uint32_t)i, (int64_t)t);
    }
    return (R)0;
}

/* one neuron, one step with its noise draw xi; returns the threshold condition */
template <typename R, int MODEL>
__device__ __forceinline__ bool hs_neuron_step(R *s, const R *k, int not_ref, R vth, R xi)
{
    if (MODEL == HS_MODEL_LIF_CUBA) {
        HsOps<R>::cuba(&s[0], &s[1], &s[2], not_ref, k);
    } else if (MODEL == HS_MODEL_LIF_NOISE) {
        s[0] = HsOps<R>::noise(s[0], xi, k);
    } else if (MODEL == HS_MODEL_HH_FAST) {
        HsOps<R>::hh_fast(s, k);
    } else {
        HsOps<R>::hh(s, k);
    }
    return (s[0] > vth) && not_ref;
}

/* ------------------------------------------------------------------------ */
/* block compaction: wave ballot + popcount, prefix over the block's waves,  */
/* one atomic per block on the list's count (indices ascending in a block)   */
/* ------------------------------------------------------------------------ */
struct HsCompactLds {
    int wave_off[16];
    int base;
    unsigned long long mbase;
};

__device__ __forceinline__ void hs_block_compact(bool pred, int idx, int32_t *list, int32_t *cnt, HsCompactLds *L,
                                                 int32_t *mon_i, int32_t *mon_t, unsigned long long *mon_n,
                                                 long long mon_cap, long long t, unsigned long long *spk_total)
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
        L->base = tot ? atomicAdd(cnt, tot) : 0;
        if (tot && spk_total) atomicAdd(spk_total, (unsigned long long)tot);
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
/* delivery of the bundles due in steps T .. T+w-1                           */
/* ------------------------------------------------------------------------ */
struct HsDeliverLds {
    int pref[HS_MAX_DELAY_SLOTS + 1];
    int wsum[16]; /* per-wave totals of the block scan (blocks of up to 1024 threads) */
};

/*
 * Work items are (step j of the window, delay slot di, spike k of the list
 * emitted at step T+j-dmin-di); their count per (j, di) comes from the spike
 * ring and is prefix-summed in LDS.  The items are dealt out CH per wavefront,
 * CH = the power of two that spreads them over every wave of the grid (1..64),
 * so that a few hundred bundles keep as many waves busy as 100k would.  Lane
 * l < CH resolves its bundle (pre, delay) to [s0, s0+len); a wave scan of len
 * flattens the CH bundles into one synapse index space that the lanes walk 64
 * at a time -- consecutive lanes read consecutive synapses of a bundle
 * (coalesced post/weight reads) and each finds its bundle by a 6-step binary
 * search over the scanned offsets (cross-lane reads).  Effects are hardware
 * atomics (global_atomic_add_f64 / _f32) into the state (plain) or the
 * accumulator slot of step T+j (fused).
 * wave_rank / n_waves: this wave's index among the waves doing delivery.
 */
template <typename R, bool ACC>
__device__ __forceinline__ void hs_deliver_path(const DPath<R> &p, long long T, int w, HsDeliverLds *L,
                                                long long wave_rank,
                                long long n_waves, unsigned long long *ev_lane, int di0 = 0, int dn = -1)
{
    /* delay slots di0 .. di0+dn-1 of the pathway (all by default); the items
       are indexed ((j * dn) + (di - di0)) * nsub + sub-list */
    const int D = p.D;
    if (dn < 0) dn = D;
    const int ns = p.nsub;
    const int npairs = w * dn * ns;
    const long long rowlen = (long long)ns * p.subcap;
    const int tq = (int)hs_mod(T, p.Q);                 /* ring slot of step T */
    const int ta = ACC ? (int)hs_mod(T, p.nacc) : 0;    /* accumulator slot of step T */
    {
        /* every thread of the block loads one count per pass: one memory
           latency for up to blockDim.x pairs, then a block scan */
        const int wid = threadIdx.x / HS_WAVE, nwv = blockDim.x / HS_WAVE;
        int carry = 0;
        for (int c = 0; c < npairs; c += blockDim.x) {
            const int q = c + threadIdx.x;
            int cnt = 0;
            if (q < npairs) {
                const int pq = q / ns, sq = q - pq * ns;
                const int j = pq / dn, di = di0 + (pq - j * dn);
                cnt = hs_ld(p.hist_cnt + ((long long)hs_wrap(tq + j - (p.dmin + di), p.Q) * ns + sq) * p.cstride);
            }
            int tot;
            const int ex = hs_wave_excl_scan(cnt, &tot);
            if (hs_lane() == 0) L->wsum[wid] = tot;
            __syncthreads();
            int before = carry, all = carry;
            for (int v = 0; v < nwv; ++v) {
                const int x = L->wsum[v];
                before += v < wid ? x : 0;
                all += x;
            }
            if (q < npairs) L->pref[q] = before + ex;
            carry = all;
            __syncthreads();
        }
        if (threadIdx.x == 0) L->pref[npairs] = carry;
    }
    __syncthreads();
    const int W = L->pref[npairs];
    long long per = ((long long)W + n_waves - 1) / n_waves;
    int CH = 1;
    while (CH < per && CH < HS_WAVE) CH <<= 1;
    for (long long c = wave_rank; c * CH < W; c += n_waves) {
        int len = 0, s0 = 0, pre = 0, jw = 0;
        if (hs_lane() < CH) {
            const long long wi = c * CH + hs_lane();
            if (wi < W) {
                int lo = 0, hi = npairs; /* pref[lo] <= wi < pref[hi] */
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (L->pref[mid] <= wi) lo = mid;
                    else hi = mid;
                }
                const int kk = (int)(wi - L->pref[lo]);
                const int pq = lo / ns, sq = lo - pq * ns;
                jw = pq / dn;
                const int di = di0 + (pq - jw * dn);
                const long long slot = hs_wrap(tq + jw - (p.dmin + di), p.Q);
                pre = hs_ld(p.hist + slot * rowlen + (long long)sq * p.subcap + kk);
                const long long b = (long long)pre * D + di;
                s0 = hs_ld(p.bstart + b);
                len = hs_ld(p.bstart + b + 1) - s0;
            }
        }
        int tot;
        const int off = hs_wave_excl_scan(len, &tot);
        for (int j0 = 0; j0 < tot; j0 += HS_WAVE) {
            const int j = j0 + hs_lane();
            int lo = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const int cand = lo + step;
                const int oc = __shfl(off, cand, HS_WAVE);
                if (oc <= j) lo = cand;
            }
            const int sb = __shfl(s0, lo, HS_WAVE);
            const int ob = __shfl(off, lo, HS_WAVE);
            const int pb = __shfl(pre, lo, HS_WAVE);
            const int jb = __shfl(jw, lo, HS_WAVE);
            if (j < tot) {
                const int syn = sb + (j - ob);
                const int tg = hs_ld(p.post + syn);
                R wv;
                if (p.wmode == HS_W_SCALAR) wv = p.wscalar;
                else if (p.wmode == HS_W_SYNAPSE) wv = hs_ld(p.w + syn);
                else wv = hs_ld(p.w + pb);
                R *target = ACC ? p.acc + (long long)hs_wrap(ta + jb, p.nacc) * p.n_tgt : p.target;
                hs_atomic_add(target + tg, wv);
            }
        }
        *ev_lane += (unsigned long long)tot;
    }
    __syncthreads();
}

/*
 * Delivery blocks split over the pathways of a launch: with at least as many
 * blocks as pathways, block b serves pathway b mod np only, its waves ranked
 * among that pathway's blocks, so a block pays one pass of count prefix and
 * bundle lookups instead of one per pathway; with fewer blocks, every block
 * serves every pathway in turn.  Loop: for (p = p0; p < np; p += pstep).
 */
struct HsSplit {
    int p0, pstep;
    long long rank, n_waves;
};

__device__ __forceinline__ HsSplit hs_split(int np, int b, int n_blk)
{
    const int nwb = blockDim.x / HS_WAVE, wv = threadIdx.x / HS_WAVE;
    if (np > 1 && n_blk >= np) {
        const int p = b % np, nb = (n_blk - p + np - 1) / np;
        return HsSplit{p, np, (long long)(b / np) * nwb + wv, (long long)nb * nwb};
    }
    return HsSplit{0, 1, (long long)b * nwb + wv, (long long)n_blk * nwb};
}

__device__ __forceinline__ void hs_flush_events(unsigned long long ev, unsigned long long *events)
{
    /* every lane of a wave holds the wave's count; lane 0 publishes it */
    if (hs_lane() == 0 && ev) hs_atomic_add(events, ev);
}

/* the same into the control block's spread counters, slot by wave rank */
__device__ __forceinline__ void hs_flush_events(unsigned long long ev, HsCtl *ctl, long long rank)
{
    if (hs_lane() == 0 && ev) hs_atomic_add(&ctl->ev_slot[(int)(rank % HS_EV_SLOTS) * HS_EV_STRIDE], ev);
}

__device__ __forceinline__ unsigned long long hs_wave_sum(unsigned long long x)
{
#pragma unroll
    for (int o = HS_WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, HS_WAVE);
    return x;
}

/* ------------------------------------------------------------------------ */
/* gather delivery of one target tile                                        */
/* ------------------------------------------------------------------------ */
struct HsGatherLds {
    int pref[HS_MAX_DELAY_SLOTS + HS_MAX_WINDOW + 1];
};

/*
 * The block owns targets [tile*HS_GTILE, +HS_GTILE) and a private LDS copy of
 * their accumulators for the w steps of the window (fused) or of their state
 * (plain, w == 1).  Every spike emitted in [T-dmax, T+w-1-dmin] can have
 * synapses arriving in the window; for each, the block reads its (pre, tile)
 * segment (on average out-degree / tiles entries, one cache line), keeps the
 * entries whose arrival step u + delay falls in the window and adds their
 * weights with LDS atomics (ds_add_f32 / ds_add_f64).  The tile is then written
 * back with ordinary coalesced stores: no global atomics, no target conflicts
 * between blocks.  Plain form: the LDS holds the state itself, so each event is
 * state += w exactly as the serial loop does it.
 */
template <typename R, bool ACC>
__device__ __forceinline__ void hs_gather_tile(const DPath<R> &p, int tile, long long T, int w, HsGatherLds *L, R *lacc,
                               unsigned long long *ev_lane)
{
    const int base = tile * HS_GTILE;
    const int ntl = min(HS_GTILE, p.n_tgt - base);
    if (ACC) {
        for (int i = threadIdx.x; i < w * HS_GTILE; i += blockDim.x) lacc[i] = (R)0;
    } else {
        for (int i = threadIdx.x; i < HS_GTILE; i += blockDim.x) lacc[i] = i < ntl ? hs_ld(p.target + base + i) : (R)0;
    }
    const int dmax = p.dmin + p.D - 1;
    const int ns = p.nsub;
    const int nU = (p.D - 1 + w) * ns; /* (candidate emission step T-dmax .. T+w-1-dmin, sub-list) */
    const long long rowlen = (long long)ns * p.subcap;
    const int tq = (int)hs_mod(T, p.Q);
    const int ta = ACC ? (int)hs_mod(T, p.nacc) : 0;
    if (threadIdx.x < HS_WAVE) {
        int carry = 0;
        for (int c = 0; c < nU; c += HS_WAVE) {
            const int q = c + hs_lane();
            const int cnt =
                q < nU ? hs_ld(p.hist_cnt + ((long long)hs_wrap(tq - dmax + q / ns, p.Q) * ns + q % ns) * p.cstride) : 0;
            int tot;
            const int ex = hs_wave_excl_scan(cnt, &tot);
            if (q < nU) L->pref[q] = carry + ex;
            carry += tot;
        }
        if (hs_lane() == 0) L->pref[nU] = carry;
    }
    __syncthreads();
    const int total = L->pref[nU];
    unsigned long long ev = 0;
    for (int c = threadIdx.x; c < total; c += blockDim.x) {
        int lo = 0, hi = nU; /* pref[lo] <= c < pref[hi] */
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (L->pref[mid] <= c) lo = mid;
            else hi = mid;
        }
        const int lu = lo / ns, ls = lo - lu * ns;
        const long long u = T - dmax + lu;
        const int pre = hs_ld(p.hist + (long long)hs_wrap(tq - dmax + lu, p.Q) * rowlen + (long long)ls * p.subcap +
                              (c - L->pref[lo]));
        const long long gi = (long long)tile * p.n_src + pre; /* tile-major: a block stays in its tile's rows */
        const int a = hs_ld(p.goff + gi), b = hs_ld(p.goff + gi + 1);
        for (int s = a; s < b; ++s) {
            const uint32_t e = hs_ld(p.gsyn + s);
            const long long arr = u + (long long)(p.dmin + (int)(e >> 16)) - T;
            if (arr >= 0 && arr < w) {
                R wv;
                if (p.wmode == HS_W_SCALAR) wv = p.wscalar;
                else if (p.wmode == HS_W_SYNAPSE) wv = hs_ld(p.w + s);
                else wv = hs_ld(p.w + pre);
                atomicAdd(&lacc[(ACC ? (int)arr * HS_GTILE : 0) + (int)(e & 0xffffu)], wv);
                ++ev;
            }
        }
    }
    *ev_lane += hs_wave_sum(ev);
    __syncthreads();
    if (ACC) {
        for (int i = threadIdx.x; i < w * HS_GTILE; i += blockDim.x) {
            const int j = i / HS_GTILE, l = i - j * HS_GTILE;
            if (l < ntl) {
                R *a = p.acc + (long long)hs_wrap(ta + j, p.nacc) * p.n_tgt + base + l;
                hs_st(a, hs_ld(a) + lacc[i]);
            }
        }
    } else {
        for (int i = threadIdx.x; i < ntl; i += blockDim.x) hs_st(p.target + base + i, lacc[i]);
    }
    __syncthreads();
}

#endif /* HS_DEVICE_H */
