/*
 * hs_engine.hip -- gfx950 kernels of the per-timestep engine.
 *
 *   update  : for every neuron, steps t .. t+w-1 with the state held in
 *             registers: pending effects (fused form) and the deferred reset,
 *             StateMonitor sample (when='start'), state update, threshold; then
 *             the w spike lists are compacted at once (wave64 ballot/popcount,
 *             one prefix over the block's waves, w parallel count reservations)
 *             into the spike ring (threshold.cu:19-45, reset.cu:3-27,
 *             common_group.cu:108-160, statemonitor.cu:74-100,
 *             spikemonitor.cu:165-196 in one pass)
 *   deliver : scatter-mode pathways -- effects of every (spike, delay) bundle due
 *             in steps t .. t+w-1, global atomics (synapses.cu:12-164 with the
 *             queue pop, spikequeue.h:504-516)
 *   gather  : gather-mode pathways -- one block per target tile, LDS accumulation
 *   fused   : update of step t and scatter delivery of step t in one launch on
 *             disjoint blocks (every delay >= 1 step: the delivery reads only
 *             spikes of earlier steps and writes only the accumulator of step t,
 *             which the update reads one step later)
 *   finalize: applies what is pending when a run ends
 *
 * Ring invariants (w <= W = the engine's window <= minimum delay):
 *   accumulators: nacc = W + 1 slots; deliver/gather of window [T, T+w) write
 *     slots T..T+w-1 while slot T-1 (written by the previous window) is still
 *     unread -- W + 1 slots keep them apart.
 *   spike ring: Q = dmax + 2W slots; an update launch clears the counts of steps
 *     T+w .. T+w+W-1 (the next window's), which alias steps older than any the
 *     current or next delivery reads.
 */
#include "hs_device.h"

struct HsCompactLdsW {
    int wave_cnt[HS_MAX_WINDOW][16];
    int base[HS_MAX_WINDOW];
    unsigned long long mbase[HS_MAX_WINDOW];
};

/* pipelined windows: the pathways a group's spikes push into */
template <typename R>
struct HsPush {
    const DPath<R> *paths;
    int np;
    unsigned long long *emit; /* events counter (this block's spread slot) */
};

/*
 * The near synapses (delay slots 0 .. nnear-1, delays < 2W) of the spikes this
 * wave emitted in steps T .. T+w-1, pushed into the accumulator slots of their
 * arrival steps -- all >= T + dmin >= T + W, none of which this launch reads.
 * Every spike's whole out-degree is counted as emitted events (a run's delivered
 * count is emitted minus what is still in flight at its end).
 */
template <typename R>
__device__ __forceinline__ void hs_push_near(const HsPush<R> &ps, int gid, int i, unsigned int bits, long long T,
                                             int w)
{
    (void)w;
    const int cnt = __popc(bits); /* this lane's spikes in the window (same neuron, same synapses) */
    if (!__ballot(cnt != 0)) return;
    constexpr int UNR = 4; /* passes whose loads are issued together */
    /* every lane at most one spike in the window (refractoriness as long as the
       window, or one-step windows): an item's spike is the lane's only one */
    const bool one = !__ballot(cnt > 1);
    unsigned long long emitted = 0;
    for (int pi = 0; pi < ps.np; ++pi) {
        const DPath<R> &p = ps.paths[pi];
        if (p.src != gid) continue;
        /* the spiking lane loads its bundle starts 0 .. nnear (nnear <= HS_MAX_WINDOW)
           and its end at once; item lanes read them by cross-lane shuffles */
        int o[HS_MAX_WINDOW + 1];
        int s_end = 0;
        if (cnt) {
            const long long base = (long long)i * p.D;
#pragma unroll
            for (int q = 0; q <= HS_MAX_WINDOW; ++q) o[q] = q <= p.nnear ? hs_ld(p.bstart + base + q) : 0;
            s_end = hs_ld(p.bstart + base + p.D);
        } else {
#pragma unroll
            for (int q = 0; q <= HS_MAX_WINDOW; ++q) o[q] = 0;
        }
        const int s0 = o[0];
        int len = 0;
#pragma unroll
        for (int q = 1; q <= HS_MAX_WINDOW; ++q)
            if (q == p.nnear) len = o[q] - s0;
        if (cnt) emitted += (unsigned long long)cnt * (unsigned long long)(s_end - s0);
        /* the wave's (spike, near synapse) items, flattened, 64 per pass */
        int tot;
        const int off = hs_wave_excl_scan(cnt * len, &tot);
        const int ta = (int)hs_mod(T, p.nacc);
        if (w == 1 && p.nnear == 1) {
            /* one step and one near delay slot (delay-0 networks): a lane holds
               at most one spike, its items are synapses s0 + k, and they all
               arrive in the one slot of step T + dmin */
            R *acc = p.acc + (long long)hs_wrap(ta + p.dmin, p.nacc) * p.n_tgt;
            for (int c0 = 0; c0 < tot; c0 += UNR * HS_WAVE) {
                int tg[UNR];
                R wv[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    tg[u] = -1;
                    if (c0 + u * HS_WAVE >= tot) continue; /* uniform: no items left */
                    const int idx = c0 + u * HS_WAVE + hs_lane();
                    int lo = 0;
#pragma unroll
                    for (int step = 32; step > 0; step >>= 1) {
                        const int cand = lo + step;
                        if (__shfl(off, cand, HS_WAVE) <= idx) lo = cand;
                    }
                    const int sy = __shfl(s0, lo, HS_WAVE) + (idx - __shfl(off, lo, HS_WAVE));
                    const int pb = __shfl(i, lo, HS_WAVE);
                    if (idx < tot) {
                        if (p.wmode == HS_W_SCALAR) wv[u] = p.wscalar;
                        else if (p.wmode == HS_W_SYNAPSE) wv[u] = hs_ld(p.w + sy);
                        else wv[u] = hs_ld(p.w + pb);
                        tg[u] = hs_ld(p.post + sy);
                    }
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u)
                    if (tg[u] >= 0) hs_atomic_add(acc + tg[u], wv[u]);
            }
            continue;
        }
        for (int c0 = 0; c0 < tot; c0 += UNR * HS_WAVE) {
            int tg[UNR], slot[UNR];
            R wv[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                tg[u] = -1;
                if (c0 + u * HS_WAVE >= tot) continue; /* uniform: no items left */
                const int idx = c0 + u * HS_WAVE + hs_lane();
                int lo = 0; /* owner: the last lane whose items start at or before idx */
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const int cand = lo + step;
                    if (__shfl(off, cand, HS_WAVE) <= idx) lo = cand;
                }
                const int ob = __shfl(off, lo, HS_WAVE);
                const int lb = __shfl(len, lo, HS_WAVE);
                const unsigned int bb = (unsigned int)__shfl((int)bits, lo, HS_WAVE);
                const int pb = __shfl(i, lo, HS_WAVE);
                int ob_q[HS_MAX_WINDOW]; /* the owner's near bundle starts (items read [0, nnear)) */
#pragma unroll
                for (int q = 0; q < HS_MAX_WINDOW; ++q) {
                    ob_q[q] = 0;
                    if (q < p.nnear) ob_q[q] = __shfl(o[q], lo, HS_WAVE);
                }
                if (idx < tot) {
                    const int k = idx - ob;
                    int r = 0; /* the owner's r-th spike of the window, synapse k - r*lb */
                    if (!one) r = k / lb;
                    const int s = ob_q[0] + (k - r * lb);
                    unsigned int bm = bb;
                    for (int q = 0; q < r; ++q) bm &= bm - 1;
                    const int j = __ffs(bm) - 1;
                    int di = 0;
#pragma unroll
                    for (int q = 1; q < HS_MAX_WINDOW; ++q) di += (q < p.nnear && ob_q[q] <= s) ? 1 : 0;
                    if (p.wmode == HS_W_SCALAR) wv[u] = p.wscalar;
                    else if (p.wmode == HS_W_SYNAPSE) wv[u] = hs_ld(p.w + s);
                    else wv[u] = hs_ld(p.w + pb);
                    slot[u] = hs_wrap(ta + j + p.dmin + di, p.nacc);
                    tg[u] = hs_ld(p.post + s);
                }
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u)
                if (tg[u] >= 0) hs_atomic_add(p.acc + (long long)slot[u] * p.n_tgt + tg[u], wv[u]);
        }
    }
    emitted = hs_wave_sum(emitted);
    if (hs_lane() == 0 && emitted) hs_atomic_add(ps.emit, emitted);
}

/* one step of one neuron inside the update window; returns the spike */
template <typename R, int MODEL, bool FUSED>
__device__ __forceinline__ bool hs_step_one(const DGroup<R> &g, R *s, const R *k, int32_t &last, long long t, int i,
                                            long long no_reset_before, R xi)
{
    constexpr int NV = HsNV<MODEL>::value;
    if constexpr (MODEL == HS_MODEL_SPIKEGEN) { /* spikegenerator.cu:43-80: no state, no refractoriness */
        const bool fire = hs_spikegen_fires(g, t, i);
        if (fire) last = (int32_t)t;
        return fire;
    } else if constexpr (MODEL == HS_MODEL_POISSON) { /* PoissonGroup: rand() < rates*dt */
        const R p = s[0] * k[0];
        const bool fire = hs_rand(g.seed, g.stream, (uint32_t)i, t) < (double)p;
        if (fire) last = (int32_t)t;
        return fire;
    } else {
        if (g.has_reset && (long long)last == t - 1 && t - 1 >= no_reset_before) {
#ifdef HS_CG_MODELS
            if constexpr (HS_IS_CODE_MODEL(MODEL))
                HsCg<R, MODEL - HS_MODEL_CODE>::reset(s, (R)g.dt, (double)(t - 1) * g.dt, i, g.seed, g.stream,
                                                      t - 1);
            else
#endif
                s[0] = g.vr;
        }
        for (int m = 0; m < g.nst; ++m) {
            /* the row of step t (uniform over the launch), then this neuron's column */
            const long long dt0 = t - g.st_t0[m];
            if (dt0 < 0) continue;
            const int every = g.st_every[m];
            const long long row = every == 1 ? dt0 : dt0 / every;
            if (row * every != dt0 || row >= g.st_rows[m]) continue;
            int col;
            if (g.st_map[m]) col = hs_ld(g.st_map[m] + i);
            else col = (i >= g.st_start[m] && i < g.st_stop[m]) ? i - g.st_start[m] : -1;
            if (col >= 0) {
                const int sv = g.st_var[m];
                R val = hs_opaque(s[0]);
#pragma unroll
                for (int v = 1; v < NV; ++v) val = sv == v ? hs_opaque(s[v]) : val;
                hs_st(g.st_buf[m] + row * (long long)g.st_width[m] + col, val);
            }
        }
        const int not_ref = (t - (long long)last) >= (long long)g.ref_steps;
        bool spk;
#ifdef HS_CG_MODELS
        if constexpr (HS_IS_CODE_MODEL(MODEL))
            spk = HsCg<R, MODEL - HS_MODEL_CODE>::step(s, not_ref, xi, (R)g.dt, (double)t * g.dt, i, g.seed, g.stream,
                                                       t) &&
                  not_ref;
        else
#endif
            spk = hs_neuron_step<R, MODEL>(s, k, not_ref, g.vth, xi);
        if (spk) last = (int32_t)t;
        return spk;
    }
}

template <typename R, int MODEL, bool FUSED, bool PUSH>
__device__ __forceinline__ void hs_update_tile(const DGroup<R> &g, int gid, int tile, long long T, int w, int wmax,
                                               long long no_reset_before, HsCompactLdsW *L,
                                               unsigned long long *spk_total, const HsPush<R> &push)
{
    constexpr int NV = HsNV<MODEL>::value;
    const int i = tile * blockDim.x + threadIdx.x;
    const bool valid = i < g.n;
    const int tq = (int)hs_mod(T, g.Q);  /* ring slot of step T */
    /* ring slots of the steps after this launch start from empty lists */
    if (tile == 0)
        for (int q = threadIdx.x; q < wmax * g.nsub; q += blockDim.x)
            hs_st(g.hist_cnt + ((long long)hs_wrap(tq + w + q / g.nsub, g.Q) * g.nsub + q % g.nsub) * HS_CNT_STRIDE,
                  0);
    /* this tile's sub-list (a tile lies inside one 256-neuron chunk) */
    const int sub = (int)(((long long)tile * blockDim.x >> 8) & (g.nsub - 1)); /* nsub: a power of two */
    const long long rowlen = (long long)g.nsub * g.subcap;
    R s[NV];
    int32_t last = HS_NEVER_SPIKED;
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = valid ? hs_ld(g.var[v] + i) : (R)0;
    if (valid) last = hs_ld(g.last + i);
    R k[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) k[j] = g.k[j];
    unsigned int bits = 0; /* bit j: spike at step T+j */
    if (valid) {
        /* the noise of the whole window does not depend on the state: its w
           draws are independent chains, evaluated together (instruction-level
           parallelism) and consumed one per step */
        constexpr bool NOISY = MODEL == HS_MODEL_LIF_NOISE || hs_cg_noisy<MODEL>();
        R xi[NOISY ? HS_MAX_WINDOW : 1];
        if (NOISY) {
            if (g.xi) {
                const int x0 = (int)hs_mod(T, g.xi_slots);
#pragma unroll
                for (int j = 0; j < HS_MAX_WINDOW; ++j)
                    xi[j] = j < w ? hs_ld(g.xi + (long long)hs_wrap(x0 + j, g.xi_slots) * g.n + i) : (R)0;
            } else {
#pragma unroll
                for (int j = 0; j < HS_MAX_WINDOW; ++j)
                    xi[j] = j < w ? hs_noise<R, MODEL>(g.seed, g.stream, i, T + j) : (R)0;
            }
        }
        /* the effects due at steps T-1 .. T+w-2 were all delivered before this
           launch: small models load the whole window's values at once (one
           memory latency per window) and consume them through a shift register;
           larger ones load one step ahead */
        constexpr int PF = (FUSED && NV <= 3) ? HS_MAX_WINDOW : 1;
        R pf[PF][NV];
        const int a0 = FUSED ? (int)hs_mod(T - 1, g.nacc) : 0; /* accumulator slot of step T-1 */
        if (FUSED) {
#pragma unroll
            for (int j = 0; j < PF; ++j) {
                const long long sj = (long long)hs_wrap(a0 + (j < w ? j : 0), g.nacc) * g.n + i;
#pragma unroll
                for (int v = 0; v < NV; ++v) pf[j][v] = (j < w && g.acc[v]) ? hs_ld(g.acc[v] + sj) : (R)0;
            }
        }
#pragma unroll 1
        for (int j = 0; j < w; ++j) {
            if (FUSED) {
                R cur[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) cur[v] = pf[0][v];
                const long long slot = (long long)hs_wrap(a0 + j, g.nacc) * g.n + i;
                if (PF > 1) {
#pragma unroll
                    for (int m = 0; m + 1 < PF; ++m)
#pragma unroll
                        for (int v = 0; v < NV; ++v) pf[m][v] = pf[m + 1][v];
                } else if (j + 1 < w) {
                    const long long sn = (long long)hs_wrap(a0 + j + 1, g.nacc) * g.n + i;
#pragma unroll
                    for (int v = 0; v < NV; ++v) pf[0][v] = g.acc[v] ? hs_ld(g.acc[v] + sn) : (R)0;
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    if (g.acc[v]) {
                        s[v] = s[v] + cur[v];
                        hs_st(g.acc[v] + slot, (R)0);
                    }
                }
            }
            R xj = (R)0;
            if (NOISY) {
                xj = xi[0];
#pragma unroll
                for (int m = 0; m + 1 < (NOISY ? HS_MAX_WINDOW : 1); ++m) xi[m] = xi[m + 1];
            }
            if (hs_step_one<R, MODEL, FUSED>(g, s, k, last, T + j, i, no_reset_before, xj)) bits |= 1u << j;
        }
    }
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) hs_st(g.var[v] + i, s[v]);
        hs_st(g.last + i, last);
    }
    /* compaction of the w spike lists: counts per wave, then w reservations */
    const int wid = threadIdx.x / HS_WAVE;
    const int nw = blockDim.x / HS_WAVE;
    for (int j = 0; j < w; ++j) {
        const unsigned long long m = __ballot((bits >> j) & 1u);
        if (hs_lane() == 0) L->wave_cnt[j][wid] = __popcll(m);
    }
    __syncthreads();
    int tot = 0;
    if ((int)threadIdx.x < w) {
        const int j = threadIdx.x;
        for (int v = 0; v < nw; ++v) {
            const int c = L->wave_cnt[j][v];
            L->wave_cnt[j][v] = tot;
            tot += c;
        }
        L->base[j] = tot ? hs_atomic_fetch_add(g.hist_cnt + ((long long)hs_wrap(tq + j, g.Q) * g.nsub + sub) *
                                                                HS_CNT_STRIDE,
                                               (int32_t)tot)
                         : 0;
        L->mbase[j] = (tot && g.smon_n) ? hs_atomic_fetch_add(g.smon_n, (unsigned long long)tot) : 0ull;
    }
    if (threadIdx.x < HS_WAVE) { /* one count per block and window */
        const unsigned long long s = hs_wave_sum((unsigned long long)tot);
        if (threadIdx.x == 0 && s && spk_total) hs_atomic_add(spk_total, s);
    }
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        const bool spk = (bits >> j) & 1u;
        const unsigned long long m = __ballot(spk);
        if (spk) {
            const int pos = L->wave_cnt[j][wid] + __popcll(m & hs_lanemask_lt());
            hs_st(g.hist + (long long)hs_wrap(tq + j, g.Q) * rowlen + (long long)sub * g.subcap + L->base[j] + pos,
                  (int32_t)i);
            if (g.smon_n) {
                const unsigned long long r = L->mbase[j] + (unsigned long long)pos;
                if ((long long)r < g.smon_cap) {
                    hs_st(g.smon_i + r, (int32_t)i);
                    hs_st(g.smon_t + r, (int32_t)(T + j));
                }
            }
        }
    }
    if (g.rmon_cnt) { /* PopulationRateMonitor: one count per wave and step */
        const bool in = valid && i >= g.rmon_start && i < g.rmon_stop;
        for (int j = 0; j < w; ++j) {
            const unsigned long long m = __ballot(in && ((bits >> j) & 1u));
            const long long row = T + j - g.rmon_t0;
            if (hs_lane() == 0 && m && row >= 0 && row < g.rmon_rows) hs_atomic_add(g.rmon_cnt + row, (int32_t)__popcll(m));
        }
    }
    if (PUSH) hs_push_near<R>(push, gid, i, bits, T, w);
    __syncthreads();
}

#ifdef HS_CG_MODELS
/* a tile of a code-model group: the update specialised for its model */
template <typename R, bool FUSED, bool PUSH>
__device__ __forceinline__ void hs_cg_update_tile(const DGroup<R> &g, int gid, int tile, long long T, int w, int wmax,
                                                  long long no_reset_before, HsCompactLdsW *L,
                                                  unsigned long long *spk_total, const HsPush<R> &push)
{
    const int c = g.model - HS_MODEL_CODE;
    if (c == 0)
        hs_update_tile<R, HS_MODEL_CODE + 0, FUSED, PUSH>(g, gid, tile, T, w, wmax, no_reset_before, L, spk_total, push);
#if HS_CG_MODELS > 1
    else if (c == 1)
        hs_update_tile<R, HS_MODEL_CODE + 1, FUSED, PUSH>(g, gid, tile, T, w, wmax, no_reset_before, L, spk_total, push);
#endif
#if HS_CG_MODELS > 2
    else if (c == 2)
        hs_update_tile<R, HS_MODEL_CODE + 2, FUSED, PUSH>(g, gid, tile, T, w, wmax, no_reset_before, L, spk_total, push);
#endif
#if HS_CG_MODELS > 3
    else if (c == 3)
        hs_update_tile<R, HS_MODEL_CODE + 3, FUSED, PUSH>(g, gid, tile, T, w, wmax, no_reset_before, L, spk_total, push);
#endif
}
#endif

/* MS: the model of every group (kernel specialised for it, so the register
   budget is that model's), or 0 for networks mixing models */
template <typename R, bool FUSED, int MS, bool PUSH = false>
__device__ __forceinline__ void hs_update_phase(const DGroup<R> *groups, int ng, int total_tiles, long long T, int w,
                                                int wmax,
                                long long no_reset_before, HsCompactLdsW *L, unsigned long long *spk_total,
                                int block_rank, int n_blocks, const HsPush<R> &push = HsPush<R>{nullptr, 0, nullptr})
{
    for (int tile = block_rank; tile < total_tiles; tile += n_blocks) {
        int gi = 0;
        while (gi + 1 < ng && tile >= groups[gi + 1].tile_begin) ++gi;
        const DGroup<R> &g = groups[gi];
        const int lt = tile - g.tile_begin;
        if (MS != 0) {
            hs_update_tile<R, (MS != 0 ? MS : 1), FUSED, PUSH>(g, gi, lt, T, w, wmax, no_reset_before, L, spk_total,
                                                               push);
        } else if (g.model == HS_MODEL_LIF_CUBA) {
            hs_update_tile<R, HS_MODEL_LIF_CUBA, FUSED, PUSH>(g, gi, lt, T, w, wmax, no_reset_before, L, spk_total,
                                                              push);
        } else if (g.model == HS_MODEL_LIF_NOISE) {
            hs_update_tile<R, HS_MODEL_LIF_NOISE, FUSED, PUSH>(g, gi, lt, T, w, wmax, no_reset_before, L, spk_total,
                                                               push);
        } else if (g.model == HS_MODEL_SPIKEGEN) {
            hs_update_tile<R, HS_MODEL_SPIKEGEN, FUSED, PUSH>(g, gi, lt, T, w, wmax, no_reset_before, L, spk_total,
                                                              push);
        } else if (g.model == HS_MODEL_POISSON) {
            hs_update_tile<R, HS_MODEL_POISSON, FUSED, PUSH>(g, gi, lt, T, w, wmax, no_reset_before, L, spk_total,
                                                             push);
        } else if (g.model == HS_MODEL_HH_FAST) {
            hs_update_tile<R, HS_MODEL_HH_FAST, FUSED, PUSH>(g, gi, lt, T, w, wmax, no_reset_before, L, spk_total,
                                                             push);
#ifdef HS_CG_MODELS
        } else if (HS_IS_CODE_MODEL(g.model)) {
            hs_cg_update_tile<R, FUSED, PUSH>(g, gi, lt, T, w, wmax, no_reset_before, L, spk_total, push);
#endif
        } else {
            hs_update_tile<R, HS_MODEL_HH_COBA, FUSED, PUSH>(g, gi, lt, T, w, wmax, no_reset_before, L, spk_total,
                                                             push);
        }
    }
}

/* noise draws of steps T0 .. T0+nsteps-1 of every group with a noise ring, by
   blocks brank of nblk: xi[(t mod xi_slots)*n + i] = draw of neuron i at step t */
template <typename R>
__device__ __forceinline__ void hs_noise_ahead(const DGroup<R> *groups, int ng, long long T0, int nsteps, int brank,
                                               int nblk)
{
    const long long nthr = (long long)nblk * blockDim.x;
    const long long tid = (long long)brank * blockDim.x + threadIdx.x;
    for (int gi = 0; gi < ng; ++gi) {
        const DGroup<R> &g = groups[gi];
        if (!g.xi) continue;
        const int s0 = (int)hs_mod(T0, g.xi_slots);
        const long long tot = (long long)g.n * nsteps;
        for (long long q = tid; q < tot; q += nthr) {
            const int j = (int)(q / g.n);
            const int i = (int)(q - (long long)j * g.n);
            hs_st(g.xi + (long long)hs_wrap(s0 + j, g.xi_slots) * g.n + i,
                  hs_noise<R, HS_MODEL_LIF_NOISE>(g.seed, g.stream, i, T0 + j));
        }
    }
}

/* the draws of the first window of a chunk (the later windows' draws are made
   by the update launch of the window before) */
template <typename R>
__global__ __launch_bounds__(256) void hs_k_noise(const DGroup<R> *__restrict__ groups, int ng, HsCtl *ctl, int koff, int nsteps)
{
    hs_noise_ahead<R>(groups, ng, ctl->base + koff, nsteps, blockIdx.x, gridDim.x);
}

/* blocks [0, n_upd): the update of steps T .. T+w-1; blocks [n_upd, grid): the
   noise draws of the next window, steps T+w .. T+w+wmax-1 (no dependence on the
   state, so they run beside the update on otherwise idle waves) */
template <typename R, bool FUSED, int MS>
__global__ __launch_bounds__(256) void hs_k_update(const DGroup<R> *__restrict__ groups, int ng, int total_tiles, HsCtl *ctl,
                                                    int koff, int w, int wmax, int n_upd)
{
    __shared__ HsCompactLdsW L;
    const long long T = ctl->base + koff;
    if ((int)blockIdx.x < n_upd)
        hs_update_phase<R, FUSED, MS>(groups, ng, total_tiles, T, w, wmax, ctl->no_reset_before, &L,
                                      ctl->spk_slot + (blockIdx.x % HS_EV_SLOTS) * HS_EV_STRIDE, blockIdx.x, n_upd);
    else
        hs_noise_ahead<R>(groups, ng, T + w, wmax, blockIdx.x - n_upd, gridDim.x - n_upd);
}

/*
 * Pipelined windows, one launch per window of steps T .. T+w-1 (W = the engine's
 * window <= every delay):
 *   blocks [0, n_upd)           update of the window, then the push of the near
 *                               synapses (delays < 2W) of its spikes
 *   blocks [n_upd, +n_far)      pull delivery of the far synapses (delays >= 2W)
 *                               arriving in T+W .. T+W+w-1, whose spikes are all
 *                               from steps <= T-1: finished by earlier launches
 *   the remaining blocks        the next window's noise draws
 * The update reads accumulator slots T-1 .. T+w-2; every write of this launch
 * goes to slots >= T+W (nacc = 3W+1 keeps them apart).  After the launch every
 * effect arriving before T+w+W has been applied or accumulated.
 */
union HsWindowLds {
    HsCompactLdsW c;
    HsDeliverLds d;
};

template <typename R, int MS>
__device__ __forceinline__ void hs_window_body(const DGroup<R> *__restrict__ groups, int ng, int total_tiles,
                                               const DPath<R> *__restrict__ paths, int np, HsCtl *ctl, int koff, int w,
                                               int W, int n_upd, int n_far, HsWindowLds &L)
{
    const long long T = ctl->base + koff;
    const int b = blockIdx.x;
    if (b < n_upd) {
        const HsPush<R> push{paths, np, ctl->ev_slot + (b % HS_EV_SLOTS) * HS_EV_STRIDE};
        hs_update_phase<R, true, MS, true>(groups, ng, total_tiles, T, w, W, ctl->no_reset_before, &L.c,
                                           ctl->spk_slot + (b % HS_EV_SLOTS) * HS_EV_STRIDE, b, n_upd, push);
    } else if (b < n_upd + n_far) {
        const HsSplit sp = hs_split(np, b - n_upd, n_far);
        unsigned long long ev = 0; /* counted at emission instead */
        for (int pi = sp.p0; pi < np; pi += sp.pstep) {
            const DPath<R> &p = paths[pi];
            if (p.nnear < p.D)
                hs_deliver_path<R, true>(p, T + W, w, &L.d, sp.rank, sp.n_waves, &ev, p.nnear, p.D - p.nnear);
        }
    } else {
        hs_noise_ahead<R>(groups, ng, T + w, W, b - n_upd - n_far, gridDim.x - n_upd - n_far);
    }
}

template <typename R, int MS>
__global__ __launch_bounds__(256) void hs_k_window(const DGroup<R> *__restrict__ groups, int ng, int total_tiles,
                                                    const DPath<R> *__restrict__ paths, int np, HsCtl *ctl, int koff,
                                                    int w, int W, int n_upd, int n_far)
{
    __shared__ HsWindowLds L;
    hs_window_body<R, MS>(groups, ng, total_tiles, paths, np, ctl, koff, w, W, n_upd, n_far, L);
}

/* the same with 8 waves per SIMD (at most 64 VGPRs): the FP32 kernels of one
   model fit without spilling, and a 512,000-neuron update (8,000 waves) then
   takes one round of the chip's 8,192 wave slots instead of 7,168 + a tail;
   launched only where the update has more waves than 7 per SIMD */
template <typename R, int MS>
__global__ __launch_bounds__(256, 8) void hs_k_window_o8(const DGroup<R> *__restrict__ groups, int ng,
                                                         int total_tiles, const DPath<R> *__restrict__ paths, int np,
                                                         HsCtl *ctl, int koff, int w, int W, int n_upd, int n_far)
{
    __shared__ HsWindowLds L;
    hs_window_body<R, MS>(groups, ng, total_tiles, paths, np, ctl, koff, w, W, n_upd, n_far, L);
}

/* events still in flight at step E = ctl->base (arrival >= E) of spikes emitted
   before E: every spike (i, u) with u >= E - dmax still owes its bundles of
   delay >= E - u.  One block; the result goes to ctl->pending. */
template <typename R>
__global__ __launch_bounds__(256) void hs_k_pending(const DPath<R> *__restrict__ paths, int np, HsCtl *ctl)
{
    __shared__ unsigned long long part[4];
    const long long E = ctl->base;
    unsigned long long acc = 0;
    for (int pi = 0; pi < np; ++pi) {
        const DPath<R> &p = paths[pi];
        const int dmax = p.dmin + p.D - 1;
        for (long long u = E - dmax; u < E; ++u) {
            const int slot = (int)hs_mod(u, p.Q);
            const int f = (int)max(0LL, E - u - p.dmin); /* first delay slot arriving at or after E */
            for (int s = 0; s < p.nsub; ++s) {
                const int cnt = p.hist_cnt[((long long)slot * p.nsub + s) * p.cstride];
                const int32_t *list = p.hist + (long long)slot * p.nsub * p.subcap + (long long)s * p.subcap;
                for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
                    const long long base = (long long)list[k] * p.D;
                    acc += (unsigned long long)(p.bstart[base + p.D] - p.bstart[base + f]);
                }
            }
        }
    }
    acc = hs_wave_sum(acc);
    if (hs_lane() == 0) part[threadIdx.x / HS_WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int v = 0; v < (int)(blockDim.x / HS_WAVE); ++v) t += part[v];
        ctl->pending = t;
    }
}

/* blocks [0, n_del): delivery of the bundles due in steps T .. T+w-1; blocks
   [n_del, grid): the noise draws of the same steps, read by the update launch
   that follows (the delivery is bound by atomic throughput and leaves the
   vector ALUs idle) */
template <typename R, bool ACC>
__global__ __launch_bounds__(1024) void hs_k_deliver(const DPath<R> *__restrict__ paths, int np, const DGroup<R> *__restrict__ groups, int ng,
                                                     HsCtl *ctl, int koff, int w, int n_del)
{
    __shared__ HsDeliverLds L;
    const long long T = ctl->base + koff;
    if ((int)blockIdx.x >= n_del) {
        hs_noise_ahead<R>(groups, ng, T, w, blockIdx.x - n_del, gridDim.x - n_del);
        return;
    }
    const HsSplit sp = hs_split(np, blockIdx.x, n_del);
    unsigned long long ev = 0;
    for (int pi = sp.p0; pi < np; pi += sp.pstep)
        if (paths[pi].dmode == HS_DELIVER_SCATTER)
            hs_deliver_path<R, ACC>(paths[pi], T, w, &L, sp.rank, sp.n_waves, &ev);
    hs_flush_events(ev, ctl, (long long)blockIdx.x * (blockDim.x / HS_WAVE) + threadIdx.x / HS_WAVE);
}

template <typename R, bool ACC>
__global__ __launch_bounds__(1024) void hs_k_gather(const DPath<R> *__restrict__ paths, int np, HsCtl *ctl, int koff, int w)
{
    __shared__ HsGatherLds L;
    __shared__ R lacc[HS_MAX_WINDOW * HS_GTILE];
    const long long T = ctl->base + koff;
    unsigned long long ev = 0;
    for (int pi = 0; pi < np; ++pi) {
        const DPath<R> &p = paths[pi];
        if (p.dmode != HS_DELIVER_GATHER) continue;
        for (int tile = blockIdx.x; tile < p.n_tiles; tile += gridDim.x)
            hs_gather_tile<R, ACC>(p, tile, T, w, &L, lacc, &ev);
    }
    hs_flush_events(ev, ctl, (long long)blockIdx.x * (blockDim.x / HS_WAVE) + threadIdx.x / HS_WAVE);
}

template <typename R, int MS>
__global__ __launch_bounds__(256) void hs_k_fused(const DGroup<R> *__restrict__ groups, int ng, int total_tiles,
                                                   const DPath<R> *__restrict__ paths, int np, HsCtl *ctl, int koff, int n_upd)
{
    __shared__ union {
        HsCompactLdsW c;
        HsDeliverLds d;
    } L;
    const long long T = ctl->base + koff;
    /* the delivery blocks come first: they do not depend on this launch's
       update, so they are dispatched at once and overlap with it instead of
       starting after the last update block */
    const int n_del = (int)gridDim.x - n_upd;
    if ((int)blockIdx.x >= n_del) {
        const int b = (int)blockIdx.x - n_del;
        hs_update_phase<R, true, MS>(groups, ng, total_tiles, T, 1, 1, ctl->no_reset_before, &L.c,
                                     ctl->spk_slot + (b % HS_EV_SLOTS) * HS_EV_STRIDE, b, n_upd);
    } else {
        const HsSplit sp = hs_split(np, blockIdx.x, n_del);
        unsigned long long ev = 0;
        for (int pi = sp.p0; pi < np; pi += sp.pstep)
            hs_deliver_path<R, true>(paths[pi], T, 1, &L.d, sp.rank, sp.n_waves, &ev);
        hs_flush_events(ev, ctl, (long long)blockIdx.x * (blockDim.x / HS_WAVE) + threadIdx.x / HS_WAVE);
    }
}

#ifdef HS_CG_MODELS
/* the reset of code-model neuron i that spiked at step ts: its variables through
   the model's reset code */
template <typename R, int C>
__device__ __forceinline__ void hs_cg_reset_c(const DGroup<R> &g, int i, long long ts)
{
    constexpr int NV = HsCg<R, C>::NV;
    R s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = g.var[v][i];
    HsCg<R, C>::reset(s, (R)g.dt, (double)ts * g.dt, i, g.seed, g.stream, ts);
#pragma unroll
    for (int v = 0; v < NV; ++v) g.var[v][i] = s[v];
}

template <typename R>
__device__ __forceinline__ void hs_cg_reset_one(const DGroup<R> &g, int i, long long ts)
{
    const int c = g.model - HS_MODEL_CODE;
    if (c == 0) hs_cg_reset_c<R, 0>(g, i, ts);
#if HS_CG_MODELS > 1
    else if (c == 1) hs_cg_reset_c<R, 1>(g, i, ts);
#endif
#if HS_CG_MODELS > 2
    else if (c == 2) hs_cg_reset_c<R, 2>(g, i, ts);
#endif
#if HS_CG_MODELS > 3
    else if (c == 3) hs_cg_reset_c<R, 3>(g, i, ts);
#endif
}
#endif

/* end of a run at step t_end = ctl->base: apply the last step's accumulated
   effects (fused form) and the resets of its spikers */
template <typename R, bool FUSED>
__global__ __launch_bounds__(256) void hs_k_finalize(const DGroup<R> *__restrict__ groups, int ng, int total_tiles, HsCtl *ctl)
{
    const long long t = ctl->base;
    for (int tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
        int gi = 0;
        while (gi + 1 < ng && tile >= groups[gi + 1].tile_begin) ++gi;
        const DGroup<R> &g = groups[gi];
        const int i = (tile - g.tile_begin) * blockDim.x + threadIdx.x;
        if (i >= g.n) continue;
        if (FUSED) {
            const long long slot = hs_mod(t - 1, g.nacc) * (long long)g.n + i;
            for (int v = 0; v < g.nvars; ++v) {
                R *a = g.acc[v];
                if (a) {
                    g.var[v][i] = g.var[v][i] + a[slot];
                    a[slot] = (R)0;
                }
            }
        }
        if (g.has_reset && (long long)g.last[i] == t - 1 && t - 1 >= ctl->no_reset_before) {
#ifdef HS_CG_MODELS
            if (HS_IS_CODE_MODEL(g.model)) {
                hs_cg_reset_one<R>(g, i, t - 1);
                continue;
            }
#endif
            g.var[0][i] = g.vr;
        }
    }
}

__global__ void hs_k_tick(HsCtl *ctl, long long k) { ctl->base += k; }
__global__ void hs_k_mark_reset(HsCtl *ctl) { ctl->no_reset_before = ctl->base; }

#ifndef __HIPCC_RTC__
/* ------------------------------------------------------------------------ */
/* launchers                                                                 */
/* ------------------------------------------------------------------------ */
template <typename R, bool FUSED, int MS>
static void hs_update_ms(const DGroup<R> *__restrict__ groups, int ng, int total_tiles, HsCtl *ctl, int koff, int w, int wmax,
                         int block, int grid, int n_noise, hipStream_t s)
{
    hipLaunchKernelGGL((hs_k_update<R, FUSED, MS>), dim3(grid + n_noise), dim3(block), 0, s, groups, ng, total_tiles,
                       ctl, koff, w, wmax, grid);
}

template <typename R>
hipError_t hs_launch_update(const DGroup<R> *groups, int ng, int total_tiles, HsCtl *ctl, int koff, int w, int wmax,
                            int fused, int ms, int block, int grid, int n_noise, hipStream_t s)
{
    if (w < 1 || w > HS_MAX_WINDOW || wmax < w || wmax > HS_MAX_WINDOW || block > 256 || grid < 1 || n_noise < 0)
        return hipErrorInvalidValue;
#define HS_UPD(F, M) hs_update_ms<R, F, M>(groups, ng, total_tiles, ctl, koff, w, wmax, block, grid, n_noise, s)
    if (fused) {
        if (ms == HS_MODEL_LIF_CUBA) HS_UPD(true, HS_MODEL_LIF_CUBA);
        else if (ms == HS_MODEL_LIF_NOISE) HS_UPD(true, HS_MODEL_LIF_NOISE);
        else if (ms == HS_MODEL_HH_COBA) HS_UPD(true, HS_MODEL_HH_COBA);
        else if (ms == HS_MODEL_HH_FAST) HS_UPD(true, HS_MODEL_HH_FAST);
        else HS_UPD(true, 0);
    } else {
        if (ms == HS_MODEL_LIF_CUBA) HS_UPD(false, HS_MODEL_LIF_CUBA);
        else if (ms == HS_MODEL_LIF_NOISE) HS_UPD(false, HS_MODEL_LIF_NOISE);
        else if (ms == HS_MODEL_HH_COBA) HS_UPD(false, HS_MODEL_HH_COBA);
        else if (ms == HS_MODEL_HH_FAST) HS_UPD(false, HS_MODEL_HH_FAST);
        else HS_UPD(false, 0);
    }
#undef HS_UPD
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_noise(const DGroup<R> *groups, int ng, HsCtl *ctl, int koff, int nsteps, int grid, hipStream_t s)
{
    if (nsteps < 1 || grid < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(hs_k_noise<R>, dim3(grid), dim3(256), 0, s, groups, ng, ctl, koff, nsteps);
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_deliver(const DPath<R> *paths, int np, const DGroup<R> *groups, int ng, HsCtl *ctl, int koff,
                             int w, int fused, int block, int grid, int n_noise, hipStream_t s)
{
    if (grid < 1 || n_noise < 0) return hipErrorInvalidValue;
    if (fused)
        hipLaunchKernelGGL((hs_k_deliver<R, true>), dim3(grid + n_noise), dim3(block), 0, s, paths, np, groups, ng, ctl,
                           koff, w, grid);
    else
        hipLaunchKernelGGL((hs_k_deliver<R, false>), dim3(grid + n_noise), dim3(block), 0, s, paths, np, groups, ng,
                           ctl, koff, w, grid);
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_gather(const DPath<R> *paths, int np, HsCtl *ctl, int koff, int w, int fused, int block,
                            int grid, hipStream_t s)
{
    if (w < 1 || w > HS_MAX_WINDOW || block < 64 || block > 1024 || block % 64) return hipErrorInvalidValue;
    if (fused) hipLaunchKernelGGL((hs_k_gather<R, true>), dim3(grid), dim3(block), 0, s, paths, np, ctl, koff, w);
    else hipLaunchKernelGGL((hs_k_gather<R, false>), dim3(grid), dim3(block), 0, s, paths, np, ctl, koff, w);
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_fused(const DGroup<R> *groups, int ng, int total_tiles, const DPath<R> *paths, int np,
                           HsCtl *ctl, int koff, int ms, int block, int n_upd, int n_del, hipStream_t s)
{
    if (block > 256) return hipErrorInvalidValue;
#define HS_FUS(M)                                                                                                 \
    hipLaunchKernelGGL((hs_k_fused<R, M>), dim3(n_upd + n_del), dim3(block), 0, s, groups, ng, total_tiles, paths, \
                       np, ctl, koff, n_upd)
    if (ms == HS_MODEL_LIF_CUBA) HS_FUS(HS_MODEL_LIF_CUBA);
    else if (ms == HS_MODEL_LIF_NOISE) HS_FUS(HS_MODEL_LIF_NOISE);
    else if (ms == HS_MODEL_HH_COBA) HS_FUS(HS_MODEL_HH_COBA);
    else if (ms == HS_MODEL_HH_FAST) HS_FUS(HS_MODEL_HH_FAST);
    else HS_FUS(0);
#undef HS_FUS
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_finalize(const DGroup<R> *groups, int ng, int total_tiles, HsCtl *ctl, int fused, int block,
                              hipStream_t s)
{
    int grid = total_tiles < 4096 ? total_tiles : 4096;
    if (grid < 1) grid = 1;
    if (fused)
        hipLaunchKernelGGL((hs_k_finalize<R, true>), dim3(grid), dim3(block), 0, s, groups, ng, total_tiles, ctl);
    else
        hipLaunchKernelGGL((hs_k_finalize<R, false>), dim3(grid), dim3(block), 0, s, groups, ng, total_tiles, ctl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hs_launch_mark_reset(ctl, s);
}

hipError_t hs_launch_mark_reset(HsCtl *ctl, hipStream_t s)
{
    hipLaunchKernelGGL(hs_k_mark_reset, dim3(1), dim3(1), 0, s, ctl);
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_window(const DGroup<R> *groups, int ng, int total_tiles, const DPath<R> *paths, int np,
                            HsCtl *ctl, int koff, int w, int W, int ms, int block, int n_upd, int n_far, int n_noise,
                            hipStream_t s)
{
    if (w < 1 || w > W || W > HS_MAX_WINDOW || block > 256 || block < 64 || block % 64 || n_upd < 1 || n_far < 0 ||
        n_noise < 0)
        return hipErrorInvalidValue;
    const dim3 grid(n_upd + n_far + n_noise);
    /* 8 waves per SIMD where the update needs more waves than 7 per SIMD hold */
    const bool o8 = (long long)n_upd * (block / HS_WAVE) > 7LL * 4 * hs_rt_cus();
#define HS_WIN(M)                                                                                                    \
    do {                                                                                                             \
        if constexpr (sizeof(R) == sizeof(float) && M != 0)                                                          \
            if (o8)                                                                                                  \
            {                                                                                                        \
                hipLaunchKernelGGL((hs_k_window_o8<R, M>), grid, dim3(block), 0, s, groups, ng, total_tiles, paths,  \
                                   np, ctl, koff, w, W, n_upd, n_far);                                               \
                break;                                                                                               \
            }                                                                                                        \
        hipLaunchKernelGGL((hs_k_window<R, M>), grid, dim3(block), 0, s, groups, ng, total_tiles, paths, np, ctl,     \
                           koff, w, W, n_upd, n_far);                                                                \
    } while (0)
    if (ms == HS_MODEL_LIF_CUBA) HS_WIN(HS_MODEL_LIF_CUBA);
    else if (ms == HS_MODEL_LIF_NOISE) HS_WIN(HS_MODEL_LIF_NOISE);
    else if (ms == HS_MODEL_HH_COBA) HS_WIN(HS_MODEL_HH_COBA);
    else if (ms == HS_MODEL_HH_FAST) HS_WIN(HS_MODEL_HH_FAST);
    else HS_WIN(0);
#undef HS_WIN
    return hipGetLastError();
}

template <typename R>
hipError_t hs_launch_pending(const DPath<R> *paths, int np, HsCtl *ctl, hipStream_t s)
{
    hipLaunchKernelGGL(hs_k_pending<R>, dim3(1), dim3(256), 0, s, paths, np, ctl);
    return hipGetLastError();
}

hipError_t hs_launch_tick(HsCtl *ctl, long long k, hipStream_t s)
{
    hipLaunchKernelGGL(hs_k_tick, dim3(1), dim3(1), 0, s, ctl, k);
    return hipGetLastError();
}

#define HS_INSTANTIATE(R)                                                                                         \
    template hipError_t hs_launch_update<R>(const DGroup<R> *, int, int, HsCtl *, int, int, int, int, int, int, int, \
                                            int, hipStream_t);                                                    \
    template hipError_t hs_launch_noise<R>(const DGroup<R> *, int, HsCtl *, int, int, int, hipStream_t);           \
    template hipError_t hs_launch_deliver<R>(const DPath<R> *, int, const DGroup<R> *, int, HsCtl *, int, int, int, \
                                             int, int, int, hipStream_t);                                         \
    template hipError_t hs_launch_gather<R>(const DPath<R> *, int, HsCtl *, int, int, int, int, int, hipStream_t);  \
    template hipError_t hs_launch_fused<R>(const DGroup<R> *, int, int, const DPath<R> *, int, HsCtl *, int, int,   \
                                           int, int, int, hipStream_t);                                           \
    template hipError_t hs_launch_finalize<R>(const DGroup<R> *, int, int, HsCtl *, int, int, hipStream_t);     \
    template hipError_t hs_launch_window<R>(const DGroup<R> *, int, int, const DPath<R> *, int, HsCtl *, int, int, \
                                            int, int, int, int, int, int, hipStream_t);                           \
    template hipError_t hs_launch_pending<R>(const DPath<R> *, int, HsCtl *, hipStream_t);
HS_INSTANTIATE(double)
HS_INSTANTIATE(float)
#endif /* !__HIPCC_RTC__ */
