/*
 * hs_conndev.h -- device side of the connectivity generated on the device
 * (hs_connect.hip holds the algorithm's description and the host launchers):
 * the draws, the batch of one row, fixed out-degree sampling and the count and
 * write passes, as templates on a filter policy F.  The library instantiates
 * them with HsConnPlain (connect(p) and fixed out-degree as they are); a
 * condition and a per-pair probability given as expressions are compiled at
 * run time into a policy of their own (hs_codegen.hip):
 *   F::COND   bool cond<R>(c, a, i, j): the pair (i, j) is kept only where it holds,
 *             tested after the candidate has been selected
 *   F::PEXPR  double p<R>(c, a, i, j): the pair's probability, replacing c.p in
 *             the one-draw-per-pair test (the geometric skips need a constant p)
 * i and j are the absolute indices of the groups; a holds the state arrays of
 * the source and target groups for expressions that read them.
 */
#ifndef HS_CONNDEV_H
#define HS_CONNDEV_H

#include "hs_device.h"
#include "hs_internal.h"
#include "hs_philox.h"

/* no condition, constant probability */
struct HsConnPlain {
    /* COPIES: the write pass of connect(n > 1) (only generated programs carry it) */
    static constexpr bool COND = false, PEXPR = false, WEXPR = false, COPIES = false;
    template <typename R>
    static __device__ __forceinline__ bool cond(const HsConnSpec &, const HsConnVars &, long long, long long)
    {
        return true;
    }
    template <typename R>
    static __device__ __forceinline__ double p(const HsConnSpec &, const HsConnVars &, long long, long long)
    {
        return 0.0;
    }
    /* the weight of copy cp of pair (i, j), when WEXPR */
    template <typename R>
    static __device__ __forceinline__ double w(const HsConnSpec &, const HsConnVars &, long long, long long, int)
    {
        return 0.0;
    }
};

#define HS_CONN_WAVES 4 /* rows in flight per block, one per wave; LDS: D ints per wave
                           (fixed out-degree: 2 waves, plus 2P ints per wave) */

__device__ __forceinline__ hs_u32x4 hs_conn_words(const HsConnSpec &c, uint32_t i, uint32_t k, uint32_t tag)
{
    hs_u32x4 ctr;
    ctr.v[0] = i;
    ctr.v[1] = k;
    ctr.v[2] = HS_CONN_CTR | tag;
    ctr.v[3] = c.stream;
    return hs_philox4x32_10(ctr, (uint32_t)c.seed, (uint32_t)(c.seed >> 32));
}

/*
 * One batch of row i: 64 draws (from draw k0, position jb) or 64 pair tests
 * (targets jb .. jb+63).  Lane l learns whether its draw selects a target and
 * which (sel, j); jb/k0 advance to the next batch and done is set (wave-uniform)
 * when the row has ended.
 */
template <typename F, typename R>
__device__ __forceinline__ void hs_conn_batch(const HsConnSpec &c, const HsConnVars &a, int i, long long &jb,
                                              uint32_t &k0, bool &done, bool &sel, long long &j)
{
    const int lane = hs_lane();
    const long long hi = c.post_hi;
    if (c.jump) {
        const uint32_t k = k0 + (uint32_t)lane;
        const hs_u32x4 wd = hs_conn_words(c, (uint32_t)i, k >> 1, HS_CONN_TAG_SKIP);
        const double r = (k & 1u) ? hs_u53(wd.v[2], wd.v[3]) : hs_u53(wd.v[0], wd.v[1]);
        const bool zero = r == 0.0;
        long long inc = 1; /* skip + 1; a skip at or past the row length only has to end the row */
        if (!zero) {
            const double skip = floor(hs_log(r) * c.pconst);
            const double cap = (double)(hi - c.post_lo);
            inc = (long long)(skip < cap ? skip : cap) + 1;
        }
        long long incl = inc;
#pragma unroll
        for (int o = 1; o < HS_WAVE; o <<= 1) {
            const long long y = __shfl_up(incl, o, HS_WAVE);
            if (lane >= o) incl += y;
        }
        const long long pos = jb + incl - 1;
        const unsigned long long stop = __ballot(zero || pos >= hi);
        const int first = stop ? __ffsll((long long)stop) - 1 : HS_WAVE;
        sel = lane < first && (c.autapses || pos != i);
        j = pos;
        if (stop) {
            done = true;
        } else {
            jb += __shfl(incl, HS_WAVE - 1, HS_WAVE);
            k0 += HS_WAVE;
        }
    } else {
        const long long jj = jb + lane;
        const uint32_t m = (uint32_t)(jj - c.post_lo);
        const hs_u32x4 wd = hs_conn_words(c, (uint32_t)i, m >> 1, HS_CONN_TAG_PAIR);
        const double r = (m & 1u) ? hs_u53(wd.v[2], wd.v[3]) : hs_u53(wd.v[0], wd.v[1]);
        double p = c.p;
        if constexpr (F::PEXPR) { /* the pair's own probability, read only for targets in range */
            p = 0.0;
            if (jj < hi) p = F::template p<R>(c, a, i, jj);
        }
        sel = jj < hi && r < p && (c.autapses || jj != i);
        j = jj;
        jb += HS_WAVE;
        if (jb >= hi) done = true;
    }
}

/* LDS ints of the sample buffer of a fixed out-degree row: the power of two >= K */
__host__ __device__ __forceinline__ int hs_conn_pow2(int k)
{
    int p = 1;
    while (p < k) p <<= 1;
    return p;
}

/* the lanes of a wave see each other's LDS writes made before this point: the
   fences keep the compiler from moving LDS accesses across it, and a wave's LDS
   accesses are performed in program order */
__device__ __forceinline__ void hs_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/*
 * Fixed out-degree: the K distinct values of row i (value x the target
 * post_lo + x), sorted ascending, into buf[0..K) (LDS of the wave,
 * P = hs_conn_pow2(K) ints; tmp: P more).  The missing values are drawn at
 * positions m..K-1 with the row's next counters, the buffer sorted (bitonic
 * over P, padded with INT_MAX), repeats dropped; until K distinct values remain
 * (about K^2/(2n) repeats in the first round).  The lanes of a wave exchange
 * values through LDS, with hs_wave_sync between the passes.
 */
__device__ void hs_conn_fixed_row(const HsConnSpec &c, int i, int32_t *buf, int32_t *tmp, int P)
{
    const int lane = hs_lane();
    const int K = c.fixed_k;
    const int n = c.post_hi - c.post_lo;
    uint32_t next = 0;
    int m = 0;
    hs_wave_sync(); /* the previous row's reads of buf are done */
    while (m < K) {
        for (int k = m + lane; k < K; k += HS_WAVE) {
            const hs_u32x4 wd = hs_conn_words(c, (uint32_t)i, next + (uint32_t)(k - m), HS_CONN_TAG_FIX);
            const int32_t v = (int32_t)floor(hs_u53(wd.v[0], wd.v[1]) * (double)n);
            buf[k] = v < n - 1 ? v : n - 1;
        }
        next += (uint32_t)(K - m);
        for (int k = K + lane; k < P; k += HS_WAVE) buf[k] = 0x7fffffff;
        hs_wave_sync();
        for (int size = 2; size <= P; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = lane; t < (P >> 1); t += HS_WAVE) {
                    const int lo = ((t & ~(stride - 1)) << 1) | (t & (stride - 1));
                    const int hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const int32_t a = buf[lo], b = buf[hi];
                    if ((a > b) == up) {
                        buf[lo] = b;
                        buf[hi] = a;
                    }
                }
                hs_wave_sync();
            }
        }
        /* drop repeats: keep buf[k] when it differs from buf[k-1] */
        int cnt = 0;
        for (int b0 = 0; b0 < K; b0 += HS_WAVE) {
            const int k = b0 + lane;
            const bool keep = k < K && (k == 0 || buf[k] != buf[k - 1]);
            const unsigned long long msk = __ballot(keep);
            if (keep) tmp[cnt + __popcll(msk & hs_lanemask_lt())] = buf[k];
            cnt += __popcll(msk);
        }
        hs_wave_sync();
        for (int k = lane; k < cnt; k += HS_WAVE) buf[k] = tmp[k];
        hs_wave_sync();
        m = cnt;
    }
}

/* delay slot of a synapse from its delay uniform: (int)(d/dt + 0.5) - dmin,
   the reference's quantisation (synapses_push_spikes.cu:375) */
__device__ __forceinline__ int hs_conn_slot(const HsConnSpec &c, double ud)
{
    return (int)((c.dlo + ud * c.dspan) / c.dt + 0.5) - c.dmin;
}

/* count pass of the rows of this block; tab: the block's dynamic LDS, [waves][D (+ 2P)] */
template <typename F, typename R>
__device__ __forceinline__ void hs_conn_count_body(const HsConnSpec &c, const HsConnVars &a, int32_t *__restrict__ cnt,
                                                   int *err, int32_t *tab)
{
    const int lane = hs_lane(), wv = threadIdx.x / HS_WAVE, nwv = blockDim.x / HS_WAVE;
    const int P = c.fixed_k > 0 ? hs_conn_pow2(c.fixed_k) : 0;
    int32_t *t = tab + wv * (c.D + 2 * P);
    int32_t *sbuf = t + c.D;
    const int rows = c.pre_hi - c.pre_lo;
    for (int r = blockIdx.x * nwv + wv; r < rows; r += gridDim.x * nwv) {
        const int i = c.pre_lo + r;
        const size_t row = (size_t)i * (size_t)c.D;
        for (int d = lane; d < c.D; d += HS_WAVE) t[d] = 0;
        long long jb = c.post_lo;
        uint32_t k0 = 0;
        bool done = jb >= c.post_hi;
        if (c.fixed_k > 0) {
            hs_conn_fixed_row(c, i, sbuf, sbuf + P, P);
            jb = 0;
            done = false;
        }
        while (!done) {
            bool sel;
            long long j;
            if (c.fixed_k > 0) {
                const int k = (int)jb + lane;
                j = k < c.fixed_k ? c.post_lo + (long long)sbuf[k] : 0;
                sel = k < c.fixed_k && (c.autapses || j != i);
                jb += HS_WAVE;
                done = jb >= c.fixed_k;
            } else {
                hs_conn_batch<F, R>(c, a, i, jb, k0, done, sel, j);
            }
            if constexpr (F::COND) sel = sel && F::template cond<R>(c, a, i, j);
            const int ncp = c.ncopy > 1 ? c.ncopy : 1;
            for (int cp = 0; sel && cp < ncp; ++cp) {
                double ud = 0.0;
                if (c.dspan != 0.0) {
                    const hs_u32x4 x = hs_conn_words(c, (uint32_t)i, (uint32_t)j, HS_CONN_TAG_COPY(cp));
                    ud = hs_u53(x.v[0], x.v[1]);
                }
                const int s = hs_conn_slot(c, ud);
                if ((unsigned)s >= (unsigned)c.D) *err = 1;
                else atomicAdd(&t[s], 1);
            }
        }
        for (int d = lane; d < c.D; d += HS_WAVE) cnt[row + d] = t[d];
    }
}

/* write of one batch's selected pairs with c.ncopy > 1 synapses each (their
   delays and weights from tags HS_CONN_TAG_COPY(cp)): positions within each
   delay slot follow creation order, lane (pair) first, then copy -- per slot, a
   lane's copies in that slot are counted and the counts prefix-summed over the
   wave */
template <typename F, typename R>
__device__ __forceinline__ void hs_conn_write_copies(const HsConnSpec &c, const HsConnVars &a, int i, long long j,
                                                     bool sel, bool draw,
                                                     size_t row, const int32_t *__restrict__ bstart,
                                                     int32_t *__restrict__ post, R *__restrict__ w, int *err,
                                                     int32_t *q)
{
    const int lane = hs_lane();
    int sl[HS_CONN_NMAX];
    double uwv[HS_CONN_NMAX];
#pragma unroll
    for (int cp = 0; cp < HS_CONN_NMAX; ++cp) {
        sl[cp] = -1;
        uwv[cp] = 0.0;
        if (sel && cp < c.ncopy) {
            double ud = 0.0;
            if (draw) {
                const hs_u32x4 x = hs_conn_words(c, (uint32_t)i, (uint32_t)j, HS_CONN_TAG_COPY(cp));
                ud = hs_u53(x.v[0], x.v[1]);
                uwv[cp] = hs_u53(x.v[2], x.v[3]);
            }
            const int s = hs_conn_slot(c, ud);
            if ((unsigned)s >= (unsigned)c.D) *err = 1;
            else sl[cp] = s;
        }
    }
    for (;;) {
        int mine = -1;
#pragma unroll
        for (int cp = 0; cp < HS_CONN_NMAX; ++cp)
            if (mine < 0 && sl[cp] >= 0) mine = sl[cp];
        const unsigned long long act = __ballot(mine >= 0);
        if (!act) break;
        const int leader = __ffsll((long long)act) - 1;
        const int s0 = __shfl(mine, leader, HS_WAVE);
        int cnt = 0;
#pragma unroll
        for (int cp = 0; cp < HS_CONN_NMAX; ++cp) cnt += sl[cp] == s0 ? 1 : 0;
        int incl = cnt;
        for (int o = 1; o < HS_WAVE; o <<= 1) {
            const int y = __shfl_up(incl, o, HS_WAVE);
            if (lane >= o) incl += y;
        }
        const int total = __shfl(incl, HS_WAVE - 1, HS_WAVE);
        int pos = q[s0] + incl - cnt;
        const int end = bstart[row + s0 + 1];
#pragma unroll
        for (int cp = 0; cp < HS_CONN_NMAX; ++cp) {
            if (sl[cp] == s0) {
                if (pos >= end) {
                    *err = 2;
                } else {
                    post[pos] = (int32_t)j;
                    if constexpr (F::WEXPR) {
                        if (w) w[pos] = (R)F::template w<R>(c, a, i, j, cp);
                    } else {
                        if (w) w[pos] = (R)(c.wlo + uwv[cp] * c.wspan);
                    }
                }
                ++pos;
                sl[cp] = -1;
            }
        }
        if (lane == leader) q[s0] += total;
    }
}

/* write pass of the rows of this block; cur: the block's dynamic LDS, [waves][D (+ 2P)], the
   next free position of each bundle of the row */
template <typename F, typename R>
__device__ __forceinline__ void hs_conn_write_body(const HsConnSpec &c, const HsConnVars &a,
                                                   const int32_t *__restrict__ bstart, int32_t *__restrict__ post,
                                                   R *__restrict__ w, int *err, int32_t *cur)
{
    const int lane = hs_lane(), wv = threadIdx.x / HS_WAVE, nwv = blockDim.x / HS_WAVE;
    const int P = c.fixed_k > 0 ? hs_conn_pow2(c.fixed_k) : 0;
    int32_t *q = cur + wv * (c.D + 2 * P);
    int32_t *sbuf = q + c.D;
    const int rows = c.pre_hi - c.pre_lo;
    const bool draw = c.dspan != 0.0 || (w && c.wspan != 0.0);
    for (int r = blockIdx.x * nwv + wv; r < rows; r += gridDim.x * nwv) {
        const int i = c.pre_lo + r;
        const size_t row = (size_t)i * (size_t)c.D;
        for (int d = lane; d < c.D; d += HS_WAVE) q[d] = bstart[row + d];
        long long jb = c.post_lo;
        uint32_t k0 = 0;
        bool done = jb >= c.post_hi;
        if (c.fixed_k > 0) {
            hs_conn_fixed_row(c, i, sbuf, sbuf + P, P);
            jb = 0;
            done = false;
        }
        while (!done) {
            bool sel;
            long long j;
            if (c.fixed_k > 0) {
                const int k = (int)jb + lane;
                j = k < c.fixed_k ? c.post_lo + (long long)sbuf[k] : 0;
                sel = k < c.fixed_k && (c.autapses || j != i);
                jb += HS_WAVE;
                done = jb >= c.fixed_k;
            } else {
                hs_conn_batch<F, R>(c, a, i, jb, k0, done, sel, j);
            }
            if constexpr (F::COND) sel = sel && F::template cond<R>(c, a, i, j);
            if constexpr (F::COPIES) {
                hs_conn_write_copies<F, R>(c, a, i, j, sel, draw, row, bstart, post, w, err, q);
                continue;
            }
            double ud = 0.0, uw = 0.0;
            int s = -1;
            if (sel) {
                if (draw) {
                    const hs_u32x4 x = hs_conn_words(c, (uint32_t)i, (uint32_t)j, HS_CONN_TAG_SYN);
                    ud = hs_u53(x.v[0], x.v[1]);
                    uw = hs_u53(x.v[2], x.v[3]);
                }
                s = hs_conn_slot(c, ud);
                if ((unsigned)s >= (unsigned)c.D) {
                    *err = 1;
                    sel = false;
                }
            }
            /* rank of each synapse among the batch's synapses of its slot, in lane
               (= creation) order: one ballot per distinct slot */
            int pos = -1;
            unsigned long long act = __ballot(sel);
            while (act) {
                const int leader = __ffsll((long long)act) - 1;
                const int s0 = __shfl(s, leader, HS_WAVE);
                const bool mine = sel && s == s0;
                const unsigned long long m = __ballot(mine);
                if (mine) pos = q[s0] + __popcll(m & hs_lanemask_lt());
                if (lane == leader) q[s0] += __popcll(m);
                act &= ~m;
            }
            if (sel) {
                if (pos >= bstart[row + s + 1]) {
                    *err = 2;
                } else {
                    post[pos] = (int32_t)j;
                    if constexpr (F::WEXPR) {
                        if (w) w[pos] = (R)F::template w<R>(c, a, i, j, 0);
                    } else {
                        if (w) w[pos] = (R)(c.wlo + uw * c.wspan);
                    }
                }
            }
        }
    }
}

#endif /* HS_CONNDEV_H */
