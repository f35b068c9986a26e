/*
 * hs_connect.hip -- random connectivity generated on the device, written straight
 * into the (pre, delay) bundle table that the delivery kernels read.
 *
 * Replaces two host stages of the reference:
 *   - Synapses.connect(p=...) generation (brian2cuda/templates/synapses_create_generator.cu:164-465):
 *     a host loop over presynaptic neurons drawing from a host buffer that cuRAND
 *     refills, followed by copies of the synapse arrays to the device;
 *   - the bundle construction of _before_run_<push_spikes>
 *     (brian2cuda/templates/synapses_push_spikes.cu:171-993): a host sort per
 *     presynaptic neuron and one copy per (pre, delay) bundle.
 *
 * Row i selects targets in increasing j with the reference's two algorithms
 * (synapses_create_generator.cu:315-333):
 *   p <  0.25: j starts at post_lo; draw k (r in [0, 1)) advances j by
 *              skip_k = floor(log(r) / log(1 - p)) and selects it, then j += 1
 *              (r == 0 ends the row; a skip past post_hi ends it too)
 *   p >= 0.25: every j is tested with its own draw and selected when r < p
 * The draws are Philox words keyed by (seed, stream, i, k) (layout in
 * hs_internal.h), so a row's synapses do not depend on launch geometry or on the
 * other rows, and the oracle regenerates any subset of rows.
 *
 * One wave per row.  The 64 lanes take 64 consecutive draws of the row at once:
 * the position of draw k is post_lo + sum_{m<=k} (skip_m + 1) - 1, an inclusive
 * wave scan, so the batch reproduces the serial chain exactly; the first draw
 * with r == 0 or a position past the end stops the row.  Three passes:
 *   count: synapses per (row, delay slot), LDS counters of the wave
 *   scan : exclusive prefix sum of the n_src*D counts -> bstart (int64 block sums)
 *   write: the rows again; synapse (i, j) goes to bstart[i*D + slot] + its rank
 *          among the row's earlier synapses of that slot (ballots per distinct
 *          slot of the batch, LDS cursors), so each bundle keeps creation
 *          (ascending j) order and a row's writes stay in the row's own region
 *          of the table (one wave, a few kB: the L2 merges them into full lines)
 * Bounds: the delay quantisation is monotone in the uniform, so every slot lies
 * in [0, D) with D taken from the quantised bounds; both passes still check it,
 * and the write pass checks each position against the end of its bundle, so a
 * disagreement between the passes raises an error flag instead of writing
 * outside the arrays.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device.h"
#include "hs_internal.h"
#include "hs_philox.h"

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
__device__ __forceinline__ void hs_conn_batch(const HsConnSpec &c, int i, long long &jb, uint32_t &k0, bool &done,
                                              bool &sel, long long &j)
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
        sel = jj < hi && r < c.p && (c.autapses || jj != i);
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

__global__ void __launch_bounds__(HS_CONN_WAVES * HS_WAVE) hs_k_conn_count(HsConnSpec c, int32_t *__restrict__ cnt,
                                                                           int *err)
{
    extern __shared__ int32_t tab[]; /* [waves][D (+ 2P)] */
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
                hs_conn_batch(c, i, jb, k0, done, sel, j);
            }
            if (sel) {
                double ud = 0.0;
                if (c.dspan != 0.0) {
                    const hs_u32x4 x = hs_conn_words(c, (uint32_t)i, (uint32_t)j, HS_CONN_TAG_SYN);
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

template <typename R>
__global__ void __launch_bounds__(HS_CONN_WAVES * HS_WAVE)
    hs_k_conn_write(HsConnSpec c, const int32_t *__restrict__ bstart, int32_t *__restrict__ post, R *__restrict__ w,
                    int *err)
{
    extern __shared__ int32_t cur[]; /* [waves][D (+ 2P)]: next free position of each bundle of the row */
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
                hs_conn_batch(c, i, jb, k0, done, sel, j);
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
                    if (w) w[pos] = (R)(c.wlo + uw * c.wspan);
                }
            }
        }
    }
}

static int hs_conn_waves(const HsConnSpec &c) { return c.fixed_k > 0 ? 2 : HS_CONN_WAVES; }

static size_t hs_conn_lds(const HsConnSpec &c)
{
    const int P = c.fixed_k > 0 ? hs_conn_pow2(c.fixed_k) : 0;
    return sizeof(int32_t) * (size_t)(c.D + 2 * P) * (size_t)hs_conn_waves(c);
}

static int hs_conn_grid(const HsConnSpec &c)
{
    const long long rows = c.pre_hi - c.pre_lo;
    const long long g = (rows + hs_conn_waves(c) - 1) / hs_conn_waves(c);
    return (int)(g < 65536 ? g : 65536);
}

hipError_t hs_launch_conn_count(const HsConnSpec &c, int32_t *cnt, int *err, hipStream_t s)
{
    if (c.pre_hi <= c.pre_lo) return hipSuccess;
    if (c.D < 1 || c.D > HS_MAX_DELAY_SLOTS || c.fixed_k < 0 || c.fixed_k > HS_CONN_KMAX) return hipErrorInvalidValue;
    hs_k_conn_count<<<hs_conn_grid(c), hs_conn_waves(c) * HS_WAVE, hs_conn_lds(c), s>>>(c, cnt, err);
    return hipGetLastError();
}

hipError_t hs_launch_conn_write(int prec, const HsConnSpec &c, const int32_t *bstart, int32_t *post, void *w, int *err,
                                hipStream_t s)
{
    if (c.pre_hi <= c.pre_lo) return hipSuccess;
    if (c.D < 1 || c.D > HS_MAX_DELAY_SLOTS || c.fixed_k < 0 || c.fixed_k > HS_CONN_KMAX) return hipErrorInvalidValue;
    const size_t lds = hs_conn_lds(c);
    const int block = hs_conn_waves(c) * HS_WAVE;
    if (prec == 64)
        hs_k_conn_write<double><<<hs_conn_grid(c), block, lds, s>>>(c, bstart, post, (double *)w, err);
    else
        hs_k_conn_write<float><<<hs_conn_grid(c), block, lds, s>>>(c, bstart, post, (float *)w, err);
    return hipGetLastError();
}

/* ------------------------------------------------------------------------ */
/* exclusive prefix sum of int32 counts (reduce, scan of block sums, write)  */
/* ------------------------------------------------------------------------ */
#define HS_SCAN_THREADS 256
#define HS_SCAN_ITEMS 16
#define HS_SCAN_TILE (HS_SCAN_THREADS * HS_SCAN_ITEMS)
#define HS_SCAN_TOP 1024

int hs_scan_blocks(long long n) { return (int)((n + HS_SCAN_TILE - 1) / HS_SCAN_TILE); }

__global__ void __launch_bounds__(HS_SCAN_THREADS)
    hs_k_scan_reduce(const int32_t *__restrict__ in, long long n, long long *__restrict__ bsum)
{
    __shared__ long long ws[HS_SCAN_THREADS / 64];
    const long long base = (long long)blockIdx.x * HS_SCAN_TILE;
    long long s = 0;
    for (int k = threadIdx.x; k < HS_SCAN_TILE; k += HS_SCAN_THREADS) {
        const long long idx = base + k;
        if (idx < n) s += in[idx];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long tot = 0;
        for (int q = 0; q < HS_SCAN_THREADS / 64; ++q) tot += ws[q];
        bsum[blockIdx.x] = tot;
    }
}

/* one block: exclusive scan of bsum[0..nblk) in place, total at bsum[nblk] */
__global__ void __launch_bounds__(HS_SCAN_TOP) hs_k_scan_top(long long *bsum, int nblk)
{
    __shared__ long long sh[HS_SCAN_TOP];
    const int t = threadIdx.x;
    long long carry = 0;
    for (int b0 = 0; b0 < nblk; b0 += HS_SCAN_TOP) {
        const int k = b0 + t;
        const long long v = k < nblk ? bsum[k] : 0;
        sh[t] = v;
        __syncthreads();
        for (int o = 1; o < HS_SCAN_TOP; o <<= 1) {
            const long long a = t >= o ? sh[t - o] : 0;
            __syncthreads();
            sh[t] += a;
            __syncthreads();
        }
        if (k < nblk) bsum[k] = carry + sh[t] - v;
        carry += sh[HS_SCAN_TOP - 1];
        __syncthreads(); /* every thread has read sh before the next chunk writes it */
    }
    if (t == 0) bsum[nblk] = carry;
}

__global__ void __launch_bounds__(HS_SCAN_THREADS)
    hs_k_scan_write(const int32_t *__restrict__ in, long long n, const long long *__restrict__ bsum, int nblk,
                    int32_t *__restrict__ out)
{
    __shared__ long long ws[HS_SCAN_THREADS / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const long long base = (long long)blockIdx.x * HS_SCAN_TILE + (long long)t * HS_SCAN_ITEMS;
    int32_t v[HS_SCAN_ITEMS];
    long long s = 0;
#pragma unroll
    for (int q = 0; q < HS_SCAN_ITEMS; ++q) {
        v[q] = base + q < n ? in[base + q] : 0;
        s += v[q];
    }
    long long inc = s;
    for (int o = 1; o < 64; o <<= 1) {
        const long long y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    long long pre = bsum[blockIdx.x] + inc - s;
    for (int q = 0; q < wave; ++q) pre += ws[q];
#pragma unroll
    for (int q = 0; q < HS_SCAN_ITEMS; ++q) {
        if (base + q < n) out[base + q] = (int32_t)pre;
        pre += v[q];
    }
    if (blockIdx.x == 0 && t == 0) out[n] = (int32_t)bsum[nblk];
}

hipError_t hs_launch_scan(const int32_t *in, long long n, long long *bsum, int32_t *out, hipStream_t s)
{
    const int nblk = hs_scan_blocks(n);
    if (nblk > 0) hs_k_scan_reduce<<<nblk, HS_SCAN_THREADS, 0, s>>>(in, n, bsum);
    hs_k_scan_top<<<1, HS_SCAN_TOP, 0, s>>>(bsum, nblk);
    if (nblk > 0) hs_k_scan_write<<<nblk, HS_SCAN_THREADS, 0, s>>>(in, n, bsum, nblk, out);
    else hs_k_scan_write<<<1, HS_SCAN_THREADS, 0, s>>>(in, 0, bsum, 0, out);
    return hipGetLastError();
}
