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
 * Row i (one thread per presynaptic neuron) selects targets in increasing j with
 * the reference's two algorithms (synapses_create_generator.cu:315-333):
 *   p <  0.25: j starts at post_lo; each draw r in [0, 1) advances j by
 *              floor(log(r) / log(1 - p)) and selects it (r == 0 ends the row;
 *              a skip past post_hi ends it too)
 *   p >= 0.25: every j is tested with its own draw and selected when r < p
 * The draws are Philox words keyed by (seed, stream, i, k) (layout in
 * hs_internal.h), so a row's synapses do not depend on launch geometry or on the
 * other rows, and the oracle regenerates any subset of rows.
 *
 * Three passes:
 *   count: synapses per (row, delay slot), accumulated per thread in LDS
 *          (D <= 128 slots) or in the row's own range of the global table
 *   scan : exclusive prefix sum of the n_src*D counts -> bstart (int64 block sums)
 *   write: the rows again; synapse (i, j) goes to bstart[i*D + slot] + its rank
 *          among the row's earlier synapses of that slot, so each bundle keeps
 *          creation (ascending j) order
 * Bounds: the delay quantisation is monotone in the uniform, so every slot lies
 * in [0, D) with D taken from the quantised bounds; both passes still check it,
 * and the write pass checks each position against the end of its bundle, so a
 * disagreement between the passes raises an error flag instead of writing
 * outside the arrays.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_internal.h"
#include "hs_philox.h"

#define HS_CONN_BLOCK 64 /* one wave: thread t owns row blockIdx.x*64 + t and LDS column t */

__device__ __forceinline__ hs_u32x4 hs_conn_words(const HsConnSpec &c, uint32_t i, uint32_t k, uint32_t tag)
{
    hs_u32x4 ctr;
    ctr.v[0] = i;
    ctr.v[1] = k;
    ctr.v[2] = HS_CONN_CTR | tag;
    ctr.v[3] = c.stream;
    return hs_philox4x32_10(ctr, (uint32_t)c.seed, (uint32_t)(c.seed >> 32));
}

/* f(j) for every target selected in row i, in increasing j */
template <class F>
__device__ __forceinline__ void hs_conn_row(const HsConnSpec &c, int i, F &f)
{
    const long long hi = c.post_hi;
    hs_u32x4 wd = {{0u, 0u, 0u, 0u}};
    if (c.jump) {
        uint32_t k = 0;
        for (long long j = c.post_lo; j < hi; ++j, ++k) {
            if (!(k & 1u)) wd = hs_conn_words(c, (uint32_t)i, k >> 1, HS_CONN_TAG_SKIP);
            const double r = (k & 1u) ? hs_u53(wd.v[2], wd.v[3]) : hs_u53(wd.v[0], wd.v[1]);
            if (r == 0.0) break;
            const double skip = floor(hs_log(r) * c.pconst);
            if (skip >= (double)(hi - j)) break;
            j += (long long)skip;
            if (c.autapses || j != i) f((int)j);
        }
    } else {
        for (long long j = c.post_lo; j < hi; ++j) {
            const uint32_t m = (uint32_t)(j - c.post_lo);
            if (!(m & 1u)) wd = hs_conn_words(c, (uint32_t)i, m >> 1, HS_CONN_TAG_PAIR);
            const double r = (m & 1u) ? hs_u53(wd.v[2], wd.v[3]) : hs_u53(wd.v[0], wd.v[1]);
            if (r < c.p && (c.autapses || j != i)) f((int)j);
        }
    }
}

/* delay slot of a synapse from its delay uniform: (int)(d/dt + 0.5) - dmin,
   the reference's quantisation (synapses_push_spikes.cu:375) */
__device__ __forceinline__ int hs_conn_slot(const HsConnSpec &c, double ud)
{
    return (int)((c.dlo + ud * c.dspan) / c.dt + 0.5) - c.dmin;
}

template <bool LDS>
__global__ void __launch_bounds__(HS_CONN_BLOCK) hs_k_conn_count(HsConnSpec c, int32_t *__restrict__ cnt, int *err)
{
    extern __shared__ int32_t tab[]; /* [D][64] */
    const int t = threadIdx.x;
    const int i = c.pre_lo + (int)blockIdx.x * HS_CONN_BLOCK + t;
    if (i >= c.pre_hi) return; /* no barrier below: each thread uses its own column only */
    const size_t row = (size_t)i * (size_t)c.D;
    if (LDS)
        for (int d = 0; d < c.D; ++d) tab[d * HS_CONN_BLOCK + t] = 0;
    auto f = [&](int j) {
        double ud = 0.0;
        if (c.dspan != 0.0) {
            const hs_u32x4 x = hs_conn_words(c, (uint32_t)i, (uint32_t)j, HS_CONN_TAG_SYN);
            ud = hs_u53(x.v[0], x.v[1]);
        }
        const int s = hs_conn_slot(c, ud);
        if ((unsigned)s >= (unsigned)c.D) {
            *err = 1;
            return;
        }
        if (LDS) tab[s * HS_CONN_BLOCK + t] += 1;
        else cnt[row + s] += 1;
    };
    hs_conn_row(c, i, f);
    if (LDS)
        for (int d = 0; d < c.D; ++d) cnt[row + d] = tab[d * HS_CONN_BLOCK + t];
}

template <typename R, bool LDS>
__global__ void __launch_bounds__(HS_CONN_BLOCK)
    hs_k_conn_write(HsConnSpec c, const int32_t *__restrict__ bstart, int32_t *__restrict__ cursor,
                    int32_t *__restrict__ post, R *__restrict__ w, int *err)
{
    extern __shared__ int32_t cur[]; /* [D][64] */
    const int t = threadIdx.x;
    const int i = c.pre_lo + (int)blockIdx.x * HS_CONN_BLOCK + t;
    if (i >= c.pre_hi) return;
    const size_t row = (size_t)i * (size_t)c.D;
    if (LDS)
        for (int d = 0; d < c.D; ++d) cur[d * HS_CONN_BLOCK + t] = bstart[row + d];
    const bool draw = c.dspan != 0.0 || (w && c.wspan != 0.0);
    auto f = [&](int j) {
        double ud = 0.0, uw = 0.0;
        if (draw) {
            const hs_u32x4 x = hs_conn_words(c, (uint32_t)i, (uint32_t)j, HS_CONN_TAG_SYN);
            ud = hs_u53(x.v[0], x.v[1]);
            uw = hs_u53(x.v[2], x.v[3]);
        }
        const int s = hs_conn_slot(c, ud);
        if ((unsigned)s >= (unsigned)c.D) {
            *err = 1;
            return;
        }
        const int pos = LDS ? cur[s * HS_CONN_BLOCK + t]++ : cursor[row + s]++;
        if (pos >= bstart[row + s + 1]) {
            *err = 2;
            return;
        }
        post[pos] = j;
        if (w) w[pos] = (R)(c.wlo + uw * c.wspan);
    };
    hs_conn_row(c, i, f);
}

hipError_t hs_launch_conn_count(const HsConnSpec &c, int32_t *cnt, int *err, hipStream_t s)
{
    const int rows = c.pre_hi - c.pre_lo;
    if (rows <= 0) return hipSuccess;
    const int grid = (rows + HS_CONN_BLOCK - 1) / HS_CONN_BLOCK;
    if (c.D <= HS_CONN_LDS_SLOTS)
        hs_k_conn_count<true><<<grid, HS_CONN_BLOCK, sizeof(int32_t) * c.D * HS_CONN_BLOCK, s>>>(c, cnt, err);
    else
        hs_k_conn_count<false><<<grid, HS_CONN_BLOCK, 0, s>>>(c, cnt, err);
    return hipGetLastError();
}

template <typename R>
static hipError_t hs_conn_write_t(const HsConnSpec &c, const int32_t *bstart, int32_t *cursor, int32_t *post, R *w,
                                  int *err, hipStream_t s)
{
    const int rows = c.pre_hi - c.pre_lo;
    if (rows <= 0) return hipSuccess;
    const int grid = (rows + HS_CONN_BLOCK - 1) / HS_CONN_BLOCK;
    if (c.D <= HS_CONN_LDS_SLOTS)
        hs_k_conn_write<R, true><<<grid, HS_CONN_BLOCK, sizeof(int32_t) * c.D * HS_CONN_BLOCK, s>>>(c, bstart, cursor,
                                                                                                   post, w, err);
    else
        hs_k_conn_write<R, false><<<grid, HS_CONN_BLOCK, 0, s>>>(c, bstart, cursor, post, w, err);
    return hipGetLastError();
}

hipError_t hs_launch_conn_write(int prec, const HsConnSpec &c, const int32_t *bstart, int32_t *cursor,
                                int32_t *post, void *w, int *err, hipStream_t s)
{
    if (prec == 64) return hs_conn_write_t<double>(c, bstart, cursor, post, (double *)w, err, s);
    return hs_conn_write_t<float>(c, bstart, cursor, post, (float *)w, err, s);
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
