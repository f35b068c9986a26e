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

#include "hs_conndev.h"

/* the library's kernels: connect(p) and fixed out-degree, no condition */
__global__ void __launch_bounds__(HS_CONN_WAVES * HS_WAVE) hs_k_conn_count(HsConnSpec c, int32_t *__restrict__ cnt,
                                                                           int *err)
{
    extern __shared__ int32_t tab[];
    const HsConnVars a = {};
    hs_conn_count_body<HsConnPlain, float>(c, a, cnt, err, tab);
}

template <typename R>
__global__ void __launch_bounds__(HS_CONN_WAVES * HS_WAVE)
    hs_k_conn_write(HsConnSpec c, const int32_t *__restrict__ bstart, int32_t *__restrict__ post, R *__restrict__ w,
                    int *err)
{
    extern __shared__ int32_t cur[];
    const HsConnVars a = {};
    hs_conn_write_body<HsConnPlain, R>(c, a, bstart, post, w, err, cur);
}

int hs_conn_waves(const HsConnSpec &c) { return c.fixed_k > 0 ? 2 : HS_CONN_WAVES; }

size_t hs_conn_lds(const HsConnSpec &c)
{
    const int P = c.fixed_k > 0 ? hs_conn_pow2(c.fixed_k) : 0;
    return sizeof(int32_t) * (size_t)(c.D + 2 * P) * (size_t)hs_conn_waves(c);
}

int hs_conn_grid(const HsConnSpec &c)
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
