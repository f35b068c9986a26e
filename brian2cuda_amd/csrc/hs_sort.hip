/*
 * hs_sort.hip -- the bundle table of a pathway built on the device from
 * synapse arrays already in HBM (hs_net_add_pathway_device).
 *
 * Replaces the host-serial bundle construction of _before_run_<push_spikes>
 * (brian2cuda/templates/synapses_push_spikes.cu:171-993: per presynaptic
 * neuron a thrust::sort_by_key of the delays on the host, :460-464, then one
 * cudaMemcpy per bundle, :537-539) with a device sort by (pre, delay):
 *   check : indices inside the groups, delays >= 0; smallest and largest delay
 *   keys  : key(s) = pre(s) * D + delay(s) - dmin, value = s
 *   sort  : stable LSD radix sort of (key, s) pairs (rocPRIM through hipCUB), so
 *           each (pre, delay) bundle keeps synapse (creation) order
 *   gather: post and weight of each sorted synapse
 *   bstart: bstart[q] = first position whose key is >= q, for q in [0, n_src*D]
 * The result is the same table the host counting sort of hs_net_add_pathway
 * builds from the same arrays.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "hs_device.h"
#include "hs_internal.h"

#define HS_SORT_BLOCK 256

/* err bits: 1 pre out of range, 2 post out of range, 4 negative delay */
__global__ void __launch_bounds__(HS_SORT_BLOCK)
    hs_k_syn_check(const int32_t *__restrict__ pre, const int32_t *__restrict__ post,
                   const int32_t *__restrict__ delay, long long n, int n_src, int n_tgt, int *minmax, int *err)
{
    int lo = 0x7fffffff, hi = -1, bad = 0;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (long long)gridDim.x * blockDim.x) {
        const int a = pre[s], b = post[s];
        if (a < 0 || a >= n_src) bad |= 1;
        if (b < 0 || b >= n_tgt) bad |= 2;
        if (delay) {
            const int d = delay[s];
            if (d < 0) bad |= 4;
            lo = min(lo, d);
            hi = max(hi, d);
        }
    }
#pragma unroll
    for (int o = HS_WAVE / 2; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, HS_WAVE));
        hi = max(hi, __shfl_xor(hi, o, HS_WAVE));
        bad |= __shfl_xor(bad, o, HS_WAVE);
    }
    if (hs_lane() == 0) {
        if (delay && hi >= 0) {
            atomicMin(&minmax[0], lo);
            atomicMax(&minmax[1], hi);
        }
        if (bad) atomicOr(err, bad);
    }
}

__global__ void __launch_bounds__(HS_SORT_BLOCK)
    hs_k_syn_keys(const int32_t *__restrict__ pre, const int32_t *__restrict__ delay, long long n, int D, int dmin,
                  uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (long long)gridDim.x * blockDim.x) {
        const int d = delay ? delay[s] - dmin : 0;
        keys[s] = (uint32_t)pre[s] * (uint32_t)D + (uint32_t)d;
        vals[s] = (int32_t)s;
    }
}

template <typename R>
__global__ void __launch_bounds__(HS_SORT_BLOCK)
    hs_k_syn_gather(const int32_t *__restrict__ order, long long n, const int32_t *__restrict__ post,
                    const double *__restrict__ w, int32_t *__restrict__ post_out, R *__restrict__ w_out)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const int s = order[k];
        post_out[k] = post[s];
        if (w_out) w_out[k] = (R)w[s];
    }
}

template <typename R>
__global__ void __launch_bounds__(HS_SORT_BLOCK)
    hs_k_syn_convert(const double *__restrict__ w, long long n, R *__restrict__ out)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        out[k] = (R)w[k];
}

/* position k takes the keys q in (keys[k-1], keys[k]] (q <= keys[0] for k = 0,
   q <= nb for k = n): every q in [0, nb] is written once */
__global__ void __launch_bounds__(HS_SORT_BLOCK)
    hs_k_syn_bstart(const uint32_t *__restrict__ keys, long long n, long long nb, int32_t *__restrict__ bstart)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k <= n;
         k += (long long)gridDim.x * blockDim.x) {
        const long long lo = k == 0 ? 0 : (long long)keys[k - 1] + 1;
        const long long hi = k == n ? nb : (long long)keys[k];
        for (long long q = lo; q <= hi; ++q) bstart[q] = (int32_t)k;
    }
}

static int hs_sort_grid(long long n)
{
    const long long g = (n + HS_SORT_BLOCK - 1) / HS_SORT_BLOCK;
    return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

hipError_t hs_launch_syn_check(const int32_t *pre, const int32_t *post, const int32_t *delay, long long n, int n_src,
                               int n_tgt, int *minmax, int *err, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hs_k_syn_check<<<hs_sort_grid(n), HS_SORT_BLOCK, 0, s>>>(pre, post, delay, n, n_src, n_tgt, minmax, err);
    return hipGetLastError();
}

hipError_t hs_launch_syn_keys(const int32_t *pre, const int32_t *delay, long long n, int D, int dmin, uint32_t *keys,
                              int32_t *vals, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hs_k_syn_keys<<<hs_sort_grid(n), HS_SORT_BLOCK, 0, s>>>(pre, delay, n, D, dmin, keys, vals);
    return hipGetLastError();
}

hipError_t hs_radix_sort_pairs(void *tmp, size_t *tmp_bytes, const uint32_t *keys, uint32_t *keys_out,
                               const int32_t *vals, int32_t *vals_out, long long n, int bits, hipStream_t s)
{
    return hipcub::DeviceRadixSort::SortPairs(tmp, *tmp_bytes, keys, keys_out, vals, vals_out, (int)n, 0, bits, s);
}

hipError_t hs_launch_syn_gather(int prec, const int32_t *order, long long n, const int32_t *post, const double *w,
                                int32_t *post_out, void *w_out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (prec == 64)
        hs_k_syn_gather<double><<<hs_sort_grid(n), HS_SORT_BLOCK, 0, s>>>(order, n, post, w, post_out,
                                                                          (double *)w_out);
    else
        hs_k_syn_gather<float><<<hs_sort_grid(n), HS_SORT_BLOCK, 0, s>>>(order, n, post, w, post_out, (float *)w_out);
    return hipGetLastError();
}

hipError_t hs_launch_syn_convert(int prec, const double *w, long long n, void *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (prec == 64) hs_k_syn_convert<double><<<hs_sort_grid(n), HS_SORT_BLOCK, 0, s>>>(w, n, (double *)out);
    else hs_k_syn_convert<float><<<hs_sort_grid(n), HS_SORT_BLOCK, 0, s>>>(w, n, (float *)out);
    return hipGetLastError();
}

hipError_t hs_launch_syn_bstart(const uint32_t *keys, long long n, long long nb, int32_t *bstart, hipStream_t s)
{
    hs_k_syn_bstart<<<hs_sort_grid(n + 1), HS_SORT_BLOCK, 0, s>>>(keys, n, nb, bstart);
    return hipGetLastError();
}
