/*
 * hs.h -- C ABI of libhs, the MI355X (gfx950) runtime behind the
 * `hip_standalone` Brian2 device.  Plain C types only: pointers, sizes, ints.
 *
 * Conventions
 *   - every entry point returns HS_OK (0) or a negative HS_ERR_* code; the text
 *     of the last error of the calling thread is returned by hs_last_error().
 *     Nothing calls exit() (the reference's CUDA_SAFE_CALL does:
 *     brian2cuda/brianlib/cuda_utils.h:20,77-104).
 *   - "device pointer" arguments are HBM addresses from hs_alloc (or any HIP
 *     allocation on the current device); "host" arguments are ordinary memory
 *     owned by the caller.  The runtime owns the memory behind an hs_net.
 *   - calls on one hs_net must come from one thread at a time.
 *   - all work runs on one HIP stream per process (default priority); hs_sync()
 *     waits for it.
 */
#ifndef HS_H
#define HS_H

#include <stddef.h>
#include <stdint.h>
#include "hs_models.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HS_OK 0
#define HS_ERR_ARG (-1)         /* bad argument (index out of range, size, mode) */
#define HS_ERR_HIP (-2)         /* a HIP runtime call failed */
#define HS_ERR_NOMEM (-3)       /* device or host allocation failed */
#define HS_ERR_STATE (-4)       /* call not valid in the object's current state */
#define HS_ERR_OVERFLOW (-5)    /* a monitor buffer was too small; see hs_net_mon_count */
#define HS_ERR_UNSUPPORTED (-6) /* model / form combination not provided */

/* ---- runtime ------------------------------------------------------------ */

/* Text of the last error on this thread ("" if none). */
const char *hs_last_error(void);

/* ABI version, bumped on incompatible change. */
int hs_version(void);

/* Select and initialise a device; creates the runtime's stream.
 * Replaces cudaSetDevice + heap-limit setup (brian2cuda/templates/main.cu:71-85). */
int hs_init(int device);

/* Device name, compute units and HBM bytes of the initialised device. */
int hs_device_info(char *name, int name_cap, int *compute_units, long long *hbm_bytes);

/* Memory: replaces the cudaMalloc/cudaMemcpy/cudaFree calls of _init_arrays /
 * _dealloc_arrays (brian2cuda/templates/objects.cu:340-489,654-726). */
int hs_alloc(size_t bytes, void **dev_ptr);
int hs_free(void *dev_ptr);
int hs_h2d(void *dev_dst, const void *host_src, size_t bytes);
int hs_d2h(void *host_dst, const void *dev_src, size_t bytes);
int hs_zero(void *dev_ptr, size_t bytes);
int hs_sync(void);

/* ---- individually testable kernels (device pointers) ---------------------- */

/*
 * Threshold + spikespace compaction over n neurons of state v (double if
 * precision == HS_F64, else float): neuron i spikes iff v[i] > vth and
 * (step - last[i]) >= ref_steps.  Spiking indices are written to spikes[0..*count)
 * and last[i] = step for them.  sorted != 0 gives ascending order (the order
 * cpp_standalone produces); sorted == 0 gives ascending order within each
 * workgroup tile, tiles in completion order.
 * Replaces threshold.cu:19-45 (+ the eventspace reset kernel :49-75 and the
 * host memset :99-105: the count is written, no O(N) -1 fill).
 * count: device int32[1].
 */
int hs_threshold_compact(int precision, const void *v, int n, double vth, int32_t *last,
                         int64_t step, int ref_steps, int sorted, int32_t *spikes, int32_t *count);

/*
 * hs_threshold_compact without any allocation: the same kernels and results,
 * with the sorted form's per-tile counts in a static device array of the
 * library (512 KiB per GPU, held until the process ends) instead of memory
 * the library allocates.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  n above HS_NOALLOC_MAX_N
 * returns HS_ERR_ARG.  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (from several threads) race.
 */
#define HS_NOALLOC_MAX_N (1 << 25) /* 33,554,432 neurons */
int hs_threshold_compact_noalloc(int precision, const void *v, int n, double vth, int32_t *last,
                                 int64_t step, int ref_steps, int sorted, int32_t *spikes, int32_t *count);

/*
 * n standard normals out[i] = N(seed, stream, i, step) (Philox4x32-10 +
 * Box-Muller, the deterministic math of the numerical contract); out is double
 * or float by precision (float: the FP32 form's draw, 24-bit uniforms and float math).
 * Replaces the cuRAND buffer refill (rand.cu:395-463, curand_buffer.h:20-160).
 */
int hs_randn(uint64_t seed, uint32_t stream, int64_t step, int n, int precision, void *out);

/* Uniform [0,1) draws with the same counter layout (first 53-bit uniform). */
int hs_rand(uint64_t seed, uint32_t stream, int64_t step, int n, void *out_f64);

/*
 * One state-update step of `model` for n neurons, in place: vars[v] are device
 * pointers of the model's state variables (host array of nvars pointers),
 * last the int32 last-spike steps (refractoriness), params the model's
 * parameter list (hs_models.h).  No threshold, no reset.
 * Replaces the generated _run_kernel_<stateupdater> (common_group.cu:108-160).
 */
int hs_stateupdate(int model, int precision, int n, void *const *vars, const int32_t *last,
                   const double *params, int nparams, double dt, uint64_t seed, int64_t step);

/*
 * Synaptic effect application for one pathway at step t: for every delay
 * d in [dmin, dmin+D) and every spike k of the spike list emitted at step t-d
 * (hist[((t-d) mod Q)*n_src + k], count hist_cnt[(t-d) mod Q]), every synapse of
 * the bundle (pre, d) -- post[bstart[pre*D+d-dmin] .. bstart[pre*D+d-dmin+1]) --
 * adds its weight to target[post] with a hardware atomic.  events (device
 * uint64[1]) is incremented by the number of synapses delivered.
 * wmode HS_W_SCALAR uses wscalar, HS_W_SYNAPSE w[synapse], HS_W_PRE w[pre].
 * Replaces _run_kernel_<synapses_pre> (synapses.cu:12-164) together with the
 * queue pop (spikequeue.h:504-516).
 */
int hs_deliver(int precision, int n_src, int dmin, int D, int Q, const int32_t *hist,
               const int32_t *hist_cnt, int64_t step, const int32_t *bstart, const int32_t *post,
               int wmode, const void *w, double wscalar, void *target, uint64_t *events);

/*
 * Effect application of a flat event list: target[idx[k]] += w for k < n
 * (device arrays; target holds n_target values of the given precision), with
 * hardware atomics and no spike list or bundle lookup.  It measures what random
 * atomics alone sustain; hs_deliver and the network kernels add the lookups.
 * An index outside [0, n_target) is skipped, and the call then returns
 * HS_ERR_ARG.  kernel_ms (nullable) receives the kernel's device time.
 * Replaces the atomic effect statement of _run_kernel_<synapses_pre> when it
 * runs with atomics (synapses.cu:202-215; _brian_atomicAdd,
 * cuda_generator.py:33-150).
 */
int hs_scatter_add(int precision, int64_t n, const int32_t *idx, int n_target, double w, void *target,
                   double *kernel_ms);

/*
 * Device evaluation of the deterministic math (for bit-equality checks):
 * fn 0 exp, 1 log, 2 sqrt, 3 cos(2 pi x), 4 expf, 5 logf, 6 sqrtf, 7 cos(2 pi x)
 * in float (x and y read/written as double).
 */
int hs_math_eval(int fn, int n, const double *x, double *y);

/* ---- networks: the per-timestep engine ------------------------------------ */

typedef struct hs_net hs_net;

/*
 * Create an empty network.  precision HS_F64 / HS_F32; engine HS_ENGINE_PLAIN
 * (state updated in place by the effects, Brian2's order, two launches per
 * step) or HS_ENGINE_FUSED (effects accumulated and applied at the start of the
 * next step; one launch per step when every delay is >= 1 step).
 * Replaces the generated objects.cu/network.cu pair (objects.cu:340-489,
 * network.cu:40-137).
 */
int hs_net_create(int precision, int engine, double dt, uint64_t seed, hs_net **out);
void hs_net_destroy(hs_net *net);

/* Add a NeuronGroup of n neurons with a model of hs_models.h. */
int hs_net_add_group(hs_net *net, int model, int n, const double *params, int nparams, int *gid);

/*
 * Add a NeuronGroup of n neurons whose model is given as code: what the
 * reference's code generator emits for any NeuronGroup's equations, threshold
 * and reset (cuda_generator.py:200-808 into the stateupdate.cu,
 * common_group.cu:108-160, threshold.cu:19-45 and reset.cu:3-27 templates).
 * The group has nvars state variables named var_names[0..nvars-1] (variable v
 * of hs_net_set_state / hs_net_get_state / monitors is var_names[v]).  Per
 * neuron and step, in the network's precision R (double or float), with the
 * variables as local R values of those names and these in scope:
 *   dt (R, seconds), t (double, the step's time), i (int, the neuron index),
 *   xi (R, a standard normal draw of (seed, stream, i, step) when noise != 0,
 *   else 0), not_refractory (int: the neuron's last spike is at least
 *   `refractory` seconds back, Brian2's timestep rule)
 *   update     C statements advancing the variables by one step
 *   threshold  C expression evaluated after update: the neuron spikes when it
 *              holds and not_refractory (Brian2 adds 'and not_refractory'
 *              when a refractory period is set); "" or NULL: never
 *   reset      C statements run for a neuron that spiked (t: the spike's
 *              time), after the effects of its spike step were applied, as
 *              Brian2's resetter runs after the synapses; "" or NULL: none
 * The code may call exp, log, log10, sqrt, abs/fabs, floor, ceil, sign and
 * clip(x, lo, hi) (the library's deterministic exp and log, IEEE sqrt: the same
 * numbers as the host oracle) and use the type R; no preprocessor lines.  The
 * engine kernels are compiled with the network's code models (hipRTC) on its
 * first run, once per distinct program in the process; a compile error fails
 * that run with HS_ERR_ARG and the compiler's messages.  At most
 * HS_MAX_CODE_MODELS code models per network; the group's model id is
 * HS_MODEL_CODE + (its index among them).
 */
int hs_net_add_group_code(hs_net *net, int n, int nvars, const char *const *var_names, const char *update,
                          const char *threshold, const char *reset, double refractory, int noise, uint32_t stream,
                          int *gid);

/* Set / get a state variable (host arrays of n doubles). */
int hs_net_set_state(hs_net *net, int gid, int var, const double *values);
int hs_net_get_state(hs_net *net, int gid, int var, double *values);
/* last spike step per neuron (HS_NEVER_SPIKED if none), host int32[n] */
int hs_net_get_last_spike(hs_net *net, int gid, int32_t *steps);

/*
 * Add a pathway (Synapses ... on_pre='var += w') from group src to group tgt.
 * pre/post: host int32[nsyn] (indices within the groups); delay_steps: host
 * int32[nsyn] or NULL for the scalar delay scalar_delay (delays are already
 * quantised: (int)(d/dt+0.5), synapses_push_spikes.cu:375).  The bundle table
 * (synapses grouped by (pre, delay), synapse order kept within a bundle) is
 * built here by a host counting sort and copied once.
 * Replaces _before_run_<push_spikes> (synapses_push_spikes.cu:171-993).
 */
int hs_net_add_pathway(hs_net *net, int src, int tgt, int var, int64_t nsyn, const int32_t *pre,
                       const int32_t *post, const int32_t *delay_steps, int scalar_delay, int wmode,
                       const double *weights, double wscalar, int *pid);

/*
 * Add a pathway with random connectivity generated on the device
 * (Synapses.connect(p=p) followed by delay = 'delay_lo + (delay_hi-delay_lo)*rand()').
 * Every source neuron i in [pre_lo, pre_hi) connects to each target j in
 * [post_lo, post_hi) with probability p: geometric skips for p < 0.25, one draw
 * per pair otherwise -- the two algorithms of the reference's generator
 * (synapses_create_generator.cu:315-333).  autapses = 0 drops j == i when src ==
 * tgt (connect(condition='i != j', p=p)).  Draws: Philox-4x32-10 keyed by
 * (seed, stream, i, draw), so the synapses are a pure function of the arguments.
 * Delays in seconds, quantised (int)(d/dt + 0.5) (synapses_push_spikes.cu:375);
 * the bundle table spans the quantised bounds [q(delay_lo), q(delay_hi)].
 * wmode HS_W_SCALAR: weight w_lo; HS_W_SYNAPSE: per-synapse uniform in
 * [w_lo, w_hi) (w_lo == w_hi: a constant array).  Bundles are written on the
 * device (scatter delivery only): no synapse array crosses PCIe.
 * Replaces the host generation + copies of synapses_create_generator.cu:164-465
 * and the bundle construction of _before_run_<push_spikes> (synapses_push_spikes.cu:171-993).
 */
int hs_net_connect_random(hs_net *net, int src, int tgt, int var, int pre_lo, int pre_hi, int post_lo, int post_hi,
                          double p, int autapses, uint64_t seed, uint32_t stream, double delay_lo, double delay_hi,
                          int wmode, double w_lo, double w_hi, int *pid);
/*
 * Fixed out-degree on the device (Synapses.connect(j='k for k in sample(N_post,
 * size=K)')): every source neuron i in [pre_lo, pre_hi) samples K distinct
 * targets of [post_lo, post_hi), uniformly over the K-subsets; autapses = 0 then
 * drops a sampled j == i when src == tgt (the generator's 'if k != i', so such
 * a row keeps K - 1).  K larger than post_hi - post_lo is an error
 * (HS_ERR_ARG, the reference's "sample size bigger than the population size"),
 * K = 0 creates nothing; K <= 2048.  Draws, delays, weights and the bundle
 * table as hs_net_connect_random.
 * Replaces the fixed-size sampling of synapses_create_generator.cu:250-312
 * (selection sampling or a tracked set, both uniform over the K-subsets).
 */
int hs_net_connect_fixed(hs_net *net, int src, int tgt, int var, int pre_lo, int pre_hi, int post_lo, int post_hi,
                         int k, int autapses, uint64_t seed, uint32_t stream, double delay_lo, double delay_hi,
                         int wmode, double w_lo, double w_hi, int *pid);
/*
 * Connectivity with a condition and a per-pair probability given as code
 * (Synapses.connect(condition='...', p=...), connect(condition='...',
 * p='<expression of i, j, x_pre, x_post>'), and the generator form
 * connect(j='k for k in sample(N_post, size=K) if <condition>')), generated on
 * the device by the passes of hs_net_connect_random / _fixed compiled at this
 * call (hipRTC) with:
 *   condition  a C expression (NULL or "": none) over i, j (long long, the
 *              indices within [pre_lo, pre_hi) and [post_lo, post_hi), as
 *              Brian2's i and j of a Synapses on those subgroups), N_pre,
 *              N_post (the range sizes), hs_pre(k) / hs_post(k) (variable k of
 *              the source / target neuron as double, read when this is
 *              called: set the state first) and the functions exp, log,
 *              log10, sqrt, abs, floor, ceil, sign, clip, hs_cg_floordiv,
 *              hs_cg_pymod, hs_cg_pow (Python's //, % and **);
 *   p_expr     the pair's probability as such an expression (NULL or "": the
 *              constant p);
 *   k          > 0: fixed out-degree K (then p and p_expr are unused);
 *   n          synapses per selected pair, 1 to 16 (connect(n=...)): copy c of
 *              pair (i, j) draws its own delay and weight (counter tag
 *              1 | c << 4; copy 0 as a single synapse), the copies consecutive
 *              in creation order;
 *   w_expr     NULL or "": weights uniform in [w_lo, w_hi] (wmode
 *              HS_W_SYNAPSE); else each synapse's weight as an expression of
 *              the same kind (S.w = '...' after connect), evaluated once per
 *              synapse; its rand() / randn() draws also key on the copy.
 * Selection is that of the two entry points: with a constant p, candidates
 * are drawn by geometric skips (p < 0.25) or one draw per pair and then kept
 * where the condition holds ('k for k in sample(N_post, p=p) if condition',
 * synapses_create_generator.cu:213-333); with p_expr each pair (i, j) is kept
 * where the condition holds and its draw is below p_expr; with K the K
 * sampled targets are filtered by the condition.  Delays, weights and the
 * bundle table as hs_net_connect_random.  An expression that does not compile
 * fails with HS_ERR_ARG and the compiler's log (hs_last_error).
 * Replaces the generated host loop of synapses_create_generator.cu:164-465 for
 * conditions and probabilities that are expressions.
 */
int hs_net_connect_code(hs_net *net, int src, int tgt, int var, int pre_lo, int pre_hi, int post_lo, int post_hi,
                        const char *condition, const char *p_expr, double p, int k, int n, int autapses,
                        uint64_t seed, uint32_t stream, double delay_lo, double delay_hi, int wmode, double w_lo,
                        double w_hi, const char *w_expr, int *pid);
/*
 * hs_net_add_pathway with the bundle table built on the device: pre, post,
 * delay_steps (or NULL: scalar_delay) and weights are DEVICE pointers (hs_alloc
 * or any allocation on the device) holding int32 / double arrays of the same
 * meaning; they are only read.  Indices and delays are checked on the device,
 * then the synapses are stably radix-sorted by (pre, delay) there, so the table
 * equals the host-built one for the same arrays.  Scatter delivery only.
 * Replaces the host sort and per-bundle copies of _before_run_<push_spikes>
 * (synapses_push_spikes.cu:441-575).
 */
int hs_net_add_pathway_device(hs_net *net, int src, int tgt, int var, int64_t nsyn, const int32_t *pre,
                              const int32_t *post, const int32_t *delay_steps, int scalar_delay, int wmode,
                              const double *weights, double wscalar, int *pid);
/* synapse count, first delay slot (steps), delay slots and delivery mode of a pathway */
int hs_net_pathway_info(hs_net *net, int pid, int64_t *nsyn, int *dmin, int *D, int *dmode);
/*
 * The synapses of a scatter-mode pathway, host arrays of cap >= nsyn entries (any
 * may be NULL), in bundle order: by (pre, delay), creation order within a bundle.
 * Replaces the write-out of the Synapses' dynamic arrays (_synaptic_pre,
 * _synaptic_post, delay, w) in _write_arrays (brian2cuda/templates/objects.cu:550-563).
 */
int hs_net_get_synapses(hs_net *net, int pid, int32_t *pre, int32_t *post, int32_t *delay_steps, double *w,
                        int64_t cap);

/*
 * SpikeGeneratorGroup pattern of a HS_MODEL_SPIKEGEN group: neuron indices[k]
 * fires at step steps[k] (host arrays of n; a (neuron, step) pair at most once),
 * repeated every period steps if the group's period parameter is > 0.  May be
 * replaced between runs.  Replaces the generator's arrays and its per-step kernel
 * (brian2cuda/templates/spikegenerator.cu:43-80): the update launch tests each
 * neuron against the step's sorted list.
 */
int hs_net_set_spikegen(hs_net *net, int gid, int64_t n, const int32_t *indices, const int64_t *steps);
/*
 * Make a pathway plastic: spike-timing-dependent plasticity with event-driven
 * traces (hs_models.h, HS_STDP_NPARAMS parameters), the synapse model of
 * brian2cuda/tests/test_cpp_cuda_consistency.py:45-61.  Needs a scatter pathway
 * with per-synapse weights; before the first run.  The pathway then runs its
 * on_pre code (target += w*scale, trace and weight update) and an on_post
 * pathway over the target group's spikes, after the step's other deliveries
 * (Brian2 orders the pre pathway before the post pathway).  In the fused form
 * the network then advances one step per update launch and the on_pre effects
 * go to the step's accumulator, as every other effect of that form does.
 * Replaces the pre- and post-pathway effect kernels of synapses.cu:12-164 for
 * such synapses.
 */
int hs_net_pathway_stdp(hs_net *net, int pid, const double *params, int nparams);
/*
 * Make a pathway's on_pre code: what the reference's code generator emits for a
 * Synapses object's on_pre statements (cuda_generator.py:393-527 into the
 * synapses.cu:12-164 effect kernel, atomics for additive target writes).  For
 * every synapse whose spike arrives at step t, in the network's precision R,
 * with these names in scope:
 *   syn_names[k]        the synapse's variable k (read and write; a variable named
 *                       w on a pathway with per-synapse weights is its weights,
 *                       others start at 0: hs_net_set_synapse_var); variable k is
 *                       HS_SYN_CODE + k for hs_net_get_synapse_var, monitors
 *   <post_names[k]>_post  starts at 0; what the code leaves in it is added to
 *                       the target neuron's variable post_vars[k] (Brian2's
 *                       'x_post += ...': additions only, as several synapses of a
 *                       step may share a target)
 *   <pre_names[k]>_pre  the source neuron's variable pre_vars[k] (read)
 *   dt (R), t (double, the step's time), i (source index), j (target index),
 *   lastupdate (double: the time of the synapse's last on_pre or on_post
 *   event, 0 at first; HS_SYN_LASTUPDATE reads its step)
 * and the functions of hs_net_add_group_code.  The pathway's own target
 * variable and weight effect are replaced by the code.  Needs a scatter
 * pathway; before the first run; compiled with the network's code models on
 * its first run.  In the fused form the network then advances one step per
 * update launch and the additions go to the step's accumulator, as STDP's do.
 */
int hs_net_pathway_code(hs_net *net, int pid, int nsvars, const char *const *syn_names, const char *on_pre, int npost,
                        const int *post_vars, const char *const *post_names, int npre, const int *pre_vars,
                        const char *const *pre_names);
/*
 * Give a pathway with on_pre code (hs_net_pathway_code) on_post code as well:
 * statements run at step t for every synapse of each target neuron that
 * spikes at t, after the step's on_pre code (Brian2 orders the pre pathway
 * before the post pathway).  In scope: the synapse variables (read and write),
 * the source variables <pre_names>_pre of hs_net_pathway_code, the target
 * variables <post_names[k]>_post (read; variable post_vars[k] of the target),
 * dt, t, i, j and lastupdate (double: the time of the synapse's last on_pre or
 * on_post; every such event sets it to t, starting from 0).  Writes go to
 * synapse variables only.  Replaces the post pathway's effect kernel
 * (synapses.cu:201-322) for such synapses; before the first run.
 */
int hs_net_pathway_code_post(hs_net *net, int pid, const char *on_post, int npost, const int *post_vars,
                             const char *const *post_names);
/* Set synapse variable var (HS_SYN_CODE + k) of a pathway with on_pre code from
   n = its synapse count values in bundle order (hs_net_get_synapses' order). */
int hs_net_set_synapse_var(hs_net *net, int pid, int var, const double *values, int64_t n);
/* a synapse variable (HS_SYN_W, _APRE, _APOST, _LASTUPDATE; HS_SYN_CODE + k of
   a pathway with on_pre code) in bundle order (the order of
   hs_net_get_synapses); cap >= nsyn */
int hs_net_get_synapse_var(hs_net *net, int pid, int var, double *out, int64_t cap);

/* SpikeMonitor on a group: capacity = device records kept (i, step) at first;
   the buffers grow before a run that could fill them. */
int hs_net_add_spike_monitor(hs_net *net, int gid, int64_t capacity, int *mid);
/*
 * StateMonitor of a synapse variable (HS_SYN_W, HS_SYN_APRE or HS_SYN_APOST) of
 * the synapses at positions indices[0..n) of a pathway (bundle order, the
 * order of hs_net_get_synapses).  Like Brian2's StateMonitor (when='start')
 * each step records the values before that step's updates; read with
 * hs_net_get_state_monitor (rows x n, time-major) or hs_net_write_monitor.
 * w needs per-synapse weights.  In the fused form the monitor is added before
 * the first run (the network then advances one step per update launch).
 * Replaces a StateMonitor on a Synapses object (statemonitor.cu:9-100, as
 * test_cpp_cuda_consistency.py:177-207 and examples/stdp.py:182 use it).
 */
int hs_net_add_synapse_monitor(hs_net *net, int pid, int var, const int32_t *indices, int n, int *mid);
/*
 * SpikeMonitor on neurons [start, stop) of a group (a Brian2 Subgroup G[start:stop]):
 * its records are those of the range, indices relative to start.  The device
 * records each group's spikes once (in the group's whole-group monitor, created
 * here if absent); a range monitor is a filter applied on read.  Any number of
 * monitors, overlapping or not, may watch one group.
 * Replaces the subgroup filter of the spike monitor (spikemonitor.cu:15-22
 * is_in_subgroup, :195 count[_idx - _source_start]).
 */
int hs_net_add_spike_monitor_range(hs_net *net, int gid, int start, int stop, int64_t capacity, int *mid);
/*
 * PopulationRateMonitor of neurons [start, stop) of a group: the update kernel
 * counts their spikes per step; hs_net_get_rate_monitor gives
 * rate = 1.0*count/dt/(stop-start) per recorded step (host double[cap]).
 * Replaces ratemonitor.cu:52-99 (a one-thread kernel per step).
 */
int hs_net_add_rate_monitor(hs_net *net, int gid, int start, int stop, int *mid);
int hs_net_get_rate_monitor(hs_net *net, int mid, double *rate, int64_t cap);
/* StateMonitor (when='start') of var over neurons [start, stop), every step. */
int hs_net_add_state_monitor(hs_net *net, int gid, int var, int start, int stop, int *mid);
/*
 * StateMonitor(G, var, record=indices, dt=every*dt) (when='start'): the values
 * of neurons indices[0..n) (distinct, in the given order: the columns) at the
 * start of every step t with t mod every == 0, from the step the monitor is
 * added (the monitor's clock ticks at multiples of its dt).  Rows are read with
 * hs_net_get_state_monitor (rows x n, time-major) or hs_net_write_monitor
 * (t = row step * dt; indices as given).
 * Replaces statemonitor.cu:9-100 with its _indices array and its own clock.
 */
int hs_net_add_state_monitor_indices(hs_net *net, int gid, int var, const int32_t *indices, int n, int every,
                                     int *mid);
/* The interval of a StateMonitor (of neurons or of synapses) in steps, as its
   own dt = every*dt: rows at steps t with t mod every == 0.  Set before the
   monitor records its first row. */
int hs_net_set_monitor_interval(hs_net *net, int mid, int every);

/* Tuning knobs: key 0 update block size, 1 scatter-delivery grid, 2 steps per
 * captured graph (0: no graph), 3 delivery blocks of the single-launch fused
 * kernel, 4 window (steps per update launch, 0: largest allowed), 5 delivery
 * layout (0 auto, 1 scatter, 2 gather), 6 gather block size, 7 kernel timing
 * (1: no graphs, every launch bracketed by HIP events; totals reset when set),
 * 8 blocks drawing a window's noise beside the delivery or update (0: 2 per CU),
 * 9 pipelined windows (1: one launch per window -- update and push of the
 * delays < 2W, pull of the longer delays for the next window; 0, the default:
 * a delivery launch and an update launch per window), 10 update grid (0: one
 * block per tile; n > 0: at most n blocks, each striding over tiles),
 * 11 fast math (FP32 only; 1: the COBA Hodgkin-Huxley update uses the device
 * exponential, reciprocal division and FMA contraction -- faster, not
 * bit-reproducible by the oracle; its error is measured against FP64).
 * brian2cuda_amd/autotune.py picks keys 0, 1, 3, 4, 8, 9, 10 by measurement. */
int hs_net_set_option(hs_net *net, int key, long long value);
int hs_net_get_option(hs_net *net, int key, long long *value);
/* With option 7 on: device time and launch count, summed over the runs since
 * the option was set, of one kernel kind: 0 update, 1 scatter delivery,
 * 2 gather delivery, 3 single-launch fused step.  No reference counterpart
 * (the reference is profiled with nvprof). */
int hs_net_kernel_time(hs_net *net, int kind, double *total_ms, int64_t *launches);

/*
 * Advance nsteps time steps.  Steps run in chunks captured as HIP graphs; no
 * host synchronisation inside a chunk (counts stay on the device).  Pending
 * effects and resets are applied at the end so that state reads see Brian2's
 * post-step state.  Replaces Network::run (network.cu:40-137).
 */
int hs_net_run(hs_net *net, int64_t nsteps);

/*
 * Everything a following hs_net_run(net, nsteps) needs before its first step,
 * without advancing the network: the device rings and accumulators, the
 * state-monitor rows, the descriptors, and the instantiated HIP graph of every
 * chunk size that run launches.  That run then only replays graphs (unless a
 * spike monitor has to grow first).  The generated code's Network::run has no
 * such split: its set-up (the _before_run_* code objects, network.cu:40-60)
 * and its step loop run in one call, and Network::_last_run_time times the loop
 * (network.cu:118-126); this call keeps graph instantiation out of the loop
 * that hs_stats.last_run_wall_s times.
 */
int hs_net_prepare(hs_net *net, int64_t nsteps);

/* Run statistics. */
typedef struct {
    int64_t steps;           /* steps simulated so far */
    uint64_t events;         /* synaptic events delivered so far */
    uint64_t spikes;         /* spikes emitted so far */
    double last_run_wall_s;  /* wall time of the last hs_net_run (host clock) */
    double last_run_gpu_ms;  /* device time of the last run's step chunks (HIP events) */
    int64_t last_run_steps;
    double launches_per_step; /* kernel launches per simulated step */
    int64_t window;           /* steps per update launch (fused form with delays >= 2 steps) */
} hs_stats;
int hs_net_stats(hs_net *net, hs_stats *out);

/* Number of records of a monitor (spikes, or recorded steps). */
int hs_net_mon_count(hs_net *net, int mid, int64_t *n);
/* Spike records sorted by (step, i); cap = room in the host arrays. */
int hs_net_get_spikes(hs_net *net, int mid, int32_t *i, int64_t *step, int64_t cap);
/* State samples, time-major [n_steps][stop-start], as double. */
int hs_net_get_state_monitor(hs_net *net, int mid, double *out, int64_t cap);
/*
 * Write a monitor's arrays as the reference's result files: one raw
 * little-endian array per file, as _write_arrays does
 * (brian2cuda/templates/objects.cu:550-563; 2-D monitor arrays time-major, one
 * row of the recorded indices per recorded step, :589-595).  NULL path: skipped.
 *   spike monitor: a = i (int32[N]), b = t (double[N], seconds), c = count
 *                  (int32 per neuron of the group), d = N (int32[1]);
 *                  records in cpp_standalone order (by step, then index)
 *   state monitor: a = t (double[rows]), b = values (the network's precision,
 *                  [rows][stop-start]), c = recorded indices (int32), d = rows (int32[1])
 *   rate monitor : a = t (double[rows]), b = rate (double[rows], Hz), c unused, d = rows
 */
int hs_net_write_monitor(hs_net *net, int mid, const char *path_a, const char *path_b, const char *path_c,
                         const char *path_d);
/* last_run_info.txt: "<wall seconds of the last run> <completed fraction>" (brian2cuda/templates/objects.cu:631-636) */
int hs_net_write_last_run_info(hs_net *net, const char *path);

/* ---- multicompartment cells: the Hines engine ----------------------------- */

typedef struct hs_hines hs_hines;

/*
 * A batch of independent SpatialNeuron trees advanced together.  Sections are
 * numbered depth-first per tree (a parent before its children); compartments
 * are contiguous per section.
 *   tree_comp[n_trees+1], tree_sec[n_trees+1]: offsets of each tree's
 *     compartments / sections in the flat arrays
 *   sec_first, sec_n, sec_parent [n_sections]: tree-local first compartment,
 *     compartment count, tree-local parent section (-1: the root)
 *   area, r_length_1, r_length_2, Cm [n_compartments]: Brian2's morphology
 *     variables (m^2, m, m, F/m^2); Ri the axial resistivity (ohm m)
 *   chan[6] = gl, El, gNa, ENa, gK, EK (S/m^2, V): HH (1952) channels,
 *     exponential Euler, in `precision`; the cable solve is always FP64
 *   v_spike: somatic (compartment 0) upward crossings are counted per tree
 * Limits: 1024 compartments and 255 sections per tree.
 * Replaces the spatial state updater (spatialstateupdate.cu:18-93 prep,
 * :108-113 channel kernel, :128-192 per-section Thomas, :202-307 coupling,
 * :316-351 combine) -- one SpatialNeuron code object per tree in the reference.
 */
int hs_hines_create(int precision, double dt, int n_trees, const int32_t *tree_comp, const int32_t *tree_sec,
                    const int32_t *sec_first, const int32_t *sec_n, const int32_t *sec_parent, const double *area,
                    const double *r_length_1, const double *r_length_2, const double *Cm, double Ri,
                    const double *chan, double v_spike, hs_hines **out);
void hs_hines_destroy(hs_hines *h);
/* var: 0 v, 1 m, 2 n, 3 h, 4 point current I (A) per compartment; host double[n_compartments] */
int hs_hines_set_state(hs_hines *h, int var, const double *values);
int hs_hines_get_state(hs_hines *h, int var, double *values);
/* key 0: steps per launch (each workgroup keeps its tree on chip for that many steps) */
int hs_hines_set_option(hs_hines *h, int key, long long value);
int hs_hines_run(hs_hines *h, int64_t nsteps);
int hs_hines_spike_counts(hs_hines *h, int32_t *counts);
int hs_hines_stats(hs_hines *h, double *last_run_wall_s, double *last_run_gpu_ms, int64_t *steps);

#ifdef __cplusplus
}
#endif

#endif /* HS_H */
