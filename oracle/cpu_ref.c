/*
 * cpu_ref.c -- THE ORACLE.  Test infrastructure only: tests/, the smoke check
 * and bench.py's CPU-baseline leg load it as the checker; the product never
 * links, loads or calls it.
 *
 * A serial C restatement of what Brian2's cpp_standalone computes per time step
 * for the networks of the hot path, following the reference's templates:
 *   schedule      network.cu:40-137 (objects called in schedule order, t = i*dt,
 *                 clocks.h:27-56); slots start, groups, thresholds, synapses, resets
 *   threshold     threshold.cu:19-45 -- but the spike list is produced in index
 *                 order, as cpp_standalone's serial loop produces it
 *                 (docs_sphinx/introduction/known_issues.rst:106-113)
 *   refractory    not_refractory = timestep(t-lastspike,dt) >= timestep(ref,dt)
 *                 (cuda_generator.py:997-1006); lastspike = t on a spike
 *                 (threshold.cu:34-41)
 *   reset         reset.cu:3-27 (spiking neurons only, after the synapses slot)
 *   delays        delay steps = (int)(d/dt + 0.5) (synapses_push_spikes.cu:375)
 *   spike queue   spikequeue.h:432-516: Q = max_delay+1 slots, a spike pushes each
 *                 of its synapses into slot (offset+delay)%Q, the current slot is
 *                 delivered in push order, cleared, and offset advances
 *   effects       synapses.cu:59-161: target[post] += w, serially, in queue order
 *   monitors      spikemonitor.cu:51-196 (i, t of every spike after thresholds),
 *                 statemonitor.cu:74-100 (when='start': before the state update)
 *
 * The per-neuron update expressions are the shared numerical contract
 * (brian2cuda_amd/csrc/hs_model_body.inc); they are pinned independently of
 * the kernels by closed forms in tests/test_oracle.py.
 *
 * Two semantics are restated:
 *   SEQ (0): Brian2's -- effects are applied to the state during delivery.
 *   ACC (1): the fused form's -- effects of step t are summed into an
 *            accumulator and added to the state at the start of step t+1,
 *            before the (deferred) reset of step t's spikers.
 * Two precisions: all state and arithmetic in double (64) or float (32).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/hs_models.h"
#include "../brian2cuda_amd/csrc/hs_model_exprs.h"
#include "cpu_ref.h"

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, written from Salmon et al. (SC'11), independent of the     */
/* product's header; pinned by the Random123 known-answer vectors.           */
/* ------------------------------------------------------------------------ */
void ref_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    int round;
    for (round = 0; round < 10; ++round) {
        uint64_t a = (uint64_t)0xD2511F53u * x0;
        uint64_t b = (uint64_t)0xCD9E8D57u * x2;
        uint32_t y0 = (uint32_t)(b >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)b;
        uint32_t y2 = (uint32_t)(a >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)a;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

static double ref_u53(uint32_t a, uint32_t b)
{
    uint64_t k = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return ldexp((double)k, -53);
}

double ref_randn(uint64_t seed, uint32_t stream, uint32_t index, int64_t step)
{
    uint32_t c[4] = {index, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32), stream};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    ref_philox4x32_10(c, k, w);
    double u1 = ref_u53(w[0], w[1]) + ldexp(1.0, -53);
    double u2 = ref_u53(w[2], w[3]);
    return hs_box_muller(u1, u2);
}

/* the FP32 form's draw: 24-bit uniforms from words 0 and 2, float Box-Muller */
float ref_randn_f32(uint64_t seed, uint32_t stream, uint32_t index, int64_t step)
{
    uint32_t c[4] = {index, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32), stream};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    ref_philox4x32_10(c, k, w);
    float u1 = ldexpf((float)((w[0] >> 8) + 1u), -24);
    float u2 = ldexpf((float)(w[2] >> 8), -24);
    return hs_box_mullerf(u1, u2);
}

double ref_rand(uint64_t seed, uint32_t stream, uint32_t index, int64_t step)
{
    uint32_t c[4] = {index, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32), stream};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    ref_philox4x32_10(c, k, w);
    return ref_u53(w[0], w[1]);
}

int ref_quantize_delay(double delay, double dt) { return (int)(delay / dt + 0.5); }

/*
 * Random connectivity, restated from the reference's generator
 * (brian2cuda/templates/synapses_create_generator.cu:315-333), outer loop over
 * presynaptic neurons, inner over targets in increasing j:
 *   p == 0    no synapses (line 315)
 *   p < 0.25  jump algorithm: r = rand(); r == 0 ends the row; j += floor(log(r) *
 *             1/log(1-p)); j past the end ends the row; else j is selected
 *   p >= 0.25 rand() >= p skips the pair
 * The draws (the product's documented layout, not the reference's cuRAND buffer):
 *   skip draw k of row i : Philox((i, k/2, 0x434E0000|0, stream)), words 0-1 (k even) or 2-3
 *   pair test of (i, j)  : Philox((i, (j-post_lo)/2, 0x434E0000|2, stream)), halves the same way
 *   synapse (i, j)       : Philox((i, j, 0x434E0000|1, stream)): delay uniform from words 0-1,
 *                          weight uniform from words 2-3
 * delay = delay_lo + u*(delay_hi - delay_lo) seconds, quantised (int)(d/dt + 0.5)
 * (synapses_push_spikes.cu:375); weight = w_lo + u*(w_hi - w_lo).
 */
static void ref_conn_words(uint64_t seed, uint32_t stream, uint32_t i, uint32_t k, uint32_t tag, uint32_t *w)
{
    uint32_t c[4] = {i, k, 0x434E0000u | tag, stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    ref_philox4x32_10(c, key, w);
}

/* a condition / per-pair probability given as code (connect(condition=...,
   p='...')): compiled for the host from the same C expressions as the device
   program, called with the absolute indices (ai, aj) and the ranges, and the
   source / target state arrays */
typedef int (*ref_conn_cond_fn)(const void *const *pre_vars, const void *const *post_vars, long long ai, long long aj,
                                long long pre_lo, long long pre_hi, long long post_lo, long long post_hi,
                                uint64_t seed, uint32_t stream);
typedef double (*ref_conn_p_fn)(const void *const *pre_vars, const void *const *post_vars, long long ai, long long aj,
                                long long pre_lo, long long pre_hi, long long post_lo, long long post_hi,
                                uint64_t seed, uint32_t stream);
/* a synapse's weight given as code (S.w = '...'), copy: its index among the
   pair's synapses (connect(n=...)) */
typedef double (*ref_conn_w_fn)(const void *const *pre_vars, const void *const *post_vars, long long ai, long long aj,
                                long long pre_lo, long long pre_hi, long long post_lo, long long post_hi,
                                uint64_t seed, uint32_t stream, int copy);
struct ref_conn_code {
    ref_conn_cond_fn cond; /* NULL: none */
    ref_conn_p_fn p;       /* NULL: the constant p */
    ref_conn_w_fn w;       /* NULL: weights uniform in [w_lo, w_hi] */
    const void *const *pre_vars, *const *post_vars;
    int pre_lo, pre_hi, post_lo, post_hi;
    uint64_t seed; /* the connect call's: keys of rand() / randn() in the code */
    uint32_t stream;
    int ncopy;     /* synapses per selected pair (connect(n=...)) */
};

/* weight of copy c of pair (i, j): the weight expression when given, else the uniform draw u */
static double ref_conn_weight(const struct ref_conn_code *cc, int i, long long j, int c, double u)
{
    if (!cc || !cc->w) return u;
    return cc->w(cc->pre_vars, cc->post_vars, i, j, cc->pre_lo, cc->pre_hi, cc->post_lo, cc->post_hi, cc->seed,
                 cc->stream, c);
}

static int ref_conn_keep(const struct ref_conn_code *cc, int i, long long j)
{
    return !cc || !cc->cond ||
           cc->cond(cc->pre_vars, cc->post_vars, i, j, cc->pre_lo, cc->pre_hi, cc->post_lo, cc->post_hi, cc->seed,
                    cc->stream) != 0;
}

static int64_t ref_connect_random_c(int pre_lo, int pre_hi, int post_lo, int post_hi, double p, int skip_self,
                                    uint64_t seed, uint32_t stream, double delay_lo, double delay_hi, double dt,
                                    double w_lo, double w_hi, const struct ref_conn_code *cc, int32_t *pre,
                                    int32_t *post, int32_t *delay_steps, double *w, int64_t cap)
{
    int64_t n = 0;
    const int pexpr = cc && cc->p;
    if (p == 0.0 && !pexpr) return 0;
    const int jump = !pexpr && p < 0.25;
    const double pconst = jump ? 1.0 / ref_log(1.0 - p) : 0.0;
    const double dspan = delay_hi - delay_lo, wspan = w_hi - w_lo;
    for (int i = pre_lo; i < pre_hi; ++i) {
        uint32_t words[4] = {0, 0, 0, 0};
        uint32_t draw = 0;
        long long j = post_lo;
        while (j < post_hi) {
            if (jump) {
                if ((draw & 1u) == 0) ref_conn_words(seed, stream, (uint32_t)i, draw / 2, 0, words);
                const double r = (draw & 1u) ? ref_u53(words[2], words[3]) : ref_u53(words[0], words[1]);
                draw++;
                if (r == 0.0) break;
                const double skip = floor(ref_log(r) * pconst);
                if (skip >= (double)(post_hi - j)) break;
                j += (long long)skip;
            } else {
                const uint32_t m = (uint32_t)(j - post_lo);
                if ((m & 1u) == 0) ref_conn_words(seed, stream, (uint32_t)i, m / 2, 2, words);
                const double r = (m & 1u) ? ref_u53(words[2], words[3]) : ref_u53(words[0], words[1]);
                const double pp =
                    pexpr ? cc->p(cc->pre_vars, cc->post_vars, i, j, cc->pre_lo, cc->pre_hi, cc->post_lo, cc->post_hi,
                                  cc->seed, cc->stream)
                          : p;
                if (!(r < pp)) {
                    ++j;
                    continue;
                }
            }
            if (!(skip_self && j == i) && ref_conn_keep(cc, i, j)) {
                /* copy c of the pair: delay and weight from tag 1 | c << 4 */
                for (int c = 0; c < (cc ? cc->ncopy : 1); ++c) {
                    if (n < cap) {
                        uint32_t sw[4];
                        ref_conn_words(seed, stream, (uint32_t)i, (uint32_t)j, 1u | ((uint32_t)c << 4), sw);
                        const double ud = ref_u53(sw[0], sw[1]), uw = ref_u53(sw[2], sw[3]);
                        if (pre) pre[n] = i;
                        if (post) post[n] = (int32_t)j;
                        if (delay_steps) delay_steps[n] = ref_quantize_delay(delay_lo + ud * dspan, dt);
                        if (w) w[n] = ref_conn_weight(cc, i, j, c, w_lo + uw * wspan);
                    }
                    n++;
                }
            }
            ++j;
        }
    }
    return n;
}

int64_t ref_connect_random(int pre_lo, int pre_hi, int post_lo, int post_hi, double p, int skip_self, uint64_t seed,
                           uint32_t stream, double delay_lo, double delay_hi, double dt, double w_lo, double w_hi,
                           int32_t *pre, int32_t *post, int32_t *delay_steps, double *w, int64_t cap)
{
    return ref_connect_random_c(pre_lo, pre_hi, post_lo, post_hi, p, skip_self, seed, stream, delay_lo, delay_hi, dt,
                                w_lo, w_hi, NULL, pre, post, delay_steps, w, cap);
}

static int ref_cmp_i32(const void *a, const void *b)
{
    const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/*
 * Fixed out-degree (connect(j='k for k in sample(N_post, size=K)'), the
 * reference's fixed-size sampling, synapses_create_generator.cu:250-312, with
 * its own draws): every source i in [pre_lo, pre_hi) samples K distinct targets
 * of the n = post_hi - post_lo candidates, uniform over the K-subsets; with
 * skip_self the sampled j == i is then dropped (the generator's condition
 * "if k != i", so such a row keeps K - 1).  Draw c of row i is
 * min(floor(u * n), n - 1), u the 53-bit uniform of words 0-1 of Philox counter
 * (i, c, tag 3); draws 0..K-1, then, after each sort, the repeated values are
 * dropped and the missing ones drawn with the next counters.  The targets are
 * listed in ascending order; delays and weights are drawn per (i, j) as for
 * connect(p).  Returns the synapse count, -1 if K > n (the reference's error
 * "sample size bigger than the population size").
 */
static int64_t ref_connect_fixed_c(int pre_lo, int pre_hi, int post_lo, int post_hi, int K, int skip_self,
                                   uint64_t seed, uint32_t stream, double delay_lo, double delay_hi, double dt,
                                   double w_lo, double w_hi, const struct ref_conn_code *cc, int32_t *pre,
                                   int32_t *post, int32_t *delay_steps, double *w, int64_t cap)
{
    int64_t n = 0;
    const int cand = post_hi - post_lo;
    if (K > cand && pre_hi > pre_lo) return -1;
    if (K <= 0 || pre_hi <= pre_lo) return 0;
    const double dspan = delay_hi - delay_lo, wspan = w_hi - w_lo;
    int32_t *x = (int32_t *)malloc(sizeof(int32_t) * (size_t)K);
    for (int i = pre_lo; i < pre_hi; ++i) {
        uint32_t c = 0;
        int m = 0; /* distinct values so far, sorted in x[0..m) */
        while (m < K) {
            for (int k = m; k < K; ++k) {
                uint32_t wd[4];
                ref_conn_words(seed, stream, (uint32_t)i, c++, 3, wd);
                const int32_t v = (int32_t)floor(ref_u53(wd[0], wd[1]) * (double)cand);
                x[k] = v < cand - 1 ? v : cand - 1;
            }
            qsort(x, (size_t)K, sizeof(int32_t), ref_cmp_i32);
            m = 0;
            for (int k = 0; k < K; ++k)
                if (k == 0 || x[k] != x[k - 1]) x[m++] = x[k];
        }
        for (int k = 0; k < K; ++k) {
            const int32_t j = post_lo + x[k];
            if ((skip_self && j == i) || !ref_conn_keep(cc, i, j)) continue;
            for (int c = 0; c < (cc ? cc->ncopy : 1); ++c) {
                if (n < cap) {
                    uint32_t sw[4];
                    ref_conn_words(seed, stream, (uint32_t)i, (uint32_t)j, 1u | ((uint32_t)c << 4), sw);
                    const double ud = ref_u53(sw[0], sw[1]), uw = ref_u53(sw[2], sw[3]);
                    if (pre) pre[n] = i;
                    if (post) post[n] = j;
                    if (delay_steps) delay_steps[n] = ref_quantize_delay(delay_lo + ud * dspan, dt);
                    if (w) w[n] = ref_conn_weight(cc, i, j, c, w_lo + uw * wspan);
                }
                n++;
            }
        }
    }
    free(x);
    return n;
}

int64_t ref_connect_fixed(int pre_lo, int pre_hi, int post_lo, int post_hi, int K, int skip_self, uint64_t seed,
                          uint32_t stream, double delay_lo, double delay_hi, double dt, double w_lo, double w_hi,
                          int32_t *pre, int32_t *post, int32_t *delay_steps, double *w, int64_t cap)
{
    return ref_connect_fixed_c(pre_lo, pre_hi, post_lo, post_hi, K, skip_self, seed, stream, delay_lo, delay_hi, dt,
                               w_lo, w_hi, NULL, pre, post, delay_steps, w, cap);
}

/*
 * connect(condition=..., p=...) with the condition and, optionally, the
 * per-pair probability compiled from code (cond_fn, p_fn: NULL for none):
 * K > 0 filters the fixed out-degree sample by the condition; otherwise the
 * candidates of connect(p) (geometric skips or one draw per pair, p_fn taking
 * the place of p in the per-pair test) are kept where the condition holds --
 * Brian2's 'k for k in sample(N_post, p=p) if condition'.  The draws are those
 * of ref_connect_random / ref_connect_fixed, so a condition that always holds
 * gives their synapses.  pre_vars / post_vars: the groups' state arrays the
 * code reads.
 */
int64_t ref_connect_code(int pre_lo, int pre_hi, int post_lo, int post_hi, double p, int K, int ncopy, int skip_self,
                         uint64_t seed, uint32_t stream, double delay_lo, double delay_hi, double dt, double w_lo,
                         double w_hi, void *cond_fn, void *p_fn, void *w_fn, const void *const *pre_vars,
                         const void *const *post_vars, int32_t *pre, int32_t *post, int32_t *delay_steps, double *w,
                         int64_t cap)
{
    struct ref_conn_code cc;
    cc.cond = (ref_conn_cond_fn)cond_fn;
    cc.p = (ref_conn_p_fn)p_fn;
    cc.w = (ref_conn_w_fn)w_fn;
    cc.pre_vars = pre_vars;
    cc.post_vars = post_vars;
    cc.pre_lo = pre_lo;
    cc.pre_hi = pre_hi;
    cc.post_lo = post_lo;
    cc.post_hi = post_hi;
    cc.seed = seed;
    cc.stream = stream;
    cc.ncopy = ncopy > 1 ? ncopy : 1;
    if (K > 0)
        return ref_connect_fixed_c(pre_lo, pre_hi, post_lo, post_hi, K, skip_self, seed, stream, delay_lo, delay_hi,
                                   dt, w_lo, w_hi, &cc, pre, post, delay_steps, w, cap);
    return ref_connect_random_c(pre_lo, pre_hi, post_lo, post_hi, p_fn ? 1.0 : p, skip_self, seed, stream, delay_lo,
                                delay_hi, dt, w_lo, w_hi, &cc, pre, post, delay_steps, w, cap);
}

long long ref_timestep(double t, double dt) { return hs_timestep(t, dt); }

double ref_exp(double x) { return hs_exp(x); }
float ref_expf(float x) { return hs_expf(x); }
double ref_log(double x) { return hs_log(x); }
double ref_sqrt(double x) { return hs_sqrt(x); }
double ref_cos2pi(double u) { return hs_cos2pi(u); }
float ref_logf(float x) { return hs_logf(x); }
float ref_sqrtf(float x) { return hs_sqrtf(x); }
float ref_cos2pif(float u) { return hs_cos2pif(u); }

int ref_model_constants(int model, const double *params, int nparams, double dt, double *k)
{
    double p[16];
    if (hs_model_params(model, params, nparams, p) < 0) return -1;
    return hs_model_constants(model, p, dt, k);
}

/* ------------------------------------------------------------------------ */
/* network                                                                   */
/* ------------------------------------------------------------------------ */
#define REF_MAX_GROUPS 16
#define REF_MAX_PATHS 32
#define REF_MAX_MONS 32
#define REF_MAX_THREADS 256

typedef struct {
    int model, n, nvars;
    double p[16];
    double k[16];
    int nk;
    double vth, vreset, refractory;
    int has_reset;
    int64_t ref_steps;
    uint32_t stream;
    void *var[HS_MAX_VARS];
    void *acc[HS_MAX_VARS];
    double *lastspike;   /* seconds, Brian2's representation */
    int32_t *spk;        /* this step's spikes, ascending */
    int64_t nspk;
    int32_t *pend;       /* spikes whose reset is pending (ACC) */
    int64_t npend;
    /* SpikeGeneratorGroup: neurons gen_idx[gen_off[s] .. gen_off[s+1]) fire at
       step s (ascending), s < gen_steps; the pattern repeats every gen_period
       steps (0: once) */
    int64_t *gen_off;
    int32_t *gen_idx;
    int64_t gen_steps, gen_period;
    /* code model (model HS_MODEL_CODE + k): the step and reset functions of the
       network's precision, compiled for the host from the model's code */
    void *cg_step, *cg_reset;
    int cg_noisy;
} ref_group;

typedef struct {
    int src, tgt, var;
    int64_t nsyn;
    int32_t *pre, *post, *delay;
    int64_t *pre_start; /* CSR by presynaptic neuron, synapse order kept */
    int64_t *by_pre;
    int wmode;
    double wscalar;
    void *w; /* REAL[nsyn] or REAL[n_src] */
    int Q, offset;
    int64_t **q;
    int64_t *qlen, *qcap;
    /* threaded form (nthr > 1): thread t owns the targets [n_tgt*t/nthr,
       n_tgt*(t+1)/nthr) and a queue of its own (tq[t*Q + slot]) holding only
       the synapses onto them, pushed in the serial order */
    int64_t **tq;
    int64_t *tqlen, *tqcap;
    /* STDP (plastic != 0): traces and last update step per synapse, and the
       synapses of each postsynaptic neuron (CSR, synapse order) for on_post */
    int plastic;
    double taupre, taupost;
    double sk[5]; /* dApre, dApost, wmin, wmax, scale */
    void *apre, *apost;
    int64_t *lastup;
    int64_t *post_start, *by_post;
    /* synaptic code (cg_fn != NULL): on_pre(sv, pre, eff, dt, t, i, j) compiled for
       the host; nsv synapse variables in synapse order (sv[w_sv] is the weight
       array); additions to target variables post_var[], reads of source
       variables pre_var[] */
    void *cg_fn;
    int nsv, npost, npre;
    void *sv[HS_MAX_SYN_VARS];
    int post_var[HS_MAX_VARS], pre_var[HS_MAX_VARS];
    /* on_post code: fn(sv, pre, post, dt, t, i, j, lastupdate), reading target
       variables post_read_var[0..npr) (post_start/by_post and lastup above) */
    void *cg_post_fn;
    int npr, post_read_var[HS_MAX_VARS];
} ref_path;

typedef struct {
    int kind; /* 0 spikes, 1 state, 2 rate, 3 synapse variable */
    int gid, var, start, stop;
    int64_t n, cap;
    int32_t *i;
    int64_t *step;
    double *data;
    int pid, nsel; /* kind 3: pathway and recorded synapse positions;
                      kind 1 with indices: nsel recorded neurons sel[] */
    int32_t *sel;
    int every;     /* kind 1: record at steps t with t mod every == 0 (StateMonitor dt) */
} ref_mon;

struct ref_net {
    int prec, sem;
    int nthr;    /* OpenMP threads of the step loop (1: the serial restatement) */
    int thr_err; /* a parallel region ran with fewer threads than nthr */
    double dt;
    uint64_t seed;
    int64_t step;
    uint64_t events;
    int ng, np, nm;
    ref_group g[REF_MAX_GROUPS];
    ref_path pw[REF_MAX_PATHS];
    ref_mon mon[REF_MAX_MONS];
};

static size_t rsz(const ref_net *net) { return net->prec == HS_F64 ? sizeof(double) : sizeof(float); }

static double rget(const ref_net *net, const void *a, int64_t i)
{
    return net->prec == HS_F64 ? ((const double *)a)[i] : (double)((const float *)a)[i];
}

ref_net *ref_net_create(int precision, int semantics, double dt, uint64_t seed)
{
    if (precision != HS_F64 && precision != HS_F32) return NULL;
    ref_net *net = (ref_net *)calloc(1, sizeof(ref_net));
    net->prec = precision;
    net->sem = semantics;
    net->dt = dt;
    net->seed = seed;
    net->nthr = 1;
    return net;
}

/*
 * The step loop over nthreads OpenMP threads, as cpp_standalone runs with
 * devices.cpp_standalone.openmp_threads (the reference's CPU benchmark,
 * brian2cuda/tests/features/cuda_configuration.py:396-418).  The results stay
 * identical to the serial loop, bit for bit: each thread updates a contiguous
 * range of neurons and the spike lists are joined in range order; each thread
 * owns a contiguous range of targets and a spike queue of its own holding the
 * synapses onto them, pushed and delivered in the serial order, so every
 * target receives its additions in the same order as with one thread.  Only
 * before the first step; STDP pathways stay serial.
 */
int ref_net_set_threads(ref_net *net, int nthreads)
{
    if (!net || nthreads < 1 || nthreads > REF_MAX_THREADS || net->step != 0) return -1;
#ifndef _OPENMP
    if (nthreads > 1) return -2;
#else
    omp_set_dynamic(0); /* exactly nthreads per region: the target ranges own queue state across steps */
#endif
    net->nthr = nthreads;
    return 0;
}

static int model_nvars(int model)
{
    if (model == HS_MODEL_LIF_CUBA) return HS_LIF_CUBA_NVARS;
    if (model == HS_MODEL_LIF_NOISE) return HS_LIF_NOISE_NVARS;
    if (model == HS_MODEL_HH_COBA) return HS_HH_NVARS;
    if (model == HS_MODEL_SPIKEGEN) return HS_SPIKEGEN_NVARS;
    if (model == HS_MODEL_POISSON) return HS_POISSON_NVARS;
    return -1;
}

int ref_net_add_group(ref_net *net, int model, int n, const double *params, int nparams)
{
    if (net->ng >= REF_MAX_GROUPS || n <= 0) return -1;
    int nv = model_nvars(model);
    if (nv < 0) return -1;
    ref_group *g = &net->g[net->ng];
    memset(g, 0, sizeof(*g));
    if (hs_model_params(model, params, nparams, g->p) < 0) return -1;
    params = g->p; /* every parameter, optional ones filled in */
    g->model = model;
    g->n = n;
    g->nvars = nv;
    g->nk = hs_model_constants(model, params, net->dt, g->k);
    if (g->nk < 0) return -1;
    if (model == HS_MODEL_LIF_CUBA) {
        g->vth = params[1]; g->vreset = params[2]; g->refractory = params[6]; g->has_reset = 1;
    } else if (model == HS_MODEL_LIF_NOISE) {
        g->vth = params[3]; g->vreset = params[4]; g->refractory = params[5]; g->has_reset = 1;
        g->stream = (uint32_t)params[6];
    } else if (model == HS_MODEL_SPIKEGEN) {
        g->refractory = 0.0; g->has_reset = 0;
        g->gen_period = hs_timestep(params[0], net->dt);
    } else if (model == HS_MODEL_POISSON) {
        g->refractory = 0.0; g->has_reset = 0;
        g->stream = (uint32_t)params[1];
    } else {
        g->vth = params[12]; g->refractory = params[13]; g->has_reset = 0;
    }
    g->ref_steps = hs_timestep(g->refractory, net->dt);
    for (int v = 0; v < nv; ++v) g->var[v] = calloc((size_t)n, rsz(net));
    if (model == HS_MODEL_POISSON) /* rates start at the rate parameter */
        for (int i = 0; i < n; ++i) {
            if (net->prec == HS_F64) ((double *)g->var[0])[i] = params[0];
            else ((float *)g->var[0])[i] = (float)params[0];
        }
    g->lastspike = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) g->lastspike[i] = -1e4; /* never spiked */
    g->spk = (int32_t *)malloc(sizeof(int32_t) * n);
    g->pend = (int32_t *)malloc(sizeof(int32_t) * n);
    return net->ng++;
}

/*
 * A group whose model is code (hs_net_add_group_code): step(s, not_ref, xi, dt,
 * t, i) advances the variables in s and returns the threshold condition;
 * reset(s, dt, t_spike, i).  The pointers take the network's precision
 * (double* or float* state, R xi and dt).
 */
int ref_net_add_group_code(ref_net *net, int n, int nvars, void *step_fn, void *reset_fn, double refractory,
                           int noise, uint32_t stream)
{
    if (net->ng >= REF_MAX_GROUPS || n <= 0 || nvars < 1 || nvars > HS_MAX_VARS || !step_fn) return -1;
    int ncode = 0;
    for (int q = 0; q < net->ng; ++q) ncode += HS_IS_CODE_MODEL(net->g[q].model);
    if (ncode >= HS_MAX_CODE_MODELS || !(refractory >= 0.0)) return -1;
    ref_group *g = &net->g[net->ng];
    memset(g, 0, sizeof(*g));
    g->model = HS_MODEL_CODE + ncode;
    g->n = n;
    g->nvars = nvars;
    g->refractory = refractory;
    g->ref_steps = hs_timestep(refractory, net->dt);
    g->has_reset = reset_fn != NULL;
    g->stream = stream;
    g->cg_step = step_fn;
    g->cg_reset = reset_fn;
    g->cg_noisy = noise != 0;
    for (int v = 0; v < nvars; ++v) g->var[v] = calloc((size_t)n, rsz(net));
    g->lastspike = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) g->lastspike[i] = -1e4; /* never spiked */
    g->spk = (int32_t *)malloc(sizeof(int32_t) * n);
    g->pend = (int32_t *)malloc(sizeof(int32_t) * n);
    return net->ng++;
}

int ref_net_set_state(ref_net *net, int gid, int var, const double *values)
{
    if (gid < 0 || gid >= net->ng || var < 0 || var >= net->g[gid].nvars) return -1;
    ref_group *g = &net->g[gid];
    for (int i = 0; i < g->n; ++i) {
        if (net->prec == HS_F64) ((double *)g->var[var])[i] = values[i];
        else ((float *)g->var[var])[i] = (float)values[i];
    }
    return 0;
}

int ref_net_get_state(ref_net *net, int gid, int var, double *out)
{
    if (gid < 0 || gid >= net->ng || var < 0 || var >= net->g[gid].nvars) return -1;
    ref_group *g = &net->g[gid];
    for (int i = 0; i < g->n; ++i) out[i] = rget(net, g->var[var], i);
    return 0;
}

int ref_net_get_lastspike(ref_net *net, int gid, double *out)
{
    if (gid < 0 || gid >= net->ng) return -1;
    memcpy(out, net->g[gid].lastspike, sizeof(double) * net->g[gid].n);
    return 0;
}

int ref_net_add_pathway(ref_net *net, int src, int tgt, int var, int64_t nsyn, const int32_t *pre,
                        const int32_t *post, const int32_t *delay_steps, int scalar_delay,
                        int wmode, const double *weights, double wscalar)
{
    if (net->np >= REF_MAX_PATHS || src < 0 || src >= net->ng || tgt < 0 || tgt >= net->ng) return -1;
    ref_group *gs = &net->g[src], *gt = &net->g[tgt];
    if (var < 0 || var >= gt->nvars) return -1;
    ref_path *p = &net->pw[net->np];
    memset(p, 0, sizeof(*p));
    p->src = src; p->tgt = tgt; p->var = var; p->nsyn = nsyn;
    p->pre = (int32_t *)malloc(sizeof(int32_t) * (nsyn ? nsyn : 1));
    p->post = (int32_t *)malloc(sizeof(int32_t) * (nsyn ? nsyn : 1));
    p->delay = (int32_t *)malloc(sizeof(int32_t) * (nsyn ? nsyn : 1));
    int maxd = 0;
    for (int64_t s = 0; s < nsyn; ++s) {
        if (pre[s] < 0 || pre[s] >= gs->n || post[s] < 0 || post[s] >= gt->n) return -2;
        p->pre[s] = pre[s];
        p->post[s] = post[s];
        p->delay[s] = delay_steps ? delay_steps[s] : scalar_delay;
        if (p->delay[s] < 0) return -3;
        if (p->delay[s] > maxd) maxd = p->delay[s];
    }
    /* CSR by pre, keeping synapse order (counting sort, stable) */
    p->pre_start = (int64_t *)calloc((size_t)gs->n + 1, sizeof(int64_t));
    for (int64_t s = 0; s < nsyn; ++s) p->pre_start[pre[s] + 1]++;
    for (int i = 0; i < gs->n; ++i) p->pre_start[i + 1] += p->pre_start[i];
    p->by_pre = (int64_t *)malloc(sizeof(int64_t) * (nsyn ? nsyn : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * gs->n);
    memcpy(fill, p->pre_start, sizeof(int64_t) * gs->n);
    for (int64_t s = 0; s < nsyn; ++s) p->by_pre[fill[pre[s]]++] = s;
    free(fill);
    p->wmode = wmode;
    p->wscalar = wscalar;
    int64_t nw = wmode == HS_W_SYNAPSE ? nsyn : (wmode == HS_W_PRE ? gs->n : 0);
    if (nw > 0) {
        p->w = malloc(rsz(net) * nw);
        for (int64_t s = 0; s < nw; ++s) {
            if (net->prec == HS_F64) ((double *)p->w)[s] = weights[s];
            else ((float *)p->w)[s] = (float)weights[s];
        }
    }
    p->Q = maxd + 1;
    p->offset = 0;
    p->q = (int64_t **)calloc(p->Q, sizeof(int64_t *));
    p->qlen = (int64_t *)calloc(p->Q, sizeof(int64_t));
    p->qcap = (int64_t *)calloc(p->Q, sizeof(int64_t));
    if (!gt->acc[var]) gt->acc[var] = calloc((size_t)gt->n, rsz(net));
    return net->np++;
}

/* spikegenerator.cu:43-80: neuron idx[k] fires at step steps[k]; a pair given
   twice is an error (one spike per neuron and step) */
int ref_net_set_spikegen(ref_net *net, int gid, int64_t n, const int32_t *idx, const int64_t *steps)
{
    if (gid < 0 || gid >= net->ng || net->g[gid].model != HS_MODEL_SPIKEGEN || n < 0) return -1;
    ref_group *g = &net->g[gid];
    int64_t smax = -1;
    for (int64_t k = 0; k < n; ++k) {
        if (idx[k] < 0 || idx[k] >= g->n || steps[k] < 0) return -2;
        if (steps[k] > smax) smax = steps[k];
    }
    if (g->gen_period > 0 && smax >= g->gen_period) return -4; /* every spike inside the period */
    free(g->gen_off);
    free(g->gen_idx);
    g->gen_steps = smax + 1;
    g->gen_off = (int64_t *)calloc((size_t)g->gen_steps + 1, sizeof(int64_t));
    g->gen_idx = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    for (int64_t k = 0; k < n; ++k) g->gen_off[steps[k] + 1]++;
    for (int64_t s = 0; s < g->gen_steps; ++s) g->gen_off[s + 1] += g->gen_off[s];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(g->gen_steps ? g->gen_steps : 1));
    memcpy(fill, g->gen_off, sizeof(int64_t) * (size_t)g->gen_steps);
    for (int64_t k = 0; k < n; ++k) {
        /* insertion into the step's ascending list */
        int64_t pos = fill[steps[k]]++;
        while (pos > g->gen_off[steps[k]] && g->gen_idx[pos - 1] > idx[k]) {
            g->gen_idx[pos] = g->gen_idx[pos - 1];
            --pos;
        }
        if (pos > g->gen_off[steps[k]] && g->gen_idx[pos - 1] == idx[k]) {
            free(fill);
            return -3;
        }
        g->gen_idx[pos] = idx[k];
    }
    free(fill);
    return 0;
}

/* STDP on a pathway with per-synapse weights (SEQ semantics only) */
int ref_net_pathway_stdp(ref_net *net, int pid, const double *params, int nparams)
{
    if (pid < 0 || pid >= net->np || nparams != HS_STDP_NPARAMS) return -1;
    ref_path *p = &net->pw[pid];
    if (p->wmode != HS_W_SYNAPSE || p->plastic) return -2;
    const ref_group *gt = &net->g[p->tgt];
    p->plastic = 1;
    p->taupre = params[0];
    p->taupost = params[1];
    for (int j = 0; j < 5; ++j) p->sk[j] = params[2 + j];
    const int64_t n = p->nsyn ? p->nsyn : 1;
    p->apre = calloc((size_t)n, rsz(net));
    p->apost = calloc((size_t)n, rsz(net));
    p->lastup = (int64_t *)calloc((size_t)n, sizeof(int64_t)); /* lastupdate = 0 s */
    p->post_start = (int64_t *)calloc((size_t)gt->n + 1, sizeof(int64_t));
    for (int64_t s = 0; s < p->nsyn; ++s) p->post_start[p->post[s] + 1]++;
    for (int i = 0; i < gt->n; ++i) p->post_start[i + 1] += p->post_start[i];
    p->by_post = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)gt->n);
    memcpy(fill, p->post_start, sizeof(int64_t) * (size_t)gt->n);
    for (int64_t s = 0; s < p->nsyn; ++s) p->by_post[fill[p->post[s]]++] = s;
    free(fill);
    return 0;
}

/*
 * A pathway whose on_pre is code (hs_net_pathway_code): fn(sv, pre, eff, dt, t,
 * i, j) of the network's precision; nsv synapse variables (w_sv >= 0: that one
 * is the pathway's per-synapse weight array), the code adds to target variables
 * post_vars[0..npost) and reads source variables pre_vars[0..npre).
 */
int ref_net_pathway_code(ref_net *net, int pid, void *fn, int nsv, int w_sv, int npost, const int *post_vars, int npre,
                         const int *pre_vars)
{
    if (pid < 0 || pid >= net->np || !fn || nsv < 0 || nsv > HS_MAX_SYN_VARS || npost < 0 || npost > HS_MAX_VARS ||
        npre < 0 || npre > HS_MAX_VARS)
        return -1;
    ref_path *p = &net->pw[pid];
    ref_group *gt = &net->g[p->tgt];
    const ref_group *gs = &net->g[p->src];
    if (p->plastic || p->cg_fn || (w_sv >= 0 && (w_sv >= nsv || p->wmode != HS_W_SYNAPSE))) return -2;
    for (int k = 0; k < npost; ++k)
        if (post_vars[k] < 0 || post_vars[k] >= gt->nvars) return -3;
    for (int k = 0; k < npre; ++k)
        if (pre_vars[k] < 0 || pre_vars[k] >= gs->nvars) return -3;
    const int64_t n = p->nsyn ? p->nsyn : 1;
    for (int k = 0; k < nsv; ++k) p->sv[k] = k == w_sv ? p->w : calloc((size_t)n, rsz(net));
    p->nsv = nsv;
    p->npost = npost;
    p->npre = npre;
    for (int k = 0; k < npost; ++k) {
        p->post_var[k] = post_vars[k];
        if (net->sem == 1 && !gt->acc[post_vars[k]]) gt->acc[post_vars[k]] = calloc((size_t)gt->n, rsz(net));
    }
    for (int k = 0; k < npre; ++k) p->pre_var[k] = pre_vars[k];
    p->lastup = (int64_t *)calloc((size_t)n, sizeof(int64_t)); /* lastupdate = 0 s */
    p->cg_fn = fn;
    return 0;
}

/* on_post code fn(sv, pre, post, dt, t, i, j, lastupdate) of a code pathway,
   reading target variables post_vars[0..npost) */
int ref_net_pathway_code_post(ref_net *net, int pid, void *fn, int npost, const int *post_vars)
{
    if (pid < 0 || pid >= net->np || !fn || npost < 0 || npost > HS_MAX_VARS) return -1;
    ref_path *p = &net->pw[pid];
    const ref_group *gt = &net->g[p->tgt];
    if (!p->cg_fn || p->cg_post_fn) return -2;
    for (int k = 0; k < npost; ++k) {
        if (post_vars[k] < 0 || post_vars[k] >= gt->nvars) return -3;
        p->post_read_var[k] = post_vars[k];
    }
    p->npr = npost;
    const int64_t n = p->nsyn ? p->nsyn : 1;
    p->post_start = (int64_t *)calloc((size_t)gt->n + 1, sizeof(int64_t));
    for (int64_t s = 0; s < p->nsyn; ++s) p->post_start[p->post[s] + 1]++;
    for (int i = 0; i < gt->n; ++i) p->post_start[i + 1] += p->post_start[i];
    p->by_post = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)gt->n);
    memcpy(fill, p->post_start, sizeof(int64_t) * (size_t)gt->n);
    for (int64_t s = 0; s < p->nsyn; ++s) p->by_post[fill[p->post[s]]++] = s;
    free(fill);
    p->cg_post_fn = fn;
    return 0;
}

/* synapse variable k (HS_SYN_CODE + k) of a code pathway, in synapse order */
int ref_net_set_synapse_var(ref_net *net, int pid, int var, const double *values)
{
    if (pid < 0 || pid >= net->np) return -1;
    ref_path *p = &net->pw[pid];
    if (!p->cg_fn || var < HS_SYN_CODE || var >= HS_SYN_CODE + p->nsv) return -2;
    for (int64_t s = 0; s < p->nsyn; ++s) {
        if (net->prec == HS_F64) ((double *)p->sv[var - HS_SYN_CODE])[s] = values[s];
        else ((float *)p->sv[var - HS_SYN_CODE])[s] = (float)values[s];
    }
    return 0;
}

/* synapse variable in synapse (creation) order: 0 w, 1 Apre, 2 Apost, 3 lastupdate (step),
   HS_SYN_CODE + k for code pathways */
int ref_net_get_synapse_var(ref_net *net, int pid, int var, double *out)
{
    if (pid < 0 || pid >= net->np) return -1;
    ref_path *p = &net->pw[pid];
    if (p->cg_fn && var >= HS_SYN_CODE && var < HS_SYN_CODE + p->nsv) {
        for (int64_t s = 0; s < p->nsyn; ++s) out[s] = rget(net, p->sv[var - HS_SYN_CODE], s);
        return 0;
    }
    if (p->cg_fn && var == HS_SYN_LASTUPDATE) {
        for (int64_t s = 0; s < p->nsyn; ++s) out[s] = (double)p->lastup[s];
        return 0;
    }
    for (int64_t s = 0; s < p->nsyn; ++s) {
        if (var == HS_SYN_W) out[s] = p->wmode == HS_W_SYNAPSE ? rget(net, p->w, s) : p->wscalar;
        else if (var == HS_SYN_APRE && p->plastic) out[s] = rget(net, p->apre, s);
        else if (var == HS_SYN_APOST && p->plastic) out[s] = rget(net, p->apost, s);
        else if (var == HS_SYN_LASTUPDATE && p->plastic) out[s] = (double)p->lastup[s];
        else return -2;
    }
    return 0;
}

int ref_net_add_spike_monitor(ref_net *net, int gid)
{
    if (net->nm >= REF_MAX_MONS || gid < 0 || gid >= net->ng) return -1;
    ref_mon *m = &net->mon[net->nm];
    memset(m, 0, sizeof(*m));
    m->kind = 0;
    m->gid = gid;
    return net->nm++;
}

int ref_net_add_state_monitor(ref_net *net, int gid, int var, int start, int stop)
{
    if (net->nm >= REF_MAX_MONS || gid < 0 || gid >= net->ng) return -1;
    if (var < 0 || var >= net->g[gid].nvars || start < 0 || stop > net->g[gid].n || stop <= start) return -1;
    ref_mon *m = &net->mon[net->nm];
    memset(m, 0, sizeof(*m));
    m->kind = 1;
    m->gid = gid; m->var = var; m->start = start; m->stop = stop;
    m->every = 1;
    return net->nm++;
}

/* StateMonitor(G, var, record=indices, dt=every*dt) (statemonitor.cu:9-100 with
   _indices): neurons idx[0..n) in the given order, at the start of every step t
   with t mod every == 0 */
/* a StateMonitor's own dt: rows at steps t with t mod every == 0 */
int ref_net_set_monitor_interval(ref_net *net, int mid, int every)
{
    if (mid < 0 || mid >= net->nm || every < 1) return -1;
    ref_mon *m = &net->mon[mid];
    if ((m->kind != 1 && m->kind != 3) || m->n) return -1;
    m->every = every;
    return 0;
}

int ref_net_add_state_monitor_indices(ref_net *net, int gid, int var, const int32_t *idx, int n, int every)
{
    if (net->nm >= REF_MAX_MONS || gid < 0 || gid >= net->ng || n <= 0 || every < 1) return -1;
    if (var < 0 || var >= net->g[gid].nvars) return -1;
    for (int k = 0; k < n; ++k)
        if (idx[k] < 0 || idx[k] >= net->g[gid].n) return -1;
    ref_mon *m = &net->mon[net->nm];
    memset(m, 0, sizeof(*m));
    m->kind = 1;
    m->gid = gid; m->var = var; m->start = 0; m->stop = n;
    m->nsel = n;
    m->sel = (int32_t *)malloc(sizeof(int32_t) * n);
    memcpy(m->sel, idx, sizeof(int32_t) * n);
    m->every = every;
    return net->nm++;
}

/* StateMonitor of a synapse variable (w, Apre or Apost) of the synapses sel[],
   when='start' (statemonitor.cu:9-100 on a Synapses source) */
int ref_net_add_synapse_monitor(ref_net *net, int pid, int var, const int32_t *sel, int n)
{
    if (net->nm >= REF_MAX_MONS || pid < 0 || pid >= net->np || n <= 0) return -1;
    const ref_path *p = &net->pw[pid];
    const int code_var = p->cg_fn && var >= HS_SYN_CODE && var < HS_SYN_CODE + p->nsv;
    if (!code_var) {
        if (var < 0 || var > 2) return -1;
        if (var == 0 && p->wmode != HS_W_SYNAPSE) return -1;
        if (var > 0 && !p->plastic) return -1;
    }
    for (int k = 0; k < n; ++k)
        if (sel[k] < 0 || sel[k] >= p->nsyn) return -1;
    ref_mon *m = &net->mon[net->nm];
    memset(m, 0, sizeof(*m));
    m->kind = 3;
    m->gid = p->tgt; m->var = var; m->pid = pid; m->nsel = n;
    m->every = 1;
    m->sel = (int32_t *)malloc(sizeof(int32_t) * n);
    memcpy(m->sel, sel, sizeof(int32_t) * n);
    return net->nm++;
}

static void mon_syn(ref_net *net, ref_mon *m)
{
    if (m->every > 1 && net->step % m->every != 0) return;
    const ref_path *p = &net->pw[m->pid];
    const void *src = m->var >= HS_SYN_CODE ? p->sv[m->var - HS_SYN_CODE]
                                            : (m->var == 0 ? p->w : (m->var == 1 ? p->apre : p->apost));
    int64_t w = m->nsel;
    if ((m->n + 1) * w > m->cap) {
        int64_t c = m->cap ? m->cap : 1024 * w;
        while (c < (m->n + 1) * w) c *= 2;
        m->data = (double *)realloc(m->data, sizeof(double) * c);
        m->cap = c;
    }
    for (int64_t j = 0; j < w; ++j) m->data[m->n * w + j] = rget(net, src, m->sel[j]);
    m->n++;
}

static void mon_spikes(ref_net *net, ref_mon *m, const ref_group *g)
{
    if (m->n + g->nspk > m->cap) {
        int64_t c = m->cap ? m->cap : 1024;
        while (c < m->n + g->nspk) c *= 2;
        m->i = (int32_t *)realloc(m->i, sizeof(int32_t) * c);
        m->step = (int64_t *)realloc(m->step, sizeof(int64_t) * c);
        m->cap = c;
    }
    for (int64_t k = 0; k < g->nspk; ++k) {
        m->i[m->n] = g->spk[k];
        m->step[m->n] = net->step;
        m->n++;
    }
}

static void mon_state(ref_net *net, ref_mon *m, const ref_group *g)
{
    if (m->every > 1 && net->step % m->every != 0) return;
    int64_t w = m->stop - m->start;
    if ((m->n + 1) * w > m->cap) {
        int64_t c = m->cap ? m->cap : 1024 * w;
        while (c < (m->n + 1) * w) c *= 2;
        m->data = (double *)realloc(m->data, sizeof(double) * c);
        m->cap = c;
    }
    for (int64_t j = 0; j < w; ++j)
        m->data[m->n * w + j] = rget(net, g->var[m->var], m->sel ? m->sel[j] : m->start + j);
    m->n++;
}

/* ratemonitor.cu:52-99: spikes of neurons [start, stop) this step (stored as a count) */
static void mon_rate(ref_net *net, ref_mon *m, const ref_group *g)
{
    (void)net;
    if (m->n + 1 > m->cap) {
        int64_t c = m->cap ? 2 * m->cap : 1024;
        m->data = (double *)realloc(m->data, sizeof(double) * c);
        m->cap = c;
    }
    int64_t c = 0;
    for (int64_t k = 0; k < g->nspk; ++k) c += g->spk[k] >= m->start && g->spk[k] < m->stop;
    m->data[m->n++] = (double)c;
}

int ref_net_add_rate_monitor(ref_net *net, int gid, int start, int stop)
{
    if (net->nm >= REF_MAX_MONS || gid < 0 || gid >= net->ng) return -1;
    if (start < 0 || stop > net->g[gid].n || stop <= start) return -1;
    ref_mon *m = &net->mon[net->nm];
    memset(m, 0, sizeof(*m));
    m->kind = 2;
    m->gid = gid; m->start = start; m->stop = stop;
    return net->nm++;
}

/* rate = 1.0*num_spikes/dt/N (ratemonitor.cu:98) */
int ref_mon_get_rate(ref_net *net, int mid, double *out)
{
    if (mid < 0 || mid >= net->nm || net->mon[mid].kind != 2) return -1;
    ref_mon *m = &net->mon[mid];
    for (int64_t k = 0; k < m->n; ++k) out[k] = 1.0 * m->data[k] / net->dt / (double)(m->stop - m->start);
    return 0;
}

static void queue_push_in(int64_t **q, int64_t *qlen, int64_t *qcap, int slot, int64_t syn)
{
    if (qlen[slot] == qcap[slot]) {
        int64_t c = qcap[slot] ? 2 * qcap[slot] : 256;
        q[slot] = (int64_t *)realloc(q[slot], sizeof(int64_t) * c);
        qcap[slot] = c;
    }
    q[slot][qlen[slot]++] = syn;
}

static void queue_push(ref_path *p, int slot, int64_t syn) { queue_push_in(p->q, p->qlen, p->qcap, slot, syn); }

/* contiguous share [*lo, *hi) of n items for thread t of T */
static void ref_share(int64_t n, int t, int T, int64_t *lo, int64_t *hi)
{
    *lo = n * t / T;
    *hi = n * (t + 1) / T;
}

/* the per-precision body: update groups, push/deliver, reset */
#define R double
#define RSFX(name) name##_f64
#include "cpu_ref_step.inc"
#undef R
#undef RSFX
#define R float
#define RSFX(name) name##_f32
#include "cpu_ref_step.inc"
#undef R
#undef RSFX

int ref_stateupdate(int model, int precision, int n, double **vars, const int32_t *last, const double *params,
                    double dt, uint64_t seed, int64_t step)
{
    if (model_nvars(model) < 0) return -1;
    if (precision == HS_F64) return ref_stateupdate_f64(model, n, vars, last, params, dt, seed, step);
    if (precision == HS_F32) return ref_stateupdate_f32(model, n, vars, last, params, dt, seed, step);
    return -1;
}

int ref_net_run(ref_net *net, int64_t nsteps)
{
    for (int64_t s = 0; s < nsteps; ++s) {
        if (net->prec == HS_F64) ref_step_f64(net);
        else ref_step_f32(net);
        if (net->thr_err) return -5;
    }
    if (net->sem == 1) {
        if (net->prec == HS_F64) ref_flush_f64(net);
        else ref_flush_f32(net);
    }
    return 0;
}

int64_t ref_net_step(ref_net *net) { return net->step; }
uint64_t ref_net_events(ref_net *net) { return net->events; }

int64_t ref_mon_count(ref_net *net, int mid)
{
    if (mid < 0 || mid >= net->nm) return -1;
    return net->mon[mid].n;
}

int ref_mon_get_spikes(ref_net *net, int mid, int32_t *i, int64_t *step)
{
    if (mid < 0 || mid >= net->nm || net->mon[mid].kind != 0) return -1;
    ref_mon *m = &net->mon[mid];
    if (m->n) {
        memcpy(i, m->i, sizeof(int32_t) * m->n);
        memcpy(step, m->step, sizeof(int64_t) * m->n);
    }
    return 0;
}

int ref_mon_get_state(ref_net *net, int mid, double *out)
{
    if (mid < 0 || mid >= net->nm || (net->mon[mid].kind != 1 && net->mon[mid].kind != 3)) return -1;
    ref_mon *m = &net->mon[mid];
    const int64_t w = m->kind == 3 ? m->nsel : m->stop - m->start;
    if (m->n) memcpy(out, m->data, sizeof(double) * m->n * w);
    return 0;
}

void ref_net_destroy(ref_net *net)
{
    if (!net) return;
    for (int g = 0; g < net->ng; ++g) {
        for (int v = 0; v < HS_MAX_VARS; ++v) { free(net->g[g].var[v]); free(net->g[g].acc[v]); }
        free(net->g[g].lastspike); free(net->g[g].spk); free(net->g[g].pend);
        free(net->g[g].gen_off); free(net->g[g].gen_idx);
    }
    for (int p = 0; p < net->np; ++p) {
        ref_path *pw = &net->pw[p];
        free(pw->pre); free(pw->post); free(pw->delay); free(pw->pre_start); free(pw->by_pre); free(pw->w);
        for (int q = 0; q < pw->Q; ++q) free(pw->q[q]);
        free(pw->q); free(pw->qlen); free(pw->qcap);
        if (pw->tq)
            for (int64_t q = 0; q < (int64_t)net->nthr * pw->Q; ++q) free(pw->tq[q]);
        free(pw->tq); free(pw->tqlen); free(pw->tqcap);
        free(pw->apre); free(pw->apost); free(pw->lastup); free(pw->post_start); free(pw->by_post);
        for (int k = 0; k < pw->nsv; ++k)
            if (pw->sv[k] != pw->w) free(pw->sv[k]); /* the weight array is freed above */
    }
    for (int m = 0; m < net->nm; ++m) {
        free(net->mon[m].i); free(net->mon[m].step); free(net->mon[m].data); free(net->mon[m].sel);
    }
    free(net);
}
