/*
 * cpu_ref.h -- C interface of the oracle (test infrastructure only).
 * See cpu_ref.c for what each part restates.
 */
#ifndef CPU_REF_H
#define CPU_REF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void ref_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
double ref_randn(uint64_t seed, uint32_t stream, uint32_t index, int64_t step);
float ref_randn_f32(uint64_t seed, uint32_t stream, uint32_t index, int64_t step);
double ref_rand(uint64_t seed, uint32_t stream, uint32_t index, int64_t step);
int ref_quantize_delay(double delay, double dt);
/* Synapses.connect(p) with per-synapse delay and weight draws, creation order;
   returns the synapse count, writes the first min(count, cap) entries */
int64_t ref_connect_random(int pre_lo, int pre_hi, int post_lo, int post_hi, double p, int skip_self, uint64_t seed,
                           uint32_t stream, double delay_lo, double delay_hi, double dt, double w_lo, double w_hi,
                           int32_t *pre, int32_t *post, int32_t *delay_steps, double *w, int64_t cap);
int64_t ref_connect_fixed(int pre_lo, int pre_hi, int post_lo, int post_hi, int K, int skip_self, uint64_t seed,
                          uint32_t stream, double delay_lo, double delay_hi, double dt, double w_lo, double w_hi,
                          int32_t *pre, int32_t *post, int32_t *delay_steps, double *w, int64_t cap);
/* the same with a condition and a per-pair probability compiled from code
   (function pointers, NULL: none) and the fixed out-degree K (> 0) */
int64_t ref_connect_code(int pre_lo, int pre_hi, int post_lo, int post_hi, double p, int K, int ncopy, int skip_self,
                         uint64_t seed, uint32_t stream, double delay_lo, double delay_hi, double dt, double w_lo,
                         double w_hi, void *cond_fn, void *p_fn, void *w_fn, const void *const *pre_vars,
                         const void *const *post_vars, int32_t *pre, int32_t *post, int32_t *delay_steps, double *w,
                         int64_t cap);
long long ref_timestep(double t, double dt);
double ref_exp(double x);
float ref_expf(float x);
double ref_log(double x);
double ref_sqrt(double x);
double ref_cos2pi(double u);
float ref_logf(float x);
float ref_sqrtf(float x);
float ref_cos2pif(float u);
int ref_model_constants(int model, const double *params, int nparams, double dt, double *k);
int ref_stateupdate(int model, int precision, int n, double **vars, const int32_t *last, const double *params,
                    double dt, uint64_t seed, int64_t step);

typedef struct ref_net ref_net;
ref_net *ref_net_create(int precision, int semantics, double dt, uint64_t seed);
/* the step loop over nthreads OpenMP threads, bit-identical to the serial one;
   before the first step (-2: built without OpenMP) */
int ref_net_set_threads(ref_net *net, int nthreads);
int ref_net_add_group(ref_net *net, int model, int n, const double *params, int nparams);
int ref_net_add_group_code(ref_net *net, int n, int nvars, void *step_fn, void *reset_fn, double refractory,
                           int noise, uint32_t stream);
int ref_net_set_state(ref_net *net, int gid, int var, const double *values);
int ref_net_get_state(ref_net *net, int gid, int var, double *out);
int ref_net_get_lastspike(ref_net *net, int gid, double *out);
int ref_net_add_pathway(ref_net *net, int src, int tgt, int var, int64_t nsyn, const int32_t *pre,
                        const int32_t *post, const int32_t *delay_steps, int scalar_delay,
                        int wmode, const double *weights, double wscalar);
int ref_net_set_spikegen(ref_net *net, int gid, int64_t n, const int32_t *idx, const int64_t *steps);
int ref_net_pathway_stdp(ref_net *net, int pid, const double *params, int nparams);
int ref_net_pathway_code(ref_net *net, int pid, void *fn, int nsv, int w_sv, int npost, const int *post_vars, int npre,
                         const int *pre_vars);
int ref_net_pathway_code_post(ref_net *net, int pid, void *fn, int npost, const int *post_vars);
int ref_net_set_synapse_var(ref_net *net, int pid, int var, const double *values);
int ref_net_get_synapse_var(ref_net *net, int pid, int var, double *out);
int ref_net_add_rate_monitor(ref_net *net, int gid, int start, int stop);
int ref_mon_get_rate(ref_net *net, int mid, double *out);
int ref_net_add_spike_monitor(ref_net *net, int gid);
int ref_net_add_state_monitor(ref_net *net, int gid, int var, int start, int stop);
int ref_net_add_state_monitor_indices(ref_net *net, int gid, int var, const int32_t *idx, int n, int every);
int ref_net_set_monitor_interval(ref_net *net, int mid, int every);
int ref_net_add_synapse_monitor(ref_net *net, int pid, int var, const int32_t *sel, int n);
int ref_net_run(ref_net *net, int64_t nsteps);
int64_t ref_net_step(ref_net *net);
uint64_t ref_net_events(ref_net *net);
int64_t ref_mon_count(ref_net *net, int mid);
int ref_mon_get_spikes(ref_net *net, int mid, int32_t *i, int64_t *step);
int ref_mon_get_state(ref_net *net, int mid, double *out);
void ref_net_destroy(ref_net *net);

/* multicompartment trees (cpu_ref_mc.c) */
typedef struct ref_mc ref_mc;
ref_mc *ref_mc_create(int precision, double dt, int ntrees, const int32_t *tree_comp, const int32_t *tree_sec,
                      const int32_t *sec_first, const int32_t *sec_n, const int32_t *sec_parent, const double *area,
                      const double *rl1, const double *rl2, const double *Cm, double Ri, const double *chan,
                      double vth);
int ref_mc_set(ref_mc *M, int var, const double *values);
int ref_mc_get(ref_mc *M, int var, double *out);
int ref_mc_run(ref_mc *M, int64_t nsteps);
int ref_mc_spike_counts(ref_mc *M, int32_t *out);
void ref_mc_destroy(ref_mc *M);

#ifdef __cplusplus
}
#endif
#endif
