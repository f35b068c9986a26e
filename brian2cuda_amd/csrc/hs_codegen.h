/*
 * hs_codegen.h -- code models: neuron groups whose update, threshold and reset
 * are given as code, compiled with hipRTC into the engine's own kernels when a
 * network holding them is built (hs_codegen.hip).  Host side only.
 */
#ifndef HS_CODEGEN_H
#define HS_CODEGEN_H

#include <string>
#include <vector>

#include "hs_internal.h"

/* one code model: variable v is names[v]; update and reset are C statements,
   threshold a C expression (empty: never) */
struct HsCgModel {
    std::vector<std::string> names;
    std::string update, threshold, reset;
    int noisy = 0;
};

/* one pathway's synaptic code: the on_pre statements over the synapse
   variables syn_names, the additions <post_names[k]>_post to target variables
   and the source variables <pre_names[k]>_pre */
struct HsSynCgModel {
    std::vector<std::string> syn_names, post_names, pre_names;
    std::string on_pre;
    /* on_post (empty: none): statements for each synapse of a spiking target
       neuron, reading the target variables <post_read_names[k]>_post */
    std::string on_post;
    std::vector<std::string> post_read_names;
};

/* the engine kernels of one network, compiled for its code models (null when
   the network has none) and its synaptic code (null when none) */
struct HsCgKernels {
    hipFunction_t update[2] = {nullptr, nullptr};   /* plain, fused */
    hipFunction_t window = nullptr;
    hipFunction_t window_o8 = nullptr;              /* 8 waves per SIMD (FP32, one model) */
    hipFunction_t fused = nullptr;
    hipFunction_t finalize[2] = {nullptr, nullptr}; /* plain, fused */
    hipFunction_t syn_code = nullptr;
    hipFunction_t syn_post = nullptr;
};

/* the two connect passes compiled with a condition and a per-pair probability */
struct HsCgConnKernels {
    hipFunction_t count = nullptr, write = nullptr;
};

/* checks a model's names and code; returns HS_OK or fails with a message */
int hs_cg_check(const HsCgModel &m);
int hs_cg_check_syn(const HsSynCgModel &m);
/* the program compiled for a network's code models and synaptic code in precision prec */
std::string hs_cg_source(int prec, const std::vector<HsCgModel> &models, const std::vector<HsSynCgModel> &syn);
/* compiled kernels for these models (compiled on first request, then cached
   for the process by precision, specialisation and program text); ms: the
   model every group of the network has (HS_MODEL_CODE + c: kernels
   specialised for it) or 0 (every model dispatched per tile); fails with the
   compiler's log */
int hs_cg_kernels(int prec, int ms, const std::vector<HsCgModel> &models, const std::vector<HsSynCgModel> &syn,
                  const HsCgKernels **out);

/* the engine launches of hs_engine.hip, through the compiled kernels (same
   arguments and checks as hs_launch_update / _window / _fused / _finalize) */
hipError_t hs_cg_launch_update(const HsCgKernels *K, const void *groups, int ng, int total_tiles, HsCtl *ctl, int koff,
                               int w, int wmax, int fused, int block, int grid, int n_noise, hipStream_t s);
hipError_t hs_cg_launch_window(const HsCgKernels *K, const void *groups, int ng, int total_tiles, const void *paths,
                               int np, HsCtl *ctl, int koff, int w, int W, int block, int n_upd, int n_far,
                               int n_noise, hipStream_t s);
hipError_t hs_cg_launch_fused(const HsCgKernels *K, const void *groups, int ng, int total_tiles, const void *paths,
                              int np, HsCtl *ctl, int koff, int block, int n_upd, int n_del, hipStream_t s);
hipError_t hs_cg_launch_finalize(const HsCgKernels *K, const void *groups, int ng, int total_tiles, HsCtl *ctl,
                                 int fused, int block, hipStream_t s);
/* the synaptic code of step ctl->base + koff, after the step's update and delivery:
   on_pre of the bundles due, then on_post of the targets' spikes */
hipError_t hs_cg_launch_syn_code(const HsCgKernels *K, const void *paths, int np, const void *groups, HsCtl *ctl,
                                 int koff, int fused, int grid, hipStream_t s);
hipError_t hs_cg_launch_syn_post(const HsCgKernels *K, const void *paths, int np, const void *groups, HsCtl *ctl,
                                 int koff, int grid, hipStream_t s);

/* connect(condition=..., p=...): cond a C expression (empty: none) over i, j
   (indices within the source and target ranges), N_pre, N_post, hs_pre(k) and
   hs_post(k) (variable k of the source / target neuron, as double); p a C
   expression of the pair's probability (empty: the constant of the spec) */
int hs_cg_conn_check(const std::string &cond, const std::string &p);
int hs_cg_conn_kernels(int prec, const std::string &cond, const std::string &p, const std::string &w, bool copies,
                       const HsCgConnKernels **out);
hipError_t hs_cg_launch_conn_count(const HsCgConnKernels *K, const HsConnSpec &c, const HsConnVars &a, int32_t *cnt,
                                   int *err, hipStream_t s);
hipError_t hs_cg_launch_conn_write(const HsCgConnKernels *K, const HsConnSpec &c, const HsConnVars &a,
                                   const int32_t *bstart, int32_t *post, void *w, int *err, hipStream_t s);

#endif /* HS_CODEGEN_H */
