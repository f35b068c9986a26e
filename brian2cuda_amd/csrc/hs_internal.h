/*
 * hs_internal.h -- device-side descriptors of the per-timestep engine and the
 * launchers shared by the translation units of libhs.
 *
 * HBM layout of one network (all arrays structure-of-arrays, 256-B aligned):
 *   group  : var[v][n] (R), last[n] (int32 step of the last spike),
 *            acc[v][nacc][n] (R, fused form only: the effects due at step t are
 *            summed in slot t mod nacc, nacc = window + 1),
 *            hist[Q][n] (int32 spike lists, one per step of a ring of
 *            Q = dmax + 2*window steps) with hist_cnt[Q]
 *   pathway: bstart[n_src*D+1] (int32 offsets of the (pre, delay) bundles),
 *            post[S] (int32 target index), w[S] or w[n_src] (R) by weight mode
 *   control: HsCtl (step base, event counter, reset guard)
 */
#ifndef HS_INTERNAL_H
#define HS_INTERNAL_H

#ifdef __HIPCC_RTC__
/* compiled at run time (generated neuron models, hs_codegen.hip): hipRTC
   provides the HIP device API; the fixed-width integer types come from here */
typedef signed char int8_t;
typedef unsigned char uint8_t;
typedef signed int int32_t;
typedef unsigned int uint32_t;
typedef signed long long int64_t;
typedef unsigned long long uint64_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif
#include "../../include/hs_models.h"

#define HS_MAX_STATE_MONS 4
/* kernel specialisation of HS_MODEL_HH_COBA in the FP32 fast form (option 11);
   groups keep their model id, the runtime selects this one for the kernels */
#define HS_MODEL_HH_FAST 103
#define HS_MAX_DELAY_SLOTS 4096 /* D per pathway (LDS prefix table) */
#define HS_MAX_WINDOW 8         /* steps per update launch of the windowed form */
#define HS_GTILE 256            /* targets per block of the gather delivery */

/* delivery modes of a pathway */
#define HS_DELIVER_SCATTER 1 /* per (pre, delay) bundles, global atomics on the targets */
#define HS_DELIVER_GATHER 2  /* per (pre, target tile) segments, LDS accumulation per tile */
#define HS_DELIVER_PLASTIC 3 /* scatter bundles with STDP (on_pre and on_post kernels, plain form) */
#define HS_DELIVER_CODE 4    /* scatter bundles whose on_pre is code (hs_net_pathway_code) */

/* The delivered-event count is split over HS_EV_SLOTS counters, one per 128-byte
   line: thousands of waves add to it per launch, and same-address atomics
   serialise in the L2 (one counter cost ~10 ns per wave). */
#define HS_EV_SLOTS 256
#define HS_EV_STRIDE 16
/* the spike count of each step of a group's ring sits on its own 128-byte line:
   the blocks of an update launch reserve list space in w counters at once */
#define HS_CNT_STRIDE 32

struct HsCtl {
    long long base;            /* simulation step of chunk offset 0 */
    long long no_reset_before; /* resets of spikes at steps < this were already applied */
    unsigned long long pending; /* pipelined windows: events of emitted spikes still in flight at base */
    unsigned long long scatter_bad; /* hs_scatter_add: indices outside the target array */
    unsigned long long pad_[12];
    unsigned long long ev_slot[HS_EV_SLOTS * HS_EV_STRIDE];  /* synaptic events delivered: sum of ev_slot[k*STRIDE] */
    unsigned long long spk_slot[HS_EV_SLOTS * HS_EV_STRIDE]; /* spikes emitted, spread the same way */
};

static inline unsigned long long hs_ctl_events(const HsCtl &c)
{
    unsigned long long e = 0;
    for (int k = 0; k < HS_EV_SLOTS; ++k) e += c.ev_slot[k * HS_EV_STRIDE];
    return e;
}

static inline unsigned long long hs_ctl_spikes(const HsCtl &c)
{
    unsigned long long e = 0;
    for (int k = 0; k < HS_EV_SLOTS; ++k) e += c.spk_slot[k * HS_EV_STRIDE];
    return e;
}

template <typename R>
struct DGroup {
    int model, n, nvars, Q;
    int has_reset, ref_steps;
    unsigned int stream;
    int nst;
    R *var[HS_MAX_VARS];
    R *acc[HS_MAX_VARS]; /* fused form: ring of nacc step slots of n values (nullptr: not a target) */
    int nacc;
    int32_t *last;
    int32_t *hist;
    int32_t *hist_cnt;
    R k[16];
    R vth, vr;
    double dt; /* the network's dt (seconds): dt and t of code models */
    unsigned long long seed;
    /* spike monitor */
    int32_t *smon_i;
    int32_t *smon_t;
    unsigned long long *smon_n;
    long long smon_cap;
    /* state monitors (when='start') */
    int st_var[HS_MAX_STATE_MONS], st_start[HS_MAX_STATE_MONS], st_stop[HS_MAX_STATE_MONS];
    long long st_t0[HS_MAX_STATE_MONS], st_rows[HS_MAX_STATE_MONS];
    R *st_buf[HS_MAX_STATE_MONS];
    /* recorded indices: st_map[m][i] = column of neuron i or -1 (nullptr: the
       range [st_start, st_stop)); a row every st_every steps from st_t0, st_width columns */
    const int32_t *st_map[HS_MAX_STATE_MONS];
    int st_every[HS_MAX_STATE_MONS], st_width[HS_MAX_STATE_MONS];
    int tile_begin; /* first tile index of this group in the concatenated tile space */
    R *xi;          /* windowed form: noise ring [xi_slots][n], draw of step t in slot t mod xi_slots (or nullptr) */
    int xi_slots;
    /* the spike list of a ring slot is split into nsub (a power of two) sub-lists of subcap
       entries: neuron i goes to sub-list (i / 256) mod nsub, so that the blocks
       of a large group reserve list space on nsub counters, not one.  List of
       (slot, s): hist[slot*nsub*subcap + s*subcap ..], count at
       hist_cnt[(slot*nsub + s) * HS_CNT_STRIDE] */
    int nsub, subcap;
    /* SpikeGeneratorGroup: neurons gen_idx[gen_off[s] .. gen_off[s+1]) (ascending)
       fire at step s < gen_steps; step t plays s = t mod gen_period (0: s = t) */
    const int32_t *gen_off;
    const int32_t *gen_idx;
    long long gen_steps, gen_period;
    /* PopulationRateMonitor: spikes of neurons [rmon_start, rmon_stop) at step t
       are added to rmon_cnt[t - rmon_t0] (rows [0, rmon_rows)), or nullptr */
    int32_t *rmon_cnt;
    long long rmon_t0, rmon_rows;
    int rmon_start, rmon_stop;
};

template <typename R>
struct DPath {
    int src, tgt, var;
    int n_src, dmin, D, Q, wmode;
    const int32_t *hist;     /* source group's spike ring (sub-lists as in DGroup) */
    const int32_t *hist_cnt; /* count of (ring slot, sub-list) at hist_cnt[(slot*nsub + s) * cstride] */
    int cstride, nsub, subcap;
    const int32_t *bstart;
    const int32_t *post;
    const R *w;
    R wscalar;
    R *target;               /* plain: the state variable */
    R *acc;                  /* fused: the target variable's accumulator ring */
    int nacc, n_tgt;
    /* gather layout: synapses sorted by (target tile, pre), synapse order kept;
       segment of (tile, pre) = gsyn[goff[tile*n_src+pre] .. goff[tile*n_src+pre+1]),
       entry = (delay - dmin) << 16 | (post - tile*HS_GTILE); weights follow the
       same order */
    int dmode, n_tiles;
    const int32_t *goff;
    const uint32_t *gsyn;
    /* pipelined windows: delay slots [0, nnear) are pushed by the emitting
       update, [nnear, D) pulled by the delivery blocks */
    int nnear;
    /* STDP (dmode HS_DELIVER_PLASTIC): weights wm (the array w), traces and
       last update step per synapse in bundle order; the synapses of target
       neuron j are qsyn[qstart[j] .. qstart[j+1]); the target group's spike
       ring for on_post */
    R *wm, *apre, *apost;
    int32_t *lastup;
    const int32_t *qstart, *qsyn;
    /* synaptic code's rand() / randn(): the network's seed and the pathway's index */
    unsigned long long seed;
    unsigned int pid;
    double taupre, taupost, dt;
    R sk[5]; /* dApre, dApost, wmin, wmax, scale */
    const int32_t *thist, *thist_cnt;
    int tQ, tnsub, tsubcap;
    /* synaptic code (dmode HS_DELIVER_CODE): code pathway cg of the network;
       nsv synapse variables sv[k][q] in bundle order; the code adds to target
       variables post_var[0..npost) and reads source variables pre_var[0..npre) */
    int cg, nsv, npost, npre;
    R *sv[HS_MAX_SYN_VARS];
    int post_var[HS_MAX_VARS], pre_var[HS_MAX_VARS];
    /* on_post: reads target variables post_read_var[0..npr); synpre[q] is the
       source neuron of synapse q (bundle order); lastup (above) the step of the
       synapse's last on_pre or on_post */
    int npr;
    int post_read_var[HS_MAX_VARS];
    const int32_t *synpre;
};

/*
 * Random connectivity built on the device (hs_connect.hip).  Row i (a
 * presynaptic neuron in [pre_lo, pre_hi)) selects targets j in [post_lo, post_hi)
 * with probability p; draw k of row i comes from the Philox counter
 * (i, k, HS_CONN_CTR | tag, stream) under key (seed_lo, seed_hi):
 *   HS_CONN_TAG_SKIP  the geometric skips of p < 0.25 (two draws per counter: words 0-1, then 2-3)
 *   HS_CONN_TAG_PAIR  one test per pair for p >= 0.25 (counter k = (j - post_lo) / 2, same halves)
 *   HS_CONN_TAG_SYN   per synapse (i, j), counter k = j: delay uniform from words 0-1,
 *                     weight uniform from words 2-3
 * The high half of word 2 marks the counter as a connectivity draw, so it never
 * meets a noise draw (index, step_lo, step_hi, stream) of a run shorter than 2^62 steps.
 */
#define HS_CONN_CTR 0x434E0000u
#define HS_CONN_TAG_SKIP 0u
#define HS_CONN_TAG_SYN 1u
#define HS_CONN_TAG_PAIR 2u
#define HS_CONN_TAG_FIX 3u
/* connect(n=ncopy): copy c of a pair draws its delay and weight from tag
   HS_CONN_TAG_SYN | c << 4 (copy 0 as a single synapse), c < HS_CONN_NMAX */
#define HS_CONN_NMAX 16
#define HS_CONN_TAG_COPY(c) (HS_CONN_TAG_SYN | ((uint32_t)(c) << 4))
/* fixed out-degree (connect(j='k for k in sample(N_post, size=K)')): row i's draw
   c (c = 0, 1, ...) is the 53-bit uniform u of words 0-1 of counter
   (i, c, HS_CONN_CTR | HS_CONN_TAG_FIX, stream), value min(floor(u * n), n - 1),
   n = post_hi - post_lo.  Draws 0..K-1 first; after sorting, repeated values
   are dropped and the missing ones drawn with the next counters, until K
   distinct values remain; value x is target post_lo + x, visited in ascending
   order, and j == i is then dropped when autapses == 0 (the condition
   "if k != i" of the generator). */
#define HS_CONN_KMAX 2048

struct HsConnSpec {
    int pre_lo, pre_hi, post_lo, post_hi;
    int autapses;    /* 0: j == i is dropped (source and target are the same group) */
    int jump;        /* p < 0.25: geometric skips; else one draw per pair */
    double p;
    double pconst;   /* 1 / log(1 - p) (jump form) */
    unsigned long long seed;
    unsigned int stream;
    double dlo, dspan, dt; /* delay = (int)((dlo + u*dspan)/dt + 0.5) steps */
    int dmin, D;           /* delay slot = delay - dmin, in [0, D) */
    double wlo, wspan;     /* weight = wlo + u*wspan */
    int fixed_k;           /* > 0: fixed out-degree K (then p and jump are unused) */
    int ncopy;             /* synapses per selected pair (connect(n=...)); 0 or 1: one */
};

/* the state arrays (precision of the network) of the source and target groups,
   for connect conditions and probabilities that read them */
struct HsConnVars {
    const void *pre[HS_MAX_VARS];
    const void *post[HS_MAX_VARS];
};

#ifndef __HIPCC_RTC__ /* host launchers */
/* count pass: cnt[i*D + slot] (int32, zeroed by the caller) for rows [pre_lo, pre_hi);
   *err becomes nonzero if a slot falls outside [0, D) */
hipError_t hs_launch_conn_count(const HsConnSpec &c, int32_t *cnt, int *err, hipStream_t s);
/* write pass: post[] and w[] (R = double or float by prec, nullptr: no weights) at
   bstart[i*D+slot] + the synapse's rank in its bundle */
hipError_t hs_launch_conn_write(int prec, const HsConnSpec &c, const int32_t *bstart, int32_t *post, void *w,
                                int *err, hipStream_t s);
/* launch shape of the two passes: waves (rows in flight) per block, dynamic LDS bytes, grid */
int hs_conn_waves(const HsConnSpec &c);
size_t hs_conn_lds(const HsConnSpec &c);
int hs_conn_grid(const HsConnSpec &c);
/* exclusive prefix sum of n int32 counts: out[0..n] (out[n] = total, also at
   bsum[hs_scan_blocks(n)] as int64); bsum holds hs_scan_blocks(n) + 1 entries */
int hs_scan_blocks(long long n);
hipError_t hs_launch_scan(const int32_t *in, long long n, long long *bsum, int32_t *out, hipStream_t s);

/* bundle table from synapse arrays in HBM (hs_sort.hip); minmax = {INT_MAX, -1}
   and err = 0 on entry to the check */
hipError_t hs_launch_syn_check(const int32_t *pre, const int32_t *post, const int32_t *delay, long long n, int n_src,
                               int n_tgt, int *minmax, int *err, hipStream_t s);
hipError_t hs_launch_syn_keys(const int32_t *pre, const int32_t *delay, long long n, int D, int dmin, uint32_t *keys,
                              int32_t *vals, hipStream_t s);
/* stable radix sort of (key, value) pairs on key bits [0, bits); tmp == nullptr: size query */
hipError_t hs_radix_sort_pairs(void *tmp, size_t *tmp_bytes, const uint32_t *keys, uint32_t *keys_out,
                               const int32_t *vals, int32_t *vals_out, long long n, int bits, hipStream_t s);
hipError_t hs_launch_syn_gather(int prec, const int32_t *order, long long n, const int32_t *post, const double *w,
                                int32_t *post_out, void *w_out, hipStream_t s);
hipError_t hs_launch_syn_convert(int prec, const double *w, long long n, void *out, hipStream_t s);
hipError_t hs_launch_syn_bstart(const uint32_t *keys, long long n, long long nb, int32_t *bstart, hipStream_t s);

/* runtime services for the translation units (hs_runtime.hip) */
hipStream_t hs_rt_stream();
int hs_rt_cus();
int hs_rt_need_init();
int hs_rt_fail(int code, const char *fmt, ...);

/* launchers (hs_ops.hip, hs_engine.hip) -- all on the runtime stream, return hipError_t */
hipError_t hs_launch_threshold(int prec, const void *v, int n, double vth, int32_t *last,
                               long long step, int ref_steps, int sorted, int32_t *spikes,
                               int32_t *count, int32_t *scratch, hipStream_t s);
hipError_t hs_launch_randn(unsigned long long seed, unsigned int stream, long long step, int n,
                           int prec, void *out, int uniform, hipStream_t s);
hipError_t hs_launch_stateupdate(int model, int prec, int n, void *const *vars, const int32_t *last,
                                 const double *k, int nk, unsigned long long seed, unsigned int stream,
                                 int ref_steps, long long step, hipStream_t s);
hipError_t hs_launch_math(int fn, int n, const double *x, double *y, hipStream_t s);
hipError_t hs_launch_synmon(int prec, const void *src, const int32_t *idx, int n, void *buf, long long rows,
                            long long t0, int every, const HsCtl *ctl, int koff, hipStream_t s);
hipError_t hs_launch_scatter_add(int prec, const int32_t *idx, long long n, int n_target, double w, void *target,
                                 unsigned long long *bad, hipStream_t s);

/* one pathway at step ctl->base, effects in place */
template <typename R>
hipError_t hs_launch_deliver_one_ctl(const DPath<R> &p, HsCtl *ctl, unsigned long long *events, int blocks,
                                     hipStream_t s);

/*
 * Engine kernels over device descriptor arrays; every kernel reads the step base
 * from ctl (t = ctl->base + koff) so a captured graph replays for any step.
 *   update : steps t .. t+w-1 of every group (plain: w == 1); wmax = the
 *            engine's window, for the ring slots it clears ahead
 *   deliver: bundles due in steps t .. t+w-1 (plain: w == 1, in place)
 *   fused  : update of step t on blocks [0, n_upd) and delivery of step t on
 *            blocks [n_upd, n_upd + n_del), one launch (every delay >= 1 step)
 */
/* ms: the model shared by every group (specialised kernel), 0 for mixed networks */
/* grid update blocks plus n_noise blocks drawing the noise of steps t+w .. t+w+wmax-1 */
template <typename R>
hipError_t hs_launch_update(const DGroup<R> *groups, int ng, int total_tiles, HsCtl *ctl, int koff, int w,
                            int wmax, int fused, int ms, int block, int grid, int n_noise, hipStream_t s);
/* noise draws of steps t .. t+nsteps-1 into the groups' noise rings */
template <typename R>
hipError_t hs_launch_noise(const DGroup<R> *groups, int ng, HsCtl *ctl, int koff, int nsteps, int grid,
                           hipStream_t s);
/* grid delivery blocks plus n_noise blocks drawing the noise of steps t .. t+w-1 */
template <typename R>
hipError_t hs_launch_deliver(const DPath<R> *paths, int np, const DGroup<R> *groups, int ng, HsCtl *ctl, int koff,
                             int w, int fused, int block, int grid, int n_noise, hipStream_t s);
template <typename R>
hipError_t hs_launch_fused(const DGroup<R> *groups, int ng, int total_tiles, const DPath<R> *paths, int np,
                           HsCtl *ctl, int koff, int ms, int block, int n_upd, int n_del, hipStream_t s);
/* gather-mode pathways (scatter-mode ones are skipped; hs_launch_deliver skips gather ones) */
template <typename R>
hipError_t hs_launch_gather(const DPath<R> *paths, int np, HsCtl *ctl, int koff, int w, int fused, int block,
                            int grid, hipStream_t s);
template <typename R>
hipError_t hs_launch_finalize(const DGroup<R> *groups, int ng, int total_tiles, HsCtl *ctl, int fused,
                              int block, hipStream_t s);
/* pipelined windows: one launch per window (update + near push | far pull | next noise) */
template <typename R>
hipError_t hs_launch_window(const DGroup<R> *groups, int ng, int total_tiles, const DPath<R> *paths, int np,
                            HsCtl *ctl, int koff, int w, int W, int ms, int block, int n_upd, int n_far, int n_noise,
                            hipStream_t s);
/* events in flight at ctl->base into ctl->pending */
template <typename R>
hipError_t hs_launch_pending(const DPath<R> *paths, int np, HsCtl *ctl, hipStream_t s);
hipError_t hs_launch_tick(HsCtl *ctl, long long k, hipStream_t s);
/* resets of spikes before ctl->base are applied (after a finalize) */
hipError_t hs_launch_mark_reset(HsCtl *ctl, hipStream_t s);
/* STDP pathways at step ctl->base + koff (plain form, after the delivery of the
   step): on_pre of the bundles due now, then on_post of the targets' spikes */
template <typename R>
hipError_t hs_launch_stdp(const DPath<R> *paths, int np, HsCtl *ctl, int koff, int grid, hipStream_t s);
#endif /* !__HIPCC_RTC__ */

#endif /* HS_INTERNAL_H */
