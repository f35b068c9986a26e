/*
 * hs_runtime.hip -- the C ABI of libhs (include/hs.h): device and memory
 * management, the bundle builder, network assembly and the graph-captured step
 * loop that replaces the generated objects.cu / network.cu of the reference.
 */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <string>
#include <vector>

#include "../../include/hs.h"
#include "hs_codegen.h"
#include "hs_internal.h"
#include "hs_model_exprs.h"

/* ------------------------------------------------------------------------ */
/* errors                                                                    */
/* ------------------------------------------------------------------------ */
static thread_local std::string g_err;

static int hs_fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HS_HIP(call)                                                                                      \
    do {                                                                                                  \
        hipError_t _e = (call);                                                                           \
        if (_e != hipSuccess)                                                                             \
            return hs_fail(HS_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                           __LINE__);                                                                     \
    } while (0)

#define HS_TRY(expr)                   \
    do {                               \
        int _r = (expr);               \
        if (_r != HS_OK) return _r;    \
    } while (0)

extern "C" const char *hs_last_error(void) { return g_err.c_str(); }
extern "C" int hs_version(void) { return 1; }

/* ------------------------------------------------------------------------ */
/* runtime                                                                   */
/* ------------------------------------------------------------------------ */
struct HsRuntime {
    int device = -1;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    long long hbm = 0;
    char name[256] = {0};
    HsCtl *ctl = nullptr;          /* control block of the standalone kernels */
    int32_t *scratch = nullptr;    /* scratch of the standalone kernels */
    size_t scratch_n = 0;
    double *kbuf = nullptr;        /* constants of hs_stateupdate */
};
static HsRuntime g_rt;

/* Device memory of the allocation-free entry points (hs_threshold_compact_noalloc):
   one int32 spike count per 256-neuron tile, HS_NOALLOC_MAX_N neurons at most
   (512 KiB).  Each GPU holds its own copy until the process ends.  In host code
   the name is not a device address: hs_init resolves the address once per device
   (g_scratch_dev) and host code uses only that. */
static __device__ __align__(256) int32_t scratch[HS_NOALLOC_MAX_N / 256];
#define HS_MAX_DEVICES 64
static int32_t *g_scratch_dev[HS_MAX_DEVICES];

static int hs_need_init()
{
    if (!g_rt.stream) return hs_fail(HS_ERR_STATE, "hs_init() has not been called");
    return HS_OK;
}

hipStream_t hs_rt_stream() { return g_rt.stream; }
int hs_rt_cus() { return g_rt.n_cu; }
int hs_rt_need_init() { return hs_need_init(); }
int hs_rt_fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" int hs_init(int device)
{
    if (g_rt.stream && g_rt.device == device) return HS_OK;
    if (g_rt.stream) return hs_fail(HS_ERR_STATE, "already initialised on device %d", g_rt.device);
    int n = 0;
    HS_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return hs_fail(HS_ERR_ARG, "device %d not present (%d devices)", device, n);
    if (device >= HS_MAX_DEVICES) return hs_fail(HS_ERR_UNSUPPORTED, "device %d: at most %d devices", device, HS_MAX_DEVICES);
    HS_HIP(hipSetDevice(device));
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        HS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return hs_fail(HS_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_dev[device] = (int32_t *)p;
    }
    hipDeviceProp_t prop;
    HS_HIP(hipGetDeviceProperties(&prop, device));
    g_rt.n_cu = prop.multiProcessorCount;
    g_rt.hbm = (long long)prop.totalGlobalMem;
    snprintf(g_rt.name, sizeof(g_rt.name), "%s", prop.gcnArchName);
    HS_HIP(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
    HS_HIP(hipMalloc((void **)&g_rt.ctl, sizeof(HsCtl)));
    HS_HIP(hipMemset(g_rt.ctl, 0, sizeof(HsCtl)));
    HS_HIP(hipMalloc((void **)&g_rt.kbuf, 16 * sizeof(double)));
    g_rt.device = device;
    return HS_OK;
}

extern "C" int hs_device_info(char *name, int name_cap, int *compute_units, long long *hbm_bytes)
{
    HS_TRY(hs_need_init());
    if (name && name_cap > 0) snprintf(name, name_cap, "%s", g_rt.name);
    if (compute_units) *compute_units = g_rt.n_cu;
    if (hbm_bytes) *hbm_bytes = g_rt.hbm;
    return HS_OK;
}

extern "C" int hs_alloc(size_t bytes, void **dev_ptr)
{
    HS_TRY(hs_need_init());
    if (!dev_ptr) return hs_fail(HS_ERR_ARG, "null output pointer");
    if (hipMalloc(dev_ptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return hs_fail(HS_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    }
    return HS_OK;
}

extern "C" int hs_free(void *p)
{
    HS_TRY(hs_need_init());
    if (p) HS_HIP(hipFree(p));
    return HS_OK;
}

extern "C" int hs_h2d(void *dst, const void *src, size_t bytes)
{
    HS_TRY(hs_need_init());
    if (!bytes) return HS_OK;
    HS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

extern "C" int hs_d2h(void *dst, const void *src, size_t bytes)
{
    HS_TRY(hs_need_init());
    if (!bytes) return HS_OK;
    HS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

extern "C" int hs_zero(void *p, size_t bytes)
{
    HS_TRY(hs_need_init());
    if (!bytes) return HS_OK;
    HS_HIP(hipMemsetAsync(p, 0, bytes, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

extern "C" int hs_sync(void)
{
    HS_TRY(hs_need_init());
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

static int hs_scratch(size_t n)
{
    if (g_rt.scratch_n >= n) return HS_OK;
    if (g_rt.scratch) HS_HIP(hipFree(g_rt.scratch));
    g_rt.scratch = nullptr;
    g_rt.scratch_n = 0;
    HS_HIP(hipMalloc((void **)&g_rt.scratch, n * sizeof(int32_t)));
    g_rt.scratch_n = n;
    return HS_OK;
}

/* ------------------------------------------------------------------------ */
/* model table                                                               */
/* ------------------------------------------------------------------------ */
struct HsModelInfo {
    int nvars, nparams, has_reset;
    int i_vth, i_vr, i_ref, i_stream; /* parameter positions (-1: none) */
};

/* refractory period in steps (Brian2's timestep()), held in int32 next to the
   int32 last-spike step; longer than 2^30 steps means "never again" */
static int hs_ref_steps(double refractory, double dt)
{
    long long r = hs_timestep(refractory, dt);
    if (r < 0) r = 0;
    if (r > (1LL << 30) - 1) r = (1LL << 30) - 1;
    return (int)r;
}

static bool hs_model_info(int model, HsModelInfo *mi)
{
    if (model == HS_MODEL_LIF_CUBA) { *mi = {HS_LIF_CUBA_NVARS, HS_LIF_CUBA_NPARAMS, 1, 1, 2, 6, -1}; return true; }
    if (model == HS_MODEL_LIF_NOISE) { *mi = {HS_LIF_NOISE_NVARS, HS_LIF_NOISE_NPARAMS, 1, 3, 4, 5, 6}; return true; }
    if (model == HS_MODEL_HH_COBA) { *mi = {HS_HH_NVARS, HS_HH_NPARAMS, 0, 12, -1, 13, -1}; return true; }
    if (model == HS_MODEL_SPIKEGEN) { *mi = {HS_SPIKEGEN_NVARS, HS_SPIKEGEN_NPARAMS, 0, -1, -1, -1, -1}; return true; }
    if (model == HS_MODEL_POISSON) { *mi = {HS_POISSON_NVARS, HS_POISSON_NPARAMS, 0, -1, -1, -1, 1}; return true; }
    return false;
}

/* ------------------------------------------------------------------------ */
/* standalone kernels                                                        */
/* ------------------------------------------------------------------------ */
extern "C" int hs_threshold_compact(int precision, const void *v, int n, double vth, int32_t *last, int64_t step,
                                    int ref_steps, int sorted, int32_t *spikes, int32_t *count)
{
    HS_TRY(hs_need_init());
    if (n < 0 || (precision != HS_F64 && precision != HS_F32) || !v || !last || !spikes || !count)
        return hs_fail(HS_ERR_ARG, "hs_threshold_compact: bad arguments");
    if (n == 0) return hs_zero(count, sizeof(int32_t));
    if (sorted) HS_TRY(hs_scratch((size_t)(n + 255) / 256 + 1));
    HS_HIP(hs_launch_threshold(precision, v, n, vth, last, step, ref_steps, sorted, spikes, count, g_rt.scratch,
                               g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

extern "C" int hs_threshold_compact_noalloc(int precision, const void *v, int n, double vth, int32_t *last,
                                            int64_t step, int ref_steps, int sorted, int32_t *spikes, int32_t *count)
{
    HS_TRY(hs_need_init());
    if (n < 0 || (precision != HS_F64 && precision != HS_F32) || !v || !last || !spikes || !count)
        return hs_fail(HS_ERR_ARG, "hs_threshold_compact_noalloc: bad arguments");
    if (n > HS_NOALLOC_MAX_N)
        return hs_fail(HS_ERR_ARG, "hs_threshold_compact_noalloc: n = %d exceeds the maximum of %d neurons", n,
                       HS_NOALLOC_MAX_N);
    int32_t *tiles = g_scratch_dev[g_rt.device];
    if (!tiles) return hs_fail(HS_ERR_STATE, "hs_threshold_compact_noalloc: scratch address not resolved");
    if (n == 0) return hs_zero(count, sizeof(int32_t));
    /* the tile counts are written in full by the count pass before the scan reads
       them: nothing left by an earlier call is read */
    HS_HIP(hs_launch_threshold(precision, v, n, vth, last, step, ref_steps, sorted, spikes, count, tiles,
                               g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

extern "C" int hs_randn(uint64_t seed, uint32_t stream, int64_t step, int n, int precision, void *out)
{
    HS_TRY(hs_need_init());
    if (n < 0 || !out || (precision != HS_F64 && precision != HS_F32)) return hs_fail(HS_ERR_ARG, "hs_randn: bad arguments");
    if (n == 0) return HS_OK;
    HS_HIP(hs_launch_randn(seed, stream, step, n, precision, out, 0, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

extern "C" int hs_rand(uint64_t seed, uint32_t stream, int64_t step, int n, void *out)
{
    HS_TRY(hs_need_init());
    if (n < 0 || !out) return hs_fail(HS_ERR_ARG, "hs_rand: bad arguments");
    if (n == 0) return HS_OK;
    HS_HIP(hs_launch_randn(seed, stream, step, n, HS_F64, out, 1, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

extern "C" int hs_stateupdate(int model, int precision, int n, void *const *vars, const int32_t *last,
                              const double *params, int nparams, double dt, uint64_t seed, int64_t step)
{
    HS_TRY(hs_need_init());
    HsModelInfo mi;
    if (!hs_model_info(model, &mi) || model == HS_MODEL_SPIKEGEN || model == HS_MODEL_POISSON)
        return hs_fail(HS_ERR_UNSUPPORTED, "no state update for model %d", model);
    double pp[16];
    if (hs_model_params(model, params, nparams, pp) < 0 || !vars || !last || n < 0 ||
        (precision != HS_F64 && precision != HS_F32))
        return hs_fail(HS_ERR_ARG, "hs_stateupdate: bad arguments (model %d expects %d params)", model, mi.nparams);
    params = pp;
    double k[16] = {0};
    if (hs_model_constants(model, params, dt, k) < 0) return hs_fail(HS_ERR_ARG, "bad model");
    int ref_steps = hs_ref_steps(params[mi.i_ref], dt);
    unsigned int stream = mi.i_stream >= 0 ? (unsigned int)params[mi.i_stream] : 0u;
    if (precision == HS_F64) {
        HS_HIP(hipMemcpyAsync(g_rt.kbuf, k, sizeof(k), hipMemcpyHostToDevice, g_rt.stream));
    } else {
        float kf[16];
        for (int j = 0; j < 16; ++j) kf[j] = (float)k[j];
        HS_HIP(hipMemcpyAsync(g_rt.kbuf, kf, sizeof(kf), hipMemcpyHostToDevice, g_rt.stream));
    }
    if (n > 0)
        HS_HIP(hs_launch_stateupdate(model, precision, n, vars, last, g_rt.kbuf, 16, seed, stream, ref_steps, step,
                                     g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

extern "C" int hs_math_eval(int fn, int n, const double *x, double *y)
{
    HS_TRY(hs_need_init());
    if (fn < 0 || fn > 7 || n < 0 || !x || !y) return hs_fail(HS_ERR_ARG, "hs_math_eval: bad arguments");
    if (n == 0) return HS_OK;
    HS_HIP(hs_launch_math(fn, n, x, y, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

template <typename R>
static int hs_deliver_t(int n_src, int dmin, int D, int Q, const int32_t *hist, const int32_t *hist_cnt, int64_t step,
                        const int32_t *bstart, const int32_t *post, int wmode, const void *w, double wscalar,
                        void *target, uint64_t *events)
{
    DPath<R> p;
    memset(&p, 0, sizeof(p));
    p.n_src = n_src;
    p.dmin = dmin;
    p.D = D;
    p.Q = Q;
    p.wmode = wmode;
    p.hist = hist;
    p.hist_cnt = hist_cnt;
    p.cstride = 1;
    p.nsub = 1;
    p.subcap = n_src;
    p.bstart = bstart;
    p.post = post;
    p.w = (const R *)w;
    p.wscalar = (R)wscalar;
    p.target = (R *)target;
    HsCtl c;
    memset(&c, 0, sizeof(c));
    c.base = step;
    HS_HIP(hipMemcpyAsync(g_rt.ctl, &c, sizeof(c), hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hs_launch_deliver_one_ctl<R>(p, g_rt.ctl, (unsigned long long *)events, g_rt.n_cu * 2, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

extern "C" int hs_deliver(int precision, int n_src, int dmin, int D, int Q, const int32_t *hist,
                          const int32_t *hist_cnt, int64_t step, const int32_t *bstart, const int32_t *post, int wmode,
                          const void *w, double wscalar, void *target, uint64_t *events)
{
    HS_TRY(hs_need_init());
    if (n_src <= 0 || dmin < 0 || D <= 0 || D > HS_MAX_DELAY_SLOTS || Q < dmin + D || !hist || !hist_cnt || !bstart ||
        !post || !target || !events || wmode < 0 || wmode > 2 || (wmode != HS_W_SCALAR && !w))
        return hs_fail(HS_ERR_ARG, "hs_deliver: bad arguments");
    if (precision == HS_F64)
        return hs_deliver_t<double>(n_src, dmin, D, Q, hist, hist_cnt, step, bstart, post, wmode, w, wscalar, target,
                                    events);
    if (precision == HS_F32)
        return hs_deliver_t<float>(n_src, dmin, D, Q, hist, hist_cnt, step, bstart, post, wmode, w, wscalar, target,
                                   events);
    return hs_fail(HS_ERR_ARG, "bad precision");
}

extern "C" int hs_scatter_add(int precision, int64_t n, const int32_t *idx, int n_target, double w, void *target,
                              double *kernel_ms)
{
    HS_TRY(hs_need_init());
    if (n < 0 || n_target <= 0 || (n > 0 && (!idx || !target)))
        return hs_fail(HS_ERR_ARG, "hs_scatter_add: bad arguments");
    if (precision != HS_F32 && precision != HS_F64) return hs_fail(HS_ERR_ARG, "bad precision");
    if (kernel_ms) *kernel_ms = 0.0;
    if (n == 0) return HS_OK;
    HsCtl c;
    memset(&c, 0, sizeof(c));
    HS_HIP(hipMemcpyAsync(g_rt.ctl, &c, sizeof(c), hipMemcpyHostToDevice, g_rt.stream));
    struct Events { /* destroyed on every return path */
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    if (kernel_ms) {
        HS_HIP(hipEventCreate(&ev.e[0]));
        HS_HIP(hipEventCreate(&ev.e[1]));
        HS_HIP(hipEventRecord(ev.e[0], g_rt.stream));
    }
    HS_HIP(hs_launch_scatter_add(precision, idx, n, n_target, w, target, &g_rt.ctl->scatter_bad, g_rt.stream));
    if (kernel_ms) HS_HIP(hipEventRecord(ev.e[1], g_rt.stream));
    unsigned long long bad = 0;
    HS_HIP(hipMemcpyAsync(&bad, &g_rt.ctl->scatter_bad, sizeof(bad), hipMemcpyDeviceToHost, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    if (kernel_ms) {
        float ms = 0.0f;
        HS_HIP(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
        *kernel_ms = ms;
    }
    if (bad) return hs_fail(HS_ERR_ARG, "hs_scatter_add: %llu indices outside [0, %d)", bad, n_target);
    return HS_OK;
}

/* ------------------------------------------------------------------------ */
/* networks                                                                  */
/* ------------------------------------------------------------------------ */
struct HGroup {
    int model, n;
    HsModelInfo mi;
    double params[16];
    double k[16];
    int ref_steps;
    unsigned int stream;
    void *var[HS_MAX_VARS] = {nullptr};
    void *acc[HS_MAX_VARS] = {nullptr}; /* fused: ring of nacc slots */
    int nacc = 0;
    int32_t *last = nullptr;
    int32_t *hist = nullptr;
    int32_t *hist_cnt = nullptr;
    int Q = 0;
    int smon = -1;
    std::vector<int> stmons;
    void *xi = nullptr; /* windowed form: noise ring [xi_slots][n], drawn one window ahead */
    int xi_slots = 0;
    int nsub = 1, subcap = 0; /* spike sub-lists per ring slot (see DGroup) */
    int32_t *gen_off = nullptr, *gen_idx = nullptr; /* SpikeGeneratorGroup pattern (see DGroup) */
    long long gen_steps = 0, gen_period = 0;
    int rmon = -1; /* PopulationRateMonitor */
    int code = -1; /* code model: index into hs_net::cg_models */
};

struct HPath {
    int src, tgt, var;
    int64_t nsyn;
    int dmin, dmax, D;
    int wmode;
    double wscalar;
    int32_t *bstart = nullptr;
    int32_t *post = nullptr;
    void *w = nullptr;
    int dmode = HS_DELIVER_SCATTER;
    int n_tiles = 0;
    int32_t *goff = nullptr;
    uint32_t *gsyn = nullptr;
    /* STDP (dmode HS_DELIVER_PLASTIC) */
    void *apre = nullptr, *apost = nullptr;
    int32_t *lastup = nullptr, *qstart = nullptr, *qsyn = nullptr;
    double stdp[HS_STDP_NPARAMS] = {0};
    /* synaptic code (dmode HS_DELIVER_CODE): index into hs_net::cg_syn, synapse
       variables in bundle order (the weight array when one is named w), target
       and source variables of the code */
    int cg = -1, nsv = 0;
    void *sv[HS_MAX_SYN_VARS] = {nullptr};
    std::vector<int> post_vars, pre_vars, post_read_vars;
    int32_t *synpre = nullptr; /* on_post: source neuron of each synapse (bundle order) */
};

struct HMon {
    int kind; /* 0 spikes, 1 state, 2 population rate (buf: int32 spike count per row),
                 3 synapse variable (buf: rows of the selected synapses' values) */
    int gid, var, start, stop;
    int64_t cap = 0;                   /* spike records */
    int32_t *mi = nullptr, *mt = nullptr;
    unsigned long long *mn = nullptr;
    void *buf = nullptr;               /* state rows */
    int64_t rows = 0;                  /* capacity in rows */
    int64_t t0 = 0;
    int rec = -1; /* spike monitor of an index range: the monitor recording the whole group
                     (the device records each group's spikes once; a range is a filter on read) */
    int pid = -1;                      /* kind 3 (synapse variable): pathway, selected synapses */
    int32_t *sidx = nullptr;
    std::vector<int32_t> hsel;
    /* kind 1: a row every `every` steps from t0 (t0 mod every == 0), `width`
       columns; recorded neurons [start, stop) or, with hidx, hidx[] in that
       order (map: device int32[n] of columns, -1 if not recorded) */
    int every = 1;
    int64_t width = 0;
    std::vector<int32_t> hidx;
    int32_t *map = nullptr;
};

/* rows a monitor holds after the network reached `step` */
static int64_t hs_mon_rows(const HMon &M, int64_t step)
{
    if (step <= M.t0) return 0;
    return (step - M.t0 + M.every - 1) / M.every;
}

struct hs_net {
    int prec, engine;
    double dt;
    uint64_t seed;
    std::vector<HGroup> g;
    std::vector<HPath> p;
    std::vector<HMon> m;
    bool built = false;     /* rings and accumulators exist: no more groups/pathways */
    bool desc_dirty = true; /* descriptors must be re-uploaded */
    HsCtl *ctl = nullptr;
    void *dgroups = nullptr, *dpaths = nullptr;
    int total_tiles = 0;
    int window = 1;        /* steps per update launch (fused form) */
    int dmin_all = 0;      /* smallest delay over all pathways */
    int mode = 0;          /* 0 plain, 1 fused per-step pair, 2 fused single launch, 3 fused windows */
    bool noise_ahead = false; /* mode 3 with a noise model: draws made one window ahead */
    int64_t step = 0;
    /* 0 update block, 1 scatter grid, 2 graph chunk, 3 fused delivery blocks,
       4 window, 5 delivery mode, 6 gather block, 7 kernel timing, 8 noise blocks,
       9 pipelined windows */
    long long opt[12] = {256, 0, 256, 0, 0, 0, 256, 0, 0, 0, 0, 0};
    /* kernel timing: event pairs of the current run, totals per kind */
    std::vector<hipEvent_t> kt_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> kt_open;
    double kt_ms[5] = {0, 0, 0, 0, 0};
    int64_t kt_n[5] = {0, 0, 0, 0, 0};
    int n_scatter = 0, n_gather = 0, gather_tiles = 0, n_plastic = 0;
    std::map<int64_t, hipGraphExec_t> graphs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_wall = 0, last_gpu_ms = 0;
    int64_t last_steps = 0;
    uint64_t events = 0;
    uint64_t spikes = 0;
    std::vector<void *> allocs;
    std::vector<HsCgModel> cg_models;  /* code models, ids HS_MODEL_CODE + index */
    std::vector<HsSynCgModel> cg_syn;  /* synaptic code of the pathways with HS_DELIVER_CODE */
    const HsCgKernels *cg = nullptr;   /* their compiled kernels (from the first run) */
};

static size_t hs_rsize(const hs_net *net) { return net->prec == HS_F64 ? sizeof(double) : sizeof(float); }

/* ring sizes are rounded up to a power of two: the kernels then take a step's
   slot as a mask (hs_mod); a ring only has to hold at least n slots */
static int hs_ring_size(int n)
{
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}

static int hs_dalloc(hs_net *net, size_t bytes, void **p, bool zero)
{
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return hs_fail(HS_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    }
    net->allocs.push_back(*p);
    if (zero) HS_HIP(hipMemsetAsync(*p, 0, bytes ? bytes : 1, g_rt.stream));
    return HS_OK;
}

static void hs_dfree(hs_net *net, void *p)
{
    if (!p) return;
    auto it = std::find(net->allocs.begin(), net->allocs.end(), p);
    if (it != net->allocs.end()) net->allocs.erase(it);
    (void)hipFree(p);
}

static void hs_drop_graphs(hs_net *net)
{
    for (auto &kv : net->graphs) (void)hipGraphExecDestroy(kv.second);
    net->graphs.clear();
}

extern "C" int hs_net_create(int precision, int engine, double dt, uint64_t seed, hs_net **out)
{
    HS_TRY(hs_need_init());
    if (!out || (precision != HS_F64 && precision != HS_F32) ||
        (engine != HS_ENGINE_PLAIN && engine != HS_ENGINE_FUSED) || !(dt > 0))
        return hs_fail(HS_ERR_ARG, "hs_net_create: bad arguments");
    hs_net *net = new hs_net();
    net->prec = precision;
    net->engine = engine;
    net->dt = dt;
    net->seed = seed;
    if (hs_dalloc(net, sizeof(HsCtl), (void **)&net->ctl, true) != HS_OK) {
        delete net;
        return HS_ERR_NOMEM;
    }
    HS_HIP(hipEventCreate(&net->ev0));
    HS_HIP(hipEventCreate(&net->ev1));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    *out = net;
    return HS_OK;
}

extern "C" void hs_net_destroy(hs_net *net)
{
    if (!net) return;
    (void)hipStreamSynchronize(g_rt.stream);
    hs_drop_graphs(net);
    for (void *p : net->allocs) (void)hipFree(p);
    if (net->ev0) (void)hipEventDestroy(net->ev0);
    if (net->ev1) (void)hipEventDestroy(net->ev1);
    for (auto &o : net->kt_open) {
        (void)hipEventDestroy(o.second.first);
        (void)hipEventDestroy(o.second.second);
    }
    for (hipEvent_t e : net->kt_pool) (void)hipEventDestroy(e);
    delete net;
}

extern "C" int hs_net_add_group(hs_net *net, int model, int n, const double *params, int nparams, int *gid)
{
    if (!net) return hs_fail(HS_ERR_ARG, "null net");
    if (net->built) return hs_fail(HS_ERR_STATE, "groups must be added before the first run");
    HsModelInfo mi;
    if (!hs_model_info(model, &mi)) return hs_fail(HS_ERR_UNSUPPORTED, "unknown model %d", model);
    HGroup g;
    if (n <= 0 || hs_model_params(model, params, nparams, g.params) < 0)
        return hs_fail(HS_ERR_ARG, "hs_net_add_group: model %d expects %d params, n > 0", model, mi.nparams);
    params = g.params; /* every parameter, optional ones filled in */
    g.model = model;
    g.n = n;
    g.mi = mi;
    memset(g.k, 0, sizeof(g.k));
    if (hs_model_constants(model, params, net->dt, g.k) < 0)
        return hs_fail(HS_ERR_ARG, "hs_net_add_group: parameters outside model %d's closed form "
                                   "(LIF_CUBA: time constants > 0, taue and taui != taum)", model);
    g.ref_steps = mi.i_ref >= 0 ? hs_ref_steps(params[mi.i_ref], net->dt) : 0;
    if (model == HS_MODEL_SPIKEGEN) g.gen_period = hs_timestep(params[0], net->dt);
    g.stream = mi.i_stream >= 0 ? (unsigned int)params[mi.i_stream] : 0u;
    for (int v = 0; v < mi.nvars; ++v) HS_TRY(hs_dalloc(net, hs_rsize(net) * n, &g.var[v], true));
    if (model == HS_MODEL_POISSON) { /* every neuron starts at the rate parameter */
        std::vector<double> r8((size_t)n, params[0]);
        std::vector<float> r4((size_t)n, (float)params[0]);
        if (net->prec == HS_F64)
            HS_HIP(hipMemcpyAsync(g.var[0], r8.data(), sizeof(double) * n, hipMemcpyHostToDevice, g_rt.stream));
        else
            HS_HIP(hipMemcpyAsync(g.var[0], r4.data(), sizeof(float) * n, hipMemcpyHostToDevice, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    }
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * n, (void **)&g.last, false));
    std::vector<int32_t> never((size_t)n, HS_NEVER_SPIKED);
    HS_HIP(hipMemcpyAsync(g.last, never.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    net->g.push_back(g);
    net->desc_dirty = true;
    if (gid) *gid = (int)net->g.size() - 1;
    return HS_OK;
}

extern "C" int hs_net_add_group_code(hs_net *net, int n, int nvars, const char *const *var_names, const char *update,
                                     const char *threshold, const char *reset, double refractory, int noise,
                                     uint32_t stream, int *gid)
{
    if (!net) return hs_fail(HS_ERR_ARG, "null net");
    if (net->built) return hs_fail(HS_ERR_STATE, "groups must be added before the first run");
    if (n <= 0 || nvars < 1 || nvars > HS_MAX_VARS || !var_names)
        return hs_fail(HS_ERR_ARG, "hs_net_add_group_code: n > 0 and 1 to %d variables", HS_MAX_VARS);
    if ((int)net->cg_models.size() >= HS_MAX_CODE_MODELS)
        return hs_fail(HS_ERR_UNSUPPORTED, "at most %d code models per network", HS_MAX_CODE_MODELS);
    if (!(refractory >= 0.0)) return hs_fail(HS_ERR_ARG, "refractory period %g < 0", refractory);
    HsCgModel m;
    for (int v = 0; v < nvars; ++v) {
        if (!var_names[v]) return hs_fail(HS_ERR_ARG, "variable name %d is null", v);
        m.names.push_back(var_names[v]);
    }
    m.update = update ? update : "";
    m.threshold = threshold ? threshold : "";
    m.reset = reset ? reset : "";
    m.noisy = noise ? 1 : 0;
    HS_TRY(hs_cg_check(m));
    HGroup g;
    g.model = HS_MODEL_CODE + (int)net->cg_models.size();
    g.code = (int)net->cg_models.size();
    g.n = n;
    g.mi = {nvars, 0, m.reset.empty() ? 0 : 1, -1, -1, -1, -1};
    memset(g.params, 0, sizeof(g.params));
    memset(g.k, 0, sizeof(g.k));
    g.ref_steps = hs_ref_steps(refractory, net->dt);
    g.stream = stream;
    for (int v = 0; v < nvars; ++v) HS_TRY(hs_dalloc(net, hs_rsize(net) * n, &g.var[v], true));
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * n, (void **)&g.last, false));
    std::vector<int32_t> never((size_t)n, HS_NEVER_SPIKED);
    HS_HIP(hipMemcpyAsync(g.last, never.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    net->cg_models.push_back(m);
    net->g.push_back(g);
    net->desc_dirty = true;
    if (gid) *gid = (int)net->g.size() - 1;
    return HS_OK;
}

extern "C" int hs_net_set_state(hs_net *net, int gid, int var, const double *values)
{
    if (!net || gid < 0 || gid >= (int)net->g.size() || !values) return hs_fail(HS_ERR_ARG, "bad group");
    HGroup &g = net->g[gid];
    if (var < 0 || var >= g.mi.nvars) return hs_fail(HS_ERR_ARG, "bad variable %d", var);
    if (net->prec == HS_F64) {
        HS_HIP(hipMemcpyAsync(g.var[var], values, sizeof(double) * g.n, hipMemcpyHostToDevice, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    } else {
        std::vector<float> f((size_t)g.n);
        for (int i = 0; i < g.n; ++i) f[i] = (float)values[i];
        HS_HIP(hipMemcpyAsync(g.var[var], f.data(), sizeof(float) * g.n, hipMemcpyHostToDevice, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    }
    return HS_OK;
}

extern "C" int hs_net_get_state(hs_net *net, int gid, int var, double *values)
{
    if (!net || gid < 0 || gid >= (int)net->g.size() || !values) return hs_fail(HS_ERR_ARG, "bad group");
    HGroup &g = net->g[gid];
    if (var < 0 || var >= g.mi.nvars) return hs_fail(HS_ERR_ARG, "bad variable %d", var);
    if (net->prec == HS_F64) {
        HS_HIP(hipMemcpyAsync(values, g.var[var], sizeof(double) * g.n, hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    } else {
        std::vector<float> f((size_t)g.n);
        HS_HIP(hipMemcpyAsync(f.data(), g.var[var], sizeof(float) * g.n, hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        for (int i = 0; i < g.n; ++i) values[i] = (double)f[i];
    }
    return HS_OK;
}

extern "C" int hs_net_get_last_spike(hs_net *net, int gid, int32_t *steps)
{
    if (!net || gid < 0 || gid >= (int)net->g.size() || !steps) return hs_fail(HS_ERR_ARG, "bad group");
    HGroup &g = net->g[gid];
    HS_HIP(hipMemcpyAsync(steps, g.last, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

/*
 * Bundle builder: counting sort of the synapses by key (pre, delay - dmin), stable,
 * so each (pre, delay) bundle is a contiguous run of post indices in synapse
 * order.  O(S) host work and one copy per array, where the reference sorts per
 * presynaptic neuron and copies once per bundle (synapses_push_spikes.cu:441-575).
 */
extern "C" int hs_net_add_pathway(hs_net *net, int src, int tgt, int var, int64_t nsyn, const int32_t *pre,
                                  const int32_t *post, const int32_t *delay_steps, int scalar_delay, int wmode,
                                  const double *weights, double wscalar, int *pid)
{
    if (!net) return hs_fail(HS_ERR_ARG, "null net");
    if (net->built) return hs_fail(HS_ERR_STATE, "pathways must be added before the first run");
    if (src < 0 || src >= (int)net->g.size() || tgt < 0 || tgt >= (int)net->g.size())
        return hs_fail(HS_ERR_ARG, "bad source/target group");
    HGroup &gs = net->g[src];
    HGroup &gt = net->g[tgt];
    if (var < 0 || var >= gt.mi.nvars) return hs_fail(HS_ERR_ARG, "bad target variable %d", var);
    if (nsyn < 0 || nsyn >= 0x7fffffffLL) return hs_fail(HS_ERR_ARG, "synapse count %lld out of range", (long long)nsyn);
    if (nsyn > 0 && (!pre || !post)) return hs_fail(HS_ERR_ARG, "null pre/post");
    if (wmode < 0 || wmode > 2) return hs_fail(HS_ERR_ARG, "bad weight mode");
    if (wmode != HS_W_SCALAR && !weights) return hs_fail(HS_ERR_ARG, "weights required for weight mode %d", wmode);
    HPath P;
    P.src = src;
    P.tgt = tgt;
    P.var = var;
    P.nsyn = nsyn;
    P.wmode = wmode;
    P.wscalar = wscalar;
    int dmin = scalar_delay, dmax = scalar_delay;
    if (delay_steps && nsyn > 0) {
        dmin = dmax = delay_steps[0];
        for (int64_t s = 0; s < nsyn; ++s) {
            dmin = std::min(dmin, delay_steps[s]);
            dmax = std::max(dmax, delay_steps[s]);
        }
    }
    if (dmin < 0) return hs_fail(HS_ERR_ARG, "negative delay");
    const int D = dmax - dmin + 1;
    if (D > HS_MAX_DELAY_SLOTS) return hs_fail(HS_ERR_UNSUPPORTED, "delay range %d steps > %d", D, HS_MAX_DELAY_SLOTS);
    P.dmin = dmin;
    P.dmax = dmax;
    P.D = D;
    for (int64_t s = 0; s < nsyn; ++s)
        if (pre[s] < 0 || pre[s] >= gs.n || post[s] < 0 || post[s] >= gt.n)
            return hs_fail(HS_ERR_ARG, "synapse %lld: index out of range", (long long)s);
    /*
     * Delivery mode.  Gather pays one (pre, tile) lookup per candidate spike and
     * target tile and reads every synapse of a spike in each window the spike's
     * delay span overlaps ((D + w - 1) / w times); scatter reads each delivered
     * synapse once and pays one global atomic per event.  On the 100k-neuron,
     * 46-delay network scatter is 2.0x (FP32, w = 5: 5.1 vs 10.3 us per step)
     * to 2.8x (FP64, w = 1) faster, so scatter is the default and gather is
     * chosen explicitly.
     */
    const int n_tiles = (gt.n + HS_GTILE - 1) / HS_GTILE;
    int dmode = HS_DELIVER_SCATTER;
    if (net->opt[5] == HS_DELIVER_GATHER) dmode = HS_DELIVER_GATHER;
    if (dmode == HS_DELIVER_GATHER && (double)gs.n * n_tiles > 2.0e9)
        return hs_fail(HS_ERR_UNSUPPORTED, "gather table of %d x %d segments too large", gs.n, n_tiles);
    P.dmode = dmode;
    const size_t rs = hs_rsize(net);
    std::vector<unsigned char> w_sorted;
    if (wmode == HS_W_SYNAPSE) w_sorted.resize(rs * (size_t)(nsyn ? nsyn : 1));
    auto put_w = [&](int32_t j, int64_t s) {
        if (wmode != HS_W_SYNAPSE) return;
        if (net->prec == HS_F64) ((double *)w_sorted.data())[j] = weights[s];
        else ((float *)w_sorted.data())[j] = (float)weights[s];
    };
    if (dmode == HS_DELIVER_SCATTER) {
        /* bundles: stable counting sort by (pre, delay) */
        const int64_t nb = (int64_t)gs.n * D;
        std::vector<int32_t> start((size_t)nb + 1, 0);
        for (int64_t s = 0; s < nsyn; ++s) {
            const int d = delay_steps ? delay_steps[s] : scalar_delay;
            start[(size_t)((int64_t)pre[s] * D + (d - dmin)) + 1]++;
        }
        for (int64_t b = 0; b < nb; ++b) start[b + 1] += start[b];
        std::vector<int32_t> pos(start.begin(), start.end() - 1);
        std::vector<int32_t> post_sorted((size_t)(nsyn ? nsyn : 1));
        for (int64_t s = 0; s < nsyn; ++s) {
            const int d = delay_steps ? delay_steps[s] : scalar_delay;
            const int32_t j = pos[(size_t)((int64_t)pre[s] * D + (d - dmin))]++;
            post_sorted[j] = post[s];
            put_w(j, s);
        }
        HS_TRY(hs_dalloc(net, sizeof(int32_t) * (size_t)(nb + 1), (void **)&P.bstart, false));
        HS_TRY(hs_dalloc(net, sizeof(int32_t) * (size_t)(nsyn ? nsyn : 1), (void **)&P.post, false));
        HS_HIP(hipMemcpyAsync(P.bstart, start.data(), sizeof(int32_t) * (size_t)(nb + 1), hipMemcpyHostToDevice,
                              g_rt.stream));
        if (nsyn)
            HS_HIP(hipMemcpyAsync(P.post, post_sorted.data(), sizeof(int32_t) * (size_t)nsyn, hipMemcpyHostToDevice,
                                  g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    } else {
        /* segments: stable counting sort by (target tile, pre) -- tile-major, so the
           block that owns a tile reads one contiguous region; entry packs the
           delay slot and the target's index inside its tile */
        P.n_tiles = n_tiles;
        const int64_t nseg = (int64_t)gs.n * n_tiles;
        std::vector<int32_t> start((size_t)nseg + 1, 0);
        for (int64_t s = 0; s < nsyn; ++s) start[(size_t)((int64_t)(post[s] / HS_GTILE) * gs.n + pre[s]) + 1]++;
        for (int64_t b = 0; b < nseg; ++b) start[b + 1] += start[b];
        std::vector<int32_t> pos(start.begin(), start.end() - 1);
        std::vector<uint32_t> ent((size_t)(nsyn ? nsyn : 1));
        for (int64_t s = 0; s < nsyn; ++s) {
            const int d = delay_steps ? delay_steps[s] : scalar_delay;
            const int tl = post[s] / HS_GTILE;
            const int32_t j = pos[(size_t)((int64_t)tl * gs.n + pre[s])]++;
            ent[j] = ((uint32_t)(d - dmin) << 16) | (uint32_t)(post[s] - tl * HS_GTILE);
            put_w(j, s);
        }
        HS_TRY(hs_dalloc(net, sizeof(int32_t) * (size_t)(nseg + 1), (void **)&P.goff, false));
        HS_TRY(hs_dalloc(net, sizeof(uint32_t) * (size_t)(nsyn ? nsyn : 1), (void **)&P.gsyn, false));
        HS_HIP(hipMemcpyAsync(P.goff, start.data(), sizeof(int32_t) * (size_t)(nseg + 1), hipMemcpyHostToDevice,
                              g_rt.stream));
        if (nsyn)
            HS_HIP(hipMemcpyAsync(P.gsyn, ent.data(), sizeof(uint32_t) * (size_t)nsyn, hipMemcpyHostToDevice,
                                  g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    }
    if (wmode == HS_W_SYNAPSE && nsyn) {
        HS_TRY(hs_dalloc(net, rs * (size_t)nsyn, &P.w, false));
        HS_HIP(hipMemcpyAsync(P.w, w_sorted.data(), rs * (size_t)nsyn, hipMemcpyHostToDevice, g_rt.stream));
    } else if (wmode == HS_W_PRE) {
        std::vector<unsigned char> wp(rs * (size_t)gs.n);
        for (int i = 0; i < gs.n; ++i) {
            if (net->prec == HS_F64) ((double *)wp.data())[i] = weights[i];
            else ((float *)wp.data())[i] = (float)weights[i];
        }
        HS_TRY(hs_dalloc(net, rs * (size_t)gs.n, &P.w, false));
        HS_HIP(hipMemcpyAsync(P.w, wp.data(), rs * (size_t)gs.n, hipMemcpyHostToDevice, g_rt.stream));
    }
    HS_HIP(hipStreamSynchronize(g_rt.stream)); /* host vectors die here */
    if (P.dmode == HS_DELIVER_GATHER) {
        net->n_gather++;
        net->gather_tiles = std::max(net->gather_tiles, P.n_tiles);
    } else {
        net->n_scatter++;
    }
    net->p.push_back(P);
    net->desc_dirty = true;
    if (pid) *pid = (int)net->p.size() - 1;
    return HS_OK;
}

/*
 * Random connectivity on the device (kernels in hs_connect.hip), connect(p) and
 * fixed out-degree: count per (pre, delay slot) -> prefix sum -> bstart ->
 * write post/w in bundle order.
 * The host touches two scalars (the total and an error flag); temporaries are
 * freed before return.
 */
static int hs_connect_q(double d, double dt, int *q)
{
    const double x = d / dt + 0.5;
    if (!(x >= 0.0) || x >= 2147483647.0) return hs_fail(HS_ERR_ARG, "delay %g s out of range", d);
    *q = (int)x;
    return HS_OK;
}

/* fixed_k > 0: K distinct targets per source neuron (p unused); else connect(p) */
/* every hs_pre(k) / hs_post(k) of a connect expression names a variable the
   source / target group has (a literal k below its count): the kernels read
   the group's array k */
static int hs_conn_expr_vars(const std::string &e, const char *fn, int nvars)
{
    const std::string pat = std::string(fn) + "(";
    for (size_t at = e.find(pat); at != std::string::npos; at = e.find(pat, at + 1)) {
        if (at > 0 && (isalnum((unsigned char)e[at - 1]) || e[at - 1] == '_')) continue;
        size_t q = at + pat.size();
        int k = 0, digits = 0;
        while (q < e.size() && isdigit((unsigned char)e[q]) && digits < 3) k = 10 * k + (e[q++] - '0'), ++digits;
        if (!digits || q >= e.size() || e[q] != ')' || k >= nvars)
            return hs_fail(HS_ERR_ARG, "connect expression: %s(...) must name one of the group's %d variables", fn,
                           nvars);
    }
    return HS_OK;
}

/* cond / pexpr: a connect condition and a per-pair probability as C expressions
   (empty: none; hs_cg_conn_kernels), compiled into the two passes */
static int hs_connect_device(hs_net *net, int src, int tgt, int var, int pre_lo, int pre_hi, int post_lo, int post_hi,
                             double p, int fixed_k, int autapses, uint64_t seed, uint32_t stream, double delay_lo,
                             double delay_hi, int wmode, double w_lo, double w_hi, int *pid,
                             const std::string &cond = std::string(), const std::string &pexpr = std::string(),
                             int ncopy = 1, const std::string &wexpr = std::string())
{
    if (!net) return hs_fail(HS_ERR_ARG, "null net");
    if (net->built) return hs_fail(HS_ERR_STATE, "pathways must be added before the first run");
    if (src < 0 || src >= (int)net->g.size() || tgt < 0 || tgt >= (int)net->g.size())
        return hs_fail(HS_ERR_ARG, "bad source/target group");
    HGroup &gs = net->g[src];
    HGroup &gt = net->g[tgt];
    if (var < 0 || var >= gt.mi.nvars) return hs_fail(HS_ERR_ARG, "bad target variable %d", var);
    if (pre_lo < 0 || pre_lo > pre_hi || pre_hi > gs.n) return hs_fail(HS_ERR_ARG, "bad source range");
    if (post_lo < 0 || post_lo > post_hi || post_hi > gt.n) return hs_fail(HS_ERR_ARG, "bad target range");
    if (fixed_k < 0) return hs_fail(HS_ERR_ARG, "out-degree %d < 0", fixed_k);
    if (ncopy < 1 || ncopy > HS_CONN_NMAX)
        return hs_fail(HS_ERR_UNSUPPORTED, "%d synapses per pair (1 to %d)", ncopy, HS_CONN_NMAX);
    if (!pexpr.empty()) {
        if (fixed_k) return hs_fail(HS_ERR_ARG, "a probability expression with a fixed out-degree");
        p = 1.0; /* rows are not skipped; every pair takes its own draw against p(i, j) */
    }
    if (!fixed_k && (!(p >= 0.0) || !(p <= 1.0)))
        return hs_fail(HS_ERR_ARG, "connection probability %g outside [0, 1]", p);
    const HsCgConnKernels *K = nullptr;
    if (!wexpr.empty() && wmode != HS_W_SYNAPSE)
        return hs_fail(HS_ERR_ARG, "a weight expression needs per-synapse weights");
    for (const std::string *x : {&cond, &pexpr, &wexpr}) {
        HS_TRY(hs_conn_expr_vars(*x, "hs_pre", gs.mi.nvars));
        HS_TRY(hs_conn_expr_vars(*x, "hs_post", gt.mi.nvars));
    }
    if (!cond.empty() || !pexpr.empty() || !wexpr.empty() || ncopy > 1)
        HS_TRY(hs_cg_conn_kernels(net->prec, cond, pexpr, wexpr, ncopy > 1, &K));
    if (!(delay_lo >= 0.0) || !(delay_hi >= delay_lo) || delay_hi > 1e300)
        return hs_fail(HS_ERR_ARG, "bad delay range [%g, %g]", delay_lo, delay_hi);
    if (wmode != HS_W_SCALAR && wmode != HS_W_SYNAPSE)
        return hs_fail(HS_ERR_UNSUPPORTED, "random connectivity takes a scalar or per-synapse weight");
    if (!(w_hi >= w_lo) || w_lo < -1e300 || w_hi > 1e300) return hs_fail(HS_ERR_ARG, "bad weight range");
    if (net->opt[5] == HS_DELIVER_GATHER)
        return hs_fail(HS_ERR_UNSUPPORTED, "random connectivity builds scatter bundles (delivery option 1)");
    HsConnSpec c = {};
    c.pre_lo = pre_lo;
    c.pre_hi = pre_hi;
    c.post_lo = post_lo;
    c.post_hi = post_hi;
    c.autapses = (autapses || src != tgt) ? 1 : 0;
    if (fixed_k > 0 && pre_hi > pre_lo) {
        if (fixed_k > post_hi - post_lo)
            return hs_fail(HS_ERR_ARG, "sample size %d bigger than the population size %d", fixed_k,
                           post_hi - post_lo);
        if (fixed_k > HS_CONN_KMAX)
            return hs_fail(HS_ERR_UNSUPPORTED, "out-degree %d > %d", fixed_k, HS_CONN_KMAX);
    }
    c.fixed_k = fixed_k;
    c.ncopy = ncopy;
    c.p = fixed_k ? 0.0 : p;
    c.jump = !fixed_k && pexpr.empty() && p < 0.25 ? 1 : 0;
    c.pconst = c.jump && p > 0.0 ? 1.0 / hs_log(1.0 - p) : 0.0;
    c.seed = seed;
    c.stream = stream;
    c.dt = net->dt;
    c.dlo = delay_lo;
    c.dspan = delay_hi - delay_lo;
    c.wlo = w_lo;
    c.wspan = wmode == HS_W_SYNAPSE ? w_hi - w_lo : 0.0;
    int dmin, dmax;
    HS_TRY(hs_connect_q(c.dlo, c.dt, &dmin));
    HS_TRY(hs_connect_q(c.dlo + 1.0 * c.dspan, c.dt, &dmax)); /* the slot formula at u = 1: a bound */
    const int D = dmax - dmin + 1;
    if (D > HS_MAX_DELAY_SLOTS) return hs_fail(HS_ERR_UNSUPPORTED, "delay range %d steps > %d", D, HS_MAX_DELAY_SLOTS);
    c.dmin = dmin;
    c.D = D;
    if (p == 0.0) c.pre_hi = c.pre_lo; /* no rows: the reference skips p == 0 (synapses_create_generator.cu:315) */

    HPath P;
    P.src = src;
    P.tgt = tgt;
    P.var = var;
    P.wmode = wmode;
    P.wscalar = w_lo;
    P.dmin = dmin;
    P.dmax = dmax;
    P.D = D;
    P.dmode = HS_DELIVER_SCATTER;
    const long long nb = (long long)gs.n * D;
    const int nblk = hs_scan_blocks(nb);
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * (size_t)(nb + 1), (void **)&P.bstart, false));
    int32_t *cnt = nullptr;
    long long *bsum = nullptr;
    int *err = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(g_rt.stream);
        if (cnt) (void)hipFree(cnt);
        if (bsum) (void)hipFree(bsum);
        if (err) (void)hipFree(err);
    };
    auto fail_hip = [&](hipError_t e, const char *what) {
        release();
        return hs_fail(HS_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc((void **)&cnt, sizeof(int32_t) * (size_t)(nb ? nb : 1))) != hipSuccess ||
        (e = hipMalloc((void **)&bsum, sizeof(long long) * (size_t)(nblk + 1))) != hipSuccess ||
        (e = hipMalloc((void **)&err, sizeof(int))) != hipSuccess) {
        (void)hipGetLastError();
        release();
        return hs_fail(HS_ERR_NOMEM, "connectivity temporaries (%lld counts): %s", nb, hipGetErrorString(e));
    }
    if ((e = hipMemsetAsync(cnt, 0, sizeof(int32_t) * (size_t)(nb ? nb : 1), g_rt.stream)) != hipSuccess ||
        (e = hipMemsetAsync(err, 0, sizeof(int), g_rt.stream)) != hipSuccess)
        return fail_hip(e, "clear");
    HsConnVars cv = {};
    for (int v = 0; v < gs.mi.nvars && v < HS_MAX_VARS; ++v) cv.pre[v] = gs.var[v];
    for (int v = 0; v < gt.mi.nvars && v < HS_MAX_VARS; ++v) cv.post[v] = gt.var[v];
    if ((e = K ? hs_cg_launch_conn_count(K, c, cv, cnt, err, g_rt.stream) : hs_launch_conn_count(c, cnt, err, g_rt.stream)) !=
        hipSuccess)
        return fail_hip(e, "count pass");
    if ((e = hs_launch_scan(cnt, nb, bsum, P.bstart, g_rt.stream)) != hipSuccess) return fail_hip(e, "prefix sum");
    long long total = 0;
    int herr = 0;
    if ((e = hipMemcpyAsync(&total, bsum + nblk, sizeof(long long), hipMemcpyDeviceToHost, g_rt.stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, g_rt.stream)) != hipSuccess ||
        (e = hipStreamSynchronize(g_rt.stream)) != hipSuccess)
        return fail_hip(e, "count readback");
    if (herr) {
        release();
        return hs_fail(HS_ERR_HIP, "count pass: delay slot outside [0, %d)", D);
    }
    if (total >= 0x7fffffffLL) {
        release();
        return hs_fail(HS_ERR_UNSUPPORTED, "%lld synapses: more than 2^31 - 1 in one pathway", total);
    }
    P.nsyn = total;
    const size_t rs = hs_rsize(net);
    int r;
    if ((r = hs_dalloc(net, sizeof(int32_t) * (size_t)(total ? total : 1), (void **)&P.post, false)) != HS_OK) {
        release();
        return r;
    }
    if (wmode == HS_W_SYNAPSE &&
        (r = hs_dalloc(net, rs * (size_t)(total ? total : 1), &P.w, false)) != HS_OK) {
        release();
        return r;
    }
    if ((e = K ? hs_cg_launch_conn_write(K, c, cv, P.bstart, P.post, wmode == HS_W_SYNAPSE ? P.w : nullptr, err,
                                         g_rt.stream)
               : hs_launch_conn_write(net->prec == HS_F64 ? 64 : 32, c, P.bstart, P.post,
                                      wmode == HS_W_SYNAPSE ? P.w : nullptr, err, g_rt.stream)) != hipSuccess)
        return fail_hip(e, "write pass");
    if ((e = hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, g_rt.stream)) != hipSuccess ||
        (e = hipStreamSynchronize(g_rt.stream)) != hipSuccess)
        return fail_hip(e, "write readback");
    release();
    if (herr) return hs_fail(HS_ERR_HIP, "write pass: count and write passes disagree (code %d)", herr);
    net->n_scatter++;
    net->p.push_back(P);
    net->desc_dirty = true;
    if (pid) *pid = (int)net->p.size() - 1;
    return HS_OK;
}

extern "C" int hs_net_connect_random(hs_net *net, int src, int tgt, int var, int pre_lo, int pre_hi, int post_lo,
                                     int post_hi, double p, int autapses, uint64_t seed, uint32_t stream,
                                     double delay_lo, double delay_hi, int wmode, double w_lo, double w_hi, int *pid)
{
    return hs_connect_device(net, src, tgt, var, pre_lo, pre_hi, post_lo, post_hi, p, 0, autapses, seed, stream,
                             delay_lo, delay_hi, wmode, w_lo, w_hi, pid);
}

extern "C" int hs_net_connect_fixed(hs_net *net, int src, int tgt, int var, int pre_lo, int pre_hi, int post_lo,
                                    int post_hi, int k, int autapses, uint64_t seed, uint32_t stream, double delay_lo,
                                    double delay_hi, int wmode, double w_lo, double w_hi, int *pid)
{
    if (k < 0) return hs_fail(HS_ERR_ARG, "out-degree %d < 0", k);
    /* K = 0 creates no synapse: the connect(p = 0) form */
    return hs_connect_device(net, src, tgt, var, pre_lo, pre_hi, post_lo, post_hi, k ? 1.0 : 0.0, k, autapses, seed,
                             stream, delay_lo, delay_hi, wmode, w_lo, w_hi, pid);
}

extern "C" int hs_net_connect_code(hs_net *net, int src, int tgt, int var, int pre_lo, int pre_hi, int post_lo,
                                   int post_hi, const char *condition, const char *p_expr, double p, int k, int n,
                                   int autapses, uint64_t seed, uint32_t stream, double delay_lo, double delay_hi,
                                   int wmode, double w_lo, double w_hi, const char *w_expr, int *pid)
{
    if (!net) return hs_fail(HS_ERR_ARG, "null net");
    const std::string cond = condition ? condition : "", pe = p_expr ? p_expr : "", we = w_expr ? w_expr : "";
    return hs_connect_device(net, src, tgt, var, pre_lo, pre_hi, post_lo, post_hi, k ? 1.0 : p, k, autapses, seed,
                             stream, delay_lo, delay_hi, wmode, w_lo, w_hi, pid, cond, pe, n, we);
}

/*
 * Bundle table from synapse arrays already in HBM (kernels in hs_sort.hip):
 * check + delay range, keys (pre, delay slot), stable radix sort, gather of
 * post and weights, bstart from the sorted keys.  Temporaries are freed before
 * return; the caller's arrays are only read.
 */
extern "C" int hs_net_add_pathway_device(hs_net *net, int src, int tgt, int var, int64_t nsyn, const int32_t *pre,
                                         const int32_t *post, const int32_t *delay_steps, int scalar_delay, int wmode,
                                         const double *weights, double wscalar, int *pid)
{
    if (!net) return hs_fail(HS_ERR_ARG, "null net");
    if (net->built) return hs_fail(HS_ERR_STATE, "pathways must be added before the first run");
    if (src < 0 || src >= (int)net->g.size() || tgt < 0 || tgt >= (int)net->g.size())
        return hs_fail(HS_ERR_ARG, "bad source/target group");
    HGroup &gs = net->g[src];
    HGroup &gt = net->g[tgt];
    if (var < 0 || var >= gt.mi.nvars) return hs_fail(HS_ERR_ARG, "bad target variable %d", var);
    if (nsyn < 0 || nsyn >= 0x7fffffffLL) return hs_fail(HS_ERR_ARG, "synapse count %lld out of range", (long long)nsyn);
    if (nsyn > 0 && (!pre || !post)) return hs_fail(HS_ERR_ARG, "null pre/post");
    if (wmode < 0 || wmode > 2) return hs_fail(HS_ERR_ARG, "bad weight mode");
    if (wmode != HS_W_SCALAR && !weights) return hs_fail(HS_ERR_ARG, "weights required for weight mode %d", wmode);
    if (!delay_steps && scalar_delay < 0) return hs_fail(HS_ERR_ARG, "negative delay");
    if (net->opt[5] == HS_DELIVER_GATHER)
        return hs_fail(HS_ERR_UNSUPPORTED, "device-built pathways use scatter bundles (delivery option 1)");
    const long long n = nsyn;
    int32_t *tmp_i = nullptr;      /* minmax[2], err */
    uint32_t *keys = nullptr, *keys_s = nullptr;
    int32_t *vals = nullptr, *vals_s = nullptr;
    void *cub = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(g_rt.stream);
        for (void *q : {(void *)tmp_i, (void *)keys, (void *)keys_s, (void *)vals, (void *)vals_s, cub})
            if (q) (void)hipFree(q);
    };
    auto fail_hip = [&](hipError_t e, const char *what) {
        (void)hipGetLastError();
        release();
        return hs_fail(HS_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc((void **)&tmp_i, 3 * sizeof(int32_t))) != hipSuccess) return fail_hip(e, "temporaries");
    const int32_t init[3] = {0x7fffffff, -1, 0};
    int32_t got[3];
    if ((e = hipMemcpyAsync(tmp_i, init, sizeof(init), hipMemcpyHostToDevice, g_rt.stream)) != hipSuccess ||
        (e = hs_launch_syn_check(pre, post, delay_steps, n, gs.n, gt.n, tmp_i, tmp_i + 2, g_rt.stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(got, tmp_i, sizeof(got), hipMemcpyDeviceToHost, g_rt.stream)) != hipSuccess ||
        (e = hipStreamSynchronize(g_rt.stream)) != hipSuccess)
        return fail_hip(e, "check");
    if (got[2]) {
        release();
        return hs_fail(HS_ERR_ARG, "%s", (got[2] & 1) ? "pre index out of range"
                                         : (got[2] & 2) ? "post index out of range" : "negative delay");
    }
    int dmin = scalar_delay, dmax = scalar_delay;
    if (delay_steps && n > 0) {
        dmin = got[0];
        dmax = got[1];
    }
    const int D = dmax - dmin + 1;
    if (D > HS_MAX_DELAY_SLOTS) {
        release();
        return hs_fail(HS_ERR_UNSUPPORTED, "delay range %d steps > %d", D, HS_MAX_DELAY_SLOTS);
    }
    const long long nb = (long long)gs.n * D;
    if (nb >= 0xffffffffLL) {
        release();
        return hs_fail(HS_ERR_UNSUPPORTED, "%lld bundles: keys beyond 32 bits", nb);
    }
    int bits = 1;
    while (bits < 32 && (1LL << bits) < nb) ++bits;
    HPath P;
    P.src = src;
    P.tgt = tgt;
    P.var = var;
    P.nsyn = nsyn;
    P.wmode = wmode;
    P.wscalar = wscalar;
    P.dmin = dmin;
    P.dmax = dmax;
    P.D = D;
    P.dmode = HS_DELIVER_SCATTER;
    const size_t rs = hs_rsize(net);
    int r;
    if ((r = hs_dalloc(net, sizeof(int32_t) * (size_t)(nb + 1), (void **)&P.bstart, false)) != HS_OK ||
        (r = hs_dalloc(net, sizeof(int32_t) * (size_t)(n ? n : 1), (void **)&P.post, false)) != HS_OK ||
        (wmode == HS_W_SYNAPSE && (r = hs_dalloc(net, rs * (size_t)(n ? n : 1), &P.w, false)) != HS_OK) ||
        (wmode == HS_W_PRE && (r = hs_dalloc(net, rs * (size_t)gs.n, &P.w, false)) != HS_OK)) {
        release();
        return r;
    }
    const size_t nn = (size_t)(n ? n : 1);
    size_t cub_bytes = 0;
    if ((e = hipMalloc((void **)&keys, sizeof(uint32_t) * nn)) != hipSuccess ||
        (e = hipMalloc((void **)&keys_s, sizeof(uint32_t) * nn)) != hipSuccess ||
        (e = hipMalloc((void **)&vals, sizeof(int32_t) * nn)) != hipSuccess ||
        (e = hipMalloc((void **)&vals_s, sizeof(int32_t) * nn)) != hipSuccess ||
        (e = hs_radix_sort_pairs(nullptr, &cub_bytes, keys, keys_s, vals, vals_s, n, bits, g_rt.stream)) !=
            hipSuccess ||
        (e = hipMalloc(&cub, cub_bytes ? cub_bytes : 1)) != hipSuccess)
        return fail_hip(e, "sort temporaries");
    if ((e = hs_launch_syn_keys(pre, delay_steps, n, D, dmin, keys, vals, g_rt.stream)) != hipSuccess)
        return fail_hip(e, "keys");
    if (n > 0 && (e = hs_radix_sort_pairs(cub, &cub_bytes, keys, keys_s, vals, vals_s, n, bits, g_rt.stream)) !=
                     hipSuccess)
        return fail_hip(e, "radix sort");
    if ((e = hs_launch_syn_gather(net->prec == HS_F64 ? 64 : 32, vals_s, n, post, weights, P.post,
                                  wmode == HS_W_SYNAPSE ? P.w : nullptr, g_rt.stream)) != hipSuccess)
        return fail_hip(e, "gather");
    if (wmode == HS_W_PRE &&
        (e = hs_launch_syn_convert(net->prec == HS_F64 ? 64 : 32, weights, gs.n, P.w, g_rt.stream)) != hipSuccess)
        return fail_hip(e, "per-pre weights");
    if ((e = hs_launch_syn_bstart(keys_s, n, nb, P.bstart, g_rt.stream)) != hipSuccess) return fail_hip(e, "bstart");
    release(); /* synchronises: the table is complete */
    net->n_scatter++;
    net->p.push_back(P);
    net->desc_dirty = true;
    if (pid) *pid = (int)net->p.size() - 1;
    return HS_OK;
}

extern "C" int hs_net_pathway_info(hs_net *net, int pid, int64_t *nsyn, int *dmin, int *D, int *dmode)
{
    if (!net || pid < 0 || pid >= (int)net->p.size()) return hs_fail(HS_ERR_ARG, "bad pathway");
    const HPath &P = net->p[pid];
    if (nsyn) *nsyn = P.nsyn;
    if (dmin) *dmin = P.dmin;
    if (D) *D = P.D;
    if (dmode) *dmode = P.dmode;
    return HS_OK;
}

/* the synapses of a scatter-mode pathway in bundle order: by (pre, delay), creation order within a bundle */
extern "C" int hs_net_get_synapses(hs_net *net, int pid, int32_t *pre, int32_t *post, int32_t *delay_steps,
                                   double *w, int64_t cap)
{
    if (!net || pid < 0 || pid >= (int)net->p.size()) return hs_fail(HS_ERR_ARG, "bad pathway");
    const HPath &P = net->p[pid];
    if (P.dmode == HS_DELIVER_GATHER) return hs_fail(HS_ERR_UNSUPPORTED, "synapse readback of gather-mode pathways");
    if (cap < P.nsyn) return hs_fail(HS_ERR_ARG, "capacity %lld < %lld synapses", (long long)cap, (long long)P.nsyn);
    const int n_src = net->g[P.src].n;
    const size_t nb = (size_t)n_src * P.D;
    std::vector<int32_t> bs(nb + 1), po((size_t)(P.nsyn ? P.nsyn : 1));
    HS_HIP(hipMemcpyAsync(bs.data(), P.bstart, sizeof(int32_t) * (nb + 1), hipMemcpyDeviceToHost, g_rt.stream));
    if (P.nsyn)
        HS_HIP(hipMemcpyAsync(po.data(), P.post, sizeof(int32_t) * (size_t)P.nsyn, hipMemcpyDeviceToHost,
                              g_rt.stream));
    const size_t rs = hs_rsize(net);
    std::vector<unsigned char> wv;
    if (w && P.wmode != HS_W_SCALAR) {
        const size_t nw = P.wmode == HS_W_SYNAPSE ? (size_t)P.nsyn : (size_t)n_src;
        wv.resize(rs * (nw ? nw : 1));
        if (nw) HS_HIP(hipMemcpyAsync(wv.data(), P.w, rs * nw, hipMemcpyDeviceToHost, g_rt.stream));
    }
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    auto wat = [&](size_t k) {
        return net->prec == HS_F64 ? ((const double *)wv.data())[k] : (double)((const float *)wv.data())[k];
    };
    for (int i = 0; i < n_src; ++i)
        for (int d = 0; d < P.D; ++d)
            for (int32_t s = bs[(size_t)i * P.D + d]; s < bs[(size_t)i * P.D + d + 1]; ++s) {
                if (pre) pre[s] = i;
                if (post) post[s] = po[s];
                if (delay_steps) delay_steps[s] = P.dmin + d;
                if (w) w[s] = P.wmode == HS_W_SCALAR ? P.wscalar : P.wmode == HS_W_SYNAPSE ? wat((size_t)s) : wat(i);
            }
    return HS_OK;
}

/* SpikeGeneratorGroup pattern: host counting sort by step, ascending neurons
   within a step, one upload; may be replaced between runs */
extern "C" int hs_net_set_spikegen(hs_net *net, int gid, int64_t n, const int32_t *indices, const int64_t *steps)
{
    if (!net || gid < 0 || gid >= (int)net->g.size()) return hs_fail(HS_ERR_ARG, "bad group");
    HGroup &g = net->g[gid];
    if (g.model != HS_MODEL_SPIKEGEN) return hs_fail(HS_ERR_ARG, "group %d is not a spike generator", gid);
    if (n < 0 || n >= 0x7fffffffLL || (n > 0 && (!indices || !steps))) return hs_fail(HS_ERR_ARG, "bad spike list");
    long long smax = -1;
    for (int64_t k = 0; k < n; ++k) {
        if (indices[k] < 0 || indices[k] >= g.n) return hs_fail(HS_ERR_ARG, "spike %lld: neuron out of range", (long long)k);
        if (steps[k] < 0 || steps[k] >= 0x7ffffffeLL) return hs_fail(HS_ERR_ARG, "spike %lld: step out of range", (long long)k);
        smax = std::max<long long>(smax, steps[k]);
    }
    if (g.gen_period > 0 && smax >= g.gen_period)
        return hs_fail(HS_ERR_ARG, "spike at step %lld is not inside the period of %lld steps", smax, g.gen_period);
    const long long ns = smax + 1;
    std::vector<int32_t> off((size_t)ns + 1, 0), idx((size_t)(n ? n : 1));
    for (int64_t k = 0; k < n; ++k) off[(size_t)steps[k] + 1]++;
    for (long long s = 0; s < ns; ++s) off[(size_t)s + 1] += off[(size_t)s];
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int64_t k = 0; k < n; ++k) idx[(size_t)fill[(size_t)steps[k]]++] = indices[k];
    for (long long s = 0; s < ns; ++s) {
        std::sort(idx.begin() + off[(size_t)s], idx.begin() + off[(size_t)s + 1]);
        for (int32_t q = off[(size_t)s] + 1; q < off[(size_t)s + 1]; ++q)
            if (idx[(size_t)q] == idx[(size_t)q - 1])
                return hs_fail(HS_ERR_ARG, "neuron %d given twice at step %lld", idx[(size_t)q], s);
    }
    HS_HIP(hipStreamSynchronize(g_rt.stream)); /* no launch still reads the old pattern */
    hs_dfree(net, g.gen_off);
    hs_dfree(net, g.gen_idx);
    g.gen_off = g.gen_idx = nullptr;
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * off.size(), (void **)&g.gen_off, false));
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * idx.size(), (void **)&g.gen_idx, false));
    HS_HIP(hipMemcpyAsync(g.gen_off, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipMemcpyAsync(g.gen_idx, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    g.gen_steps = ns;
    net->desc_dirty = true;
    hs_drop_graphs(net);
    return HS_OK;
}

/* STDP on a pathway: traces and last-update steps allocated (zero: Brian2's
   initial values), the target-indexed synapse table built from the bundles */
/* the synapses of each target neuron j: qsyn[qstart[j] .. qstart[j+1]) in bundle
   order (the post pathway's table, built once from the bundles) */
static int hs_post_index(hs_net *net, HPath &P)
{
    const HGroup &gt = net->g[P.tgt];
    const size_t ns = (size_t)(P.nsyn ? P.nsyn : 1);
    std::vector<int32_t> po(ns), qs((size_t)gt.n + 1, 0), qy(ns);
    if (P.nsyn)
        HS_HIP(hipMemcpyAsync(po.data(), P.post, sizeof(int32_t) * (size_t)P.nsyn, hipMemcpyDeviceToHost, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    for (int64_t s = 0; s < P.nsyn; ++s) qs[(size_t)po[(size_t)s] + 1]++;
    for (int j = 0; j < gt.n; ++j) qs[(size_t)j + 1] += qs[(size_t)j];
    std::vector<int32_t> fill(qs.begin(), qs.end() - 1);
    for (int64_t s = 0; s < P.nsyn; ++s) qy[(size_t)fill[(size_t)po[(size_t)s]]++] = (int32_t)s;
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * qs.size(), (void **)&P.qstart, false));
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * ns, (void **)&P.qsyn, false));
    HS_HIP(hipMemcpyAsync(P.qstart, qs.data(), sizeof(int32_t) * qs.size(), hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipMemcpyAsync(P.qsyn, qy.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream)); /* host vectors die here */
    return HS_OK;
}

extern "C" int hs_net_pathway_stdp(hs_net *net, int pid, const double *params, int nparams)
{
    if (!net || pid < 0 || pid >= (int)net->p.size()) return hs_fail(HS_ERR_ARG, "bad pathway");
    if (net->built) return hs_fail(HS_ERR_STATE, "plasticity must be set before the first run");
    if (!params || nparams != HS_STDP_NPARAMS) return hs_fail(HS_ERR_ARG, "STDP takes %d parameters", HS_STDP_NPARAMS);
    HPath &P = net->p[pid];
    if (P.dmode != HS_DELIVER_SCATTER) return hs_fail(HS_ERR_UNSUPPORTED, "STDP needs a scatter-mode pathway");
    if (P.wmode != HS_W_SYNAPSE) return hs_fail(HS_ERR_ARG, "STDP needs per-synapse weights (HS_W_SYNAPSE)");
    if (!(params[0] > 0.0) || !(params[1] > 0.0)) return hs_fail(HS_ERR_ARG, "trace time constants must be > 0");
    if (!(params[4] <= params[5])) return hs_fail(HS_ERR_ARG, "wmin > wmax");
    const size_t ns = (size_t)(P.nsyn ? P.nsyn : 1);
    const size_t rs = hs_rsize(net);
    HS_TRY(hs_dalloc(net, rs * ns, &P.apre, true));
    HS_TRY(hs_dalloc(net, rs * ns, &P.apost, true));
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * ns, (void **)&P.lastup, true));
    HS_TRY(hs_post_index(net, P));
    for (int j = 0; j < HS_STDP_NPARAMS; ++j) P.stdp[j] = params[j];
    P.dmode = HS_DELIVER_PLASTIC;
    net->n_scatter--;
    net->n_plastic++;
    net->desc_dirty = true;
    return HS_OK;
}

extern "C" int hs_net_pathway_code(hs_net *net, int pid, int nsvars, const char *const *syn_names,
                                   const char *on_pre, int npost, const int *post_vars,
                                   const char *const *post_names, int npre, const int *pre_vars,
                                   const char *const *pre_names)
{
    if (!net || pid < 0 || pid >= (int)net->p.size()) return hs_fail(HS_ERR_ARG, "bad pathway");
    if (net->built) return hs_fail(HS_ERR_STATE, "synaptic code must be set before the first run");
    HPath &P = net->p[pid];
    if (P.dmode != HS_DELIVER_SCATTER) return hs_fail(HS_ERR_UNSUPPORTED, "synaptic code needs a scatter-mode pathway");
    if ((int)net->cg_syn.size() >= HS_MAX_CODE_MODELS)
        return hs_fail(HS_ERR_UNSUPPORTED, "at most %d pathways with synaptic code per network", HS_MAX_CODE_MODELS);
    if (nsvars < 0 || nsvars > HS_MAX_SYN_VARS || npost < 0 || npost > HS_MAX_VARS || npre < 0 ||
        npre > HS_MAX_VARS || (nsvars && !syn_names) || (npost && (!post_vars || !post_names)) ||
        (npre && (!pre_vars || !pre_names)))
        return hs_fail(HS_ERR_ARG, "hs_net_pathway_code: bad variable lists");
    const HGroup &gs = net->g[P.src], &gt = net->g[P.tgt];
    HsSynCgModel m;
    m.on_pre = on_pre ? on_pre : "";
    for (int k = 0; k < nsvars; ++k) {
        if (!syn_names[k]) return hs_fail(HS_ERR_ARG, "synapse variable name %d is null", k);
        m.syn_names.push_back(syn_names[k]);
    }
    for (int k = 0; k < npost; ++k) {
        if (!post_names[k] || post_vars[k] < 0 || post_vars[k] >= gt.mi.nvars)
            return hs_fail(HS_ERR_ARG, "target variable %d of the synaptic code", k);
        for (int q = 0; q < k; ++q)
            if (post_vars[q] == post_vars[k]) return hs_fail(HS_ERR_ARG, "target variable %d twice", post_vars[k]);
        m.post_names.push_back(post_names[k]);
    }
    for (int k = 0; k < npre; ++k) {
        if (!pre_names[k] || pre_vars[k] < 0 || pre_vars[k] >= gs.mi.nvars)
            return hs_fail(HS_ERR_ARG, "source variable %d of the synaptic code", k);
        m.pre_names.push_back(pre_names[k]);
    }
    HS_TRY(hs_cg_check_syn(m));
    const size_t ns = (size_t)(P.nsyn ? P.nsyn : 1);
    for (int k = 0; k < nsvars; ++k) {
        if (m.syn_names[k] == "w" && P.wmode == HS_W_SYNAPSE) P.sv[k] = P.w; /* the pathway's weights */
        else HS_TRY(hs_dalloc(net, hs_rsize(net) * ns, &P.sv[k], true));
    }
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * ns, (void **)&P.lastup, true)); /* lastupdate = 0 s */
    P.nsv = nsvars;
    P.post_vars.assign(post_vars, post_vars + npost);
    P.pre_vars.assign(pre_vars, pre_vars + npre);
    P.cg = (int)net->cg_syn.size();
    P.dmode = HS_DELIVER_CODE;
    net->cg_syn.push_back(m);
    net->n_scatter--;
    net->desc_dirty = true;
    return HS_OK;
}

extern "C" int hs_net_pathway_code_post(hs_net *net, int pid, const char *on_post, int npost,
                                        const int *post_vars, const char *const *post_names)
{
    if (!net || pid < 0 || pid >= (int)net->p.size()) return hs_fail(HS_ERR_ARG, "bad pathway");
    if (net->built) return hs_fail(HS_ERR_STATE, "synaptic code must be set before the first run");
    HPath &P = net->p[pid];
    if (P.dmode != HS_DELIVER_CODE) return hs_fail(HS_ERR_ARG, "on_post needs a pathway with on_pre code");
    if (P.qstart) return hs_fail(HS_ERR_STATE, "the pathway has on_post code already");
    if (!on_post || !*on_post) return hs_fail(HS_ERR_ARG, "empty on_post code");
    if (npost < 0 || npost > HS_MAX_VARS || (npost && (!post_vars || !post_names)))
        return hs_fail(HS_ERR_ARG, "hs_net_pathway_code_post: bad variable list");
    const HGroup &gt = net->g[P.tgt];
    HsSynCgModel m = net->cg_syn[(size_t)P.cg];
    m.on_post = on_post;
    m.post_read_names.clear();
    for (int k = 0; k < npost; ++k) {
        if (!post_names[k] || post_vars[k] < 0 || post_vars[k] >= gt.mi.nvars)
            return hs_fail(HS_ERR_ARG, "target variable %d of the on_post code", k);
        m.post_read_names.push_back(post_names[k]);
    }
    HS_TRY(hs_cg_check_syn(m));
    /* the source neuron of each synapse, from the bundle starts */
    const HGroup &gs = net->g[P.src];
    const size_t nb = (size_t)gs.n * (size_t)P.D + 1, ns = (size_t)(P.nsyn ? P.nsyn : 1);
    std::vector<int32_t> bs(nb), sp(ns, 0);
    HS_HIP(hipMemcpyAsync(bs.data(), P.bstart, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    for (size_t r = 0; r + 1 < nb; ++r)
        for (int32_t q = bs[r]; q < bs[r + 1]; ++q) sp[(size_t)q] = (int32_t)(r / (size_t)P.D);
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * ns, (void **)&P.synpre, false));
    HS_HIP(hipMemcpyAsync(P.synpre, sp.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    HS_TRY(hs_post_index(net, P));
    P.post_read_vars.assign(post_vars, post_vars + npost);
    net->cg_syn[(size_t)P.cg] = m;
    net->desc_dirty = true;
    return HS_OK;
}

/* a synapse variable of a code pathway (HS_SYN_CODE + k) from a host array in bundle order */
extern "C" int hs_net_set_synapse_var(hs_net *net, int pid, int var, const double *values, int64_t n)
{
    if (!net || pid < 0 || pid >= (int)net->p.size() || !values) return hs_fail(HS_ERR_ARG, "bad pathway");
    HPath &P = net->p[pid];
    if (P.dmode != HS_DELIVER_CODE || var < HS_SYN_CODE || var >= HS_SYN_CODE + P.nsv)
        return hs_fail(HS_ERR_ARG, "synapse variable %d of a pathway with synaptic code", var);
    if (n != P.nsyn) return hs_fail(HS_ERR_ARG, "%lld values for %lld synapses", (long long)n, (long long)P.nsyn);
    if (!n) return HS_OK;
    void *dst = P.sv[var - HS_SYN_CODE];
    if (net->prec == HS_F64) {
        HS_HIP(hipMemcpyAsync(dst, values, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    } else {
        std::vector<float> f((size_t)n);
        for (int64_t s = 0; s < n; ++s) f[(size_t)s] = (float)values[s];
        HS_HIP(hipMemcpyAsync(dst, f.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    }
    return HS_OK;
}

/* a synapse variable (HS_SYN_*) of a pathway in bundle order (hs_net_get_synapses' order) */
extern "C" int hs_net_get_synapse_var(hs_net *net, int pid, int var, double *out, int64_t cap)
{
    if (!net || pid < 0 || pid >= (int)net->p.size() || !out) return hs_fail(HS_ERR_ARG, "bad pathway");
    const HPath &P = net->p[pid];
    if (P.dmode == HS_DELIVER_GATHER) return hs_fail(HS_ERR_UNSUPPORTED, "synapse readback of gather-mode pathways");
    if (cap < P.nsyn) return hs_fail(HS_ERR_ARG, "capacity %lld < %lld synapses", (long long)cap, (long long)P.nsyn);
    if (P.dmode == HS_DELIVER_CODE && var == HS_SYN_LASTUPDATE) {
        std::vector<int32_t> v((size_t)(P.nsyn ? P.nsyn : 1));
        if (P.nsyn)
            HS_HIP(hipMemcpyAsync(v.data(), P.lastup, sizeof(int32_t) * (size_t)P.nsyn, hipMemcpyDeviceToHost,
                                  g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        for (int64_t q = 0; q < P.nsyn; ++q) out[q] = (double)v[(size_t)q];
        return HS_OK;
    }
    if (P.dmode == HS_DELIVER_CODE && var >= HS_SYN_CODE && var < HS_SYN_CODE + P.nsv) {
        if (P.nsyn == 0) return HS_OK;
        const size_t n = (size_t)P.nsyn;
        const void *src = P.sv[var - HS_SYN_CODE];
        if (net->prec == HS_F64) {
            HS_HIP(hipMemcpyAsync(out, src, sizeof(double) * n, hipMemcpyDeviceToHost, g_rt.stream));
            HS_HIP(hipStreamSynchronize(g_rt.stream));
        } else {
            std::vector<float> v(n);
            HS_HIP(hipMemcpyAsync(v.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost, g_rt.stream));
            HS_HIP(hipStreamSynchronize(g_rt.stream));
            for (size_t s = 0; s < n; ++s) out[s] = (double)v[s];
        }
        return HS_OK;
    }
    if (var < HS_SYN_W || var > HS_SYN_LASTUPDATE) return hs_fail(HS_ERR_ARG, "bad synapse variable %d", var);
    if (var != HS_SYN_W && P.dmode != HS_DELIVER_PLASTIC)
        return hs_fail(HS_ERR_ARG, "synapse variable %d exists on STDP pathways only", var);
    if (P.nsyn == 0) return HS_OK;
    if (var == HS_SYN_W && P.wmode != HS_W_SYNAPSE)
        return hs_net_get_synapses(net, pid, nullptr, nullptr, nullptr, out, cap);
    const size_t n = (size_t)P.nsyn;
    if (var == HS_SYN_LASTUPDATE) {
        std::vector<int32_t> v(n);
        HS_HIP(hipMemcpyAsync(v.data(), P.lastup, sizeof(int32_t) * n, hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        for (size_t s = 0; s < n; ++s) out[s] = (double)v[s];
        return HS_OK;
    }
    const void *src = var == HS_SYN_W ? P.w : (var == HS_SYN_APRE ? P.apre : P.apost);
    if (net->prec == HS_F64) {
        HS_HIP(hipMemcpyAsync(out, src, sizeof(double) * n, hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    } else {
        std::vector<float> v(n);
        HS_HIP(hipMemcpyAsync(v.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        for (size_t s = 0; s < n; ++s) out[s] = (double)v[s];
    }
    return HS_OK;
}

/* the group's recording monitor (created on first use), then for a range other
   than the whole group a view on it */
static int hs_add_spike_monitor(hs_net *net, int gid, int start, int stop, int64_t capacity, int *mid)
{
    if (!net || gid < 0 || gid >= (int)net->g.size()) return hs_fail(HS_ERR_ARG, "bad group");
    HGroup &g = net->g[gid];
    if (start < 0 || stop > g.n || stop <= start) return hs_fail(HS_ERR_ARG, "bad spike monitor range");
    if (capacity <= 0) capacity = 1 << 20;
    if (g.smon < 0) {
        HMon M;
        M.kind = 0;
        M.gid = gid;
        M.var = 0;
        M.start = 0;
        M.stop = g.n;
        M.cap = capacity;
        HS_TRY(hs_dalloc(net, sizeof(int32_t) * capacity, (void **)&M.mi, false));
        HS_TRY(hs_dalloc(net, sizeof(int32_t) * capacity, (void **)&M.mt, false));
        HS_TRY(hs_dalloc(net, sizeof(unsigned long long), (void **)&M.mn, true));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        net->m.push_back(M);
        g.smon = (int)net->m.size() - 1;
        net->desc_dirty = true;
        hs_drop_graphs(net);
        if (start == 0 && stop == g.n) {
            if (mid) *mid = g.smon;
            return HS_OK;
        }
    }
    HMon V;
    V.kind = 0;
    V.gid = gid;
    V.start = start;
    V.stop = stop;
    V.rec = g.smon;
    net->m.push_back(V);
    if (mid) *mid = (int)net->m.size() - 1;
    return HS_OK;
}

extern "C" int hs_net_add_spike_monitor(hs_net *net, int gid, int64_t capacity, int *mid)
{
    if (!net || gid < 0 || gid >= (int)net->g.size()) return hs_fail(HS_ERR_ARG, "bad group");
    return hs_add_spike_monitor(net, gid, 0, net->g[gid].n, capacity, mid);
}

extern "C" int hs_net_add_spike_monitor_range(hs_net *net, int gid, int start, int stop, int64_t capacity, int *mid)
{
    return hs_add_spike_monitor(net, gid, start, stop, capacity, mid);
}

extern "C" int hs_net_add_synapse_monitor(hs_net *net, int pid, int var, const int32_t *indices, int n, int *mid)
{
    if (!net || pid < 0 || pid >= (int)net->p.size()) return hs_fail(HS_ERR_ARG, "bad pathway");
    const HPath &P = net->p[pid];
    if (net->engine != HS_ENGINE_PLAIN && net->built && net->mode != 1)
        return hs_fail(HS_ERR_UNSUPPORTED, "synapse StateMonitors run in the per-step forms: add them before the "
                                           "first run of a fused network");
    const bool code_var = P.dmode == HS_DELIVER_CODE && var >= HS_SYN_CODE && var < HS_SYN_CODE + P.nsv;
    if (!code_var) {
        if (var < HS_SYN_W || var > HS_SYN_APOST) return hs_fail(HS_ERR_ARG, "bad synapse variable %d", var);
        if (var == HS_SYN_W && P.wmode != HS_W_SYNAPSE)
            return hs_fail(HS_ERR_ARG, "w is recorded on pathways with per-synapse weights");
        if (var != HS_SYN_W && P.dmode != HS_DELIVER_PLASTIC)
            return hs_fail(HS_ERR_ARG, "synapse variable %d exists on STDP pathways only", var);
    }
    if (P.dmode == HS_DELIVER_GATHER) return hs_fail(HS_ERR_UNSUPPORTED, "gather-mode pathways");
    if (n <= 0 || !indices) return hs_fail(HS_ERR_ARG, "no synapses to record");
    HMon M;
    M.kind = 3;
    M.gid = P.tgt;
    M.var = var;
    M.pid = pid;
    M.hsel.assign(indices, indices + n);
    for (int32_t x : M.hsel)
        if (x < 0 || (int64_t)x >= P.nsyn) return hs_fail(HS_ERR_ARG, "synapse %d outside [0, %lld)", x, (long long)P.nsyn);
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * (size_t)n, (void **)&M.sidx, false));
    HS_HIP(hipMemcpyAsync(M.sidx, M.hsel.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    M.t0 = net->step;
    net->m.push_back(M);
    net->desc_dirty = true;
    hs_drop_graphs(net);
    if (mid) *mid = (int)net->m.size() - 1;
    return HS_OK;
}

extern "C" int hs_net_add_state_monitor(hs_net *net, int gid, int var, int start, int stop, int *mid)
{
    if (!net || gid < 0 || gid >= (int)net->g.size()) return hs_fail(HS_ERR_ARG, "bad group");
    HGroup &g = net->g[gid];
    if (var < 0 || var >= g.mi.nvars || start < 0 || stop > g.n || stop <= start)
        return hs_fail(HS_ERR_ARG, "bad state monitor range");
    if ((int)g.stmons.size() >= HS_MAX_STATE_MONS) return hs_fail(HS_ERR_UNSUPPORTED, "at most %d state monitors per group", HS_MAX_STATE_MONS);
    HMon M;
    M.kind = 1;
    M.gid = gid;
    M.var = var;
    M.start = start;
    M.stop = stop;
    M.width = stop - start;
    M.t0 = net->step;
    net->m.push_back(M);
    g.stmons.push_back((int)net->m.size() - 1);
    net->desc_dirty = true;
    hs_drop_graphs(net);
    if (mid) *mid = (int)net->m.size() - 1;
    return HS_OK;
}

extern "C" int hs_net_add_state_monitor_indices(hs_net *net, int gid, int var, const int32_t *indices, int n,
                                                int every, int *mid)
{
    if (!net || gid < 0 || gid >= (int)net->g.size()) return hs_fail(HS_ERR_ARG, "bad group");
    HGroup &g = net->g[gid];
    if (var < 0 || var >= g.mi.nvars || n <= 0 || !indices || every < 1)
        return hs_fail(HS_ERR_ARG, "bad state monitor (variable, %d indices, every %d steps)", n, every);
    if ((int)g.stmons.size() >= HS_MAX_STATE_MONS)
        return hs_fail(HS_ERR_UNSUPPORTED, "at most %d state monitors per group", HS_MAX_STATE_MONS);
    std::vector<int32_t> map((size_t)g.n, -1);
    for (int k = 0; k < n; ++k) {
        const int32_t x = indices[k];
        if (x < 0 || x >= g.n) return hs_fail(HS_ERR_ARG, "record index %d outside [0, %d)", x, g.n);
        if (map[(size_t)x] >= 0) return hs_fail(HS_ERR_ARG, "record index %d given twice", x);
        map[(size_t)x] = k;
    }
    HMon M;
    M.kind = 1;
    M.gid = gid;
    M.var = var;
    M.start = 0;
    M.stop = g.n;
    M.width = n;
    M.every = every;
    M.hidx.assign(indices, indices + n);
    M.t0 = (net->step + every - 1) / every * every; /* the monitor's clock ticks at multiples of its dt */
    HS_TRY(hs_dalloc(net, sizeof(int32_t) * (size_t)g.n, (void **)&M.map, false));
    HS_HIP(hipMemcpyAsync(M.map, map.data(), sizeof(int32_t) * (size_t)g.n, hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    net->m.push_back(M);
    g.stmons.push_back((int)net->m.size() - 1);
    net->desc_dirty = true;
    hs_drop_graphs(net);
    if (mid) *mid = (int)net->m.size() - 1;
    return HS_OK;
}

extern "C" int hs_net_set_monitor_interval(hs_net *net, int mid, int every)
{
    if (!net || mid < 0 || mid >= (int)net->m.size() || every < 1) return hs_fail(HS_ERR_ARG, "bad monitor or interval");
    HMon &M = net->m[mid];
    if (M.kind != 1 && M.kind != 3)
        return hs_fail(HS_ERR_UNSUPPORTED, "only StateMonitors (of neurons or synapses) take an interval");
    if (hs_mon_rows(M, net->step) > 0)
        return hs_fail(HS_ERR_STATE, "the interval is set before the monitor records its first row");
    M.every = every;
    M.t0 = (net->step + every - 1) / every * every;
    net->desc_dirty = true;
    hs_drop_graphs(net);
    return HS_OK;
}

/* PopulationRateMonitor (ratemonitor.cu): spikes per step of neurons [start, stop) */
extern "C" int hs_net_add_rate_monitor(hs_net *net, int gid, int start, int stop, int *mid)
{
    if (!net || gid < 0 || gid >= (int)net->g.size()) return hs_fail(HS_ERR_ARG, "bad group");
    HGroup &g = net->g[gid];
    if (start < 0 || stop > g.n || stop <= start) return hs_fail(HS_ERR_ARG, "bad rate monitor range");
    if (g.rmon >= 0) return hs_fail(HS_ERR_UNSUPPORTED, "one rate monitor per group");
    HMon M;
    M.kind = 2;
    M.gid = gid;
    M.var = 0;
    M.start = start;
    M.stop = stop;
    M.t0 = net->step;
    net->m.push_back(M);
    g.rmon = (int)net->m.size() - 1;
    net->desc_dirty = true;
    hs_drop_graphs(net);
    if (mid) *mid = (int)net->m.size() - 1;
    return HS_OK;
}

/* rate[k] = count / dt / (stop - start) for each recorded step, as ratemonitor.cu:98 computes it */
extern "C" int hs_net_get_rate_monitor(hs_net *net, int mid, double *rate, int64_t cap)
{
    if (!net || mid < 0 || mid >= (int)net->m.size() || net->m[mid].kind != 2 || !rate)
        return hs_fail(HS_ERR_ARG, "bad rate monitor");
    const HMon &M = net->m[mid];
    const int64_t rows = std::max<int64_t>(0, net->step - M.t0);
    if (cap < rows) return hs_fail(HS_ERR_ARG, "output too small");
    if (!rows) return HS_OK;
    std::vector<int32_t> c((size_t)rows);
    HS_HIP(hipMemcpyAsync(c.data(), M.buf, sizeof(int32_t) * (size_t)rows, hipMemcpyDeviceToHost, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    for (int64_t k = 0; k < rows; ++k) rate[k] = 1.0 * c[(size_t)k] / net->dt / (double)(M.stop - M.start);
    return HS_OK;
}

extern "C" int hs_net_set_option(hs_net *net, int key, long long value)
{
    if (!net || key < 0 || key > 11) return hs_fail(HS_ERR_ARG, "bad option key");
    if (key == 7 && (value < 0 || value > 1)) return hs_fail(HS_ERR_ARG, "kernel timing: 0 off, 1 on");
    if (key == 11 && (value < 0 || value > 1)) return hs_fail(HS_ERR_ARG, "fast math: 0 off, 1 on");
    if (key == 11 && value && net->prec != HS_F32)
        return hs_fail(HS_ERR_UNSUPPORTED, "the fast form of the HH update is an FP32 form");
    if (key == 8 && (value < 0 || value > (1 << 16))) return hs_fail(HS_ERR_ARG, "bad noise block count");
    if (key == 9 && (value < 0 || value > 1)) return hs_fail(HS_ERR_ARG, "pipelined windows: 0 off, 1 on");
    if (key == 9 && net->built) return hs_fail(HS_ERR_STATE, "the engine form is fixed at the first run");
    if (key == 7)
        for (int q = 0; q < 5; ++q) {
            net->kt_ms[q] = 0;
            net->kt_n[q] = 0;
        }
    if (key == 6 && (value < 64 || value > 1024 || value % 64))
        return hs_fail(HS_ERR_ARG, "gather block size must be a multiple of 64 in [64,1024]");
    if (key == 5 && (value < 0 || value > 2)) return hs_fail(HS_ERR_ARG, "delivery mode: 0 auto, 1 scatter, 2 gather");
    if (key == 0 && value != 64 && value != 128 && value != 256)
        return hs_fail(HS_ERR_ARG, "update block size must be 64, 128 or 256 (a divisor of the 256-neuron chunk)");
    if ((key == 1 || key == 3 || key == 10) && (value < 0 || value > (1 << 20)))
        return hs_fail(HS_ERR_ARG, "bad grid size");
    if (key == 2 && (value < 0 || value > 100000)) return hs_fail(HS_ERR_ARG, "bad graph chunk");
    if (key == 4 && (value < 0 || value > HS_MAX_DELAY_SLOTS)) return hs_fail(HS_ERR_ARG, "bad window");
    if (key == 4 && net->built) return hs_fail(HS_ERR_STATE, "the window is fixed at the first run");
    net->opt[key] = value;
    net->desc_dirty = true;
    hs_drop_graphs(net);
    return HS_OK;
}

extern "C" int hs_net_get_option(hs_net *net, int key, long long *value)
{
    if (!net || key < 0 || key > 11 || !value) return hs_fail(HS_ERR_ARG, "bad option key");
    *value = net->opt[key];
    return HS_OK;
}

/* the model id the kernels are specialised for: the HH model runs its fast
   form when option 11 is on (FP32) */
static int hs_kernel_model(const hs_net *net, int model)
{
    return (model == HS_MODEL_HH_COBA && net->opt[11] && net->prec == HS_F32) ? HS_MODEL_HH_FAST : model;
}

/* allocate rings/accumulators once (first run) */
static int hs_build(hs_net *net)
{
    if (net->built) return HS_OK;
    if (net->n_plastic && net->engine != HS_ENGINE_PLAIN && net->n_gather)
        return hs_fail(HS_ERR_UNSUPPORTED, "STDP pathways in the fused form need scatter delivery");
    /* form of the step loop */
    int dmin_all = 1 << 30, dspan = 1;
    for (const HPath &P : net->p) {
        dmin_all = std::min(dmin_all, P.dmin);
        dspan = std::max(dspan, P.D);
    }
    if (net->p.empty()) dmin_all = 1;
    net->dmin_all = dmin_all;
    net->window = 1;
    if (net->engine == HS_ENGINE_PLAIN) {
        net->mode = 0;
    } else if (net->n_plastic || !net->cg_syn.empty() ||
               std::any_of(net->m.begin(), net->m.end(), [](const HMon &M) { return M.kind == 3; })) {
        /* STDP: its pre and post code run after each step's update (the post
           code needs the step's spikes), effects into the step's accumulator;
           synapse StateMonitors record at the start of each step */
        net->mode = 1;
    } else if (net->p.empty() || dmin_all < 1) {
        /* effects due in the same step: update, then deliver -- or, pipelined,
           one launch per step whose update blocks push every synapse of delay
           0 and 1 of their own spikes (all due after this launch's update) */
        net->mode = (!net->p.empty() && net->opt[9] && net->n_gather == 0) ? 4 : 1;
    } else {
        /* the window cannot exceed the smallest delay, and its (step, delay) pairs
           must fit the delivery kernel's prefix table */
        int w = net->opt[4] > 0 ? (int)net->opt[4] : HS_MAX_WINDOW;
        w = std::min(w, HS_MAX_WINDOW);
        w = std::min(w, dmin_all);
        w = std::min(w, HS_MAX_DELAY_SLOTS / dspan);
        w = std::max(w, 1);
        net->window = w;
        /* the single-launch kernels carry scatter delivery only */
        if (net->n_gather == 0 && net->opt[9]) net->mode = 4;
        else if (w == 1 && net->n_gather == 0) net->mode = 2;
        else net->mode = 3;
    }
    const int W = net->window;
    for (HGroup &g : net->g) {
        const int gid = (int)(&g - &net->g[0]);
        int dmax = 0;
        for (const HPath &P : net->p)
            if (P.src == gid) dmax = std::max(dmax, P.dmax);
        g.Q = hs_ring_size(dmax + 2 * W); /* see the ring invariants in hs_engine.hip */
        /* large groups split each step's spike list into sub-lists so that
           thousands of blocks do not reserve list space on one counter; the
           delivery prefix tables bound the (step, delay, sub-list) items */
        const int nchunks = (g.n + 255) / 256;
        int nsub = 1;
        if (nchunks >= 512)
            while (nsub < 32 && nsub * 2 * 128 <= nchunks) nsub *= 2;
        for (const HPath &P : net->p) {
            if (P.src != gid) continue;
            const int items = P.dmode == HS_DELIVER_GATHER ? P.D - 1 + W : W * P.D;
            while (nsub > 1 && items * nsub > HS_MAX_DELAY_SLOTS) nsub /= 2;
        }
        g.nsub = nsub;
        g.subcap = (nchunks + nsub - 1) / nsub * 256;
        HS_TRY(hs_dalloc(net, sizeof(int32_t) * (size_t)g.Q * nsub * g.subcap, (void **)&g.hist, false));
        HS_TRY(hs_dalloc(net, sizeof(int32_t) * (size_t)g.Q * nsub * HS_CNT_STRIDE, (void **)&g.hist_cnt, true));
    }
    if (net->engine == HS_ENGINE_FUSED) {
        for (const HPath &P : net->p) {
            HGroup &gt = net->g[P.tgt];
            /* pipelined windows write up to 3W-2 steps ahead of the window they update */
            gt.nacc = hs_ring_size(net->mode == 4 ? 3 * W + 1 : W + 1);
            std::vector<int> vars = P.dmode == HS_DELIVER_CODE ? P.post_vars : std::vector<int>{P.var};
            for (int v : vars)
                if (!gt.acc[v]) HS_TRY(hs_dalloc(net, hs_rsize(net) * (size_t)gt.n * gt.nacc, &gt.acc[v], true));
        }
    }
    /* windowed form: the noise of a window is drawn one window ahead, by extra
       blocks of the previous update launch (the update's own blocks leave most
       of the chip idle), into a ring of 2W steps that the update reads */
    if (net->mode == 3 || net->mode == 4)
        for (HGroup &g : net->g)
            if (g.model == HS_MODEL_LIF_NOISE || (g.code >= 0 && net->cg_models[(size_t)g.code].noisy)) {
                g.xi_slots = hs_ring_size(2 * W);
                HS_TRY(hs_dalloc(net, hs_rsize(net) * (size_t)g.n * g.xi_slots, &g.xi, false));
                net->noise_ahead = true;
            }
    /* code models: the engine kernels compiled with them (cached per program) */
    if (!net->cg_models.empty() || !net->cg_syn.empty()) {
        /* one code model in every group: kernels specialised for it */
        int ms = net->cg_models.empty() ? 0 : net->g[0].model;
        for (const HGroup &g : net->g)
            if (g.model != ms) ms = 0;
        if (!HS_IS_CODE_MODEL(ms)) ms = 0;
        HS_TRY(hs_cg_kernels(net->prec, ms, net->cg_models, net->cg_syn, &net->cg));
    }
    HsCtl c;
    memset(&c, 0, sizeof(c));
    c.base = net->step;
    HS_HIP(hipMemcpyAsync(net->ctl, &c, sizeof(c), hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    net->built = true;
    return HS_OK;
}

template <typename R>
static int hs_upload_desc_t(hs_net *net)
{
    const int block = (int)net->opt[0];
    std::vector<DGroup<R>> dg(net->g.size());
    int tiles = 0;
    for (size_t gi = 0; gi < net->g.size(); ++gi) {
        const HGroup &g = net->g[gi];
        DGroup<R> &d = dg[gi];
        memset(&d, 0, sizeof(d));
        d.model = hs_kernel_model(net, g.model);
        d.n = g.n;
        d.nvars = g.mi.nvars;
        d.Q = g.Q;
        d.has_reset = g.mi.has_reset;
        d.ref_steps = g.ref_steps;
        d.stream = g.stream;
        for (int v = 0; v < HS_MAX_VARS; ++v) {
            d.var[v] = (R *)g.var[v];
            d.acc[v] = (R *)g.acc[v];
        }
        d.nacc = g.nacc > 0 ? g.nacc : 1;
        d.last = g.last;
        d.hist = g.hist;
        d.hist_cnt = g.hist_cnt;
        d.nsub = g.nsub;
        d.subcap = g.subcap;
        for (int j = 0; j < 16; ++j) d.k[j] = (R)g.k[j];
        d.vth = g.mi.i_vth >= 0 ? (R)g.params[g.mi.i_vth] : (R)0;
        d.gen_off = g.gen_off;
        d.gen_idx = g.gen_idx;
        d.gen_steps = g.gen_steps;
        d.gen_period = g.gen_period;
        if (g.rmon >= 0) {
            const HMon &M = net->m[g.rmon];
            d.rmon_cnt = (int32_t *)M.buf;
            d.rmon_t0 = M.t0;
            d.rmon_rows = M.rows;
            d.rmon_start = M.start;
            d.rmon_stop = M.stop;
        }
        d.vr = g.mi.i_vr >= 0 ? (R)g.params[g.mi.i_vr] : (R)0;
        d.dt = net->dt;
        d.seed = net->seed;
        if (g.smon >= 0) {
            const HMon &M = net->m[g.smon];
            d.smon_i = M.mi;
            d.smon_t = M.mt;
            d.smon_n = M.mn;
            d.smon_cap = M.cap;
        }
        d.nst = (int)g.stmons.size();
        for (int s = 0; s < d.nst; ++s) {
            const HMon &M = net->m[g.stmons[s]];
            d.st_var[s] = M.var;
            d.st_start[s] = M.start;
            d.st_stop[s] = M.stop;
            d.st_t0[s] = M.t0;
            d.st_rows[s] = M.rows;
            d.st_buf[s] = (R *)M.buf;
            d.st_map[s] = M.map;
            d.st_every[s] = M.every;
            d.st_width[s] = (int)M.width;
        }
        d.tile_begin = tiles;
        d.xi = (R *)g.xi;
        d.xi_slots = g.xi_slots;
        tiles += (g.n + block - 1) / block;
    }
    std::vector<DPath<R>> dp(net->p.size());
    for (size_t pi = 0; pi < net->p.size(); ++pi) {
        const HPath &P = net->p[pi];
        const HGroup &gs = net->g[P.src];
        const HGroup &gt = net->g[P.tgt];
        DPath<R> &d = dp[pi];
        memset(&d, 0, sizeof(d));
        d.src = P.src;
        d.tgt = P.tgt;
        d.var = P.var;
        d.pid = (unsigned int)pi;
        d.seed = net->seed;
        d.n_src = gs.n;
        d.dmin = P.dmin;
        d.D = P.D;
        d.Q = gs.Q;
        d.wmode = P.wmode;
        d.hist = gs.hist;
        d.hist_cnt = gs.hist_cnt;
        d.cstride = HS_CNT_STRIDE;
        d.nsub = gs.nsub;
        d.subcap = gs.subcap;
        d.bstart = P.bstart;
        d.post = P.post;
        d.w = (const R *)P.w;
        d.wscalar = (R)P.wscalar;
        d.target = (R *)gt.var[P.var];
        d.acc = (R *)gt.acc[P.var];
        d.nacc = gt.nacc;
        d.n_tgt = gt.n;
        d.dmode = P.dmode;
        d.n_tiles = P.n_tiles;
        d.goff = P.goff;
        d.gsyn = P.gsyn;
        /* pipelined windows: delays < 2W are pushed at emission */
        d.nnear = net->mode == 4 ? std::max(0, std::min(P.D, 2 * net->window - P.dmin)) : 0;
        if (P.dmode == HS_DELIVER_PLASTIC) {
            d.wm = (R *)P.w;
            d.apre = (R *)P.apre;
            d.apost = (R *)P.apost;
            d.lastup = P.lastup;
            d.qstart = P.qstart;
            d.qsyn = P.qsyn;
            d.taupre = P.stdp[0];
            d.taupost = P.stdp[1];
            d.dt = net->dt;
            for (int j = 0; j < 5; ++j) d.sk[j] = (R)P.stdp[2 + j];
            d.thist = gt.hist;
            d.thist_cnt = gt.hist_cnt;
            d.tQ = gt.Q;
            d.tnsub = gt.nsub;
            d.tsubcap = gt.subcap;
        }
        d.dt = net->dt;
        d.cg = P.cg;
        d.nsv = P.nsv;
        for (int k = 0; k < P.nsv; ++k) d.sv[k] = (R *)P.sv[k];
        d.npost = (int)P.post_vars.size();
        for (int k = 0; k < d.npost; ++k) d.post_var[k] = P.post_vars[(size_t)k];
        d.npre = (int)P.pre_vars.size();
        for (int k = 0; k < d.npre; ++k) d.pre_var[k] = P.pre_vars[(size_t)k];
        if (P.dmode == HS_DELIVER_CODE) {
            d.lastup = P.lastup;
            d.synpre = P.synpre;
            d.npr = (int)P.post_read_vars.size();
            for (int k = 0; k < d.npr; ++k) d.post_read_var[k] = P.post_read_vars[(size_t)k];
            if (P.qstart) {
                d.qstart = P.qstart;
                d.qsyn = P.qsyn;
                d.thist = gt.hist;
                d.thist_cnt = gt.hist_cnt;
                d.tQ = gt.Q;
                d.tnsub = gt.nsub;
                d.tsubcap = gt.subcap;
            }
        }
    }
    if (!net->dgroups) HS_TRY(hs_dalloc(net, sizeof(DGroup<R>) * std::max<size_t>(1, dg.size()), &net->dgroups, false));
    if (!net->dpaths) HS_TRY(hs_dalloc(net, sizeof(DPath<R>) * std::max<size_t>(1, dp.size()), &net->dpaths, false));
    HS_HIP(hipMemcpyAsync(net->dgroups, dg.data(), sizeof(DGroup<R>) * dg.size(), hipMemcpyHostToDevice, g_rt.stream));
    if (!dp.empty())
        HS_HIP(hipMemcpyAsync(net->dpaths, dp.data(), sizeof(DPath<R>) * dp.size(), hipMemcpyHostToDevice, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    net->total_tiles = tiles;
    net->desc_dirty = false;
    return HS_OK;
}

static int hs_upload_desc(hs_net *net)
{
    return net->prec == HS_F64 ? hs_upload_desc_t<double>(net) : hs_upload_desc_t<float>(net);
}

/* grow state-monitor buffers to hold rows up to step `until` */
static int hs_grow_state_monitors(hs_net *net, int64_t until)
{
    for (HMon &M : net->m) {
        if (M.kind == 0) continue;
        const int64_t need = hs_mon_rows(M, until);
        if (need <= M.rows) continue;
        int64_t rows = std::max<int64_t>(need, M.rows * 2);
        /* state rows: one value per recorded neuron; rate rows: one int32 count, zeroed */
        const size_t rowb = M.kind == 1   ? hs_rsize(net) * (size_t)M.width
                            : M.kind == 3 ? hs_rsize(net) * M.hsel.size()
                                          : sizeof(int32_t);
        void *nb = nullptr;
        HS_TRY(hs_dalloc(net, rowb * (size_t)rows, &nb, M.kind == 2));
        if (M.buf && M.rows)
            HS_HIP(hipMemcpyAsync(nb, M.buf, rowb * (size_t)M.rows, hipMemcpyDeviceToDevice, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        hs_dfree(net, M.buf);
        M.buf = nb;
        M.rows = rows;
        net->desc_dirty = true;
    }
    return HS_OK;
}

/* grid of the delivery / fused kernels */
static int hs_deliver_grid(const hs_net *net)
{
    if (net->opt[1] > 0) return (int)net->opt[1];
    return g_rt.n_cu * 4;
}

/* pipelined windows: does any pathway have synapses of delay >= 2W, which
   the pull blocks deliver (the update blocks push the nearer ones)? */
static bool hs_has_far(const hs_net *net)
{
    for (const HPath &P : net->p)
        if (std::max(0, std::min(P.D, 2 * net->window - P.dmin)) < P.D) return true;
    return false;
}

/* delivery blocks of the single-launch fused kernel */
static int hs_fused_deliver_blocks(const hs_net *net)
{
    if (net->opt[3] > 0) return (int)net->opt[3];
    return g_rt.n_cu * 2;
}

/* blocks drawing the next window's noise beside the update blocks */
static int hs_noise_grid(const hs_net *net)
{
    if (net->opt[8] > 0) return (int)net->opt[8];
    return g_rt.n_cu * 2;
}

/* one block per tile, or at most opt[10] blocks striding over the tiles */
static int hs_update_grid(const hs_net *net)
{
    const int cap = net->opt[10] > 0 ? (int)net->opt[10] : (1 << 20);
    return std::max(1, std::min(net->total_tiles, cap));
}

/* kernel timing (option 7): an event pair around one launch of kind `kind` */
static hipEvent_t hs_kt_event(hs_net *net)
{
    if (!net->kt_pool.empty()) {
        hipEvent_t e = net->kt_pool.back();
        net->kt_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

template <typename F>
static int hs_kt_launch(hs_net *net, int kind, F launch)
{
    if (!net->opt[7]) return launch();
    hipEvent_t a = hs_kt_event(net), b = hs_kt_event(net);
    if (!a || !b) return hs_fail(HS_ERR_HIP, "hipEventCreate failed");
    HS_HIP(hipEventRecord(a, g_rt.stream));
    const int r = launch();
    HS_HIP(hipEventRecord(b, g_rt.stream));
    net->kt_open.push_back({kind, {a, b}});
    return r;
}

/* after the stream has drained: add the open pairs to the totals */
static int hs_kt_collect(hs_net *net)
{
    for (auto &o : net->kt_open) {
        float ms = 0;
        HS_HIP(hipEventElapsedTime(&ms, o.second.first, o.second.second));
        net->kt_ms[o.first] += ms;
        net->kt_n[o.first] += 1;
        net->kt_pool.push_back(o.second.first);
        net->kt_pool.push_back(o.second.second);
    }
    net->kt_open.clear();
    return HS_OK;
}

extern "C" int hs_net_kernel_time(hs_net *net, int kind, double *total_ms, int64_t *launches)
{
    if (!net || kind < 0 || kind > 4)
        return hs_fail(HS_ERR_ARG, "kernel kind: 0 update, 1 scatter, 2 gather, 3 fused, 4 window");
    if (total_ms) *total_ms = net->kt_ms[kind];
    if (launches) *launches = net->kt_n[kind];
    return HS_OK;
}

/*
 * enqueue k steps (chunk offsets 0..k-1) and the tick.
 *   mode 0 plain          : update(t), deliver(t) in place
 *   mode 1 fused, delay 0 : update(t), deliver(t) into the accumulator
 *   mode 2 fused, W = 1   : one kernel per step
 *   mode 3 fused, W >= 2  : deliver(window), update(window) per W steps
 */
template <typename R>
static int hs_enqueue_steps_t(hs_net *net, int64_t k)
{
    const int block = (int)net->opt[0];
    const int ng = (int)net->g.size(), np = (int)net->p.size();
    const DGroup<R> *dg = (const DGroup<R> *)net->dgroups;
    const DPath<R> *dp = (const DPath<R> *)net->dpaths;
    const int fused = net->engine == HS_ENGINE_FUSED;
    int ms = net->g.empty() ? 0 : hs_kernel_model(net, net->g[0].model); /* one model for all groups: specialised kernels */
    for (const HGroup &g : net->g)
        if (hs_kernel_model(net, g.model) != ms) ms = 0;
    /* code models: the kernels compiled for them (every model dispatched) */
    const HsCgKernels *cg = net->cg_models.empty() ? nullptr : net->cg;
    const int ggrid = std::max(1, std::min(net->gather_tiles, 1 << 20));
    /* where a window's noise is drawn: by extra blocks of its scatter-delivery
       launch (two-launch windows), else one window ahead by the update or window
       launch, after a first launch for the chunk's first window */
    const int n_noise = net->noise_ahead ? hs_noise_grid(net) : 0;
    const bool noise_in_deliver = n_noise > 0 && net->mode == 3 && net->n_scatter > 0;
    const int n_noise_del = noise_in_deliver ? n_noise : 0;
    const int n_noise_ahead = noise_in_deliver ? 0 : n_noise;
    auto deliver = [&](int koff, int w) -> int {
        if (net->n_scatter)
            HS_TRY(hs_kt_launch(net, 1, [&]() -> int {
                HS_HIP(hs_launch_deliver<R>(dp, np, dg, ng, net->ctl, koff, w, fused, 256, hs_deliver_grid(net),
                                            n_noise_del, g_rt.stream));
                return HS_OK;
            }));
        if (net->n_gather)
            HS_TRY(hs_kt_launch(net, 2, [&]() -> int {
                HS_HIP(hs_launch_gather<R>(dp, np, net->ctl, koff, w, fused, (int)net->opt[6], ggrid, g_rt.stream));
                return HS_OK;
            }));
        return HS_OK;
    };
    auto update = [&](int koff, int w, int wmax, int fu) -> int {
        return hs_kt_launch(net, 0, [&]() -> int {
            if (cg)
                HS_HIP(hs_cg_launch_update(cg, dg, ng, net->total_tiles, net->ctl, koff, w, wmax, fu, block,
                                           hs_update_grid(net), n_noise_ahead, g_rt.stream));
            else
                HS_HIP(hs_launch_update<R>(dg, ng, net->total_tiles, net->ctl, koff, w, wmax, fu, ms, block,
                                           hs_update_grid(net), n_noise_ahead, g_rt.stream));
            return HS_OK;
        });
    };
    if (net->mode == 3 || net->mode == 4) {
        if (n_noise_ahead) HS_HIP(hs_launch_noise<R>(dg, ng, net->ctl, 0, net->window, n_noise, g_rt.stream));
        for (int64_t j = 0; j < k; j += net->window) {
            const int w = (int)std::min<int64_t>(net->window, k - j);
            if (net->mode == 4) {
                HS_TRY(hs_kt_launch(net, 4, [&]() -> int {
                    const int n_far = hs_has_far(net) ? hs_deliver_grid(net) : 0;
                    if (cg)
                        HS_HIP(hs_cg_launch_window(cg, dg, ng, net->total_tiles, dp, np, net->ctl, (int)j, w,
                                                   net->window, block, hs_update_grid(net), n_far, n_noise,
                                                   g_rt.stream));
                    else
                        HS_HIP(hs_launch_window<R>(dg, ng, net->total_tiles, dp, np, net->ctl, (int)j, w, net->window,
                                                   ms, block, hs_update_grid(net), n_far, n_noise, g_rt.stream));
                    return HS_OK;
                }));
            } else {
                HS_TRY(deliver((int)j, w));
                HS_TRY(update((int)j, w, net->window, 1));
            }
        }
    } else {
        for (int64_t j = 0; j < k; ++j) {
            if (net->mode == 2) {
                HS_TRY(hs_kt_launch(net, 3, [&]() -> int {
                    if (cg)
                        HS_HIP(hs_cg_launch_fused(cg, dg, ng, net->total_tiles, dp, np, net->ctl, (int)j, block,
                                                  hs_update_grid(net), hs_fused_deliver_blocks(net), g_rt.stream));
                    else
                        HS_HIP(hs_launch_fused<R>(dg, ng, net->total_tiles, dp, np, net->ctl, (int)j, ms, block,
                                                  hs_update_grid(net), hs_fused_deliver_blocks(net), g_rt.stream));
                    return HS_OK;
                }));
            } else {
                for (const HMon &M : net->m) /* synapse StateMonitors, when='start' */
                    if (M.kind == 3) {
                        const HPath &P = net->p[M.pid];
                        const void *src = M.var >= HS_SYN_CODE ? P.sv[M.var - HS_SYN_CODE]
                                          : (M.var == HS_SYN_W ? P.w : (M.var == HS_SYN_APRE ? P.apre : P.apost));
                        HS_HIP(hs_launch_synmon(net->prec, src, M.sidx, (int)M.hsel.size(), M.buf, M.rows, M.t0,
                                                M.every, net->ctl, (int)j, g_rt.stream));
                    }
                HS_TRY(update((int)j, 1, 1, fused));
                if (np > 0) HS_TRY(deliver((int)j, 1));
                if (net->n_plastic)
                    HS_HIP(hs_launch_stdp<R>(dp, np, net->ctl, (int)j, std::max(1, hs_rt_cus()), g_rt.stream));
                if (!net->cg_syn.empty()) {
                    HS_HIP(hs_cg_launch_syn_code(net->cg, dp, np, dg, net->ctl, (int)j, fused,
                                                 std::max(1, hs_rt_cus()), g_rt.stream));
                    if (std::any_of(net->cg_syn.begin(), net->cg_syn.end(),
                                    [](const HsSynCgModel &c) { return !c.on_post.empty(); }))
                        HS_HIP(hs_cg_launch_syn_post(net->cg, dp, np, dg, net->ctl, (int)j,
                                                     std::max(1, hs_rt_cus()), g_rt.stream));
                }
            }
        }
    }
    HS_HIP(hs_launch_tick(net->ctl, k, g_rt.stream));
    return HS_OK;
}

static int hs_enqueue_steps(hs_net *net, int64_t k)
{
    return net->prec == HS_F64 ? hs_enqueue_steps_t<double>(net, k) : hs_enqueue_steps_t<float>(net, k);
}

/* the instantiated graph of a chunk of k steps, captured on first use */
static int hs_chunk_graph(hs_net *net, int64_t k, hipGraphExec_t *out)
{
    auto it = net->graphs.find(k);
    if (it == net->graphs.end()) {
        hipGraph_t graph;
        HS_HIP(hipStreamBeginCapture(g_rt.stream, hipStreamCaptureModeThreadLocal));
        int r = hs_enqueue_steps(net, k);
        hipError_t e = hipStreamEndCapture(g_rt.stream, &graph);
        if (r != HS_OK) return r;
        if (e != hipSuccess) return hs_fail(HS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        hipGraphExec_t exec;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hs_fail(HS_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
        it = net->graphs.emplace(k, exec).first;
    }
    *out = it->second;
    return HS_OK;
}

static int hs_chunk(hs_net *net, int64_t k)
{
    if (net->opt[2] <= 0 || net->opt[7]) return hs_enqueue_steps(net, k);
    hipGraphExec_t exec;
    HS_TRY(hs_chunk_graph(net, k, &exec));
    HS_HIP(hipGraphLaunch(exec, g_rt.stream));
    return HS_OK;
}

/* device state, state-monitor rows and descriptors for a run of nsteps steps */
static int hs_prepare_run(hs_net *net, int64_t nsteps)
{
    HS_TRY(hs_build(net));
    HS_TRY(hs_grow_state_monitors(net, net->step + nsteps));
    if (net->desc_dirty) {
        hs_drop_graphs(net);
        HS_TRY(hs_upload_desc(net));
    }
    return HS_OK;
}

extern "C" int hs_net_prepare(hs_net *net, int64_t nsteps)
{
    if (!net) return hs_fail(HS_ERR_ARG, "null net");
    if (nsteps < 0) return hs_fail(HS_ERR_ARG, "negative step count");
    if (net->g.empty()) return hs_fail(HS_ERR_STATE, "network has no groups");
    HS_TRY(hs_prepare_run(net, nsteps));
    if (net->opt[2] <= 0 || net->opt[7] || nsteps == 0) return HS_OK;
    const int64_t K = net->opt[2];
    hipGraphExec_t exec;
    if (nsteps >= K) HS_TRY(hs_chunk_graph(net, K, &exec));
    if (nsteps % K) HS_TRY(hs_chunk_graph(net, nsteps % K, &exec));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    return HS_OK;
}

/* before a chunk of k steps: make sure the spike monitors have room */
static int hs_check_spike_monitors(hs_net *net, int64_t k, bool *resized)
{
    *resized = false;
    for (HMon &M : net->m) {
        if (M.kind != 0 || M.rec >= 0) continue;
        unsigned long long n = 0;
        HS_HIP(hipMemcpyAsync(&n, M.mn, sizeof(n), hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        if ((int64_t)n > M.cap)
            return hs_fail(HS_ERR_OVERFLOW, "spike monitor overflow: %llu records for capacity %lld", n,
                           (long long)M.cap);
        /* a chunk of k steps adds at most n_group * k records: keep that much room */
        const int64_t headroom = (int64_t)net->g[M.gid].n * k;
        if ((int64_t)n + headroom > M.cap) {
            int64_t cap = std::max<int64_t>(2 * M.cap, (int64_t)n + headroom);
            int32_t *ni = nullptr, *nt = nullptr;
            HS_TRY(hs_dalloc(net, sizeof(int32_t) * cap, (void **)&ni, false));
            HS_TRY(hs_dalloc(net, sizeof(int32_t) * cap, (void **)&nt, false));
            if (n) {
                HS_HIP(hipMemcpyAsync(ni, M.mi, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, g_rt.stream));
                HS_HIP(hipMemcpyAsync(nt, M.mt, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, g_rt.stream));
            }
            HS_HIP(hipStreamSynchronize(g_rt.stream));
            hs_dfree(net, M.mi);
            hs_dfree(net, M.mt);
            M.mi = ni;
            M.mt = nt;
            M.cap = cap;
            *resized = true;
        }
    }
    return HS_OK;
}

extern "C" int hs_net_run(hs_net *net, int64_t nsteps)
{
    if (!net) return hs_fail(HS_ERR_ARG, "null net");
    if (nsteps < 0) return hs_fail(HS_ERR_ARG, "negative step count");
    if (net->g.empty()) return hs_fail(HS_ERR_STATE, "network has no groups");
    HS_TRY(hs_prepare_run(net, nsteps));
    bool have_smon = false;
    for (const HMon &M : net->m) have_smon |= (M.kind == 0);
    const int64_t K = net->opt[2] > 0 ? net->opt[2] : 256;
    auto t0 = std::chrono::steady_clock::now();
    HS_HIP(hipEventRecord(net->ev0, g_rt.stream));
    int64_t done = 0;
    while (done < nsteps) {
        const int64_t k = std::min<int64_t>(K, nsteps - done);
        if (have_smon) {
            bool resized = false;
            HS_TRY(hs_check_spike_monitors(net, k, &resized));
            if (resized) {
                hs_drop_graphs(net);
                HS_TRY(hs_upload_desc(net));
            }
        }
        HS_TRY(hs_chunk(net, k));
        done += k;
    }
    HS_HIP(hipEventRecord(net->ev1, g_rt.stream));
    const int block = (int)net->opt[0];
    const bool fused = net->engine == HS_ENGINE_FUSED;
    if (net->cg && !net->cg_models.empty())
        HS_HIP(hs_cg_launch_finalize(net->cg, net->dgroups, (int)net->g.size(), net->total_tiles, net->ctl, fused,
                                     block, g_rt.stream));
    else if (net->prec == HS_F64)
        HS_HIP(hs_launch_finalize<double>((const DGroup<double> *)net->dgroups, (int)net->g.size(), net->total_tiles,
                                          net->ctl, fused, block, g_rt.stream));
    else
        HS_HIP(hs_launch_finalize<float>((const DGroup<float> *)net->dgroups, (int)net->g.size(), net->total_tiles,
                                         net->ctl, fused, block, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    auto t1 = std::chrono::steady_clock::now();
    HS_TRY(hs_kt_collect(net));
    if (net->mode == 4 && !net->p.empty()) {
        /* pipelined windows count events at emission: subtract those still in flight */
        if (net->prec == HS_F64)
            HS_HIP(hs_launch_pending<double>((const DPath<double> *)net->dpaths, (int)net->p.size(), net->ctl,
                                             g_rt.stream));
        else
            HS_HIP(hs_launch_pending<float>((const DPath<float> *)net->dpaths, (int)net->p.size(), net->ctl,
                                            g_rt.stream));
    }
    float ms = 0;
    HS_HIP(hipEventElapsedTime(&ms, net->ev0, net->ev1));
    net->last_gpu_ms = ms;
    net->last_wall = std::chrono::duration<double>(t1 - t0).count();
    net->last_steps = nsteps;
    net->step += nsteps;
    std::vector<HsCtl> cbuf(1);
    HsCtl &c = cbuf[0];
    HS_HIP(hipMemcpyAsync(&c, net->ctl, sizeof(c), hipMemcpyDeviceToHost, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    net->events = hs_ctl_events(c) - (net->mode == 4 ? c.pending : 0ull);
    net->spikes = hs_ctl_spikes(c);
    if (c.base != net->step) return hs_fail(HS_ERR_STATE, "device step %lld != host step %lld", c.base, (long long)net->step);
    for (const HMon &M : net->m) {
        if (M.kind != 0 || M.rec >= 0) continue;
        unsigned long long n = 0;
        HS_HIP(hipMemcpyAsync(&n, M.mn, sizeof(n), hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        if ((int64_t)n > M.cap)
            return hs_fail(HS_ERR_OVERFLOW, "spike monitor overflow: %llu records for capacity %lld", n,
                           (long long)M.cap);
    }
    return HS_OK;
}

extern "C" int hs_net_stats(hs_net *net, hs_stats *out)
{
    if (!net || !out) return hs_fail(HS_ERR_ARG, "null argument");
    out->steps = net->step;
    out->events = net->events;
    out->spikes = net->spikes;
    out->last_run_wall_s = net->last_wall;
    out->last_run_gpu_ms = net->last_gpu_ms;
    out->last_run_steps = net->last_steps;
    if (net->p.empty() || net->mode == 2) out->launches_per_step = 1.0;
    else if (net->mode == 4) out->launches_per_step = 1.0 / net->window;
    else {
        const double per = 1.0 + (net->n_scatter ? 1 : 0) + (net->n_gather ? 1 : 0);
        out->launches_per_step = net->mode == 3 ? per / net->window : per;
    }
    out->window = net->window;
    return HS_OK;
}

/* every record of a group's recording monitor, as (step << 32 | index) keys in
   the order cpp_standalone writes them (by step, then index) */
static int hs_spike_keys(hs_net *net, const HMon &M, std::vector<uint64_t> &key)
{
    unsigned long long c = 0;
    HS_HIP(hipMemcpyAsync(&c, M.mn, sizeof(c), hipMemcpyDeviceToHost, g_rt.stream));
    HS_HIP(hipStreamSynchronize(g_rt.stream));
    const int64_t n = (int64_t)c;
    if (n > M.cap) return hs_fail(HS_ERR_OVERFLOW, "spike monitor overflow: %lld records for capacity %lld", (long long)n, (long long)M.cap);
    std::vector<int32_t> hi((size_t)n), ht((size_t)n);
    if (n) {
        HS_HIP(hipMemcpyAsync(hi.data(), M.mi, sizeof(int32_t) * n, hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipMemcpyAsync(ht.data(), M.mt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    }
    /* records of one step arrive tile by tile; order them as cpp_standalone does */
    key.resize((size_t)n);
    for (int64_t r = 0; r < n; ++r) key[r] = ((uint64_t)(uint32_t)ht[r] << 32) | (uint32_t)hi[r];
    std::sort(key.begin(), key.end());
    return HS_OK;
}

extern "C" int hs_net_mon_count(hs_net *net, int mid, int64_t *n)
{
    if (!net || mid < 0 || mid >= (int)net->m.size() || !n) return hs_fail(HS_ERR_ARG, "bad monitor");
    const HMon &M = net->m[mid];
    if (M.kind == 0 && M.rec >= 0) {
        std::vector<uint64_t> key;
        HS_TRY(hs_spike_keys(net, net->m[M.rec], key));
        int64_t c = 0;
        for (uint64_t k : key) {
            const int32_t i = (int32_t)(k & 0xffffffffu);
            c += (i >= M.start && i < M.stop);
        }
        *n = c;
    } else if (M.kind == 0) {
        unsigned long long c = 0;
        HS_HIP(hipMemcpyAsync(&c, M.mn, sizeof(c), hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        *n = (int64_t)c;
    } else {
        *n = hs_mon_rows(M, net->step);
    }
    return HS_OK;
}

extern "C" int hs_net_get_spikes(hs_net *net, int mid, int32_t *i, int64_t *step, int64_t cap)
{
    if (!net || mid < 0 || mid >= (int)net->m.size() || net->m[mid].kind != 0) return hs_fail(HS_ERR_ARG, "bad monitor");
    const HMon &M = net->m[mid];
    std::vector<uint64_t> key;
    HS_TRY(hs_spike_keys(net, M.rec >= 0 ? net->m[M.rec] : M, key));
    /* a range monitor keeps the records of its indices, relative to its start
       (spikemonitor.cu:15-22 is_in_subgroup, :195 count[_idx - _source_start]) */
    int64_t n = 0;
    for (uint64_t k : key) {
        const int32_t ii = (int32_t)(k & 0xffffffffu);
        n += (ii >= M.start && ii < M.stop);
    }
    if (cap < n) return hs_fail(HS_ERR_ARG, "output arrays too small (%lld < %lld)", (long long)cap, (long long)n);
    int64_t r = 0;
    for (uint64_t k : key) {
        const int32_t ii = (int32_t)(k & 0xffffffffu);
        if (ii < M.start || ii >= M.stop) continue;
        step[r] = (int64_t)(k >> 32);
        i[r] = ii - M.start;
        ++r;
    }
    return HS_OK;
}

/* raw little-endian array file, as _write_arrays writes each array (objects.cu:550-563) */
static int hs_write_raw(const char *path, const void *data, size_t bytes)
{
    if (!path) return HS_OK;
    FILE *f = fopen(path, "wb");
    if (!f) return hs_fail(HS_ERR_ARG, "cannot open %s for writing", path);
    const size_t put = bytes ? fwrite(data, 1, bytes, f) : 0;
    const int closed = fclose(f);
    if (put != bytes || closed != 0) return hs_fail(HS_ERR_ARG, "short write to %s", path);
    return HS_OK;
}

extern "C" int hs_net_write_monitor(hs_net *net, int mid, const char *path_a, const char *path_b,
                                    const char *path_c, const char *path_d)
{
    if (!net || mid < 0 || mid >= (int)net->m.size()) return hs_fail(HS_ERR_ARG, "bad monitor");
    const HMon &M = net->m[mid];
    int64_t n = 0;
    HS_TRY(hs_net_mon_count(net, mid, &n));
    if (M.kind == 0) {
        /* SpikeMonitor: i (int32), t (double seconds), count (int32 per neuron), N */
        std::vector<int32_t> i((size_t)(n ? n : 1));
        std::vector<int64_t> st((size_t)(n ? n : 1));
        HS_TRY(hs_net_get_spikes(net, mid, i.data(), st.data(), n));
        std::vector<double> t((size_t)(n ? n : 1));
        std::vector<int32_t> count((size_t)(M.stop - M.start), 0);
        for (int64_t r = 0; r < n; ++r) {
            t[(size_t)r] = (double)st[(size_t)r] * net->dt;
            count[(size_t)i[(size_t)r]]++;
        }
        const int32_t N = (int32_t)n;
        HS_TRY(hs_write_raw(path_a, i.data(), sizeof(int32_t) * (size_t)n));
        HS_TRY(hs_write_raw(path_b, t.data(), sizeof(double) * (size_t)n));
        HS_TRY(hs_write_raw(path_c, count.data(), sizeof(int32_t) * count.size()));
        HS_TRY(hs_write_raw(path_d, &N, sizeof(N)));
        return HS_OK;
    }
    if (M.kind == 2) {
        /* PopulationRateMonitor: t (double), rate (double), -, N (rows) */
        std::vector<double> t((size_t)(n ? n : 1)), rate((size_t)(n ? n : 1));
        for (int64_t r = 0; r < n; ++r) t[(size_t)r] = (double)(M.t0 + r) * net->dt;
        HS_TRY(hs_net_get_rate_monitor(net, mid, rate.data(), n));
        const int32_t N = (int32_t)n;
        HS_TRY(hs_write_raw(path_a, t.data(), sizeof(double) * (size_t)n));
        HS_TRY(hs_write_raw(path_b, rate.data(), sizeof(double) * (size_t)n));
        HS_TRY(hs_write_raw(path_d, &N, sizeof(N)));
        return HS_OK;
    }
    /* StateMonitor: t (double), values (the variable's type, time-major rows of
       the recorded indices, objects.cu:589-595), indices (int32), N (rows); a
       synapse StateMonitor records its synapses' positions as indices */
    const int64_t w = M.kind == 3 ? (int64_t)M.hsel.size() : M.width;
    std::vector<double> t((size_t)(n ? n : 1));
    for (int64_t r = 0; r < n; ++r) t[(size_t)r] = (double)(M.t0 + r * M.every) * net->dt;
    std::vector<unsigned char> vals(hs_rsize(net) * (size_t)(n * w ? n * w : 1));
    if (n * w) {
        HS_HIP(hipMemcpyAsync(vals.data(), M.buf, hs_rsize(net) * (size_t)(n * w), hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    }
    std::vector<int32_t> idx((size_t)(w ? w : 1));
    for (int64_t k = 0; k < w; ++k)
        idx[(size_t)k] = M.kind == 3 ? M.hsel[(size_t)k] : (!M.hidx.empty() ? M.hidx[(size_t)k] : (int32_t)(M.start + k));
    const int32_t N = (int32_t)n;
    HS_TRY(hs_write_raw(path_a, t.data(), sizeof(double) * (size_t)n));
    HS_TRY(hs_write_raw(path_b, vals.data(), hs_rsize(net) * (size_t)(n * w)));
    HS_TRY(hs_write_raw(path_c, idx.data(), sizeof(int32_t) * (size_t)w));
    HS_TRY(hs_write_raw(path_d, &N, sizeof(N)));
    return HS_OK;
}

extern "C" int hs_net_write_last_run_info(hs_net *net, const char *path)
{
    if (!net || !path) return hs_fail(HS_ERR_ARG, "bad arguments");
    FILE *f = fopen(path, "w");
    if (!f) return hs_fail(HS_ERR_ARG, "cannot open %s for writing", path);
    /* Network::_last_run_time and _last_run_completed_fraction as std::ostream prints them */
    fprintf(f, "%g %g\n", net->last_wall, net->last_steps > 0 || net->step > 0 ? 1.0 : 0.0);
    if (fclose(f) != 0) return hs_fail(HS_ERR_ARG, "short write to %s", path);
    return HS_OK;
}

extern "C" int hs_net_get_state_monitor(hs_net *net, int mid, double *out, int64_t cap)
{
    if (!net || mid < 0 || mid >= (int)net->m.size() || (net->m[mid].kind != 1 && net->m[mid].kind != 3))
        return hs_fail(HS_ERR_ARG, "bad monitor");
    const HMon &M = net->m[mid];
    const int64_t rows = hs_mon_rows(M, net->step);
    const int64_t w = M.kind == 3 ? (int64_t)M.hsel.size() : M.width;
    if (cap < rows * w) return hs_fail(HS_ERR_ARG, "output too small");
    if (!rows) return HS_OK;
    if (net->prec == HS_F64) {
        HS_HIP(hipMemcpyAsync(out, M.buf, sizeof(double) * rows * w, hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
    } else {
        std::vector<float> f((size_t)(rows * w));
        HS_HIP(hipMemcpyAsync(f.data(), M.buf, sizeof(float) * rows * w, hipMemcpyDeviceToHost, g_rt.stream));
        HS_HIP(hipStreamSynchronize(g_rt.stream));
        for (int64_t r = 0; r < rows * w; ++r) out[r] = f[r];
    }
    return HS_OK;
}
