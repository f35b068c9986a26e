"""The oracle: ctypes wrapper of libcpuref.so (oracle/cpu_ref.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, always as the checker; the product (brian2cuda_amd)
never imports it.  Build with `make -C oracle` (done by __graft_entry__.build()).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpuref.so")

SEQ, ACC = 0, 1

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    U32P = C.POINTER(C.c_uint32)
    I32P = C.POINTER(C.c_int32)
    I64P = C.POINTER(C.c_int64)
    DP = C.POINTER(C.c_double)
    V = C.c_void_p
    sig = {
        "ref_philox4x32_10": (None, [U32P, U32P, U32P]),
        "ref_randn": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64]),
        "ref_randn_f32": (C.c_float, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64]),
        "ref_rand": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64]),
        "ref_quantize_delay": (C.c_int, [C.c_double, C.c_double]),
        "ref_connect_random": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint64,
                                           C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                           I32P, I32P, I32P, DP, C.c_int64]),
        "ref_connect_fixed": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                          C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                          I32P, I32P, I32P, DP, C.c_int64]),
        "ref_connect_code": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                         C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         I32P, I32P, I32P, DP, C.c_int64]),
        "ref_timestep": (C.c_longlong, [C.c_double, C.c_double]),
        "ref_exp": (C.c_double, [C.c_double]),
        "ref_expf": (C.c_float, [C.c_float]),
        "ref_log": (C.c_double, [C.c_double]),
        "ref_sqrt": (C.c_double, [C.c_double]),
        "ref_cos2pi": (C.c_double, [C.c_double]),
        "ref_logf": (C.c_float, [C.c_float]),
        "ref_sqrtf": (C.c_float, [C.c_float]),
        "ref_cos2pif": (C.c_float, [C.c_float]),
        "ref_model_constants": (C.c_int, [C.c_int, DP, C.c_int, C.c_double, DP]),
        "ref_net_set_spikegen": (C.c_int, [V, C.c_int, C.c_int64, I32P, I64P]),
        "ref_net_pathway_stdp": (C.c_int, [V, C.c_int, DP, C.c_int]),
        "ref_net_get_synapse_var": (C.c_int, [V, C.c_int, C.c_int, DP]),
        "ref_stateupdate": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(DP), I32P, DP, C.c_double,
                                      C.c_uint64, C.c_int64]),
        "ref_net_create": (V, [C.c_int, C.c_int, C.c_double, C.c_uint64]),
        "ref_net_add_group": (C.c_int, [V, C.c_int, C.c_int, DP, C.c_int]),
        "ref_net_add_group_code": (C.c_int, [V, C.c_int, C.c_int, V, V, C.c_double, C.c_int, C.c_uint32]),
        "ref_net_pathway_code": (C.c_int, [V, C.c_int, V, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int,
                                           C.POINTER(C.c_int)]),
        "ref_net_set_synapse_var": (C.c_int, [V, C.c_int, C.c_int, DP]),
        "ref_net_pathway_code_post": (C.c_int, [V, C.c_int, V, C.c_int, C.POINTER(C.c_int)]),
        "ref_net_set_state": (C.c_int, [V, C.c_int, C.c_int, DP]),
        "ref_net_get_state": (C.c_int, [V, C.c_int, C.c_int, DP]),
        "ref_net_get_lastspike": (C.c_int, [V, C.c_int, DP]),
        "ref_net_add_pathway": (C.c_int, [V, C.c_int, C.c_int, C.c_int, C.c_int64, I32P, I32P, I32P, C.c_int,
                                          C.c_int, DP, C.c_double]),
        "ref_net_add_spike_monitor": (C.c_int, [V, C.c_int]),
        "ref_net_add_rate_monitor": (C.c_int, [V, C.c_int, C.c_int, C.c_int]),
        "ref_mon_get_rate": (C.c_int, [V, C.c_int, DP]),
        "ref_net_add_state_monitor": (C.c_int, [V, C.c_int, C.c_int, C.c_int, C.c_int]),
        "ref_net_add_state_monitor_indices": (C.c_int, [V, C.c_int, C.c_int, C.POINTER(C.c_int32), C.c_int, C.c_int]),
        "ref_net_set_monitor_interval": (C.c_int, [V, C.c_int, C.c_int]),
        "ref_net_add_synapse_monitor": (C.c_int, [V, C.c_int, C.c_int, C.POINTER(C.c_int32), C.c_int]),
        "ref_net_run": (C.c_int, [V, C.c_int64]),
        "ref_net_set_threads": (C.c_int, [V, C.c_int]),
        "ref_net_step": (C.c_int64, [V]),
        "ref_net_events": (C.c_uint64, [V]),
        "ref_mon_count": (C.c_int64, [V, C.c_int]),
        "ref_mon_get_spikes": (C.c_int, [V, C.c_int, I32P, I64P]),
        "ref_mon_get_state": (C.c_int, [V, C.c_int, DP]),
        "ref_net_destroy": (None, [V]),
        "ref_mc_create": (V, [C.c_int, C.c_double, C.c_int, I32P, I32P, I32P, I32P, I32P, DP, DP, DP, DP,
                              C.c_double, DP, C.c_double]),
        "ref_mc_set": (C.c_int, [V, C.c_int, DP]),
        "ref_mc_get": (C.c_int, [V, C.c_int, DP]),
        "ref_mc_run": (C.c_int, [V, C.c_int64]),
        "ref_mc_spike_counts": (C.c_int, [V, I32P]),
        "ref_mc_destroy": (None, [V]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().ref_philox4x32_10(c, k, o)
    return tuple(o)


def connect_random(pre_range, post_range, p, skip_self, seed, stream, delay, dt, weight):
    """Synapses of Synapses.connect(p) with drawn delays and weights, creation
    order (pre ascending, post ascending): dict of pre, post, delay_steps (int32)
    and w (float64).  delay, weight: (lo, hi) ranges."""
    L = lib()
    args = (int(pre_range[0]), int(pre_range[1]), int(post_range[0]), int(post_range[1]), float(p),
            int(bool(skip_self)), int(seed), int(stream), float(delay[0]), float(delay[1]), float(dt),
            float(weight[0]), float(weight[1]))
    n = L.ref_connect_random(*args, None, None, None, None, 0)
    out = {"pre": np.empty(n, np.int32), "post": np.empty(n, np.int32), "delay_steps": np.empty(n, np.int32),
           "w": np.empty(n, np.float64)}
    I32P, DP = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    m = L.ref_connect_random(*args, out["pre"].ctypes.data_as(I32P), out["post"].ctypes.data_as(I32P),
                             out["delay_steps"].ctypes.data_as(I32P), out["w"].ctypes.data_as(DP), n)
    assert m == n
    return out


def connect_fixed(pre_range, post_range, K, skip_self, seed, stream, delay, dt, weight):
    """Synapses of a fixed out-degree connect (K distinct targets sampled per
    source, connect(j='k for k in sample(N_post, size=K)'); skip_self drops a
    sampled j == i), creation order (pre ascending, targets ascending): dict of
    pre, post, delay_steps, w."""
    L = lib()
    args = (int(pre_range[0]), int(pre_range[1]), int(post_range[0]), int(post_range[1]), int(K),
            int(bool(skip_self)), int(seed), int(stream), float(delay[0]), float(delay[1]), float(dt),
            float(weight[0]), float(weight[1]))
    n = L.ref_connect_fixed(*args, None, None, None, None, 0)
    if n < 0:
        raise ValueError("sample size K bigger than the population size")
    out = {"pre": np.empty(n, np.int32), "post": np.empty(n, np.int32), "delay_steps": np.empty(n, np.int32),
           "w": np.empty(n, np.float64)}
    I32P, DP = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    m = L.ref_connect_fixed(*args, out["pre"].ctypes.data_as(I32P), out["post"].ctypes.data_as(I32P),
                            out["delay_steps"].ctypes.data_as(I32P), out["w"].ctypes.data_as(DP), n)
    assert m == n
    return out


def rand(seed, stream, index, step):
    """One uniform draw in [0, 1) (the same in both precisions)."""
    return lib().ref_rand(seed, stream, index, step)


def randn(seed, stream, index, step, precision=64):
    """One draw; precision 32 gives the FP32 form's draw (24-bit uniforms, float math)."""
    if precision == 32:
        return lib().ref_randn_f32(seed, stream, index, step)
    return lib().ref_randn(seed, stream, index, step)


def model_constants(model, params, dt):
    p = np.ascontiguousarray(params, dtype=np.float64)
    k = np.zeros(16)
    n = lib().ref_model_constants(model, p.ctypes.data_as(C.POINTER(C.c_double)), p.size, dt,
                                  k.ctypes.data_as(C.POINTER(C.c_double)))
    if n < 0:
        raise ValueError("bad model or parameter count")
    return k[:n]


def stateupdate(model, precision, arrays, last, params, dt, seed, step):
    """One oracle state-update step on copies of `arrays`; returns the new arrays."""
    arrays = [np.array(a, dtype=np.float64, copy=True) for a in arrays]
    DP = C.POINTER(C.c_double)
    ptrs = (DP * len(arrays))(*[a.ctypes.data_as(DP) for a in arrays])
    last = np.ascontiguousarray(last, dtype=np.int32)
    p = np.zeros(16)  # optional trailing parameters default to 0
    p[:len(params)] = params
    r = lib().ref_stateupdate(model, precision, arrays[0].size, ptrs, last.ctypes.data_as(C.POINTER(C.c_int32)),
                              p.ctypes.data_as(DP), dt, seed, step)
    if r < 0:
        raise ValueError("bad model")
    return arrays


# ---- code models ------------------------------------------------------------
# A code model's statements (hs_net_add_group_code) compiled for the host: the
# same wrapping (variables as locals of the network's precision, dt, t, i, xi,
# not_refractory in scope; the threshold evaluated after the update) and the same
# math (the deterministic exp and log of hs_detmath.h, IEEE sqrt), with g++ under
# the oracle's IEEE flags, so the host and device results agree bit for bit.
_CG_MATH = r"""
#include <math.h>
#include "hs_detmath.h"
#include "hs_cgfuncs.h"
static inline double hs_cg_exp(double x) { return hs_exp(x); }
static inline float hs_cg_exp(float x) { return hs_expf(x); }
static inline double hs_cg_exp(int x) { return hs_exp((double)x); }
static inline double hs_cg_log(double x) { return hs_log(x); }
static inline float hs_cg_log(float x) { return hs_logf(x); }
static inline double hs_cg_log(int x) { return hs_log((double)x); }
static inline double hs_cg_log10(double x) { return hs_log(x) / 2.302585092994046; }
static inline float hs_cg_log10(float x) { return hs_logf(x) / 2.3025851f; }
static inline double hs_cg_sqrt(double x) { return hs_sqrt(x); }
static inline float hs_cg_sqrt(float x) { return hs_sqrtf(x); }
static inline double hs_cg_sqrt(int x) { return hs_sqrt((double)x); }
static inline double hs_cg_abs(double x) { return x < 0.0 ? -x : x; }
static inline float hs_cg_abs(float x) { return x < 0.0f ? -x : x; }
static inline int hs_cg_abs(int x) { return x < 0 ? -x : x; }
static inline long long hs_cg_abs(long long x) { return x < 0 ? -x : x; }
static inline double hs_cg_exp(long long x) { return hs_exp((double)x); }
static inline double hs_cg_log(long long x) { return hs_log((double)x); }
static inline double hs_cg_sqrt(long long x) { return hs_sqrt((double)x); }
static inline double hs_cg_floor(long long x) { return (double)x; }
static inline double hs_cg_ceil(long long x) { return (double)x; }
static inline double hs_cg_floor(double x) { return floor(x); }
static inline float hs_cg_floor(float x) { return floorf(x); }
static inline double hs_cg_ceil(double x) { return ceil(x); }
static inline float hs_cg_ceil(float x) { return ceilf(x); }
template <typename T> static inline T hs_cg_sign(T x) { return (T)((x > (T)0) - (x < (T)0)); }
template <typename A, typename B, typename C>
static inline auto hs_cg_clip(A x, B lo, C hi) -> decltype(x + lo + hi)
{
    typedef decltype(x + lo + hi) T;
    const T v = (T)x;
    return v < (T)lo ? (T)lo : (v > (T)hi ? (T)hi : v);
}
#define exp hs_cg_exp
#define log hs_cg_log
#define log10 hs_cg_log10
#define sqrt hs_cg_sqrt
#define abs hs_cg_abs
#define fabs hs_cg_abs
#define floor hs_cg_floor
#define ceil hs_cg_ceil
#define sign hs_cg_sign
#define clip hs_cg_clip
#define sin hs_cg_sin
#define cos hs_cg_cos
#define tan hs_cg_tan
#define sinh hs_cg_sinh
#define cosh hs_cg_cosh
#define tanh hs_cg_tanh
#define arcsin hs_cg_arcsin
#define arccos hs_cg_arccos
#define arctan hs_cg_arctan
#define expm1 hs_cg_expm1
#define log1p hs_cg_log1p
#define exprel hs_cg_exprel
#define hs_rand_u(k) hs_rand(hs_seed, hs_cg_rnd_stream(hs_stream, k, 0), (uint32_t)i, hs_step)
#define hs_rand_n(k) hs_randn(hs_seed, hs_cg_rnd_stream(hs_stream, k, 1), (uint32_t)i, hs_step)
#define hs_rand_xi(k) hs_randn(hs_seed, hs_cg_rnd_stream(hs_stream, k, 2), (uint32_t)i, hs_step)
#define hs_srand_u(k) hs_rand(hs_seed, hs_cg_syn_stream(hs_path, k, 0), (uint32_t)i, hs_cg_syn_ctr(hs_step, j))
#define hs_srand_n(k) hs_randn(hs_seed, hs_cg_syn_stream(hs_path, k, 1), (uint32_t)i, hs_cg_syn_ctr(hs_step, j))
#define hs_crand_u(k) hs_rand(hs_seed, hs_cg_conn_stream(hs_cstream, k, 0), (uint32_t)i, hs_cg_conn_ctr(j, hs_ccopy))
#define hs_crand_n(k) hs_randn(hs_seed, hs_cg_conn_stream(hs_cstream, k, 1), (uint32_t)i, hs_cg_conn_ctr(j, hs_ccopy))
"""
_cg_libs = {}


def code_model_source(model, precision):
    """The host program of one code model: extern "C" hs_cg_step / hs_cg_reset."""
    R = "double" if precision == 64 else "float"
    load = "".join(f"    R {n} = hs_s[{k}];\n" for k, n in enumerate(model.names))
    store = "".join(f"    hs_s[{k}] = {n};\n" for k, n in enumerate(model.names))
    thr = f"({model.threshold}) ? 1 : 0" if model.threshold else "0"
    rnd = "unsigned long long hs_seed, unsigned int hs_stream, long long hs_step"
    unused = "(void)hs_seed; (void)hs_stream; (void)hs_step;"
    return (_CG_MATH + f"typedef {R} R;\n"
            f'extern "C" int hs_cg_step(R *hs_s, int not_refractory, R xi, R dt, double t, int i, {rnd})\n{{\n'
            f"    (void)not_refractory; (void)xi; (void)dt; (void)t; (void)i; {unused}\n{load}"
            f"    {{\n{model.update}\n    }}\n{store}    return {thr};\n}}\n"
            f'extern "C" void hs_cg_reset(R *hs_s, R dt, double t, int i, {rnd})\n{{\n'
            f"    (void)dt; (void)t; (void)i; {unused}\n{load}    {{\n{model.reset or ''}\n    }}\n{store}}}\n")


def synapse_code_source(code, precision):
    """The host program of one pathway's synaptic code: extern "C" hs_scg_on_pre
    and hs_scg_on_post (a no-op without on_post code)."""
    R = "double" if precision == 64 else "float"
    head = [f"    R {n} = hs_sv[{k}];" for k, n in enumerate(code.syn_names)]
    head += [f"    const R {n}_pre = hs_pre[{k}];" for k, n in enumerate(code.pre_names)]
    tail = [f"    hs_sv[{k}] = {n};" for k, n in enumerate(code.syn_names)]
    tail += [f"    (void){n}_pre;" for n in code.pre_names]
    pre = head + [f"    R {n}_post = (R)0;" for n in code.post_names] + ["    {", code.on_pre, "    }"] + tail
    pre += [f"    hs_eff[{k}] = {n}_post;" for k, n in enumerate(code.post_names)]
    post = head + [f"    const R {n}_post = hs_post[{k}];" for k, n in enumerate(code.post_read_names)]
    post += ["    {", code.on_post or "", "    }"] + tail
    post += [f"    (void){n}_post;" for n in code.post_read_names]
    args = ("R dt, double t, int i, int j, double lastupdate, unsigned long long hs_seed, unsigned int hs_path, "
            "long long hs_step")
    unused = "(void)dt; (void)t; (void)i; (void)j; (void)lastupdate; (void)hs_seed; (void)hs_path; (void)hs_step;"
    return (_CG_MATH + f"typedef {R} R;\n"
            f'extern "C" void hs_scg_on_pre(R *hs_sv, const R *hs_pre, R *hs_eff, {args})\n{{\n'
            f"    (void)hs_sv; (void)hs_pre; (void)hs_eff; {unused}\n" + "\n".join(pre) + "\n}\n"
            f'extern "C" void hs_scg_on_post(R *hs_sv, const R *hs_pre, const R *hs_post, {args})\n{{\n'
            f"    (void)hs_sv; (void)hs_pre; (void)hs_post; {unused}\n" + "\n".join(post) + "\n}\n")


def _host_lib(src):
    """ctypes handle of src compiled for the host (g++, the oracle's IEEE flags), cached by text."""
    import hashlib
    import tempfile
    key = hashlib.sha1(src.encode()).hexdigest()[:20]
    if key not in _cg_libs:
        d = os.path.join(tempfile.gettempdir(), f"hs_oracle_cg_{os.getuid()}")
        os.makedirs(d, exist_ok=True)
        so = os.path.join(d, key + ".so")
        if not os.path.exists(so):
            cpp = os.path.join(d, key + ".cpp")
            with open(cpp, "w") as f:
                f.write(src)
            tmp = so + f".{os.getpid()}.tmp"
            r = subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                                "-I", os.path.join(HERE, "..", "brian2cuda_amd", "csrc"), "-o", tmp, cpp],
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise ValueError("code does not compile for the host:\n" + r.stderr[-3000:])
            os.replace(tmp, so)
        _cg_libs[key] = C.CDLL(so)
    return _cg_libs[key]


def connect_code_source(cond, p_expr, precision, w_expr=""):
    """The host program of a connect condition / probability: extern "C"
    hs_cn_cond and hs_cn_p over the absolute indices (ai, aj), the ranges and
    the groups' state arrays, the expressions of the device program."""
    R = "double" if precision == 64 else "float"
    args = ("(const void *const *hs_pre_v, const void *const *hs_post_v, long long hs_ai, long long hs_aj, "
            "long long pre_lo, long long pre_hi, long long post_lo, long long post_hi, unsigned long long hs_seed, "
            "unsigned int hs_cstream)")
    head = ("{\n    const int hs_ccopy = 0;\n    const long long i = hs_ai - pre_lo, j = hs_aj - post_lo, N_pre = pre_hi - pre_lo, "
            "N_post = post_hi - post_lo;\n    (void)hs_pre_v; (void)hs_post_v; (void)i; (void)j; (void)N_pre; "
            "(void)N_post; (void)hs_seed; (void)hs_cstream;\n")
    return ('#include "hs_cgconn.h"\n' + _CG_MATH + f"typedef {R} R;\n"
            "#define hs_pre(k) ((double)((const R *)hs_pre_v[k])[hs_ai])\n"
            "#define hs_post(k) ((double)((const R *)hs_post_v[k])[hs_aj])\n"
            f'extern "C" int hs_cn_cond{args}\n{head}    return ({cond or "true"}) ? 1 : 0;\n}}\n'
            f'extern "C" double hs_cn_p{args}\n{head}    return (double)({p_expr or "0.0"});\n}}\n'
            f'extern "C" double hs_cn_w{args[:-1]}, int hs_ccopy)\n'
            f'{head.replace("    const int hs_ccopy = 0;", "    (void)hs_ccopy;")}'
            f'    return (double)({w_expr or "0.0"});\n}}\n')


def connect_code(pre_range, post_range, cond, p_expr, p, K, skip_self, seed, stream, delay, dt, weight,
                 pre_vars=(), post_vars=(), precision=64, n=1, w_expr=""):
    """Synapses of connect(condition=cond, p=p or p_expr) or of a fixed
    out-degree K filtered by cond, the C expressions of the device program
    (brian2cuda_amd.codegen.connect_expression), compiled for the host;
    pre_vars / post_vars: the source / target groups' state arrays (read in
    precision); n: synapses per selected pair (consecutive in creation order);
    w_expr: each synapse's weight as such an expression ("": uniform in weight).
    Creation order; dict of pre, post, delay_steps, w."""
    L = lib()
    fns = _host_lib(connect_code_source(cond, p_expr, precision, w_expr))
    dt_np = np.float64 if precision == 64 else np.float32
    keep = [np.ascontiguousarray(v, dtype=dt_np) for v in list(pre_vars) + list(post_vars)]
    npre = len(pre_vars)
    arr_pre = (C.c_void_p * max(1, npre))(*[a.ctypes.data for a in keep[:npre]])
    arr_post = (C.c_void_p * max(1, len(keep) - npre))(*[a.ctypes.data for a in keep[npre:]])
    cf = C.cast(fns.hs_cn_cond, C.c_void_p).value if cond else None
    pf = C.cast(fns.hs_cn_p, C.c_void_p).value if p_expr else None
    wf = C.cast(fns.hs_cn_w, C.c_void_p).value if w_expr else None
    args = (int(pre_range[0]), int(pre_range[1]), int(post_range[0]), int(post_range[1]), float(p), int(K), int(n),
            int(bool(skip_self)), int(seed), int(stream), float(delay[0]), float(delay[1]), float(dt),
            float(weight[0]), float(weight[1]), cf, pf, wf, C.cast(arr_pre, C.c_void_p), C.cast(arr_post, C.c_void_p))
    if K > 0 and K > int(post_range[1]) - int(post_range[0]) and int(pre_range[1]) > int(pre_range[0]):
        raise ValueError("sample size K bigger than the population size")
    n = L.ref_connect_code(*args, None, None, None, None, 0)
    out = {"pre": np.empty(n, np.int32), "post": np.empty(n, np.int32), "delay_steps": np.empty(n, np.int32),
           "w": np.empty(n, np.float64)}
    I32P, DP = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    m = L.ref_connect_code(*args, out["pre"].ctypes.data_as(I32P), out["post"].ctypes.data_as(I32P),
                           out["delay_steps"].ctypes.data_as(I32P), out["w"].ctypes.data_as(DP), n)
    assert m == n
    del keep
    return out


def synapse_code_functions(code, precision):
    """(on_pre, on_post) addresses of a pathway's synaptic code compiled for the
    host (on_post None without on_post code)."""
    L = _host_lib(synapse_code_source(code, precision))
    post = C.cast(L.hs_scg_on_post, C.c_void_p).value if code.on_post else None
    return C.cast(L.hs_scg_on_pre, C.c_void_p).value, post


def code_model_functions(model, precision):
    """(step, reset) function addresses of a code model compiled for the host
    (reset None when the model has no reset); compiled once per source."""
    L = _host_lib(code_model_source(model, precision))
    step = C.cast(L.hs_cg_step, C.c_void_p).value
    reset = C.cast(L.hs_cg_reset, C.c_void_p).value if model.reset else None
    return step, reset


class RefNet:
    """Serial cpp_standalone-semantics network (SEQ) or the fused form's semantics (ACC)."""

    def __init__(self, precision=64, semantics=SEQ, dt=1e-4, seed=1234):
        self.L = lib()
        self.h = self.L.ref_net_create(precision, semantics, dt, seed)
        if not self.h:
            raise ValueError("bad precision")
        self.precision = precision
        self._wmode = {}
        self.sizes = []

    def close(self):
        if self.h:
            self.L.ref_net_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_group(self, model, n, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        g = self.L.ref_net_add_group(self.h, model, n, p.ctypes.data_as(C.POINTER(C.c_double)), p.size)
        if g < 0:
            raise ValueError("bad group")
        self.sizes.append(n)
        return g

    def add_group_code(self, n, model):
        """A group whose model is code (brian2cuda_amd.codegen.CodeModel), its
        statements compiled for the host (code_model_functions)."""
        step, reset = code_model_functions(model, self.precision)
        g = self.L.ref_net_add_group_code(self.h, n, len(model.names), step, reset, float(model.refractory),
                                          int(bool(model.noise)), int(model.stream))
        if g < 0:
            raise ValueError("bad code group")
        self.sizes.append(n)
        return g

    def set_state(self, gid, var, values):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), (self.sizes[gid],)))
        if self.L.ref_net_set_state(self.h, gid, var, v.ctypes.data_as(C.POINTER(C.c_double))) < 0:
            raise ValueError("bad state")

    def get_state(self, gid, var):
        out = np.empty(self.sizes[gid])
        if self.L.ref_net_get_state(self.h, gid, var, out.ctypes.data_as(C.POINTER(C.c_double))) < 0:
            raise ValueError("bad state")
        return out

    def get_lastspike(self, gid):
        out = np.empty(self.sizes[gid])
        self.L.ref_net_get_lastspike(self.h, gid, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def add_pathway(self, src, tgt, var, pre, post, delay_steps=None, scalar_delay=0, wmode=0, weights=None,
                    wscalar=0.0):
        I32P = C.POINTER(C.c_int32)
        pre = np.ascontiguousarray(pre, dtype=np.int32)
        post = np.ascontiguousarray(post, dtype=np.int32)
        d = None if delay_steps is None else np.ascontiguousarray(delay_steps, dtype=np.int32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        r = self.L.ref_net_add_pathway(self.h, src, tgt, var, pre.size, pre.ctypes.data_as(I32P),
                                       post.ctypes.data_as(I32P), None if d is None else d.ctypes.data_as(I32P),
                                       int(scalar_delay), int(wmode),
                                       None if w is None else w.ctypes.data_as(C.POINTER(C.c_double)),
                                       float(wscalar))
        if r < 0:
            raise ValueError(f"bad pathway ({r})")
        self._wmode[r] = int(wmode)
        return r

    def set_spikegen(self, gid, indices, steps):
        i = np.ascontiguousarray(indices, dtype=np.int32)
        s = np.ascontiguousarray(steps, dtype=np.int64)
        r = self.L.ref_net_set_spikegen(self.h, gid, i.size, i.ctypes.data_as(C.POINTER(C.c_int32)),
                                        s.ctypes.data_as(C.POINTER(C.c_int64)))
        if r < 0:
            raise ValueError(f"bad spike generator pattern ({r})")

    def pathway_code(self, pid, code, post_vars, pre_vars, post_read_vars=()):
        """Synaptic code (brian2cuda_amd.codegen.SynapseCode) on pathway pid: on_pre
        adds to target variables post_vars, both read source variables pre_vars,
        on_post reads target variables post_read_vars (indices, in the order of
        code.post_names / pre_names / post_read_names)."""
        fn, fn_post = synapse_code_functions(code, self.precision)
        w_sv = code.syn_names.index("w") if "w" in code.syn_names and self._wmode[pid] == 1 else -1
        pv = (C.c_int * max(1, len(post_vars)))(*post_vars)
        qv = (C.c_int * max(1, len(pre_vars)))(*pre_vars)
        r = self.L.ref_net_pathway_code(self.h, pid, fn, len(code.syn_names), w_sv, len(post_vars), pv,
                                        len(pre_vars), qv)
        if r < 0:
            raise ValueError(f"bad synaptic code ({r})")
        if fn_post is not None:
            rv = (C.c_int * max(1, len(post_read_vars)))(*post_read_vars)
            r = self.L.ref_net_pathway_code_post(self.h, pid, fn_post, len(post_read_vars), rv)
            if r < 0:
                raise ValueError(f"bad on_post code ({r})")

    def set_synapse_var(self, pid, var, values, nsyn):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), (nsyn,)))
        if self.L.ref_net_set_synapse_var(self.h, pid, var, v.ctypes.data_as(C.POINTER(C.c_double))) < 0:
            raise ValueError("bad synapse variable")

    def pathway_stdp(self, pid, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        r = self.L.ref_net_pathway_stdp(self.h, pid, p.ctypes.data_as(C.POINTER(C.c_double)), p.size)
        if r < 0:
            raise ValueError(f"bad plasticity ({r})")

    def synapse_var(self, pid, var, nsyn):
        """a synapse variable (HS_SYN_*) in synapse (creation) order"""
        out = np.empty(nsyn)
        if self.L.ref_net_get_synapse_var(self.h, pid, var, out.ctypes.data_as(C.POINTER(C.c_double))) < 0:
            raise ValueError("bad synapse variable")
        return out

    def add_spike_monitor(self, gid):
        return self.L.ref_net_add_spike_monitor(self.h, gid)

    def add_rate_monitor(self, gid, start, stop):
        m = self.L.ref_net_add_rate_monitor(self.h, gid, start, stop)
        if m < 0:
            raise ValueError("bad rate monitor")
        return m

    def rate_monitor(self, mid):
        """population rate (Hz) of every step since the monitor was added"""
        n = self.L.ref_mon_count(self.h, mid)
        out = np.empty(max(n, 1))
        self.L.ref_mon_get_rate(self.h, mid, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[:n]

    def add_state_monitor(self, gid, var, start, stop):
        m = self.L.ref_net_add_state_monitor(self.h, gid, var, start, stop)
        if m < 0:
            raise ValueError("bad state monitor")
        return m

    def set_monitor_interval(self, mid, every):
        if self.L.ref_net_set_monitor_interval(self.h, int(mid), int(every)) < 0:
            raise ValueError("bad monitor interval")

    def add_state_monitor_indices(self, gid, var, indices, every=1):
        """StateMonitor(record=indices, dt=every*dt): rows at steps t with t % every == 0"""
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        m = self.L.ref_net_add_state_monitor_indices(self.h, gid, var, idx.ctypes.data_as(C.POINTER(C.c_int32)),
                                                     idx.size, int(every))
        if m < 0:
            raise ValueError("bad state monitor")
        return m

    def add_synapse_monitor(self, pid, var, indices):
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        m = self.L.ref_net_add_synapse_monitor(self.h, pid, var, idx.ctypes.data_as(C.POINTER(C.c_int32)), idx.size)
        if m < 0:
            raise ValueError("bad synapse monitor")
        return m

    def run(self, nsteps):
        r = self.L.ref_net_run(self.h, int(nsteps))
        if r < 0:
            raise RuntimeError(f"ref_net_run failed ({r})")

    def set_threads(self, n):
        """Run the step loop over n OpenMP threads (before the first step);
        the results stay bit-identical to the serial loop."""
        r = self.L.ref_net_set_threads(self.h, int(n))
        if r < 0:
            raise ValueError(f"ref_net_set_threads({n}) failed ({r})")

    @property
    def step(self):
        return self.L.ref_net_step(self.h)

    @property
    def events(self):
        return self.L.ref_net_events(self.h)

    def spikes(self, mid):
        n = self.L.ref_mon_count(self.h, mid)
        i = np.empty(max(n, 1), np.int32)
        t = np.empty(max(n, 1), np.int64)
        self.L.ref_mon_get_spikes(self.h, mid, i.ctypes.data_as(C.POINTER(C.c_int32)),
                                  t.ctypes.data_as(C.POINTER(C.c_int64)))
        return i[:n], t[:n]

    def state_monitor(self, mid, width):
        rows = self.L.ref_mon_count(self.h, mid)
        out = np.empty(max(rows * width, 1))
        self.L.ref_mon_get_state(self.h, mid, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[: rows * width].reshape(rows, width)


def initial_state(g):
    """A group spec's state arrays before the first step: zeros, then its init values."""
    out = [np.zeros(g.n) for _ in _var_names(g)]
    for var, val in g.init.items():
        out[var][:] = val
    return out


def random_pathway(spec, path, rows=None, precision=64):
    """The synapses of a RandomConnect pathway (creation order); rows: a
    sub-range of source rows (rows are generated independently of each other).
    A condition or probability given as code reads the groups' initial state
    in precision."""
    r = path.random
    pre = r.pre if rows is None else rows
    skip_self = (not r.autapses) and path.src == path.tgt
    if getattr(r, "cond", "") or getattr(r, "p_expr", "") or getattr(r, "n", 1) > 1 or getattr(r, "w_expr", ""):
        return connect_code(pre, r.post, r.cond, r.p_expr, r.p, r.k, skip_self, r.seed, r.stream, r.delay, spec.dt,
                            r.w, initial_state(spec.groups[path.src]), initial_state(spec.groups[path.tgt]),
                            precision, getattr(r, "n", 1), getattr(r, "w_expr", ""))
    if getattr(r, "k", 0):
        return connect_fixed(pre, r.post, r.k, skip_self, r.seed, r.stream, r.delay, spec.dt, r.w)
    return connect_random(pre, r.post, r.p, skip_self, r.seed, r.stream, r.delay, spec.dt, r.w)


HS_SYN_CODE = 16  # synapse variable k of a pathway with on_pre code
_BUILTIN_VARS = {1: ["v", "ge", "gi"], 2: ["V"], 3: ["v", "m", "n", "h", "ge", "gi"], 4: ["_placeholder"],
                 5: ["rates"]}


def _var_names(g):
    """State-variable names of a group spec (code models: their own)."""
    return list(g.code.names) if getattr(g, "code", None) is not None else _BUILTIN_VARS[g.model]


def build_ref(spec, precision=64, semantics=SEQ, threads=1):
    """Instantiate a brian2cuda_amd.networks.NetSpec on the oracle (its step
    loop over `threads` OpenMP threads)."""
    net = RefNet(precision, semantics, spec.dt, spec.seed)
    if threads > 1:
        net.set_threads(threads)
    for g in spec.groups:
        gid = net.add_group_code(g.n, g.code) if getattr(g, "code", None) is not None else \
            net.add_group(g.model, g.n, g.params)
        for var, val in g.init.items():
            net.set_state(gid, var, val)
        if getattr(g, "spikes", None) is not None:
            net.set_spikegen(gid, *g.spikes)
    for p in spec.paths:
        if p.random is not None:
            syn = random_pathway(spec, p, precision=precision)
            pid = net.add_pathway(p.src, p.tgt, p.var, syn["pre"], syn["post"], syn["delay_steps"], 0,
                                  p.random.wmode, syn["w"], p.random.w[0])
        else:
            pid = net.add_pathway(p.src, p.tgt, p.var, p.pre, p.post, p.delay_steps, p.scalar_delay, p.wmode,
                                  p.weights, p.wscalar)
        if getattr(p, "stdp", None) is not None:
            net.pathway_stdp(pid, p.stdp)
        if getattr(p, "code", None) is not None:
            tn = _var_names(spec.groups[p.tgt])
            sn = _var_names(spec.groups[p.src])
            net.pathway_code(pid, p.code, [tn.index(n) for n in p.code.post_names],
                             [sn.index(n) for n in p.code.pre_names],
                             [tn.index(n) for n in getattr(p.code, "post_read_names", [])])
            nsyn = len(syn["pre"]) if p.random is not None else len(p.pre)
            for name, val in (p.code.init or {}).items():
                if name == "w" and (p.random.wmode if p.random is not None else p.wmode) == 1:
                    continue  # the pathway's weights
                net.set_synapse_var(pid, HS_SYN_CODE + p.code.syn_names.index(name), val, nsyn)
    smons = [net.add_spike_monitor(g) for g in spec.spike_monitors]
    stmons = [net.add_state_monitor_indices(m.gid, m.var, m.indices, m.every) if hasattr(m, "indices")
              else net.add_state_monitor(*m) for m in spec.state_monitors]
    net.rate_mids = [net.add_rate_monitor(*m) for m in getattr(spec, "rate_monitors", [])]
    return net, smons, stmons


class RefMC:
    """Batch of multicompartment trees, advanced serially as the reference's
    spatial state updater does (cpu_ref_mc.c).  precision: channel arithmetic."""

    def __init__(self, tb, precision=64):
        self.L = lib()
        I32P = C.POINTER(C.c_int32)
        DP = C.POINTER(C.c_double)
        self._keep = [np.ascontiguousarray(a, dtype=np.int32) for a in
                      (tb.tree_comp, tb.tree_sec, tb.sec_first, tb.sec_n, tb.sec_parent)]
        self._keepf = [np.ascontiguousarray(a, dtype=np.float64) for a in (tb.area, tb.rl1, tb.rl2, tb.Cm, tb.chan)]
        i32 = [a.ctypes.data_as(I32P) for a in self._keep]
        f64 = [a.ctypes.data_as(DP) for a in self._keepf]
        self.h = self.L.ref_mc_create(precision, tb.dt, tb.n_trees, *i32, f64[0], f64[1], f64[2], f64[3], tb.Ri,
                                      f64[4], tb.vth)
        if not self.h:
            raise ValueError("bad tree batch")
        self.n_comp = tb.n_comp
        self.n_trees = tb.n_trees
        for var, val in tb.init.items():
            self.set_state(var, val)

    def set_state(self, var, values):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float64), (self.n_comp,)))
        if self.L.ref_mc_set(self.h, var, v.ctypes.data_as(C.POINTER(C.c_double))) < 0:
            raise ValueError("bad variable")

    def get_state(self, var):
        out = np.empty(self.n_comp)
        if self.L.ref_mc_get(self.h, var, out.ctypes.data_as(C.POINTER(C.c_double))) < 0:
            raise ValueError("bad variable")
        return out

    def run(self, nsteps):
        self.L.ref_mc_run(self.h, int(nsteps))

    def spike_counts(self):
        out = np.empty(self.n_trees, np.int32)
        self.L.ref_mc_spike_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def close(self):
        if self.h:
            self.L.ref_mc_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
