"""ctypes binding of libhs (include/hs.h).

The library is built in-tree by ``python -m brian2cuda_amd.build``.  There is no
fallback: if the shared object is missing or a call fails, an exception says why.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhs.so")

HS_OK = 0
ERRORS = {
    -1: "HS_ERR_ARG",
    -2: "HS_ERR_HIP",
    -3: "HS_ERR_NOMEM",
    -4: "HS_ERR_STATE",
    -5: "HS_ERR_OVERFLOW",
    -6: "HS_ERR_UNSUPPORTED",
}


class HsError(RuntimeError):
    """A libhs entry point returned a negative code."""

    def __init__(self, code, func, message):
        self.code = code
        super().__init__(f"{func} -> {ERRORS.get(code, code)}: {message}")


class HsStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("events", C.c_uint64),
        ("spikes", C.c_uint64),
        ("last_run_wall_s", C.c_double),
        ("last_run_gpu_ms", C.c_double),
        ("last_run_steps", C.c_int64),
        ("launches_per_step", C.c_double),
        ("window", C.c_int64),
    ]


P = C.c_void_p
I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
DP = C.POINTER(C.c_double)

# name: (restype, argtypes); restype int means "status code"
SIGNATURES = {
    "hs_last_error": (C.c_char_p, []),
    "hs_version": (C.c_int, []),
    "hs_init": (C.c_int, [C.c_int]),
    "hs_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
    "hs_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "hs_free": (C.c_int, [P]),
    "hs_h2d": (C.c_int, [P, P, C.c_size_t]),
    "hs_d2h": (C.c_int, [P, P, C.c_size_t]),
    "hs_zero": (C.c_int, [P, C.c_size_t]),
    "hs_sync": (C.c_int, []),
    "hs_threshold_compact": (C.c_int, [C.c_int, P, C.c_int, C.c_double, P, C.c_int64, C.c_int, C.c_int, P, P]),
    "hs_threshold_compact_noalloc": (C.c_int, [C.c_int, P, C.c_int, C.c_double, P, C.c_int64, C.c_int, C.c_int, P,
                                                P]),
    "hs_randn": (C.c_int, [C.c_uint64, C.c_uint32, C.c_int64, C.c_int, C.c_int, P]),
    "hs_rand": (C.c_int, [C.c_uint64, C.c_uint32, C.c_int64, C.c_int, P]),
    "hs_stateupdate": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p), P, DP, C.c_int, C.c_double,
                                 C.c_uint64, C.c_int64]),
    "hs_deliver": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, C.c_int64, P, P, C.c_int, P,
                             C.c_double, P, P]),
    "hs_math_eval": (C.c_int, [C.c_int, C.c_int, P, P]),
    "hs_scatter_add": (C.c_int, [C.c_int, C.c_int64, P, C.c_int, C.c_double, P, C.POINTER(C.c_double)]),
    "hs_net_create": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_uint64, C.POINTER(C.c_void_p)]),
    "hs_net_destroy": (None, [P]),
    "hs_net_add_group": (C.c_int, [P, C.c_int, C.c_int, DP, C.c_int, C.POINTER(C.c_int)]),
    "hs_net_pathway_code": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_char_p), C.c_char_p, C.c_int,
                                      C.POINTER(C.c_int), C.POINTER(C.c_char_p), C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_char_p)]),
    "hs_net_set_synapse_var": (C.c_int, [P, C.c_int, C.c_int, DP, C.c_int64]),
    "hs_net_pathway_code_post": (C.c_int, [P, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int),
                                           C.POINTER(C.c_char_p)]),
    "hs_net_add_group_code": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_char_p), C.c_char_p, C.c_char_p,
                                        C.c_char_p, C.c_double, C.c_int, C.c_uint32, C.POINTER(C.c_int)]),
    "hs_net_set_state": (C.c_int, [P, C.c_int, C.c_int, DP]),
    "hs_net_get_state": (C.c_int, [P, C.c_int, C.c_int, DP]),
    "hs_net_get_last_spike": (C.c_int, [P, C.c_int, I32P]),
    "hs_net_add_pathway": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int64, I32P, I32P, I32P, C.c_int, C.c_int,
                                     DP, C.c_double, C.POINTER(C.c_int)]),
    "hs_net_add_spike_monitor": (C.c_int, [P, C.c_int, C.c_int64, C.POINTER(C.c_int)]),
    "hs_net_add_synapse_monitor": (C.c_int, [P, C.c_int, C.c_int, P, C.c_int, C.POINTER(C.c_int)]),
    "hs_net_add_spike_monitor_range": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int64, C.POINTER(C.c_int)]),
    "hs_net_add_state_monitor": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "hs_net_add_state_monitor_indices": (C.c_int, [P, C.c_int, C.c_int, I32P, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "hs_net_set_monitor_interval": (C.c_int, [P, C.c_int, C.c_int]),
    "hs_net_set_option": (C.c_int, [P, C.c_int, C.c_longlong]),
    "hs_net_get_option": (C.c_int, [P, C.c_int, C.POINTER(C.c_longlong)]),
    "hs_net_run": (C.c_int, [P, C.c_int64]),
    "hs_net_prepare": (C.c_int, [P, C.c_int64]),
    "hs_net_stats": (C.c_int, [P, C.POINTER(HsStats)]),
    "hs_net_kernel_time": (C.c_int, [P, C.c_int, DP, I64P]),
    "hs_net_connect_random": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_double, C.c_int, C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_int,
                                        C.c_double, C.c_double, C.POINTER(C.c_int)]),
    "hs_net_connect_fixed": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_int,
                                        C.c_double, C.c_double, C.POINTER(C.c_int)]),
    "hs_net_connect_code": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_char_p, C.c_char_p, C.c_double, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                       C.c_uint32,
                                       C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.c_char_p,
                                       C.POINTER(C.c_int)]),
    "hs_net_add_pathway_device": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int64, P, P, P, C.c_int, C.c_int, P,
                                            C.c_double, C.POINTER(C.c_int)]),
    "hs_net_pathway_info": (C.c_int, [P, C.c_int, I64P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "hs_net_get_synapses": (C.c_int, [P, C.c_int, I32P, I32P, I32P, DP, C.c_int64]),
    "hs_net_set_spikegen": (C.c_int, [P, C.c_int, C.c_int64, I32P, I64P]),
    "hs_net_pathway_stdp": (C.c_int, [P, C.c_int, DP, C.c_int]),
    "hs_net_get_synapse_var": (C.c_int, [P, C.c_int, C.c_int, DP, C.c_int64]),
    "hs_net_add_rate_monitor": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "hs_net_get_rate_monitor": (C.c_int, [P, C.c_int, DP, C.c_int64]),
    "hs_net_write_monitor": (C.c_int, [P, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p]),
    "hs_net_write_last_run_info": (C.c_int, [P, C.c_char_p]),
    "hs_net_mon_count": (C.c_int, [P, C.c_int, I64P]),
    "hs_net_get_spikes": (C.c_int, [P, C.c_int, I32P, I64P, C.c_int64]),
    "hs_net_get_state_monitor": (C.c_int, [P, C.c_int, DP, C.c_int64]),
    "hs_hines_create": (C.c_int, [C.c_int, C.c_double, C.c_int, I32P, I32P, I32P, I32P, I32P, DP, DP, DP, DP,
                                  C.c_double, DP, C.c_double, C.POINTER(C.c_void_p)]),
    "hs_hines_destroy": (None, [P]),
    "hs_hines_set_state": (C.c_int, [P, C.c_int, DP]),
    "hs_hines_get_state": (C.c_int, [P, C.c_int, DP]),
    "hs_hines_set_option": (C.c_int, [P, C.c_int, C.c_longlong]),
    "hs_hines_run": (C.c_int, [P, C.c_int64]),
    "hs_hines_spike_counts": (C.c_int, [P, I32P]),
    "hs_hines_stats": (C.c_int, [P, DP, DP, I64P]),
}

_lib = None


def load_library(path=None):
    """Load libhs.so and declare every entry point; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run `python -m brian2cuda_amd.build` (hipcc, gfx950). "
            "The hip_standalone device has no CPU fallback."
        )
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code, func):
    if code != HS_OK:
        msg = _lib.hs_last_error().decode(errors="replace") if _lib is not None else ""
        raise HsError(code, func, msg)
    return code


def call(name, *args):
    lib = load_library()
    return check(getattr(lib, name)(*args), name)


_initialised = None


def init(device=0):
    """hs_init once per process."""
    global _initialised
    if _initialised is None:
        call("hs_init", device)
        _initialised = device
    return _initialised


def device_info():
    init()
    name = C.create_string_buffer(256)
    cu = C.c_int()
    hbm = C.c_longlong()
    call("hs_device_info", name, 256, C.byref(cu), C.byref(hbm))
    return {"arch": name.value.decode(), "compute_units": cu.value, "hbm_bytes": hbm.value}
