"""Python host side of libhs: device buffers, the standalone kernels and networks.

This mirrors what the reference's generated C++ does around the hot path
(brian2cuda/templates/objects.cu: arrays; network.cu: the step loop;
synapses_push_spikes.cu: connectivity preparation) with numpy arrays on the
host and HBM buffers owned by libhs.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import HsStats, call, init

HS_F64, HS_F32 = 64, 32
HS_ENGINE_PLAIN, HS_ENGINE_FUSED = 0, 1
HS_W_SCALAR, HS_W_SYNAPSE, HS_W_PRE = 0, 1, 2

HS_MODEL_LIF_CUBA, HS_MODEL_LIF_NOISE, HS_MODEL_HH_COBA, HS_MODEL_SPIKEGEN, HS_MODEL_POISSON = 1, 2, 3, 4, 5
HS_MODEL_CODE, HS_MAX_CODE_MODELS = 64, 4  # code models: ids HS_MODEL_CODE + k (hs_net_add_group_code)
HS_SYN_CODE = 16  # synapse variable k of a pathway with on_pre code (hs_net_pathway_code)
MODEL_NVARS = {HS_MODEL_LIF_CUBA: 3, HS_MODEL_LIF_NOISE: 1, HS_MODEL_HH_COBA: 6, HS_MODEL_SPIKEGEN: 1,
               HS_MODEL_POISSON: 1}
MODEL_NPARAMS = {HS_MODEL_LIF_CUBA: 7, HS_MODEL_LIF_NOISE: 7, HS_MODEL_HH_COBA: 14, HS_MODEL_SPIKEGEN: 1,
                 HS_MODEL_POISSON: 1}
MODEL_VARS = {
    HS_MODEL_LIF_CUBA: ("v", "ge", "gi"),
    HS_MODEL_LIF_NOISE: ("V",),
    HS_MODEL_HH_COBA: ("v", "m", "n", "h", "ge", "gi"),
    HS_MODEL_SPIKEGEN: ("_unused",),
    HS_MODEL_POISSON: ("rates",),
}


def spikegen_steps(times, dt):
    """Time bins of SpikeGeneratorGroup spike times: Brian2's timestep(),
    (int)((t + 1e-3*dt)/dt) (brian2cuda/cuda_generator.py:997-1006), so a time
    just below a bin edge by rounding error still lands in that bin."""
    return ((np.asarray(times, dtype=np.float64) + 1e-3 * dt) / dt).astype(np.int64)


HS_STDP_NPARAMS = 7  # taupre, taupost, dApre, dApost, wmin, wmax, scale
HS_SYN_W, HS_SYN_APRE, HS_SYN_APOST, HS_SYN_LASTUPDATE = 0, 1, 2, 3
NEVER_SPIKED = -1073741824

OPT_UPDATE_BLOCK, OPT_DELIVER_GRID, OPT_GRAPH_CHUNK, OPT_FUSED_GRID, OPT_WINDOW, OPT_DELIVERY = 0, 1, 2, 3, 4, 5
OPT_GATHER_BLOCK, OPT_KERNEL_TIMING, OPT_NOISE_BLOCKS, OPT_PIPELINED, OPT_UPDATE_GRID = 6, 7, 8, 9, 10
OPT_FAST_MATH = 11
DELIVERY_AUTO, DELIVERY_SCATTER, DELIVERY_GATHER = 0, 1, 2
KT_UPDATE, KT_SCATTER, KT_GATHER, KT_FUSED, KT_WINDOW = 0, 1, 2, 3, 4
KERNEL_NAMES = {KT_UPDATE: "hs_k_update", KT_SCATTER: "hs_k_deliver", KT_GATHER: "hs_k_gather", KT_FUSED: "hs_k_fused",
                KT_WINDOW: "hs_k_window"}


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _i32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class DeviceArray:
    """An HBM buffer owned by libhs (hs_alloc) with numpy-typed copies."""

    def __init__(self, n, dtype):
        init()
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        call("hs_alloc", self.nbytes, C.byref(p))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError(f"size {a.size} != {self.n}")
        call("hs_h2d", self.ptr, _ptr(a), self.nbytes)

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        call("hs_d2h", _ptr(out), self.ptr, self.nbytes)
        return out

    def zero(self):
        call("hs_zero", self.ptr, self.nbytes)

    def free(self):
        if self.ptr is not None and self.ptr.value:
            call("hs_free", self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            if self.ptr is not None and self.ptr.value and _lib._lib is not None:
                _lib._lib.hs_free(self.ptr)
        except Exception:
            pass


# ---- standalone kernels ---------------------------------------------------------

def threshold_compact(v, vth, last, step, ref_steps, sorted_=True):
    """hs_threshold_compact on device arrays v (f64/f32) and last (int32).

    Returns the spike indices (numpy int32)."""
    prec = HS_F64 if v.dtype == np.float64 else HS_F32
    spikes = DeviceArray(max(v.n, 1), np.int32)
    count = DeviceArray(1, np.int32)
    call("hs_threshold_compact", prec, v.ptr, v.n, float(vth), last.ptr, int(step), int(ref_steps),
         1 if sorted_ else 0, spikes.ptr, count.ptr)
    n = int(count.numpy()[0])
    return spikes.numpy()[:n]


def randn(seed, stream, step, n, precision=HS_F64):
    out = DeviceArray(n, np.float64 if precision == HS_F64 else np.float32)
    call("hs_randn", int(seed), int(stream), int(step), int(n), precision, out.ptr)
    return out.numpy()


def rand(seed, stream, step, n):
    out = DeviceArray(n, np.float64)
    call("hs_rand", int(seed), int(stream), int(step), int(n), out.ptr)
    return out.numpy()


def math_eval(fn, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    dx = DeviceArray.from_numpy(x)
    dy = DeviceArray(x.size, np.float64)
    call("hs_math_eval", int(fn), x.size, dx.ptr, dy.ptr)
    return dy.numpy()


def stateupdate(model, precision, vars_dev, last_dev, params, dt, seed, step):
    """One in-place state-update step of `model` on device arrays."""
    arr = (C.c_void_p * len(vars_dev))(*[v.ptr.value for v in vars_dev])
    p = np.ascontiguousarray(params, dtype=np.float64)
    call("hs_stateupdate", model, precision, vars_dev[0].n, arr, last_dev.ptr, _f64(p), p.size, float(dt),
         int(seed), int(step))


def deliver(precision, n_src, dmin, D, Q, hist, hist_cnt, step, bstart, post, wmode, w, wscalar, target, events):
    call("hs_deliver", precision, n_src, dmin, D, Q, hist.ptr, hist_cnt.ptr, int(step), bstart.ptr, post.ptr,
         wmode, (w.ptr if w is not None else None), float(wscalar), target.ptr, events.ptr)


def scatter_add(idx, w, target):
    """hs_scatter_add: target[idx[k]] += w on device arrays (idx int32; target
    f64 or f32).  Returns the kernel's device time in ms."""
    prec = HS_F64 if target.dtype == np.float64 else HS_F32
    ms = C.c_double(0.0)
    call("hs_scatter_add", prec, int(idx.n), idx.ptr, int(target.n), float(w), target.ptr, C.byref(ms))
    return ms.value


# ---- networks -------------------------------------------------------------------

class Net:
    """A network on the device: NeuronGroups, pathways, monitors and the step loop.

    precision: 64 or 32.  engine: "plain" (Brian2's in-place order, two launches
    per step) or "fused" (one launch per step; effects applied at the start of the
    next step).
    """

    def __init__(self, precision=64, engine="plain", dt=1e-4, seed=1234):
        init()
        self.precision = int(precision)
        self.engine = {"plain": HS_ENGINE_PLAIN, "fused": HS_ENGINE_FUSED}[engine]
        self.dt = float(dt)
        self.seed = int(seed)
        h = C.c_void_p()
        call("hs_net_create", self.precision, self.engine, self.dt, self.seed, C.byref(h))
        self.h = h
        self.groups = []

    def close(self):
        if self.h is not None and self.h.value:
            _lib.load_library().hs_net_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_group(self, model, n, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        gid = C.c_int()
        call("hs_net_add_group", self.h, int(model), int(n), _f64(p), p.size, C.byref(gid))
        self.groups.append((model, n))
        return gid.value

    def add_group_code(self, n, model):
        """A group of n neurons whose model is code (a brian2cuda_amd.codegen.CodeModel:
        variable names, update statements, threshold expression, reset
        statements, refractory period, noise); compiled for the device with the
        network's other code models on its first run.  Returns the group id."""
        names = [s.encode() for s in model.names]
        arr = (C.c_char_p * len(names))(*names)
        gid = C.c_int()
        call("hs_net_add_group_code", self.h, int(n), len(names), arr, model.update.encode(),
             (model.threshold or "").encode(), (model.reset or "").encode(), float(model.refractory),
             int(bool(model.noise)), int(model.stream), C.byref(gid))
        self.groups.append((HS_MODEL_CODE + sum(1 for m, _ in self.groups if m >= HS_MODEL_CODE), n))
        return gid.value

    def pathway_code(self, pid, code, post_vars, pre_vars, post_read_vars=()):
        """Synaptic code (a brian2cuda_amd.codegen.SynapseCode) on pathway pid:
        on_pre adds to target variables post_vars, both codes read source
        variables pre_vars, on_post reads target variables post_read_vars
        (indices, in the order of code.post_names / pre_names / post_read_names)."""
        enc = lambda names: (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
        pv = (C.c_int * max(1, len(post_vars)))(*post_vars)
        qv = (C.c_int * max(1, len(pre_vars)))(*pre_vars)
        call("hs_net_pathway_code", self.h, pid, len(code.syn_names), enc(code.syn_names), code.on_pre.encode(),
             len(post_vars), pv, enc(code.post_names), len(pre_vars), qv, enc(code.pre_names))
        if code.on_post:
            rv = (C.c_int * max(1, len(post_read_vars)))(*post_read_vars)
            call("hs_net_pathway_code_post", self.h, pid, code.on_post.encode(), len(post_read_vars), rv,
                 enc(code.post_read_names))

    def set_synapse_var(self, pid, var, values):
        """Synapse variable var (HS_SYN_CODE + k) of a code pathway, values in bundle order."""
        n = self.pathway_info(pid)["nsyn"]
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), (n,)))
        call("hs_net_set_synapse_var", self.h, pid, int(var), _f64(v), n)

    def set_state(self, gid, var, values):
        n = self.groups[gid][1]
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), (n,)))
        call("hs_net_set_state", self.h, gid, var, _f64(v))

    def get_state(self, gid, var):
        out = np.empty(self.groups[gid][1], np.float64)
        call("hs_net_get_state", self.h, gid, var, _f64(out))
        return out

    def get_last_spike(self, gid):
        out = np.empty(self.groups[gid][1], np.int32)
        call("hs_net_get_last_spike", self.h, gid, _i32(out))
        return out

    def add_pathway(self, src, tgt, var, pre, post, delay_steps=None, scalar_delay=0, wmode=HS_W_SCALAR,
                    weights=None, wscalar=0.0):
        pre = np.ascontiguousarray(pre, dtype=np.int32)
        post = np.ascontiguousarray(post, dtype=np.int32)
        if pre.size != post.size:
            raise ValueError("pre/post size mismatch")
        d = None
        if delay_steps is not None:
            d = np.ascontiguousarray(delay_steps, dtype=np.int32)
            if d.size != pre.size:
                raise ValueError("delay size mismatch")
        w = None
        if wmode != HS_W_SCALAR:
            w = np.ascontiguousarray(weights, dtype=np.float64)
        pid = C.c_int()
        call("hs_net_add_pathway", self.h, src, tgt, var, pre.size, _i32(pre), _i32(post),
             (_i32(d) if d is not None else None), int(scalar_delay), int(wmode),
             (_f64(w) if w is not None else None), float(wscalar), C.byref(pid))
        return pid.value

    def connect_random(self, src, tgt, var, p, pre=None, post=None, autapses=True, seed=None, stream=0,
                       delay=0.0, wmode=HS_W_SCALAR, w=0.0):
        """Synapses.connect(p=p) generated on the device, with delays uniform in
        delay=(lo, hi) seconds (or one value) and weights w (scalar, or (lo, hi)
        uniform per synapse with wmode HS_W_SYNAPSE).  pre/post: (lo, hi) index
        ranges inside the groups (default: whole group).  Returns the pathway id."""
        n_src, n_tgt = self.groups[src][1], self.groups[tgt][1]
        pre = (0, n_src) if pre is None else pre
        post = (0, n_tgt) if post is None else post
        dlo, dhi = (delay, delay) if np.isscalar(delay) else delay
        wlo, whi = (w, w) if np.isscalar(w) else w
        pid = C.c_int()
        call("hs_net_connect_random", self.h, src, tgt, var, int(pre[0]), int(pre[1]), int(post[0]), int(post[1]),
             float(p), int(bool(autapses)), int(self.seed if seed is None else seed), int(stream), float(dlo),
             float(dhi), int(wmode), float(wlo), float(whi), C.byref(pid))
        return pid.value

    def connect_fixed(self, src, tgt, var, K, pre=None, post=None, autapses=True, seed=None, stream=0,
                      delay=0.0, wmode=HS_W_SCALAR, w=0.0):
        """A fixed out-degree connect generated on the device: every source
        neuron of pre samples K distinct targets of post, uniformly
        (connect(j='k for k in sample(N_post, size=K)')); autapses=False drops
        a sampled j == i; delays and weights as connect_random.  Returns the
        pathway id."""
        n_src, n_tgt = self.groups[src][1], self.groups[tgt][1]
        pre = (0, n_src) if pre is None else pre
        post = (0, n_tgt) if post is None else post
        dlo, dhi = (delay, delay) if np.isscalar(delay) else delay
        wlo, whi = (w, w) if np.isscalar(w) else w
        pid = C.c_int()
        call("hs_net_connect_fixed", self.h, src, tgt, var, int(pre[0]), int(pre[1]), int(post[0]), int(post[1]),
             int(K), int(bool(autapses)), int(self.seed if seed is None else seed), int(stream), float(dlo),
             float(dhi), int(wmode), float(wlo), float(whi), C.byref(pid))
        return pid.value

    def connect_code(self, src, tgt, var, cond="", p_expr="", p=1.0, K=0, pre=None, post=None, autapses=True,
                     seed=None, stream=0, delay=0.0, wmode=HS_W_SCALAR, w=0.0, n=1, w_expr=""):
        """connect(condition=..., p=..., n=...) generated on the device with the
        condition and the per-pair probability compiled from C expressions
        (brian2cuda_amd.codegen.connect_expression; "" for none): with p_expr
        each pair is kept where cond holds and its draw is below p_expr, else
        the candidates of connect(p) (K > 0: of the fixed out-degree sample)
        are kept where cond holds.  The expressions read the groups' state as
        it is when this is called.  Delays and weights as connect_random; n > 1:
        n synapses per selected pair, each with its own delay and weight;
        w_expr: each synapse's weight as such an expression (wmode HS_W_SYNAPSE).
        Returns the pathway id."""
        n_src, n_tgt = self.groups[src][1], self.groups[tgt][1]
        pre = (0, n_src) if pre is None else pre
        post = (0, n_tgt) if post is None else post
        dlo, dhi = (delay, delay) if np.isscalar(delay) else delay
        wlo, whi = (w, w) if np.isscalar(w) else w
        pid = C.c_int()
        call("hs_net_connect_code", self.h, src, tgt, var, int(pre[0]), int(pre[1]), int(post[0]), int(post[1]),
             (cond or "").encode(), (p_expr or "").encode(), float(p), int(K), int(n), int(bool(autapses)),
             int(self.seed if seed is None else seed), int(stream), float(dlo), float(dhi), int(wmode), float(wlo),
             float(whi), (w_expr or "").encode(), C.byref(pid))
        return pid.value

    def set_spikegen(self, gid, indices, steps):
        """SpikeGeneratorGroup pattern: neuron indices[k] fires at step steps[k]."""
        i = np.ascontiguousarray(indices, dtype=np.int32)
        s = np.ascontiguousarray(steps, dtype=np.int64)
        if i.size != s.size:
            raise ValueError("indices/steps size mismatch")
        call("hs_net_set_spikegen", self.h, gid, i.size, _i32(i), s.ctypes.data_as(C.POINTER(C.c_int64)))

    def pathway_stdp(self, pid, taupre, taupost, dApre, dApost, wmin=-np.inf, wmax=np.inf, scale=1.0):
        """STDP with event-driven traces on a per-synapse-weight pathway (plain form)."""
        p = np.array([taupre, taupost, dApre, dApost, wmin, wmax, scale], dtype=np.float64)
        call("hs_net_pathway_stdp", self.h, pid, _f64(p), p.size)

    def synapse_var(self, pid, var):
        """A synapse variable (HS_SYN_*) in bundle order (the order of synapses())."""
        n = self.pathway_info(pid)["nsyn"]
        out = np.empty(max(n, 1))
        call("hs_net_get_synapse_var", self.h, pid, int(var), _f64(out), n)
        return out[:n]

    def write_monitor(self, mid, paths):
        """Write a monitor's result files (raw little-endian arrays, the
        reference's format): paths = (a, b, c, d) as in hs_net_write_monitor,
        None to skip one."""
        enc = [None if p is None else str(p).encode() for p in paths]
        call("hs_net_write_monitor", self.h, mid, *enc)

    def write_last_run_info(self, path):
        call("hs_net_write_last_run_info", self.h, str(path).encode())

    def pathway_info(self, pid):
        n = C.c_int64()
        dmin, D, mode = C.c_int(), C.c_int(), C.c_int()
        call("hs_net_pathway_info", self.h, pid, C.byref(n), C.byref(dmin), C.byref(D), C.byref(mode))
        return {"nsyn": n.value, "dmin": dmin.value, "D": D.value, "delivery": mode.value}

    def synapses(self, pid):
        """(pre, post, delay_steps, w) of a scatter-mode pathway in bundle order:
        by (pre, delay), creation order within a bundle."""
        n = self.pathway_info(pid)["nsyn"]
        pre, post, d = (np.empty(n, np.int32) for _ in range(3))
        w = np.empty(n, np.float64)
        call("hs_net_get_synapses", self.h, pid, _i32(pre), _i32(post), _i32(d), _f64(w), n)
        return pre, post, d, w

    def add_pathway_device(self, src, tgt, var, pre, post, delay_steps=None, scalar_delay=0, wmode=HS_W_SCALAR,
                           weights=None, wscalar=0.0):
        """add_pathway with the bundle table built on the device: the arrays
        (numpy, or DeviceArray already in HBM) are sorted by (pre, delay) there."""
        held = []

        def dev(a, dtype):
            if a is None:
                return None
            if isinstance(a, DeviceArray):
                return a.ptr
            d = DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dtype))
            held.append(d)
            return d.ptr

        n = pre.n if isinstance(pre, DeviceArray) else int(np.asarray(pre).size)
        pid = C.c_int()
        try:
            call("hs_net_add_pathway_device", self.h, src, tgt, var, n, dev(pre, np.int32), dev(post, np.int32),
                 dev(delay_steps, np.int32), int(scalar_delay), int(wmode),
                 dev(weights, np.float64) if wmode != HS_W_SCALAR else None, float(wscalar), C.byref(pid))
        finally:
            for d in held:
                d.free()
        return pid.value

    def add_spike_monitor(self, gid, capacity=1 << 20, start=None, stop=None):
        """SpikeMonitor of a group, or of its neurons [start, stop) (a Subgroup:
        recorded indices relative to start)."""
        mid = C.c_int()
        if start is None and stop is None:
            call("hs_net_add_spike_monitor", self.h, gid, int(capacity), C.byref(mid))
        else:
            start = 0 if start is None else start
            stop = self.groups[gid][1] if stop is None else stop
            call("hs_net_add_spike_monitor_range", self.h, gid, int(start), int(stop), int(capacity), C.byref(mid))
        return mid.value

    def add_synapse_monitor(self, pid, var, indices):
        """StateMonitor of synapse variable var (HS_SYN_W, _APRE, _APOST) of the
        synapses at `indices` (bundle order); read with state_monitor(mid, len(indices))."""
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        mid = C.c_int()
        call("hs_net_add_synapse_monitor", self.h, int(pid), int(var), _i32(idx), idx.size, C.byref(mid))
        return mid.value

    def add_rate_monitor(self, gid, start=0, stop=None):
        """PopulationRateMonitor over neurons [start, stop) of a group."""
        stop = self.groups[gid][1] if stop is None else stop
        mid = C.c_int()
        call("hs_net_add_rate_monitor", self.h, gid, int(start), int(stop), C.byref(mid))
        return mid.value

    def rate_monitor(self, mid):
        """population rate (Hz) of every step since the monitor was added"""
        n = self.mon_count(mid)
        out = np.empty(max(n, 1))
        call("hs_net_get_rate_monitor", self.h, mid, _f64(out), n)
        return out[:n]

    def add_state_monitor(self, gid, var, start, stop):
        mid = C.c_int()
        call("hs_net_add_state_monitor", self.h, gid, var, int(start), int(stop), C.byref(mid))
        return mid.value

    def set_monitor_interval(self, mid, every):
        """a StateMonitor's own dt as every*dt (before its first row)"""
        call("hs_net_set_monitor_interval", self.h, int(mid), int(every))

    def add_state_monitor_indices(self, gid, var, indices, every=1):
        """StateMonitor(record=indices, dt=every*dt): columns in the given order,
        a row at every step t with t % every == 0."""
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        mid = C.c_int()
        call("hs_net_add_state_monitor_indices", self.h, gid, var, _i32(idx), idx.size, int(every), C.byref(mid))
        return mid.value

    def set_option(self, key, value):
        call("hs_net_set_option", self.h, int(key), int(value))

    def get_option(self, key):
        v = C.c_longlong()
        call("hs_net_get_option", self.h, int(key), C.byref(v))
        return v.value

    def run(self, nsteps):
        call("hs_net_run", self.h, int(nsteps))

    def prepare(self, nsteps):
        """Instantiate everything a following run(nsteps) launches (hs_net_prepare)."""
        call("hs_net_prepare", self.h, int(nsteps))

    def stats(self):
        s = HsStats()
        call("hs_net_stats", self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in HsStats._fields_}

    def kernel_time(self, kind):
        """(total device ms, launches) of one kernel kind (KT_*) while OPT_KERNEL_TIMING is on."""
        ms, n = C.c_double(), C.c_int64()
        call("hs_net_kernel_time", self.h, int(kind), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def mon_count(self, mid):
        n = C.c_int64()
        call("hs_net_mon_count", self.h, mid, C.byref(n))
        return n.value

    def spikes(self, mid):
        """(i, step) of every recorded spike, sorted by (step, i)."""
        n = self.mon_count(mid)
        i = np.empty(max(n, 1), np.int32)
        t = np.empty(max(n, 1), np.int64)
        call("hs_net_get_spikes", self.h, mid, _i32(i), t.ctypes.data_as(C.POINTER(C.c_int64)), n)
        return i[:n], t[:n]

    def state_monitor(self, mid, width):
        rows = self.mon_count(mid)
        out = np.empty(max(rows * width, 1), np.float64)
        call("hs_net_get_state_monitor", self.h, mid, _f64(out), rows * width)
        return out[: rows * width].reshape(rows, width)


class Hines:
    """A batch of multicompartment trees on the device (hs_hines_*): one
    workgroup per tree, the tree kept on chip for `steps per launch` steps.
    precision: channel arithmetic (64 or 32); the cable solve is FP64."""

    V, M, N, H, I = 0, 1, 2, 3, 4

    def __init__(self, tb, precision=64):
        init()
        self.precision = int(precision)
        self.n_comp = tb.n_comp
        self.n_trees = tb.n_trees
        ints = [np.ascontiguousarray(a, dtype=np.int32) for a in
                (tb.tree_comp, tb.tree_sec, tb.sec_first, tb.sec_n, tb.sec_parent)]
        flts = [np.ascontiguousarray(a, dtype=np.float64) for a in (tb.area, tb.rl1, tb.rl2, tb.Cm, tb.chan)]
        h = C.c_void_p()
        call("hs_hines_create", self.precision, float(tb.dt), tb.n_trees, *[_i32(a) for a in ints],
             _f64(flts[0]), _f64(flts[1]), _f64(flts[2]), _f64(flts[3]), float(tb.Ri), _f64(flts[4]),
             float(tb.vth), C.byref(h))
        self.h = h

    def close(self):
        if self.h is not None and self.h.value:
            _lib.load_library().hs_hines_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, var, values):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float64), (self.n_comp,)))
        call("hs_hines_set_state", self.h, int(var), _f64(v))

    def get_state(self, var):
        out = np.empty(self.n_comp)
        call("hs_hines_get_state", self.h, int(var), _f64(out))
        return out

    def set_option(self, key, value):
        call("hs_hines_set_option", self.h, int(key), int(value))

    def run(self, nsteps):
        call("hs_hines_run", self.h, int(nsteps))

    def spike_counts(self):
        out = np.empty(self.n_trees, np.int32)
        call("hs_hines_spike_counts", self.h, _i32(out))
        return out

    def stats(self):
        w, g, s = C.c_double(), C.c_double(), C.c_int64()
        call("hs_hines_stats", self.h, C.byref(w), C.byref(g), C.byref(s))
        return {"last_run_wall_s": w.value, "last_run_gpu_ms": g.value, "steps": s.value}
