"""Synthetic workloads of the hot path, generated from a seed as plain arrays.

Every generator returns a :class:`NetSpec` -- groups (model, size, parameters,
initial state), pathways (pre, post, quantised delays, weights) and monitors --
that both the device runtime (:func:`build_hs`) and the oracle consume, so "same
seed and inputs" holds by construction.

Models and parameters follow the reference's examples and benchmarks:
  cuba      examples/cuba.py:118-143                      (CUBA LIF, exact)
  cobahh    examples/cobahh.py:140-190, tests/features/speed.py:73-113
  brunel    examples/brunelhakim.py:140-163                (Euler-Maruyama, noise)
  batch     many CUBA networks concatenated into one index space
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .runtime import (HS_MODEL_HH_COBA, HS_MODEL_LIF_CUBA, HS_MODEL_LIF_NOISE, HS_MODEL_POISSON, HS_MODEL_SPIKEGEN,
                      HS_SYN_CODE, HS_W_PRE, HS_W_SCALAR, HS_W_SYNAPSE, spikegen_steps)

mV, ms, nS = 1e-3, 1e-3, 1e-9


@dataclass
class GroupSpec:
    model: int
    n: int
    params: list
    init: dict = field(default_factory=dict)  # var index -> array (or scalar)
    spikes: Optional[tuple] = None            # spike generator: (indices, steps)
    code: object = None                       # a codegen.CodeModel: the group's model is code (model unused)


@dataclass
class RandomConnect:
    """Synapses.connect(p=p) generated where the pathway is built (on the device
    by the runtime, by the oracle's restatement for the oracle), with delays
    uniform in delay=(lo, hi) seconds and weights w=(lo, hi) (wmode SYNAPSE:
    uniform per synapse; SCALAR: w[0]).  k > 0: a fixed out-degree instead,
    K distinct targets per source (connect(j='k for k in sample(N_post,
    size=K)'))."""
    p: float
    pre: tuple
    post: tuple
    seed: int
    stream: int = 0
    autapses: bool = True
    delay: tuple = (0.0, 0.0)
    wmode: int = HS_W_SCALAR
    w: tuple = (0.0, 0.0)
    k: int = 0  # > 0: fixed out-degree K instead of p (sample(N_post, size=K))
    # connect(condition=..., p='expression'): C expressions of
    # codegen.connect_expression ("" for none), compiled where the pathway is built
    cond: str = ""
    p_expr: str = ""
    n: int = 1  # synapses per selected pair (connect(n=...))
    w_expr: str = ""  # S.w = '<expression>': each synapse's weight as code (wmode SYNAPSE)


@dataclass
class PathSpec:
    src: int
    tgt: int
    var: int
    pre: Optional[np.ndarray]
    post: Optional[np.ndarray]
    delay_steps: Optional[np.ndarray] = None
    scalar_delay: int = 0
    wmode: int = HS_W_SCALAR
    weights: Optional[np.ndarray] = None
    wscalar: float = 0.0
    random: Optional[RandomConnect] = None
    stdp: Optional[tuple] = None  # taupre, taupost, dApre, dApost, wmin, wmax, scale
    code: object = None           # a codegen.SynapseCode: the pathway's on_pre is code (var unused)

    @property
    def nsyn(self):
        """synapse count (None for a random pathway: known once built)"""
        return None if self.pre is None else int(self.pre.size)


@dataclass
class StateMon:
    """StateMonitor(G, var, record=indices, dt=every*dt): the columns are the
    given neurons in the given order, a row at every step t with t % every == 0.
    (A plain tuple (gid, var, start, stop) in NetSpec.state_monitors is the
    contiguous range [start, stop) recorded every step.)"""
    gid: int
    var: int
    indices: np.ndarray
    every: int = 1

    @property
    def width(self):
        return int(np.asarray(self.indices).size)


def state_monitor_width(m):
    """columns of a NetSpec state monitor entry"""
    return m.width if isinstance(m, StateMon) else m[3] - m[2]


@dataclass
class NetSpec:
    name: str
    dt: float
    seed: int
    groups: List[GroupSpec]
    paths: List[PathSpec]
    spike_monitors: List[int] = field(default_factory=list)         # group ids
    state_monitors: List[object] = field(default_factory=list)      # (gid, var, start, stop) or StateMon
    rate_monitors: List[tuple] = field(default_factory=list)        # (gid, start, stop)
    info: dict = field(default_factory=dict)

    @property
    def n_neurons(self):
        return sum(g.n for g in self.groups)

    @property
    def n_synapses(self):
        n = [p.nsyn for p in self.paths]
        return None if None in n else sum(n)


def quantise_delay(d, dt):
    """Delay in steps as the reference computes it: (int)(d/dt + 0.5)
    (brian2cuda/templates/synapses_push_spikes.cu:375)."""
    return (np.asarray(d, dtype=np.float64) / dt + 0.5).astype(np.int32)


#: state-variable names of the built-in models, in slot order (include/hs_models.h)
MODEL_VAR_NAMES = {
    HS_MODEL_LIF_CUBA: ["v", "ge", "gi"],
    HS_MODEL_LIF_NOISE: ["V"],
    HS_MODEL_HH_COBA: ["v", "m", "n", "h", "ge", "gi"],
    HS_MODEL_SPIKEGEN: ["_placeholder"],
    HS_MODEL_POISSON: ["rates"],
}


def group_var_names(g):
    """State-variable names of a group spec, in slot order."""
    return list(g.code.names) if g.code is not None else MODEL_VAR_NAMES[g.model]


def code_var_indices(spec, p):
    """(target, source, target-read) variable indices of a code pathway's
    post_names / pre_names / post_read_names."""
    tn, sn = group_var_names(spec.groups[p.tgt]), group_var_names(spec.groups[p.src])
    for n in list(p.code.post_names) + list(p.code.post_read_names):
        if n not in tn:
            raise ValueError(f"the target group has no variable {n!r}")
    for n in p.code.pre_names:
        if n not in sn:
            raise ValueError(f"the source group has no variable {n!r}")
    return ([tn.index(n) for n in p.code.post_names], [sn.index(n) for n in p.code.pre_names],
            [tn.index(n) for n in p.code.post_read_names])


def bundle_order(p):
    """Creation-order indices of an explicit pathway's synapses in bundle order
    (stable by presynaptic neuron, then delay): the order the device keeps
    synapse variables in."""
    d = np.zeros(len(p.pre), np.int64) if p.delay_steps is None else np.asarray(p.delay_steps, np.int64)
    return np.lexsort((d, np.asarray(p.pre, np.int64)))


def bernoulli_connect(rng, n_pre, n_post, p, pre_offset=0, post_offset=0):
    """Independent Bernoulli(p) synapses for every (pre, post) pair.

    Gaps between successes of a Bernoulli process are geometric, so the
    successes over the flattened n_pre*n_post pairs are a cumulative sum of
    geometric draws (the skip method the reference's generator uses per row,
    synapses_create_generator.cu:315-333).  Returns pre, post (int32), ordered
    by (pre, post)."""
    total = int(n_pre) * int(n_post)
    if p <= 0 or total == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.int32)
    if p >= 1:
        flat = np.arange(total, dtype=np.int64)
    else:
        mean = total * p
        m = int(mean + 8 * np.sqrt(mean) + 16)
        pos = np.cumsum(rng.geometric(p, size=m).astype(np.int64)) - 1
        while pos[-1] < total:
            more = np.cumsum(rng.geometric(p, size=m).astype(np.int64)) + pos[-1]
            pos = np.concatenate([pos, more])
        flat = pos[pos < total]
    pre = (flat // n_post).astype(np.int32) + np.int32(pre_offset)
    post = (flat % n_post).astype(np.int32) + np.int32(post_offset)
    return pre, post


# ---- C1: CUBA --------------------------------------------------------------------

def cuba_params(El=-49 * mV, Vt=-50 * mV, Vr=-60 * mV, taum=20 * ms, taue=5 * ms, taui=10 * ms, refractory=5 * ms):
    return [El, Vt, Vr, taum, taue, taui, refractory]


def cuba(N=4000, seed=1234, p=0.02, delay_steps=1, dt=0.1 * ms, record=100, we=1.62 * mV, wi=-9 * mV,
         monitors=True):
    """C1: 4000 LIF (3200 exc / 800 inh), p=0.02, delay one dt."""
    rng = np.random.default_rng(seed)
    Ne = int(0.8 * N)
    Vt, Vr = -50 * mV, -60 * mV
    v0 = Vr + rng.random(N) * (Vt - Vr)
    g = GroupSpec(HS_MODEL_LIF_CUBA, N, cuba_params(), {0: v0, 1: 0.0, 2: 0.0})
    pe, qe = bernoulli_connect(rng, Ne, N, p)
    pi, qi = bernoulli_connect(rng, N - Ne, N, p, pre_offset=Ne)
    paths = [
        PathSpec(0, 0, 1, pe, qe, scalar_delay=delay_steps, wmode=HS_W_SCALAR, wscalar=we),
        PathSpec(0, 0, 2, pi, qi, scalar_delay=delay_steps, wmode=HS_W_SCALAR, wscalar=wi),
    ]
    spec = NetSpec("cuba", dt, seed, [g], paths, info={"N": N, "p": p})
    if monitors:
        spec.spike_monitors = [0]
        if record:
            spec.state_monitors = [(0, 0, 0, min(record, N))]
    return spec


CUBA_EQS = """
dv/dt  = (ge+gi-(v-El))/taum : volt (unless refractory)
dge/dt = -ge/taue : volt
dgi/dt = -gi/taui : volt
"""


def cuba_conditional(N=4000, seed=9, dt=0.1 * ms, we=1.62 * mV, wi=-9 * mV, monitors=True):
    """C1's neurons with connectivity given as code, generated where the network
    is built: excitatory sources (the first 80%) connect with the probability
    expression p = '0.5*exp(-abs(i*N_post/N_pre - j)/50)' (about 50 targets
    around each source's place), inhibitory ones where the condition
    'i != j and abs(i - j) % 41 == 0' holds, with p = 0.7 (about 68 targets);
    one step of delay."""
    from .codegen import connect_expression
    spec = cuba(N=N, seed=seed, dt=dt, we=we, wi=wi, monitors=monitors)
    ne = int(0.8 * N)
    pe, _ = connect_expression("0.5*exp(-abs(i*N_post/N_pre - j)/50)")
    ci, _ = connect_expression("i != j and abs(i - j) % 41 == 0")
    d = (dt, dt)
    spec.paths = [
        PathSpec(0, 0, 1, None, None, wmode=HS_W_SCALAR, wscalar=we,
                 random=RandomConnect(1.0, (0, ne), (0, N), seed + 1, 0, True, d, HS_W_SCALAR, (we, we), p_expr=pe)),
        PathSpec(0, 0, 2, None, None, wmode=HS_W_SCALAR, wscalar=wi,
                 random=RandomConnect(0.7, (ne, N), (0, N), seed + 2, 1, True, d, HS_W_SCALAR, (wi, wi), cond=ci)),
    ]
    spec.name = "cuba_conditional"
    return spec


def code_cuba(N=4000, seed=1234, p=0.02, delay_steps=1, dt=0.1 * ms, method=None, we=1.62 * mV, wi=-9 * mV,
              monitors=True):
    """C1 with the model given as code: examples/cuba.py's equations, threshold
    and reset through the code generator (brian2cuda_amd.codegen, method None:
    'exact', as Brian2 picks for this linear system), same connectivity as cuba()."""
    from .codegen import neuron_model
    spec = cuba(N=N, seed=seed, p=p, delay_steps=delay_steps, dt=dt, record=0, we=we, wi=wi, monitors=monitors)
    ns = dict(El=-49 * mV, taum=20 * ms, taue=5 * ms, taui=10 * ms, Vt=-50 * mV, Vr=-60 * mV)
    model, names = neuron_model(CUBA_EQS, "v > Vt", "v = Vr", 5 * ms, method, ns, dt=dt)
    g = spec.groups[0]
    spec.groups[0] = GroupSpec(0, g.n, [], g.init, code=model)
    spec.name = "code_cuba"
    return spec


ADEX_EQS = """
dv/dt = (gL*(EL - v) + gL*DeltaT*exp((v - VT)/DeltaT) + (ge + gr)*(Ee - v) + gi*(Ei - v) - w + I)/C : volt
dw/dt = (a*(v - EL) - w)/tauw : amp
dge/dt = -ge/taue : siemens
dgi/dt = -gi/taui : siemens
dgr/dt = -gr/taue : siemens
I : amp (constant)
"""
NOISY_LIF_EQS = """
dV/dt = (-V + mu + sigma*sqrt(tau)*xi)/tau : volt
dg/dt = -g/taug : volt
"""


def code_mix(N=3000, seed=7, dt=0.1 * ms, method="rk4", monitors=True):
    """Three code models in one network with a built-in Poisson input: adaptive
    exponential integrate-and-fire neurons (AdEx, conductance inputs, reset
    'v = Vr; w += b'), noise-driven LIF neurons (Euler-Maruyama, '(unless
    refractory)'-free, a second variable) and a CUBA LIF inhibitory population
    (exact), with random delays of 1 to 30 steps between them.  Each target
    variable has one incoming pathway of one weight, so the effects a neuron
    receives in a step are equal and their sum does not depend on the order the
    device adds them in."""
    from .codegen import neuron_model
    rng = np.random.default_rng(seed)
    ne, nn, ni, npo = int(0.5 * N), int(0.25 * N), N - int(0.5 * N) - int(0.25 * N), 500
    adex, _ = neuron_model(ADEX_EQS, "v > -40*mV", "v = Vr\nw += b", 2 * ms, method,
                           dict(gL=30 * nS, EL=-70.6 * mV, DeltaT=2 * mV, VT=-50.4 * mV, C=281e-12, a=4 * nS,
                                tauw=144 * ms, b=0.0805e-9, Vr=-70.6 * mV, Ee=0 * mV, Ei=-80 * mV, taue=5 * ms,
                                taui=10 * ms), dt=dt)
    noisy, _ = neuron_model(NOISY_LIF_EQS.replace("-V + mu", "-V + mu + g"), "V > theta", "V = Vr", 2 * ms, "euler",
                            dict(mu=18 * mV, sigma=3 * mV, tau=20 * ms, taug=5 * ms, theta=20 * mV, Vr=10 * mV),
                            dt=dt, stream=3)
    inh, _ = neuron_model(CUBA_EQS, "v > Vt", "v = Vr", 1 * ms, None,
                          dict(El=-49 * mV, taum=10 * ms, taue=5 * ms, taui=10 * ms, Vt=-50 * mV, Vr=-60 * mV), dt=dt)
    groups = [
        GroupSpec(0, ne, [], {0: -70.6 * mV + 10 * mV * rng.random(ne), 5: 0.4e-9 + 0.2e-9 * rng.random(ne)},
                  code=adex),
        GroupSpec(0, nn, [], {0: 10 * mV + 10 * mV * rng.random(nn)}, code=noisy),
        GroupSpec(0, ni, [], {0: -60 * mV + 10 * mV * rng.random(ni)}, code=inh),
        GroupSpec(HS_MODEL_POISSON, npo, [20.0, 5.0]),
    ]

    def rnd(src_n, tgt_n, p, src, tgt, var, w, dmax):
        pre, post = bernoulli_connect(rng, src_n, tgt_n, p)
        d = rng.integers(1, dmax + 1, pre.size).astype(np.int32)
        return PathSpec(src, tgt, var, pre, post, delay_steps=d, wmode=HS_W_SCALAR, wscalar=w)

    paths = [
        rnd(npo, ne, 0.05, 3, 0, 2, 3 * nS, 5),       # Poisson -> AdEx ge
        rnd(ne, ne, 0.02, 0, 0, 4, 1 * nS, 30),       # AdEx -> AdEx gr
        rnd(ni, ne, 0.05, 2, 0, 3, 4 * nS, 10),       # inh -> AdEx gi
        rnd(ne, nn, 0.02, 0, 1, 1, 0.5 * mV, 20),     # AdEx -> noisy g
        rnd(nn, ni, 0.05, 1, 2, 1, 1.0 * mV, 15),     # noisy -> inh ge
        rnd(ni, ni, 0.05, 2, 2, 2, -2.0 * mV, 8),     # inh -> inh gi
    ]
    spec = NetSpec("code_mix", dt, seed, groups, paths, info={"N": N})
    if monitors:
        spec.spike_monitors = [0, 1, 2]
    return spec


HH_EQS = """
dv/dt = (gl*(El-v)+ge*(Ee-v)+gi*(Ei-v)-gNa*(m*m*m)*h*(v-ENa)-gK*(n*n*n*n)*(v-EK))/Cm : volt
dm/dt = alpha_m*(1-m)-beta_m*m : 1
dn/dt = alpha_n*(1-n)-beta_n*n : 1
dh/dt = alpha_h*(1-h)-beta_h*h : 1
dge/dt = -ge*(1./taue) : siemens
dgi/dt = -gi*(1./taui) : siemens
alpha_m = 0.32*(mV**-1)*(13.*mV-v+VT)/(exp((13.*mV-v+VT)/(4.*mV))-1.)/ms : Hz
beta_m = 0.28*(mV**-1.)*(v-VT-40.*mV)/(exp((v-VT-40.*mV)/(5.*mV))-1.)/ms : Hz
alpha_h = 0.128*exp((17.*mV-v+VT)/(18.*mV))/ms : Hz
beta_h = 4./(1.+exp((40.*mV-v+VT)/(5.*mV)))/ms : Hz
alpha_n = 0.032*(mV**-1.)*(15.*mV-v+VT)/(exp((15.*mV-v+VT)/(5.*mV))-1.)/ms : Hz
beta_n = .5*exp((10.*mV-v+VT)/(40.*mV))/ms : Hz
"""


def code_cobahh(N=4000, seed=1234, method="exponential_euler", **kw):
    """C2 with the model given as code: examples/cobahh.py's equations
    (exponential Euler, as the example asks) through the code generator, same
    connectivity and initial state as cobahh()."""
    from .codegen import neuron_model
    spec = cobahh(N=N, seed=seed, **kw)
    p = cobahh_params()
    ns = dict(Cm=p[0], gl=p[1], El=p[2], EK=p[3], ENa=p[4], gNa=p[5], gK=p[6], VT=p[7], Ee=p[8], Ei=p[9],
              taue=p[10], taui=p[11])
    model, names = neuron_model(HH_EQS, "v > -20*mV", None, p[13], method, ns, dt=spec.dt)
    assert names == ["v", "m", "n", "h", "ge", "gi"]
    g = spec.groups[0]
    spec.groups[0] = GroupSpec(0, g.n, [], g.init, code=model)
    spec.name = "code_cobahh"
    return spec


def code_synapses(seed=3, npo=400, nl=60, p=0.2, release=False):
    """Poisson inputs onto code-model LIF neurons through synaptic code: each
    arrival adds 1.5 mV to ge, counts itself in 'hits' and moves w with the
    source's rate ('w = clip(w + 1e-3*rates_pre*w - 0.002, 0, 1)'); per-synapse
    delays of 0 to 11 steps, w initially uniform.
    release=True: stochastic release instead -- an arrival is released with
    probability 0.5 ('released = rand() < 0.5'), adds 1.5 mV only then and counts
    in 'nrel'; w takes a multiplicative normal step ('w = clip(w + 0.01*randn()*w,
    0, 1)'), and each spike of the target lowers w by up to 0.002 (on_post
    'w = clip(w - 0.002*rand(), 0, 1)')."""
    from .codegen import neuron_model, synapse_code
    rng = np.random.default_rng(seed)
    lif, names = neuron_model("dv/dt = (ge - v)/(10*ms) : volt\ndge/dt = -ge/(5*ms) : volt",
                              "v > 5*mV", "v = 0*volt", 2e-3, "euler", {})
    pre, post = bernoulli_connect(rng, npo, nl, p)
    d = rng.integers(0, 12, pre.size).astype(np.int32)
    if release:
        code = synapse_code("released = rand() < 0.5\nge_post += 1.5*mV*released\nhits += 1\nnrel += released\n"
                            "w = clip(w + 0.01*randn()*w, 0, 1)", ["w", "hits", "released", "nrel"], names, [],
                            on_post="w = clip(w - 0.002*rand(), 0, 1)")
        code.init = {"w": rng.random(pre.size), "hits": 0.0, "released": 0.0, "nrel": 0.0}
    else:
        code = synapse_code("ge_post += 1.5*mV\nhits += 1\nw = clip(w + 1e-3*rates_pre*w - 0.002, 0, 1)",
                            ["w", "hits"], names, ["rates"])
        code.init = {"w": rng.random(pre.size), "hits": 0.0}
    rates = 5.0 + 30.0 * rng.random(npo)
    groups = [GroupSpec(HS_MODEL_POISSON, npo, [20.0, 2.0], {0: rates}), GroupSpec(0, nl, [], {}, code=lif)]
    spec = NetSpec("code_synapses", 0.1 * ms, seed, groups, [PathSpec(0, 1, 0, pre, post, delay_steps=d, code=code)])
    spec.spike_monitors = [1]
    return spec


STDP_SYNAPSE_MODEL = """
w : 1
dApre/dt = -Apre / taupre : 1 (event-driven)
dApost/dt = -Apost / taupost : 1 (event-driven)
"""


def stdp_feedforward(delay_steps=0, n_in=100, n_out=60, seed=5, code=False):
    """A spike generator (at most one spike per step) onto LIF neurons, all to
    all, with STDP (test_cpp_cuda_consistency.py:45-61's synapse model):
    built in (hs_net_pathway_stdp), or with code=True the same synapse model
    written as Brian2 code -- event-driven traces, on_pre 'ge += w*scale;
    Apre += dApre; w = clip(w + Apost, wmin, wmax)', on_post 'Apost += dApost;
    w = clip(w + Apre, wmin, wmax)' -- through the code generator."""
    from .codegen import synapse_code
    rs = np.random.RandomState(seed)
    times = rs.choice(np.arange(3000), 1500, replace=False)
    src = rs.randint(0, n_in, times.size)
    G = GroupSpec(HS_MODEL_SPIKEGEN, n_in, [0.0], spikes=(src.astype(np.int32), times.astype(np.int64)))
    P = GroupSpec(HS_MODEL_LIF_CUBA, n_out, [-49e-3, -50e-3, -60e-3, 20e-3, 5e-3, 5e-3, 5e-3, 1.0],
                  {0: -60e-3 + 10e-3 * rs.rand(n_out), 1: 0.0})
    ii, jj = np.meshgrid(np.arange(n_in), np.arange(n_out), indexing="ij")
    w = 1.5 + 2.0 * rs.rand(n_in * n_out)
    params = dict(taupre=20e-3, taupost=20e-3, dApre=0.05, dApost=-0.055, wmin=0.0, wmax=4.0, scale=1e-3)
    S = PathSpec(0, 1, 1, ii.ravel().astype(np.int32), jj.ravel().astype(np.int32), scalar_delay=delay_steps,
                 wmode=HS_W_SYNAPSE, weights=w)
    if code:
        S.code = synapse_code("ge += w*scale\nApre += dApre\nw = clip(w + Apost, wmin, wmax)", [], ["v", "ge", "gi"],
                              ["_placeholder"], params, on_post="Apost += dApost\nw = clip(w + Apre, wmin, wmax)",
                              model=STDP_SYNAPSE_MODEL)
    else:
        S.stdp = tuple(params[k] for k in ("taupre", "taupost", "dApre", "dApost", "wmin", "wmax", "scale"))
    return NetSpec("stdp_ff" + ("_code" if code else ""), 0.1 * ms, seed, [G, P], [S], spike_monitors=[1],
                   state_monitors=[(1, 0, 0, 20)])


# ---- C2: COBAHH ------------------------------------------------------------------

def cobahh_params():
    area = 20000e-12  # m^2
    Cm = 1e-2 * area   # 1 uF/cm^2
    gl = 0.5 * area    # 5e-5 S/cm^2
    gNa = 1000.0 * area
    gK = 300.0 * area
    return [Cm, gl, -60 * mV, -90 * mV, 50 * mV, gNa, gK, -63 * mV, 0 * mV, -80 * mV, 5 * ms, 10 * ms,
            -20 * mV, 3 * ms]


def cobahh(N=4000, seed=1234, p=None, dt=0.1 * ms, we=6 * nS, wi=67 * nS, delay_steps=0, monitors=True):
    """C2: COBAHH, 80/20 exc/inh, conductance synapses; p=0.02 at N=4000, else mean in-degree 80."""
    rng = np.random.default_rng(seed)
    if p is None:
        p = 0.02 if N <= 4000 else 80.0 / N
    Ne = int(0.8 * N)
    v0 = -60 * mV + 5 * mV * rng.standard_normal(N)
    g = GroupSpec(HS_MODEL_HH_COBA, N, cobahh_params(), {0: v0, 1: 0.0, 2: 0.0, 3: 0.0, 4: 0.0, 5: 0.0})
    pe, qe = bernoulli_connect(rng, Ne, N, p)
    pi, qi = bernoulli_connect(rng, N - Ne, N, p, pre_offset=Ne)
    paths = [
        PathSpec(0, 0, 4, pe, qe, scalar_delay=delay_steps, wmode=HS_W_SCALAR, wscalar=we),
        PathSpec(0, 0, 5, pi, qi, scalar_delay=delay_steps, wmode=HS_W_SCALAR, wscalar=wi),
    ]
    spec = NetSpec("cobahh", dt, seed, [g], paths, info={"N": N, "p": p})
    if monitors:
        spec.spike_monitors = [0]
    return spec


# ---- C3: Brunel-Hakim --------------------------------------------------------------

def brunel_params(tau=20 * ms, mu=25 * mV, sigma=1 * mV, theta=20 * mV, Vr=10 * mV, refractory=2 * ms, stream=0):
    return [tau, mu, sigma, theta, Vr, refractory, float(stream)]


def brunel(N=100_000, K=1000, seed=1234, dt=0.1 * ms, J=-0.1 * mV, dlo=0.5 * ms, dhi=5.0 * ms,
           per_synapse_weights=True, monitors=False, sigma=1 * mV, mu=25 * mV, device_connect=False,
           fixed_outdegree=False):
    """C3: sparse inhibitory LIF network with noise and heterogeneous delays.

    connect(p=K/N) (examples/brunelhakim.py:144-156), delays U[dlo, dhi] quantised
    to dt, on_pre 'V += -|J|' with a per-synapse weight array.  device_connect:
    the synapses are generated where the network is built (RandomConnect)
    instead of by numpy here; with fixed_outdegree, exactly K targets per neuron
    (connect(j='k for k in sample(N, size=K)')) in place of connect(p)."""
    g = GroupSpec(HS_MODEL_LIF_NOISE, N, brunel_params(sigma=sigma, mu=mu), {0: 10 * mV})
    if fixed_outdegree and not device_connect:
        raise ValueError("fixed_outdegree is generated where the network is built (device_connect=True)")
    if device_connect:
        rc = RandomConnect(K / N, (0, N), (0, N), seed, stream=1, delay=(dlo, dhi),
                           wmode=HS_W_SYNAPSE if per_synapse_weights else HS_W_SCALAR, w=(J, J),
                           k=K if fixed_outdegree else 0)
        ps = PathSpec(0, 0, 0, None, None, wmode=rc.wmode, wscalar=J, random=rc)
        spec = NetSpec("brunel", dt, seed, [g], [ps], info={"N": N, "K": K})
        if monitors:
            spec.spike_monitors = [0]
        return spec
    rng = np.random.default_rng(seed)
    pre, post = bernoulli_connect(rng, N, N, K / N)
    d = quantise_delay(dlo + (dhi - dlo) * rng.random(pre.size), dt)
    if per_synapse_weights:
        ps = PathSpec(0, 0, 0, pre, post, delay_steps=d, wmode=HS_W_SYNAPSE,
                      weights=np.full(pre.size, J, dtype=np.float32).astype(np.float64))
    else:
        ps = PathSpec(0, 0, 0, pre, post, delay_steps=d, wmode=HS_W_SCALAR, wscalar=J)
    spec = NetSpec("brunel", dt, seed, [g], [ps], info={"N": N, "K": K})
    if monitors:
        spec.spike_monitors = [0]
    return spec


def code_brunel(N=100_000, K=1000, seed=1234, **kw):
    """C3 with the model given as code: examples/brunelhakim.py's equation
    'dV/dt = (-V+muext + sigmaext*sqrt(tau)*xi)/tau' (Euler-Maruyama), threshold,
    reset and refractory period through the code generator; same connectivity,
    delays and initial state as brunel()."""
    from .codegen import neuron_model
    spec = brunel(N=N, K=K, seed=seed, **kw)
    g = spec.groups[0]
    tau, mu, sigma, theta, vr, refr, stream = g.params
    model, _ = neuron_model("dV/dt = (-V + muext + sigmaext*sqrt(tau)*xi)/tau : volt", "V > theta", "V = Vr", refr,
                            "euler", dict(tau=tau, muext=mu, sigmaext=sigma, theta=theta, Vr=vr), stream=int(stream))
    spec.groups[0] = GroupSpec(0, g.n, [], g.init, code=model)
    spec.name = "code_brunel"
    return spec


# ---- C4: batched parameter sweep ----------------------------------------------------

def batch_cuba(n_nets=256, seed=1234, nmin=500, nmax=16000, p=0.02, dt=0.1 * ms, wi=-9 * mV, delay_steps=1,
               monitors=True):
    """C4: n_nets independent CUBA networks of log-uniform sizes in one index space.

    Net b (seed+b) has its own w_e ~ U[0.5, 3] mV; its neurons occupy
    [offset_b, offset_b + N_b) and its synapses stay inside that range, so the
    concatenated group advances all networks in the same launches.  The
    excitatory pathway carries w_e per presynaptic neuron (weight mode PRE)."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n_nets))).astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N = int(offsets[-1])
    Vt, Vr = -50 * mV, -60 * mV
    v0 = np.empty(N)
    wpre = np.zeros(N)
    we_net = np.empty(n_nets)
    pre_e, post_e, pre_i, post_i = [], [], [], []
    for b in range(n_nets):
        rb = np.random.default_rng(seed + b)
        nb, o = int(sizes[b]), int(offsets[b])
        ne = int(0.8 * nb)
        v0[o:o + nb] = Vr + rb.random(nb) * (Vt - Vr)
        we_net[b] = rb.uniform(0.5, 3.0) * mV
        wpre[o:o + nb] = we_net[b]
        a, c = bernoulli_connect(rb, ne, nb, p, pre_offset=o, post_offset=o)
        pre_e.append(a); post_e.append(c)
        a, c = bernoulli_connect(rb, nb - ne, nb, p, pre_offset=o + ne, post_offset=o)
        pre_i.append(a); post_i.append(c)
    g = GroupSpec(HS_MODEL_LIF_CUBA, N, cuba_params(), {0: v0, 1: 0.0, 2: 0.0})
    paths = [
        PathSpec(0, 0, 1, np.concatenate(pre_e), np.concatenate(post_e), scalar_delay=delay_steps,
                 wmode=HS_W_PRE, weights=wpre),
        PathSpec(0, 0, 2, np.concatenate(pre_i), np.concatenate(post_i), scalar_delay=delay_steps,
                 wmode=HS_W_SCALAR, wscalar=wi),
    ]
    spec = NetSpec("batch_cuba", dt, seed, [g], paths,
                   info={"n_nets": n_nets, "sizes": sizes, "offsets": offsets, "we": we_net})
    if monitors:
        spec.spike_monitors = [0]
    return spec


# ---- plasticity: the reference's STDP consistency networks ---------------------------

def stdp_poisson_example(N=10_000, K=1000, seed=1234, dt=0.1 * ms, delay=0.0, n_lif=None):
    """The network of examples/stdp.py:143-182 (post_effects, delays 'none' or a
    homogeneous delay in seconds):

      input  : PoissonGroup(N, rates=F), F = 15 Hz
      neurons: N/K LIF, dv/dt = (ge*(Ee-vr) + El - v)/taum, dge/dt = -ge/taue,
               threshold v > vt, reset v = vr, method 'exact'.  The equation is
               linear ((Ee - vr) is a constant): it is the CUBA model with
               ge_CUBA = ge*(Ee - vr), so the pathway scales w by Ee - vr
      S      : input -> neurons, connect(p = K/N), w = rand()*gmax, STDP
               (taupre = taupost = 20 ms, dApre = 0.01 gmax,
               dApost = -1.05 dApre, w clipped to [0, gmax])
    Brian2's default v = 0 starts every neuron above threshold."""
    taum, taupre, taupost = 10 * ms, 20 * ms, 20 * ms
    Ee, vt, vr, El, taue = 0 * mV, -54 * mV, -60 * mV, -74 * mV, 5 * ms
    F, gmax = 15.0, 0.01
    dApre = 0.01
    dApost = -dApre * taupre / taupost * 1.05
    dApost *= gmax
    dApre *= gmax
    n_lif = N // K if n_lif is None else n_lif
    inp = GroupSpec(HS_MODEL_POISSON, N, [F, 1.0])
    # taui is unused (gi = 0); it must only differ from taum
    lif = GroupSpec(HS_MODEL_LIF_CUBA, n_lif, [El, vt, vr, taum, taue, taue, 0.0], {0: 0.0, 1: 0.0, 2: 0.0})
    rc = RandomConnect(p=K / N, pre=(0, N), post=(0, n_lif), seed=seed, stream=2, delay=(delay, delay),
                       wmode=HS_W_SYNAPSE, w=(0.0, gmax))
    S = PathSpec(0, 1, 1, None, None, wmode=HS_W_SYNAPSE, random=rc,
                 stdp=(taupre, taupost, dApre, dApost, 0.0, gmax, Ee - vr))
    return NetSpec("stdp_poisson", dt, seed, [inp, lif], [S], spike_monitors=[1],
                   info={"gmax": gmax, "N": N, "K": K})


def stdp_example(heterog_delays=False, n_cells=100, seed=42, dt=0.1 * ms):
    """The networks of brian2cuda/tests/test_cpp_cuda_consistency.py:10-89
    (test_stdp_example) and :92-172 (test_stdp_heterog_delays_example), with the
    test's own numpy draws (legacy RandomState(42), same order):

      P: n_cells LIF, dv/dt = (g-(v-El))/taum, dg/dt = -g/taus, refractory 5 ms
         without "(unless refractory)" -> CUBA with gi unused and free = 1
      Q: SpikeGeneratorGroup(n_cells, sources, times), one spike per ms at most
      S: P -> P all-to-all, STDP (event-driven traces), w = fac*randn, on_pre
         g += w*mV; heterog_delays: delay = U[0, 2) ms per synapse
      T: Q -> P one-to-one, w = 10*fac, on_pre g += w*mV
    Pathway order: T then S (Brian2 sorts the synapses slot by order, then by
    name: synapses_1_pre before synapses_pre), S's post pathway after both."""
    rs = np.random.RandomState(seed)
    taum, taus = 20 * ms, 5 * ms
    Vt, Vr, El = -50 * mV, -60 * mV, -49 * mV
    fac = 60 * 0.27 / 10
    gmax = 20 * fac
    dApre, taupre = 0.01, 20 * ms
    taupost = taupre
    dApost = -dApre * taupre / taupost * 1.05
    dApost *= 0.1 * gmax
    dApre *= 0.1 * gmax
    connectivity = rs.randn(n_cells, n_cells)
    sources = rs.randint(0, n_cells, 10 * n_cells)
    times_ms = rs.choice(np.arange(10 * n_cells), 10 * n_cells, replace=False)
    v_init = Vr + rs.rand(n_cells) * (Vt - Vr)
    delay = rs.rand(n_cells ** 2) * 2 * ms if heterog_delays else None
    P = GroupSpec(HS_MODEL_LIF_CUBA, n_cells, [El, Vt, Vr, taum, taus, taus, 5 * ms, 1.0], {0: v_init, 1: 0.0})
    Q = GroupSpec(HS_MODEL_SPIKEGEN, n_cells, [0.0],
                  spikes=(sources.astype(np.int32), spikegen_steps(times_ms * ms, dt)))
    ii, jj = np.meshgrid(np.arange(n_cells), np.arange(n_cells), indexing="ij")
    T = PathSpec(1, 0, 1, np.arange(n_cells, dtype=np.int32), np.arange(n_cells, dtype=np.int32),
                 wmode=HS_W_SCALAR, wscalar=10 * fac * mV)
    S = PathSpec(0, 0, 1, ii.ravel().astype(np.int32), jj.ravel().astype(np.int32),
                 delay_steps=None if delay is None else quantise_delay(delay, dt),
                 wmode=HS_W_SYNAPSE, weights=fac * connectivity.ravel(),
                 stdp=(taupre, taupost, dApre, dApost, -np.inf, np.inf, mV))
    spec = NetSpec("stdp" + ("_heterog" if heterog_delays else ""), dt, seed, [P, Q], [T, S],
                   spike_monitors=[0], state_monitors=[(0, 0, 0, 10)], rate_monitors=[(0, 0, n_cells)],
                   info={"n_cells": n_cells, "record_synapses": list(range(10)), "record_every_steps": 1000})
    return spec


# ---- instantiate on the device --------------------------------------------------------

def build_hs(spec, precision=64, engine="plain", spike_capacity=1 << 20, options=None, device_sort=False):
    """Create a runtime.Net from a spec.  Returns (net, spike_monitor_ids, state_monitor_ids).
    device_sort: array pathways are uploaded and bundled on the device
    (add_pathway_device) instead of by the host counting sort."""
    from .runtime import Net
    net = Net(precision, engine, spec.dt, spec.seed)
    for k, v in (options or {}).items():  # before the pathways: the delivery mode is chosen per pathway
        net.set_option(k, v)
    for g in spec.groups:
        gid = net.add_group_code(g.n, g.code) if g.code is not None else net.add_group(g.model, g.n, g.params)
        for var, val in g.init.items():
            net.set_state(gid, var, val)
        if g.spikes is not None:
            net.set_spikegen(gid, *g.spikes)
    for p in spec.paths:
        if p.random is not None:
            r = p.random
            if r.cond or r.p_expr or r.n > 1 or r.w_expr:
                pid = net.connect_code(p.src, p.tgt, p.var, r.cond, r.p_expr, r.p, r.k, r.pre, r.post, r.autapses,
                                       r.seed, r.stream, r.delay, r.wmode, r.w, r.n, r.w_expr)
            elif r.k:
                pid = net.connect_fixed(p.src, p.tgt, p.var, r.k, r.pre, r.post, r.autapses, r.seed, r.stream,
                                        r.delay, r.wmode, r.w)
            else:
                pid = net.connect_random(p.src, p.tgt, p.var, r.p, r.pre, r.post, r.autapses, r.seed, r.stream,
                                         r.delay, r.wmode, r.w)
        else:
            add = net.add_pathway_device if device_sort else net.add_pathway
            pid = add(p.src, p.tgt, p.var, p.pre, p.post, p.delay_steps, p.scalar_delay, p.wmode, p.weights,
                      p.wscalar)
        if p.stdp is not None:
            net.pathway_stdp(pid, *p.stdp)
        if p.code is not None:
            post_i, pre_i, read_i = code_var_indices(spec, p)
            net.pathway_code(pid, p.code, post_i, pre_i, read_i)
            for name, val in (p.code.init or {}).items():
                if name == "w" and p.wmode == HS_W_SYNAPSE:
                    continue  # the pathway's weights
                a = np.asarray(val, np.float64)
                if a.size > 1:
                    if p.random is not None:
                        raise ValueError("per-synapse initial values need explicit synapses")
                    a = a[bundle_order(p)]
                net.set_synapse_var(pid, HS_SYN_CODE + p.code.syn_names.index(name), a)
    smons = [net.add_spike_monitor(g, spike_capacity) for g in spec.spike_monitors]
    stmons = [net.add_state_monitor_indices(m.gid, m.var, m.indices, m.every) if isinstance(m, StateMon)
              else net.add_state_monitor(*m) for m in spec.state_monitors]
    net.rate_mids = [net.add_rate_monitor(*m) for m in spec.rate_monitors]
    return net, smons, stmons
