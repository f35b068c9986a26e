"""CPU checks of the hip_standalone device's mapping from object descriptions
(equation text, threshold/reset, namespace, connect, on_pre, monitors) onto the
engine's network spec: the reference examples' models are recognised with their
parameters, unsupported code is refused with NotImplementedError, and the
module imports with and without Brian2.  (No GPU.)"""
import os
import subprocess
import sys

import numpy as np
import pytest

from brian2cuda_amd import device as dv
from brian2cuda_amd import networks as nw
from brian2cuda_amd.runtime import (HS_MODEL_HH_COBA, HS_MODEL_LIF_CUBA, HS_MODEL_LIF_NOISE, HS_MODEL_POISSON,
                                    HS_MODEL_SPIKEGEN, HS_W_SCALAR, HS_W_SYNAPSE)

mV, ms, nS = 1e-3, 1e-3, 1e-9
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CUBA_EQS = """
dv/dt  = (ge+gi-(v-El))/taum : volt (unless refractory)
dge/dt = -ge/taue : volt
dgi/dt = -gi/taui : volt
"""
CUBA_NS = {"taum": 20 * ms, "taue": 5 * ms, "taui": 10 * ms, "Vt": -50 * mV, "Vr": -60 * mV, "El": -49 * mV,
           "we": 1.62 * mV, "wi": -9 * mV, "Ne": 3200}

COBAHH_EQS = """
dv/dt = (gl*(El-v)+ge*(Ee-v)+gi*(Ei-v)-
         g_na*(m*m*m)*h*(v-ENa)-
         g_kd*(n*n*n*n)*(v-EK))/Cm : volt
dm/dt = alpha_m*(1-m)-beta_m*m : 1
dn/dt = alpha_n*(1-n)-beta_n*n : 1
dh/dt = alpha_h*(1-h)-beta_h*h : 1
dge/dt = -ge*(1./taue) : siemens
dgi/dt = -gi*(1./taui) : siemens
alpha_m = 0.32*(mV**-1)*(13.*mV-v+VT)/
         (exp((13.*mV-v+VT)/(4.*mV))-1.)/ms : Hz
beta_m = 0.28*(mV**-1.)*(v-VT-40.*mV)/
        (exp((v-VT-40.*mV)/(5.*mV))-1.)/ms : Hz
alpha_h = 0.128*exp((17.*mV-v+VT)/(18.*mV))/ms : Hz
beta_h = 4./(1.+exp((40.*mV-v+VT)/(5.*mV)))/ms : Hz
alpha_n = 0.032*(mV**-1.)*(15.*mV-v+VT)/
         (exp((15.*mV-v+VT)/(5.*mV))-1.)/ms : Hz
beta_n = .5*exp((10.*mV-v+VT)/(40.*mV))/ms : Hz
"""


def _cobahh_ns():
    Cm, gl, El, EK, ENa, gNa, gK, VT, Ee, Ei, taue, taui, vth, refr = nw.cobahh_params()
    return {"Cm": Cm, "gl": gl, "El": El, "EK": EK, "ENa": ENa, "g_na": gNa, "g_kd": gK, "VT": VT, "Ee": Ee,
            "Ei": Ei, "taue": taue, "taui": taui}


def test_cuba_example_maps_to_the_c1_model():
    """examples/cuba.py:118-143: the model, its parameters, connect('i<Ne', p=0.02)
    as a pre range, on_pre 'ge += we' as a scalar weight, monitors."""
    v0 = np.linspace(-60 * mV, -50 * mV, 4000)
    g = dv.GroupDesc("neurongroup", 4000, equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr",
                     refractory=5 * ms, method="exact", namespace=CUBA_NS, init={"v": v0, "ge": 0.0, "gi": 0.0})
    ce = dv.SynapsesDesc("synapses", "neurongroup", "neurongroup", "ge += we", namespace=CUBA_NS, p=0.02,
                         condition="i<Ne", delay=0.1 * ms, seed=7)
    ci = dv.SynapsesDesc("synapses_1", "neurongroup", "neurongroup", "gi += wi", namespace=CUBA_NS, p=0.02,
                         condition="i>=Ne", delay=0.1 * ms, seed=7)
    mons = [dv.MonitorDesc("spikemonitor", "spike", "neurongroup"),
            dv.MonitorDesc("statemonitor", "state", "neurongroup", variable="v", record=range(100))]
    spec, names = dv.network_spec([g], [ce, ci], mons, dt=0.1 * ms, seed=7)
    gs = spec.groups[0]
    assert gs.model == HS_MODEL_LIF_CUBA and gs.params == nw.cuba_params()
    np.testing.assert_array_equal(gs.init[0], v0)
    assert gs.init[1] == 0.0 and gs.init[2] == 0.0
    pe, pi = spec.paths
    assert (pe.var, pi.var) == (1, 2)
    assert pe.random.pre == (0, 3200) and pe.random.post == (0, 4000) and pe.wscalar == 1.62 * mV
    assert pi.random.pre == (3200, 4000) and pi.wscalar == -9 * mV
    assert pe.random.delay == (0.1 * ms, 0.1 * ms) and pe.random.wmode == HS_W_SCALAR
    assert spec.spike_monitors == [0] and spec.state_monitors == [(0, 0, 0, 100)]
    assert names["paths"] == ["synapses", "synapses_1"] and names["state_monitors"] == ["statemonitor"]


def test_cuba_without_unless_refractory_sets_free():
    eqs = CUBA_EQS.replace("(unless refractory)", "")
    g = dv.GroupDesc("g", 10, equations=eqs, threshold="v > Vt", reset="v=Vr", refractory=2 * ms,
                     namespace=CUBA_NS)
    spec, _ = dv.group_spec(g)
    assert spec.params == nw.cuba_params(refractory=2 * ms) + [1.0]


def test_brunel_hakim_example_maps_to_the_c3_model():
    """examples/brunelhakim.py:140-183 (V += -J, connect(p), delay as an array)."""
    eqs = "\ndV/dt = (-V+muext + sigmaext * sqrt(tau) * xi)/tau : volt\n"
    ns = {"Vr": 10 * mV, "theta": 20 * mV, "tau": 20 * ms, "J": 0.1 * mV, "muext": 25 * mV, "sigmaext": 1 * mV}
    g = dv.GroupDesc("neurongroup", 500, equations=eqs, threshold="V>theta", reset="V=Vr", refractory=2 * ms,
                     method="euler", namespace=ns, init={"V": 10 * mV})
    rng = np.random.default_rng(1)
    i, j = rng.integers(0, 500, 3000), rng.integers(0, 500, 3000)
    d = rng.uniform(0.5 * ms, 5 * ms, 3000)
    s = dv.SynapsesDesc("synapses", "neurongroup", "neurongroup", "V += -J", namespace=ns, i=i, j=j, delay=d)
    spec, _ = dv.network_spec([g], [s], [], dt=0.1 * ms)
    gs, ps = spec.groups[0], spec.paths[0]
    assert gs.model == HS_MODEL_LIF_NOISE and gs.params == nw.brunel_params()
    assert ps.wmode == HS_W_SCALAR and ps.wscalar == -0.1 * mV
    np.testing.assert_array_equal(ps.pre, i)
    np.testing.assert_array_equal(ps.delay_steps, nw.quantise_delay(d, 0.1 * ms))


def test_fixed_outdegree_connect_maps_to_the_device_sampler():
    """connect(condition='i != j', j='k for k in sample(N, size=K)') with drawn
    delays and weights: a RandomConnect with k = K that drops j == i."""
    eqs = "\ndV/dt = (-V+muext + sigmaext * sqrt(tau) * xi)/tau : volt\n"
    ns = {"Vr": 10 * mV, "theta": 20 * mV, "tau": 20 * ms, "J": 0.1 * mV, "muext": 25 * mV, "sigmaext": 1 * mV}
    g = dv.GroupDesc("neurongroup", 500, equations=eqs, threshold="V>theta", reset="V=Vr", refractory=2 * ms,
                     method="euler", namespace=ns, init={"V": 10 * mV})
    s = dv.SynapsesDesc("synapses", "neurongroup", "neurongroup", "V += w", k=50, condition="i != j",
                        values={"w": None}, weight_range=(0.1 * mV, 0.2 * mV), delay=(0.5 * ms, 5 * ms), seed=3)
    spec, _ = dv.network_spec([g], [s], [], dt=0.1 * ms)
    r = spec.paths[0].random
    assert (r.k, r.autapses, r.pre, r.post, r.seed) == (50, False, (0, 500), (0, 500), 3)
    assert r.wmode == HS_W_SYNAPSE and r.w == (0.1 * mV, 0.2 * mV) and r.delay == (0.5 * ms, 5 * ms)
    with pytest.raises(ValueError):
        dv.network_spec([g], [dv.SynapsesDesc("synapses", "neurongroup", "neurongroup", "V += J", namespace=ns,
                                              k=5, p=0.1)], [], dt=0.1 * ms)


def test_cobahh_example_maps_to_the_c2_model():
    """examples/cobahh.py:160-190 with subgroup sources Pe = P[:Ne], Pi = P[Ne:]
    and per-synapse constant weights 'we : siemens (constant)'."""
    g = dv.GroupDesc("neurongroup", 400, equations=COBAHH_EQS, threshold="v>-20*mV", refractory=3 * ms,
                     method="exponential_euler", namespace=_cobahh_ns())
    rng = np.random.default_rng(2)
    i, j = rng.integers(0, 320, 2000), rng.integers(0, 400, 2000)
    ce = dv.SynapsesDesc("synapses", "neurongroup", "neurongroup", "ge+=we", source_range=(0, 320), i=i, j=j,
                         values={"we": 6 * nS})
    ci = dv.SynapsesDesc("synapses_1", "neurongroup", "neurongroup", "gi+=wi", source_range=(320, 400),
                         i=i[:500] % 80, j=j[:500], values={"wi": np.full(500, 67 * nS)})
    spec, _ = dv.network_spec([g], [ce, ci], [], dt=0.1 * ms)
    assert spec.groups[0].model == HS_MODEL_HH_COBA and spec.groups[0].params == nw.cobahh_params()
    pe, pi = spec.paths
    assert (pe.var, pi.var) == (4, 5) and pe.wmode == HS_W_SYNAPSE
    np.testing.assert_array_equal(pe.weights, np.full(2000, 6 * nS))
    np.testing.assert_array_equal(pi.pre, i[:500] % 80 + 320)
    assert pe.scalar_delay == 0


def test_poisson_and_spike_generator_groups():
    p = dv.GroupDesc("poissongroup", 100, kind="poisson", rates=15.0)
    sg = dv.GroupDesc("spikegeneratorgroup", 5, kind="spikegenerator", indices=[0, 4, 2],
                      times=[0.0, 0.5 * ms, 1.0 * ms])
    spec, _ = dv.network_spec([p, sg], [], [], dt=0.1 * ms)
    assert spec.groups[0].model == HS_MODEL_POISSON and spec.groups[0].params == [15.0]
    assert spec.groups[1].model == HS_MODEL_SPIKEGEN
    np.testing.assert_array_equal(spec.groups[1].spikes[1], [0, 5, 10])


@pytest.mark.parametrize("change,msg", [
    (dict(equations="dv/dt = -v/tau : volt"), "no value for 'tau'"),
    (dict(equations="dv/dt = erf(v)/taum : volt"), "function 'erf'"),
    (dict(reset="v = Vr; w += 1"), "'w' is not a state variable"),
    (dict(method="gsl"), "method 'gsl'"),
    (dict(namespace={k: v for k, v in CUBA_NS.items() if k != "taum"}), "taum"),
])
def test_unsupported_groups_are_refused(change, msg):
    kw = dict(equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr", refractory=5 * ms, namespace=CUBA_NS)
    kw.update(change)
    with pytest.raises(NotImplementedError, match=msg):
        dv.group_spec(dv.GroupDesc("g", 10, **kw), dt=0.1 * ms)


@pytest.mark.parametrize("on_pre,kw,msg", [
    ("w += 1", {}, "not a synapse or target variable"),
    ("w += v_post", dict(model="w : 1"), "reads the target variable"),
    ("ge += we", dict(p=0.1, condition="rand(3) < 0.5"), "no arguments"),
    ("ge += we", dict(p="0.1 * unknown_name"), "unknown_name"),
    ("ge += w", dict(p=0.1, values={"w": 1.0}), "'lo \\+ \\(hi-lo\\)\\*rand\\(\\)'"),
])
def test_unsupported_synapses_are_refused(on_pre, kw, msg):
    g = dv.GroupDesc("g", 10, equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr", namespace=CUBA_NS)
    s = dv.SynapsesDesc("s", "g", "g", on_pre, namespace=CUBA_NS, **kw)
    with pytest.raises(NotImplementedError, match=msg):
        dv.network_spec([g], [s], [], dt=0.1 * ms)


def test_connect_conditions_as_code():
    """connect(condition=..., p=...) beyond index ranges and i != j: the
    condition and a probability that reads indices or variables become the C
    expressions of the device generator (codegen.connect_expression), i and j
    within the Synapses' subgroups; constant probabilities fold to numbers;
    connect(condition=...) without p connects every pair where it holds."""
    g = dv.GroupDesc("g", 100, equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr", namespace=CUBA_NS)
    syns = [dv.SynapsesDesc("s0", "g", "g", "ge += we", namespace=CUBA_NS, p=0.1, condition="abs(i - j) < 5"),
            dv.SynapsesDesc("s1", "g", "g", "gi += wi", namespace=dict(CUBA_NS, sigma=20.0),
                            source_range=(80, 100), target_range=(10, 60),
                            p="exp(-(i - j)**2 / (2 * sigma**2))", condition="i != j"),
            dv.SynapsesDesc("s2", "g", "g", "ge += we", namespace=CUBA_NS, condition="v_pre > v_post"),
            dv.SynapsesDesc("s3", "g", "g", "ge += we", namespace=CUBA_NS, p="2 * 0.05", condition="i < 50"),
            dv.SynapsesDesc("s4", "g", "g", "ge += we", namespace=CUBA_NS, k=10, condition="j % 3 == 0")]
    spec, _ = dv.network_spec([g], syns, [], dt=0.1 * ms)
    r = [p.random for p in spec.paths]
    assert r[0].cond == "((abs((i - j)) < 5LL))" and r[0].p == 0.1 and r[0].p_expr == ""
    assert r[1].pre == (80, 100) and r[1].post == (10, 60) and r[1].cond == "((i != j))" and r[1].autapses
    assert r[1].p_expr.startswith("exp(") and "(800.0)" in r[1].p_expr and r[1].p == 1.0
    assert r[2].cond == "((hs_pre(0) > hs_post(0)))" and r[2].p == 1.0
    assert r[3].cond == "" and r[3].pre == (0, 50) and r[3].p == pytest.approx(0.1)  # ranges stay ranges
    assert r[4].k == 10 and r[4].cond == "((hs_cg_pymod(j, 3LL) == 0LL))"


def test_generator_syntax():
    """connect(j='...') in Brian2's generator syntax: one-to-one and other
    expressions of i, range() with bounds of i and a step, sample(p=...) with
    if clauses, sample(size=K) over a sub-range (the loop variable counts in the
    Synapses' own j); the synapses the oracle generates for each are the pairs
    Python enumerates from the same generator."""
    import oracle
    n = 60
    g = dv.GroupDesc("g", n, equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr", namespace=CUBA_NS)
    forms = ["i", "(i + 7) % N_post", "k for k in range(i - 3, i + 4) if k != i",
             "k for k in range(N_post - 1, 0, -3) if k > i", "k for k in sample(N_post, p=1.0) if k % 4 == i % 4",
             "k for k in sample(10, 50, size=8) if k != i"]
    syns = [dv.SynapsesDesc(f"s{q}", "g", "g", "ge += we", namespace=CUBA_NS, j=f, seed=3) for q, f in enumerate(forms)]
    spec, _ = dv.network_spec([g], syns, [], dt=0.1 * ms)
    got = [set(zip(*(lambda s: (s["pre"].tolist(), s["post"].tolist()))(oracle.random_pathway(spec, p))))
           for p in spec.paths]
    assert got[0] == {(i, i) for i in range(n)}
    assert got[1] == {(i, (i + 7) % n) for i in range(n)}
    assert got[2] == {(i, k) for i in range(n) for k in range(i - 3, i + 4) if k != i and 0 <= k < n}
    assert got[3] == {(i, k) for i in range(n) for k in range(n - 1, 0, -3) if k > i}
    assert got[4] == {(i, k) for i in range(n) for k in range(n) if k % 4 == i % 4}
    r5 = spec.paths[5].random
    assert r5.k == 8 and r5.post == (10, 50) and r5.cond
    rows = {}
    for i, j in got[5]:
        rows.setdefault(i, []).append(j)
    assert set(rows) == set(range(n)) and all(10 <= j < 50 and j != i for i, js in rows.items() for j in js)
    # 8 sampled per row; the if clause drops j == i, which only a source inside the range can lose
    assert all(len(js) == 8 or (len(js) == 7 and 10 <= i < 50) for i, js in rows.items())
    with pytest.raises(NotImplementedError, match="element"):
        dv.network_spec([g], [dv.SynapsesDesc("s", "g", "g", "ge += we", namespace=CUBA_NS,
                                              j="k + 1 for k in range(3)")], [], dt=0.1 * ms)


def test_state_monitor_indices_and_dt():
    """StateMonitor(S, 'v', record=[...], dt=...) as the reference's consistency
    tests use it (test_cpp_cuda_consistency.py:72): arbitrary indices and a
    monitor dt become an indexed monitor; a contiguous range every step stays a
    range; repeated indices and a dt that is not a multiple are refused."""
    g = dv.GroupDesc("g", 10, equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr", namespace=CUBA_NS)
    ms_ = [dv.MonitorDesc("m0", "state", "g", variable="v", record=[0, 4, 2]),
           dv.MonitorDesc("m1", "state", "g", variable="ge", record=range(3), dt=1 * ms),
           dv.MonitorDesc("m2", "state", "g", variable="gi", record=range(2, 5))]
    spec, _ = dv.network_spec([g], [], ms_, dt=0.1 * ms)
    a, b, c = spec.state_monitors
    assert isinstance(a, nw.StateMon) and list(a.indices) == [0, 4, 2] and a.every == 1 and a.var == 0
    assert isinstance(b, nw.StateMon) and list(b.indices) == [0, 1, 2] and b.every == 10 and b.var == 1
    assert c == (0, 2, 2, 5)
    for bad, msg in [(dict(record=[1, 1]), "distinct"), (dict(record=[1], dt=0.25 * ms), "multiple")]:
        m = dv.MonitorDesc("m", "state", "g", variable="v", **bad)
        with pytest.raises(NotImplementedError, match=msg):
            dv.network_spec([g], [], [m], dt=0.1 * ms)


def test_device_module_imports_with_and_without_brian2():
    for code in ["import sys; sys.modules['brian2'] = None\n",
                 "import sys, types; sys.modules['brian2'] = types.ModuleType('brian2')\n"]:
        r = subprocess.run([sys.executable, "-c", code + "import brian2cuda_amd.device as d\n"
                            "print(hasattr(d, 'HipStandaloneDevice'))"], cwd=ROOT, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0 and r.stdout.strip() == "False", r.stderr


@pytest.mark.parametrize("change", [dict(threshold="v > Vt and v < 0*mV"), dict(method="euler"),
                                    dict(reset="v = Vr + 1*mV"),
                                    dict(equations=CUBA_EQS.replace("-ge/taue", "-ge/taue + 0*v"))])
def test_other_models_become_code_models(change):
    """What is not a built-in model (another threshold, reset, method or equation
    set) maps onto a code model through the generator."""
    kw = dict(equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr", refractory=5 * ms, namespace=CUBA_NS)
    kw.update(change)
    spec, names = dv.group_spec(dv.GroupDesc("g", 10, init={"v": -60 * mV}, **kw), dt=0.1 * ms)
    assert spec.code is not None and names == ["v", "ge", "gi"] and spec.init == {0: -60 * mV}
    assert spec.code.refractory == 5 * ms and spec.code.reset


def test_adex_example_maps_to_a_code_model():
    """An adaptive exponential integrate-and-fire population (Brian2's AdEx form)
    with a two-variable reset: a code model, rk4 as asked."""
    eqs = """
    dvm/dt = (gL*(EL - vm) + gL*DeltaT*exp((vm - VT)/DeltaT) + I - w)/C : volt
    dw/dt = (a*(vm - EL) - w)/tauw : amp
    I : amp
    """
    ns = dict(gL=30 * nS, EL=-70.6 * mV, DeltaT=2 * mV, VT=-50.4 * mV, C=281e-12, a=4 * nS, tauw=144 * ms,
              b=0.0805e-9, Vr=-70.6 * mV)
    g = dv.GroupDesc("adex", 100, equations=eqs, threshold="vm > -40*mV", reset="vm = Vr; w += b", method="rk4",
                     namespace=ns, init={"I": np.linspace(0, 1e-9, 100)})
    spec, _ = dv.network_spec([g], [], [dv.MonitorDesc("m", "spike", "adex")], dt=0.1 * ms)
    c = spec.groups[0].code
    assert c.names == ["vm", "w", "I"] and "_hs_k4_vm" in c.update and c.reset.count(";") == 2
    np.testing.assert_array_equal(spec.groups[0].init[2], np.linspace(0, 1e-9, 100))


def test_synaptic_code_from_the_description():
    """The reference's sorted-indices network (test_cpp_cuda_consistency.py:177-205)
    described: NeuronGroup(5, '', threshold='True'), Synapses(P, P, 'w : 1',
    on_pre='w += 1'), connect(j='i'), delays k*dt -> a code model and a code
    pathway; an on_pre that reads a source variable and adds to a target one."""
    g = dv.GroupDesc("neurongroup", 5, equations="", threshold="True")
    s = dv.SynapsesDesc("synapses", "neurongroup", "neurongroup", "w += 1", model="w : 1", i=np.arange(5),
                        j=np.arange(5), delay=np.arange(5) * 0.1 * ms)
    spec, _ = dv.network_spec([g], [s], [], dt=0.1 * ms)
    p = spec.paths[0]
    assert spec.groups[0].code.threshold == "true" and p.code.on_pre == "w += ((R)1);"
    np.testing.assert_array_equal(p.delay_steps, np.arange(5))
    lif = dv.GroupDesc("lif", 10, equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr", namespace=CUBA_NS)
    s2 = dv.SynapsesDesc("s2", "lif", "lif", "ge_post += w*(1 + 0*v_pre)\nc += 1", model="w : volt\nc : 1",
                         i=[0, 1], j=[2, 3], values={"w": np.array([1e-3, 2e-3])})
    spec2, _ = dv.network_spec([lif], [s2], [], dt=0.1 * ms)
    q = spec2.paths[0]
    assert q.code.syn_names == ["w", "c"] and q.code.post_names == ["ge"] and q.code.pre_names == ["v"]
    assert q.wmode == 1 and list(q.weights) == [1e-3, 2e-3]


def test_stdp_example_synapses_map_to_synaptic_code():
    """examples/stdp.py's Synapses (event-driven Apre/Apost, on_pre and on_post)
    described: a code pathway whose statements advance the traces first."""
    g = dv.GroupDesc("neurons", 10, equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr", refractory=5 * ms,
                     namespace=CUBA_NS)
    ns = dict(taupre=20 * ms, taupost=20 * ms, dApre=0.01, dApost=-0.0105, gmax=0.01)
    s = dv.SynapsesDesc("synapses", "neurons", "neurons", "ge += w\nApre += dApre\nw = clip(w + Apost, 0, gmax)",
                        model=nw.STDP_SYNAPSE_MODEL, on_post="Apost += dApost\nw = clip(w + Apre, 0, gmax)",
                        namespace=ns, i=np.arange(5), j=np.arange(5)[::-1], values={"w": np.full(5, 0.005)})
    spec, _ = dv.network_spec([g], [s], [], dt=0.1 * ms)
    c = spec.paths[0].code
    assert c.syn_names == ["w", "Apre", "Apost"] and c.post_names == ["ge"]
    assert c.on_pre.startswith("const R _hs_el = (R)(t - lastupdate);") and "exp(" in c.on_post
    assert spec.paths[0].wmode == 1
