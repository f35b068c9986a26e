"""GPU: neuron groups whose model is code (hs_net_add_group_code) -- Brian2
equations through the code generator (brian2cuda_amd.codegen), compiled with
hipRTC into the engine's kernels -- bit for bit against the oracle running the
same statements compiled for the host, in every engine form."""
import numpy as np
import pytest

from brian2cuda_amd import codegen as cg
from brian2cuda_amd import networks as nw
from brian2cuda_amd import runtime as rt
from brian2cuda_amd._lib import HsError

pytestmark = pytest.mark.gpu

FORMS = [("plain", 0, {}), ("fused", 1, {}), ("fused", 1, {9: 1})]


def _compare(spec, prec, engine, sem, opts, oracle_lib, steps):
    net, gsm, gst = nw.build_hs(spec, prec, engine, options=opts)
    ref, rsm, rst = oracle_lib.build_ref(spec, prec, sem)
    for k in steps:
        net.run(k)
        ref.run(k)
    for a, b in zip(gsm, rsm):
        gi, gt = net.spikes(a)
        ri, rt_ = ref.spikes(b)
        assert len(gi) == len(ri) > 0
        assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    for gid, g in enumerate(spec.groups):
        nv = len(g.code.names) if g.code is not None else 1
        for v in range(nv):
            assert np.array_equal(net.get_state(gid, v), ref.get_state(gid, v), equal_nan=True), (gid, v)
    assert net.stats()["events"] == ref.events
    for a, b, m in zip(gst, rst, spec.state_monitors):
        w = nw.state_monitor_width(m)
        assert np.array_equal(net.state_monitor(a, w), ref.state_monitor(b, w))
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem,opts", FORMS)
def test_code_cuba_vs_oracle(hs, oracle_lib, prec, engine, sem, opts):
    spec = nw.code_cuba(N=4000, seed=3)
    spec.state_monitors = [(0, 0, 0, 50), nw.StateMon(0, 1, np.array([7, 3999, 100]), every=3)]
    _compare(spec, prec, engine, sem, opts, oracle_lib, (1, 333, 1166))


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem,opts", FORMS)
def test_code_mix_vs_oracle(hs, oracle_lib, prec, engine, sem, opts):
    """AdEx (rk4, reset of two variables), noise-driven LIF (Euler-Maruyama) and
    an exact CUBA population, plus a built-in Poisson input, delays 1..30 steps."""
    spec = nw.code_mix(N=3000, seed=7)
    spec.state_monitors = [nw.StateMon(0, 1, np.array([0, 5, 1499]), every=1)]
    _compare(spec, prec, engine, sem, opts, oracle_lib, (250, 1250))


@pytest.mark.parametrize("method", ["euler", "rk2"])
def test_code_methods_vs_oracle(hs, oracle_lib, method):
    spec = nw.code_mix(N=2000, seed=11, method=method)
    _compare(spec, 64, "fused", 1, {9: 1}, oracle_lib, (700,))


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem,opts", FORMS[:2])
def test_code_cobahh_vs_oracle(hs, oracle_lib, prec, engine, sem, opts):
    """examples/cobahh.py's equations, exponential Euler, through the code
    generator (delay 0: update then delivery in each step)."""
    spec = nw.code_cobahh(N=4000, seed=2)
    _compare(spec, prec, engine, sem, opts, oracle_lib, (200, 300))


def test_code_model_errors(hs):
    net = rt.Net(64, "plain", 1e-4, 1)
    bad = cg.CodeModel(["v"], "v = v + ;", "v > 1", "")
    net.add_group_code(10, bad)
    with pytest.raises(HsError, match="do not compile"):
        net.run(1)
    net.close()
    net = rt.Net(64, "plain", 1e-4, 1)
    for names, upd in [(["dt"], ""), (["v", "v"], ""), (["hs_x"], ""), (["v"], "#define X 1"), (["v"], "{"),
                       ([f"x{k}" for k in range(9)], ""), (["v"], "v = hs_s[7];"), (["v"], "v = __builtin_inf();")]:
        with pytest.raises(HsError):
            net.add_group_code(10, cg.CodeModel(names, upd))
    for k in range(4):
        net.add_group_code(10, cg.CodeModel(["v"], "v = v + dt;"))
    with pytest.raises(HsError):
        net.add_group_code(10, cg.CodeModel(["v"], "v = v + dt;"))
    net.close()


def test_code_cuba_matches_the_builtin_model(hs, oracle_lib):
    """The code form of the CUBA model and the built-in one integrate the same
    linear system exactly: over 300 subthreshold steps (no spikes) their states
    agree to FP64 rounding."""
    a = nw.code_cuba(N=400, seed=5, we=0.0, wi=0.0)
    b = nw.cuba(N=400, seed=5, we=0.0, wi=0.0, record=0)
    for s in (a, b):
        s.groups[0].init[0] = -55e-3 + 4e-3 * np.random.default_rng(1).random(400)
        s.groups[0].init[1] = 2e-3
        s.groups[0].init[2] = -1e-3
    na, _, _ = nw.build_hs(a, 64, "plain")
    nb, _, _ = nw.build_hs(b, 64, "plain")
    na.run(300)
    nb.run(300)
    for v in range(3):
        np.testing.assert_allclose(na.get_state(0, v), nb.get_state(0, v), rtol=0, atol=1e-12)
    na.close()
    nb.close()


# ---- synaptic code --------------------------------------------------------------------------
def _sorted_indices_spec(n=5):
    m, names = cg.neuron_model("", "True", None, 0.0, None, {})
    code = cg.synapse_code("w += 1", cg.synapse_variables("w : 1"), names, names)
    i = np.arange(n, dtype=np.int32)
    return nw.NetSpec("sorted_indices", 1e-4, 1, [nw.GroupSpec(0, n, [], {}, code=m)],
                      [nw.PathSpec(0, 0, 0, i, i, delay_steps=np.arange(n, dtype=np.int32), code=code)])


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_sorted_indices_statemonitor(hs, prec, engine):
    """The reference's test_sorted_indices_statemonitor (test_cpp_cuda_consistency.py:177-205):
    NeuronGroup(5, model='', threshold='True'), on_pre 'w += 1', connect(j='i'),
    delay k*dt; StateMonitor(S, 'w') at the start of step s holds max(0, s - k)."""
    n = 5
    net, _, _ = nw.build_hs(_sorted_indices_spec(n), prec, engine)
    mid = net.add_synapse_monitor(0, rt.HS_SYN_CODE, np.arange(n))
    net.run(n + 1)
    s, k = np.meshgrid(np.arange(n + 1), np.arange(n), indexing="ij")
    np.testing.assert_array_equal(net.state_monitor(mid, n), np.maximum(0, s - k))
    np.testing.assert_array_equal(net.synapse_var(0, rt.HS_SYN_CODE), n + 1 - np.arange(n))
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_additive_code_equals_the_builtin_effect(hs, prec, engine):
    """C1 with 'ge += we' / 'gi += wi' as synaptic code: the built-in pathways' bits."""
    a = nw.cuba(N=4000, seed=4, record=0)
    b = nw.cuba(N=4000, seed=4, record=0)
    for p, var in zip(b.paths, ("ge", "gi")):
        p.code = cg.synapse_code(f"{var} += {p.wscalar!r}", [], ["v", "ge", "gi"], ["v", "ge", "gi"])
    na, sa, _ = nw.build_hs(a, prec, engine)
    nb, sb, _ = nw.build_hs(b, prec, engine)
    na.run(800)
    nb.run(800)
    assert np.array_equal(na.spikes(sa[0])[0], nb.spikes(sb[0])[0])
    assert na.stats()["events"] == nb.stats()["events"] > 0
    for v in range(3):
        assert np.array_equal(na.get_state(0, v), nb.get_state(0, v))
    na.close()
    nb.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
def test_synaptic_code_vs_oracle(hs, oracle_lib, prec, engine, sem):
    spec = nw.code_synapses()
    p = spec.paths[0]
    net, gsm, _ = nw.build_hs(spec, prec, engine)
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, sem)
    mid = net.add_synapse_monitor(0, rt.HS_SYN_CODE, np.arange(20))
    order = nw.bundle_order(p)
    rmid = ref.add_synapse_monitor(0, rt.HS_SYN_CODE, order[:20])
    for k in (300, 900):
        net.run(k)
        ref.run(k)
    gi, gt = net.spikes(gsm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(gi) == len(ri) > 0 and np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert net.stats()["events"] == ref.events
    for k in range(2):
        g = net.synapse_var(0, rt.HS_SYN_CODE + k)
        r = ref.synapse_var(0, rt.HS_SYN_CODE + k, len(p.pre))[order]
        assert np.array_equal(g, r), k
    assert np.array_equal(net.state_monitor(mid, 20), ref.state_monitor(rmid, 20))
    for v in range(2):
        assert np.array_equal(net.get_state(1, v), ref.get_state(1, v))
    net.close()


def test_synaptic_code_errors(hs):
    net = rt.Net(64, "plain", 1e-4, 1)
    g = net.add_group(rt.HS_MODEL_LIF_CUBA, 10, nw.cuba_params())
    pid = net.add_pathway(g, g, 1, np.arange(10), np.arange(10)[::-1], scalar_delay=1)
    for args in [(["dt"], "", [], []), (["w", "w"], "", [], []), (["w"], "{", [], []),
                 (["w"], "", [(7, "v")], []), (["w"], "", [], [(-1, "v")])]:
        names, code, post, pre = args
        c = cg.SynapseCode(names, code, [n for _, n in post], [n for _, n in pre])
        with pytest.raises(HsError):
            net.pathway_code(pid, c, [k for k, _ in post], [k for k, _ in pre])
    net.pathway_code(pid, cg.SynapseCode(["w"], "w += ;", [], []), [], [])
    with pytest.raises(HsError, match="do not compile"):
        net.run(1)
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
@pytest.mark.parametrize("delay", [0, 3])
def test_stdp_as_code_vs_oracle(hs, oracle_lib, prec, engine, sem, delay):
    """The STDP synapse model as Brian2 code (event-driven traces, on_pre and
    on_post through the generator) on the feed-forward network: spikes,
    weights, traces, last updates, voltages and events bit for bit."""
    spec = nw.stdp_feedforward(delay, code=True)
    S = spec.paths[0]
    perm = nw.bundle_order(S)
    net, sm, st = nw.build_hs(spec, prec, engine)
    ref, rsm, rst = oracle_lib.build_ref(spec, prec, sem)
    for k in (700, 1100, 1300):
        net.run(k)
        ref.run(k)
        for var in [rt.HS_SYN_CODE + k for k in range(3)] + [rt.HS_SYN_LASTUPDATE]:
            assert np.array_equal(net.synapse_var(0, var), ref.synapse_var(0, var, len(S.pre))[perm]), var
    gi, gt = net.spikes(sm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(gi) > 50 and np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert np.array_equal(net.state_monitor(st[0], 20), ref.state_monitor(rst[0], 20))
    assert net.stats()["events"] == ref.events
    net.close()


def test_stdp_as_code_matches_the_builtin_rule_gpu(hs):
    """On the GPU, the code form and the built-in STDP agree: same spikes,
    weights and traces to FP64 rounding."""
    a, b = nw.stdp_feedforward(3), nw.stdp_feedforward(3, code=True)
    na, sa, _ = nw.build_hs(a, 64, "plain")
    nb, sb, _ = nw.build_hs(b, 64, "plain")
    na.run(3000)
    nb.run(3000)
    assert np.array_equal(na.spikes(sa[0])[0], nb.spikes(sb[0])[0])
    for k, var in enumerate((rt.HS_SYN_W, rt.HS_SYN_APRE, rt.HS_SYN_APOST)):
        np.testing.assert_allclose(nb.synapse_var(0, rt.HS_SYN_CODE + k), na.synapse_var(0, var), rtol=1e-12,
                                   atol=1e-15)
    na.close()
    nb.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("opts", [{}, {9: 1}, {4: 3}])
def test_code_brunel_vs_oracle(hs, oracle_lib, prec, opts):
    """C3's noise-driven LIF as code (Euler-Maruyama through the generator), one
    code model in every group (kernels specialised for it), in the fused,
    pipelined and windowed forms (the windowed forms draw the noise one window
    ahead into a ring, as the built-in model does)."""
    spec = nw.code_brunel(N=4000, K=400, seed=8, monitors=True)
    net, gsm, _ = nw.build_hs(spec, prec, "fused", options=opts)
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, 1)
    for k in (600, 900):
        net.run(k)
        ref.run(k)
    gi, gt = net.spikes(gsm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(gi) == len(ri) > 0 and np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert np.array_equal(net.get_state(0, 0), ref.get_state(0, 0))
    assert net.stats()["events"] == ref.events
    net.close()


def test_code_models_with_gather_delivery_and_builtin_stdp(hs, oracle_lib):
    """Code models beside the other engine parts: C1 as code with gather-mode
    delivery (plain form), and a code-model target under the built-in STDP rule."""
    spec = nw.code_cuba(N=4000, seed=12)
    _compare(spec, 64, "plain", 0, {5: 2}, oracle_lib, (700,))
    ff = nw.stdp_feedforward(2)
    lif, names = cg.neuron_model(nw.CUBA_EQS.replace(" (unless refractory)", ""), "v > Vt", "v = Vr", 5e-3, None,
                                 dict(El=-49e-3, taum=20e-3, taue=5e-3, taui=5e-3, Vt=-50e-3, Vr=-60e-3), dt=1e-4)
    g = ff.groups[1]
    ff.groups[1] = nw.GroupSpec(0, g.n, [], g.init, code=lif)
    net, sm, _ = nw.build_hs(ff, 64, "plain")
    ref, rsm, _ = oracle_lib.build_ref(ff, 64, 0)
    net.run(2500)
    ref.run(2500)
    assert np.array_equal(net.spikes(sm[0])[0], ref.spikes(rsm[0])[0]) and len(ref.spikes(rsm[0])[0]) > 20
    perm = nw.bundle_order(ff.paths[0])
    assert np.array_equal(net.synapse_var(0, rt.HS_SYN_W), ref.synapse_var(0, rt.HS_SYN_W, len(perm))[perm])
    net.close()


def test_synaptic_code_on_device_connectivity(hs, oracle_lib):
    """Synaptic code on a pathway generated on the device (connect(p) with drawn
    per-synapse weights and delays): 'V_post += w', with w the pathway's weight
    array, gives the built-in pathway's bits; a second variable counts arrivals."""
    a = nw.brunel(N=4000, K=400, seed=5, monitors=True, device_connect=True)
    b = nw.brunel(N=4000, K=400, seed=5, monitors=True, device_connect=True)
    b.paths[0].code = cg.synapse_code("V_post += w\nn += 1", ["w", "n"], ["V"], ["V"])
    for prec, engine, sem in ((64, "plain", 0), (32, "fused", 1)):
        na, sa, _ = nw.build_hs(a, prec, engine)
        nb, sb, _ = nw.build_hs(b, prec, engine)
        rb, rsb, _ = oracle_lib.build_ref(b, prec, sem)
        for net in (na, nb, rb):
            net.run(1200)
        assert np.array_equal(na.spikes(sa[0])[0], nb.spikes(sb[0])[0])
        assert np.array_equal(nb.spikes(sb[0])[0], rb.spikes(rsb[0])[0])
        assert np.array_equal(na.get_state(0, 0), nb.get_state(0, 0))
        n = nb.synapse_var(0, rt.HS_SYN_CODE + 1)
        assert n.sum() == nb.stats()["events"] == rb.events
        na.close()
        nb.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem,opts", [("plain", 0, {}), ("fused", 1, {9: 1})])
def test_trig_and_hyperbolic_functions_vs_oracle(hs, oracle_lib, prec, engine, sem, opts):
    """A code model calling every function added to exp / log / sqrt (sin, cos,
    tan, sinh, cosh, tanh, arcsin, arccos, arctan, expm1, log1p, exprel):
    the device and the oracle evaluate them with the same operations, so
    spikes and state agree bit for bit."""
    eqs = """
    dv/dt = (I*(1 + sin(2*pi*f*t + phase)) + 3*mV*tanh(v/(10*mV)) + mV*arctan(v/mV) - v)/tau : volt
    dw/dt = (0.1*(cosh(0.5*tanh(w)) - 1) + 0.1*expm1(-w*w) + 0.05*log1p(w*w) + 0.1*(exprel(-w) - 1)
             + 0.2*arcsin(0.5*sin(w)) + 0.1*(arccos(0.5*cos(w)) - 1) + 0.1*tan(0.1*tanh(w))
             + 0.1*sinh(0.2*tanh(w)) - w)/tau_w : 1
    phase : 1 (constant)
    """
    ns = {"I": 12e-3, "f": 40.0, "tau": 10e-3, "tau_w": 30e-3, "pi": np.pi}
    model, names = cg.neuron_model(eqs, "v > 10*mV", "v = 0*mV; w = w + 0.1", 2e-3, "euler", ns, dt=1e-4)
    n = 2000
    rng = np.random.default_rng(5)
    init = {names.index("v"): 10e-3 * rng.random(n), names.index("w"): rng.random(n),
            names.index("phase"): 2 * np.pi * rng.random(n)}
    spec = nw.NetSpec("trig", 1e-4, 3, [nw.GroupSpec(0, n, [], init, code=model)], [])
    spec.spike_monitors = [0]
    _compare(spec, prec, engine, sem, opts, oracle_lib, (150, 350))


def _rnd_stream(stream, k, normal):
    return (stream & 0xFFFFFF) | ((1 + k + 16 * normal) << 24)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem,opts", FORMS)
def test_rand_randn_code_vs_oracle(hs, oracle_lib, prec, engine, sem, opts):
    """rand() and randn() in a code model's update (euler), threshold and reset:
    each call site a Philox draw keyed by seed, stream, call site, neuron and
    step, the same on the device and in the oracle: spikes and state bit for bit."""
    eqs = "dv/dt = (I + 2*mV*randn() - v)/tau : volt"
    model, names = cg.neuron_model(eqs, "v > 10*mV and rand() < 0.8", "v = rand()*2*mV", 2e-3, "euler",
                                   {"I": 12e-3, "tau": 10e-3}, dt=1e-4, stream=4)
    n = 3000
    init = {0: 10e-3 * np.random.default_rng(8).random(n)}
    spec = nw.NetSpec("rand_code", 1e-4, 21, [nw.GroupSpec(0, n, [], init, code=model)], [])
    spec.spike_monitors = [0]
    spec.state_monitors = [(0, 0, 0, 50)]
    _compare(spec, prec, engine, sem, opts, oracle_lib, (200, 400))


@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_rand_threshold_equals_builtin_poisson_gpu(hs, engine):
    """threshold='rand() < rates*dt' equals the built-in PoissonGroup drawing
    from the call site's stream: identical spikes on the device (FP64)."""
    dt = 1e-4
    rates = 5.0 + 80.0 * np.random.default_rng(2).random(400)
    m, _ = cg.neuron_model("rates : Hz", "rand() < rates*dt", None, 0.0, None, {}, dt=dt, stream=3)
    code = nw.NetSpec("poisson_code", dt, 11, [nw.GroupSpec(0, 400, [], {0: rates}, code=m)], [])
    builtin = nw.NetSpec("poisson", dt, 11, [nw.GroupSpec(nw.HS_MODEL_POISSON, 400, [0.0, _rnd_stream(3, 0, 0)],
                                                          {0: rates})], [])
    out = []
    for spec in (code, builtin):
        spec.spike_monitors = [0]
        net, sm, _ = nw.build_hs(spec, 64, engine)
        net.run(2000)
        out.append(net.spikes(sm[0]))
        net.close()
    assert len(out[0][0]) > 1000
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_rand_reference_test_rand_gpu(hs):
    """brian2cuda's test_rand and test_random_number_generation_with_multiple_runs
    (tests/test_random_number_generation.py:194-224) on the device."""
    m, _ = cg.neuron_model("dv/dt = rand() : second", None, None, 0.0, None, {}, dt=1e-4)
    spec = nw.NetSpec("test_rand", 1e-4, 1, [nw.GroupSpec(0, 1000, [], {}, code=m)], [])
    spec.state_monitors = [(0, 0, 0, 1000)]
    net, _, st = nw.build_hs(spec, 64, "plain")
    net.run(1)
    net.run(2)
    v = net.state_monitor(st[0], 1000)
    assert v.shape == (3, 1000) and not np.all(v[-1] == 0) and not np.array_equal(v[-2], v[-1])
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem,opts", FORMS)
@pytest.mark.parametrize("method", ["heun", "milstein"])
def test_stochastic_methods_vs_oracle(hs, oracle_lib, prec, engine, sem, opts, method):
    """heun (two noise variables, one multiplicative) and milstein
    (multiplicative noise) on a spiking LIF: the engine's xi, the second
    variable's own draw, the support values and the corrections give the
    oracle's bits in every form and precision."""
    if method == "heun":
        eqs = ("dv/dt = (I - v)/tau + s1*(1 + v/(20*mV))*xi_1*mV/sqrt(ms) + s2*xi_2*mV/sqrt(ms) : volt")
    else:
        eqs = "dv/dt = (I - v)/tau + s1*(1 + v/(20*mV))*xi*mV/sqrt(ms) : volt"
    model, names = cg.neuron_model(eqs, "v > 10*mV", "v = 0*mV", 2e-3, method,
                                   {"I": 11e-3, "tau": 10e-3, "s1": 0.5, "s2": 0.3}, dt=1e-4, stream=6)
    assert model.noise
    n = 2500
    init = {0: 10e-3 * np.random.default_rng(12).random(n)}
    spec = nw.NetSpec(f"sde_{method}", 1e-4, 31, [nw.GroupSpec(0, n, [], init, code=model)], [])
    spec.spike_monitors = [0]
    spec.state_monitors = [(0, 0, 0, 40)]
    _compare(spec, prec, engine, sem, opts, oracle_lib, (150, 350))


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
def test_stochastic_release_vs_oracle(hs, oracle_lib, prec, engine, sem):
    """rand() and randn() in synaptic code (networks.code_synapses(release=True)):
    'released = rand() < 0.5' gates each arrival's 1.5 mV, w takes a randn() step,
    on_post draws too; each call site a Philox draw keyed by seed, pathway, i, j
    and step, the same on the device and in the oracle: spikes, every synapse
    variable and the target state bit for bit."""
    spec = nw.code_synapses(release=True)
    p = spec.paths[0]
    print(f"stochastic release prec {prec} engine {engine}", flush=True)
    net, gsm, _ = nw.build_hs(spec, prec, engine)
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, sem)
    for k in (300, 900):
        net.run(k)
        ref.run(k)
    gi, gt = net.spikes(gsm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(gi) == len(ri) > 0 and np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert net.stats()["events"] == ref.events
    order = nw.bundle_order(p)
    for k in range(4):
        g = net.synapse_var(0, rt.HS_SYN_CODE + k)
        r = ref.synapse_var(0, rt.HS_SYN_CODE + k, len(p.pre))[order]
        assert np.array_equal(g, r), k
    for v in range(2):
        assert np.array_equal(net.get_state(1, v), ref.get_state(1, v))
    net.close()
