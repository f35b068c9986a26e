"""CPU tests of the code generator (brian2cuda_amd.codegen: Brian2 equations ->
the statements of hs_net_add_group_code) run through the oracle's code models
(the same statements compiled for the host), pinned against what the
mathematics fixes:

  - each integration method on dv/dt = (a - v)/tau is the recurrence
    v <- a + (v - a) q with q = 1 - h (euler), 1 - h + h^2/2 (rk2),
    1 - h + h^2/2 - h^3/6 + h^4/24 (rk4), exp(-h) (exponential_euler, exact);
  - 'exact' on the coupled CUBA system equals repeated products with the
    matrix exponential, and the built-in CUBA model (its own closed form);
  - Euler-Maruyama equals a numpy recurrence over the oracle's normal draws;
  - threshold, reset (two variables) and refractoriness with and without
    '(unless refractory)' give the spike steps derived by hand;
  - what the generator cannot express raises NotImplementedError.
"""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle
from brian2cuda_amd import codegen as cg
from brian2cuda_amd import networks as nw

DT = 1e-4


def _run(model, init, steps, prec=64, sem=oracle.SEQ, n=None, monitor=False):
    n = n or len(next(iter(init.values())))
    spec = nw.NetSpec("t", DT, 5, [nw.GroupSpec(0, n, [], init, code=model)], [])
    if monitor:
        spec.spike_monitors = [0]
    ref, sm, _ = oracle.build_ref(spec, prec, sem)
    ref.run(steps)
    return ref, sm


@pytest.mark.parametrize("method,q", [
    ("euler", lambda h: 1 - h), ("rk2", lambda h: 1 - h + h * h / 2),
    ("rk4", lambda h: 1 - h + h ** 2 / 2 - h ** 3 / 6 + h ** 4 / 24),
    ("exponential_euler", lambda h: np.exp(-h)), ("exact", lambda h: np.exp(-h))])
def test_methods_follow_their_recurrence(method, q):
    a, tau = 0.7, 7e-3
    m, names = cg.neuron_model("dv/dt = (a - v)/tau : 1", None, None, 0.0, method, dict(a=a, tau=tau), dt=DT)
    v0 = np.array([-1.0, 0.0, 0.5, 2.0])
    ref, _ = _run(m, {0: v0}, 500)
    expect = a + (v0 - a) * q(DT / tau) ** 500
    np.testing.assert_allclose(ref.get_state(0, 0), expect, rtol=1e-11, atol=1e-14)


def test_exact_coupled_system_and_the_builtin_cuba():
    ns = dict(El=-49e-3, taum=20e-3, taue=5e-3, taui=10e-3, Vt=1.0, Vr=-60e-3)
    m, names = cg.neuron_model(nw.CUBA_EQS, "v > Vt", "v = Vr", 0.0, None, ns, dt=DT)
    assert names == ["v", "ge", "gi"]
    x0 = np.array([[-55e-3, -52e-3], [2e-3, 0.0], [-1e-3, -3e-3]])
    ref, _ = _run(m, {0: x0[0], 1: x0[1], 2: x0[2]}, 400)
    # dv/dt = (ge + gi - v + El)/taum, dge/dt = -ge/taue, dgi/dt = -gi/taui, augmented by the constant
    M = np.array([[-1 / 20e-3, 1 / 20e-3, 1 / 20e-3, -49e-3 / 20e-3],
                  [0, -1 / 5e-3, 0, 0], [0, 0, -1 / 10e-3, 0], [0, 0, 0, 0]])
    P = np.linalg.matrix_power(expm(M * DT), 400)
    X = P @ np.vstack([x0, np.ones((1, 2))])
    for v in range(3):
        np.testing.assert_allclose(ref.get_state(0, v), X[v], rtol=1e-10, atol=1e-15)
    # the built-in CUBA model (a closed form of its own) on the same states
    b = nw.cuba(N=2, seed=1, we=0.0, wi=0.0, record=0, monitors=False)
    b.groups[0].params[1] = 1.0
    b.groups[0].init = {0: x0[0], 1: x0[1], 2: x0[2]}
    rb, _, _ = oracle.build_ref(b, 64, oracle.SEQ)
    rb.run(400)
    for v in range(3):
        np.testing.assert_allclose(ref.get_state(0, v), rb.get_state(0, v), rtol=0, atol=1e-13)


def test_euler_maruyama_matches_the_draws():
    mu, sigma, tau = 0.02, 0.004, 0.02
    m, _ = cg.neuron_model("dV/dt = (-V + mu + sigma*sqrt(tau)*xi)/tau : volt", None, None, 0.0, None,
                           dict(mu=mu, sigma=sigma, tau=tau), stream=4)
    assert m.noise and m.stream == 4
    n, steps = 3, 200
    ref, _ = _run(m, {0: np.zeros(n)}, steps)
    v = np.zeros(n)
    for s in range(steps):
        xi = np.array([oracle.randn(5, 4, i, s) for i in range(n)])
        v = v + DT * (mu - v) / tau + sigma * np.sqrt(tau) / tau * np.sqrt(DT) * xi
    np.testing.assert_allclose(ref.get_state(0, 0), v, rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("unless,period,first", [(False, 50, 40), (True, 90, 40)])
def test_threshold_reset_and_refractoriness(unless, period, first):
    """dv/dt = (I - v)/tau, euler, I = 3, h = dt/tau = 0.01: from v = 0 the
    threshold 1 is crossed after 41 updates (first spike at step 40).  With a
    5 ms (50-step) refractory period v keeps integrating and the next spike
    waits for step 90 (period 50); held by '(unless refractory)' it restarts
    from 0 at step 90 and spikes 41 updates later (period 90).  The reset also
    counts spikes in a second variable."""
    eqs = "dv/dt = (I - v)/tau : 1" + (" (unless refractory)" if unless else "") + "\nc : 1\nI : 1"
    m, names = cg.neuron_model(eqs, "v > 1", "v = 0\nc += 1", 5e-3, "euler", dict(tau=10e-3))
    assert names == ["v", "c", "I"]
    ref, sm = _run(m, {0: np.zeros(1), 2: np.full(1, 3.0)}, 1000, monitor=True)
    i, t = ref.spikes(sm[0])
    expect = np.arange(first, 1000, period)
    np.testing.assert_array_equal(t, expect)
    assert ref.get_state(0, 1)[0] == expect.size


@pytest.mark.parametrize("prec,sem", [(64, oracle.SEQ), (32, oracle.ACC)])
def test_code_mix_runs_in_both_semantics(prec, sem):
    spec = nw.code_mix(N=1200, seed=2)
    ref, sm, _ = oracle.build_ref(spec, prec, sem)
    ref.run(600)
    assert all(len(ref.spikes(s)[0]) > 0 for s in sm)
    assert np.all(np.isfinite(ref.get_state(0, 0)))


@pytest.mark.parametrize("eqs,kw,msg", [
    ("dv/dt = erf(v)/tau : 1", {}, "function"),
    ("dv/dt = v*(xi_1 + xi_2)/tau : 1", {"method": "milstein"}, "one noise variable"),
    ("dv/dt = v*xi/tau : 1", {"method": "euler"}, "additive"),
    ("dv/dt = v*v/tau : 1", {"method": "exponential_euler"}, "not linear"),
    ("dv/dt = -v/tau + xi/tau : 1", {"method": "rk4"}, "no noise"),
    ("dv/dt = -v/foo : 1", {}, "no value"),
    ("dv/dt = -v/tau : 1 (event-driven)", {}, "belongs to Synapses"),
    ("dv/dt = -v/tau : 1 (clock-driven)", {}, "flag"),
    ("dv/dt = -v/tau : 1", {"method": "gsl"}, "not supported"),
    ("dv/dt = v*t/tau : 1", {"method": "exact"}, "linear system"),
])
def test_unsupported_forms_are_refused(eqs, kw, msg):
    with pytest.raises(NotImplementedError, match=msg):
        cg.neuron_model(eqs, None, None, 0.0, kw.get("method"), dict(tau=0.01), dt=DT)


def test_units_subexpressions_and_printing():
    m, names = cg.neuron_model("""
        dv/dt = (I_syn - v)/(10*ms) : volt
        I_syn = g*(E - v) : volt
        g : 1
    """, "v > -40*mV and v < 1*volt", "v = -65*mV", 2 * 1e-3, "euler", dict(E=0.0))
    assert names == ["v", "g"]
    # I_syn substituted (E = 0 drops out), 10*ms folded: -v/tau - g*v/tau with 1/tau = 100
    assert "I_syn" not in m.update and "(((R)-100.0)*g*v)" in m.update
    assert "&&" in m.threshold and "((R)-0.04)" in m.threshold
    assert m.reset == "v = ((R)-0.065);"


def test_code_cobahh_matches_the_builtin_model():
    """examples/cobahh.py's equations through the generator (exponential Euler)
    against the built-in HH model on the same network: the same 3,949 spikes
    over 300 steps and states within 1e-11 of each other (FP64)."""
    a = nw.code_cobahh(N=4000, seed=2)
    b = nw.cobahh(N=4000, seed=2)
    ra, sa, _ = oracle.build_ref(a, 64, oracle.SEQ)
    rb, sb, _ = oracle.build_ref(b, 64, oracle.SEQ)
    ra.run(300)
    rb.run(300)
    (ia, ta), (ib, tb) = ra.spikes(sa[0]), rb.spikes(sb[0])
    assert len(ia) > 3000 and np.array_equal(ia, ib) and np.array_equal(ta, tb)
    assert ra.events == rb.events
    for v in range(6):
        np.testing.assert_allclose(ra.get_state(0, v), rb.get_state(0, v), rtol=0, atol=1e-11)


# ---- synaptic code ---------------------------------------------------------------------
def sorted_indices_spec(n=5, prec_dt=1e-4):
    """The network of the reference's test_sorted_indices_statemonitor
    (test_cpp_cuda_consistency.py:177-205): NeuronGroup(n, model='',
    threshold='True'), Synapses(P, P, 'w : 1', on_pre='w += 1'),
    connect(j='i'), delay of synapse k = k*dt."""
    m, names = cg.neuron_model("", "True", None, 0.0, None, {})
    code = cg.synapse_code("w += 1", cg.synapse_variables("w : 1"), names, names)
    i = np.arange(n, dtype=np.int32)
    return nw.NetSpec("sorted_indices", prec_dt, 1, [nw.GroupSpec(0, n, [], {}, code=m)],
                      [nw.PathSpec(0, 0, 0, i, i, delay_steps=np.arange(n, dtype=np.int32), code=code)])


@pytest.mark.parametrize("prec,sem", [(64, oracle.SEQ), (32, oracle.ACC)])
def test_sorted_indices_known_answer(prec, sem):
    """Every neuron fires every step; synapse k's spikes arrive k steps later,
    so at the start of step s its w is max(0, s - k)."""
    n = 5
    ref, _, _ = oracle.build_ref(sorted_indices_spec(n), prec, sem)
    mid = ref.add_synapse_monitor(0, nw.HS_SYN_CODE, np.arange(n))
    ref.run(n + 1)
    s, k = np.meshgrid(np.arange(n + 1), np.arange(n), indexing="ij")
    np.testing.assert_array_equal(ref.state_monitor(mid, n), np.maximum(0, s - k))


def test_additive_code_equals_the_builtin_effect():
    """'ge += we' as synaptic code on C1 gives the built-in pathway's bits."""
    a = nw.cuba(N=1000, seed=4, record=0)
    b = nw.cuba(N=1000, seed=4, record=0)
    for p, var in zip(b.paths, ("ge", "gi")):
        p.code = cg.synapse_code(f"{var} += {p.wscalar!r}", [], ["v", "ge", "gi"], ["v", "ge", "gi"])
    for prec, sem in ((64, oracle.SEQ), (32, oracle.ACC)):
        ra, sa, _ = oracle.build_ref(a, prec, sem)
        rb, sb, _ = oracle.build_ref(b, prec, sem)
        ra.run(500)
        rb.run(500)
        assert np.array_equal(ra.spikes(sa[0])[0], rb.spikes(sb[0])[0]) and ra.events == rb.events
        for v in range(3):
            assert np.array_equal(ra.get_state(0, v), rb.get_state(0, v))


@pytest.mark.parametrize("code,msg", [
    ("w += v_post", "reads the target"), ("v = w", "additions only"), ("q += 1", "not a synapse or target"),
    ("w += foo", "no value"), ("w + 1", "not an assignment")])
def test_synaptic_code_refusals(code, msg):
    with pytest.raises(NotImplementedError, match=msg):
        cg.synapse_code(code, ["w"], ["v"], ["v"])
    with pytest.raises(NotImplementedError, match="only 'name : unit'"):
        cg.synapse_variables("dw/dt = -w/tau : 1")
    with pytest.raises(NotImplementedError, match="constant coefficients|only"):
        cg.synapse_code("w += 1", ["w"], ["v"], ["v"], {"tau": 1.0}, model="dA/dt = -A*w/tau : 1 (event-driven)")


@pytest.mark.parametrize("delay", [0, 3])
def test_stdp_as_code_matches_the_builtin_rule(delay):
    """The STDP synapse model written as Brian2 code (event-driven traces,
    on_pre and on_post) against the built-in rule on the same feed-forward
    network: identical spikes, weights and traces to FP64 rounding (the two
    compute the decay exponent differently), equal last-update steps."""
    a = nw.stdp_feedforward(delay)
    b = nw.stdp_feedforward(delay, code=True)
    assert b.paths[0].code.syn_names == ["w", "Apre", "Apost"] and b.paths[0].code.on_post
    ra, sa, _ = oracle.build_ref(a, 64, oracle.SEQ)
    rb, sb, _ = oracle.build_ref(b, 64, oracle.SEQ)
    ra.run(3000)
    rb.run(3000)
    (ia, ta), (ib, tb) = ra.spikes(sa[0]), rb.spikes(sb[0])
    assert len(ia) > 50 and np.array_equal(ia, ib) and np.array_equal(ta, tb)
    n = len(a.paths[0].pre)
    for k, var in enumerate((0, 1, 2)):
        np.testing.assert_allclose(rb.synapse_var(0, nw.HS_SYN_CODE + k, n), ra.synapse_var(0, var, n),
                                   rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(rb.synapse_var(0, 3, n), ra.synapse_var(0, 3, n))


def test_event_driven_trace_known_answer():
    """One synapse, source spikes at steps 10 and 60, target at 30 (generators):
    Apre after the second pre spike is dApre*exp(-50 dt/tau) + dApre; Apost,
    updated at 60 from its value set at 30, is dApost*exp(-30 dt/tau)."""
    tau, dA = 20e-3, 0.5
    G = nw.GroupSpec(nw.HS_MODEL_SPIKEGEN, 1, [0.0], spikes=(np.array([0, 0], np.int32), np.array([10, 60])))
    H = nw.GroupSpec(nw.HS_MODEL_SPIKEGEN, 1, [0.0], spikes=(np.array([0], np.int32), np.array([30])))
    code = cg.synapse_code("Apre += dA\nn_pre += 1", ["n_pre"], ["_placeholder"], [], dict(tau=tau, dA=dA),
                           on_post="Apost += dA", model="dApre/dt = -Apre/tau : 1 (event-driven)\n"
                                                        "dApost/dt = -Apost/tau : 1 (event-driven)")
    S = nw.PathSpec(0, 1, 0, np.array([0], np.int32), np.array([0], np.int32), code=code)
    spec = nw.NetSpec("ed", DT, 1, [G, H], [S])
    ref, _, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(100)
    names = code.syn_names
    val = {n: ref.synapse_var(0, nw.HS_SYN_CODE + k, 1)[0] for k, n in enumerate(names)}
    assert val["n_pre"] == 2
    np.testing.assert_allclose(val["Apre"], dA * np.exp(-50 * DT / tau) + dA, rtol=1e-12)
    np.testing.assert_allclose(val["Apost"], dA * np.exp(-30 * DT / tau), rtol=1e-12)  # decayed at 60
    assert ref.synapse_var(0, 3, 1)[0] == 60


# ---- rand() / randn() in code models ------------------------------------------------
def _rnd_stream(stream, k, normal):
    """the stream of call site k (hs_cg_rnd_stream in hs_cgfuncs.h)"""
    return (stream & 0xFFFFFF) | ((1 + k + 16 * normal) << 24)


def test_rand_draws_known_answer():
    """dv/dt = rand()/ms (euler): after 3 steps v is the sum of the draws of
    call site 0, each (1000 u) dt in the update's own order; the reset
    'v = 0.5*randn()' of a spike at step s holds the normal draw of call site
    0 at (neuron, s).  The draws are the oracle's Philox words under the
    documented keys (network seed, the group's stream with the call site's
    index above bit 24, neuron, step)."""
    dt = 1e-4
    m, names = cg.neuron_model("dv/dt = rand()/ms : 1", None, None, 0.0, "euler", {}, dt=dt, stream=5)
    spec = nw.NetSpec("rand", dt, 77, [nw.GroupSpec(0, 20, [], {}, code=m)], [])
    ref, _, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(3)
    L = oracle.lib()
    want = np.zeros(20)
    for t in range(3):
        for i in range(20):
            want[i] = want[i] + 1000.0 * L.ref_rand(77, _rnd_stream(5, 0, 0), i, t) * dt
    assert np.array_equal(ref.get_state(0, 0), want)
    # a threshold that always fires and a reset that draws
    m, names = cg.neuron_model("v : 1", "True", "v = 0.5*randn()", 0.0, None, {}, dt=dt, stream=9)
    spec = nw.NetSpec("randn_reset", dt, 77, [nw.GroupSpec(0, 10, [], {}, code=m)], [])
    ref, _, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(4)  # the spike of step 3 is reset at the run's end
    want = np.array([0.5 * oracle.randn(77, _rnd_stream(9, 0, 1), i, 3) for i in range(10)])
    assert np.array_equal(ref.get_state(0, 0), want)


def test_rand_threshold_equals_builtin_poisson():
    """threshold='rand() < rates*dt' on a group of rates is the PoissonGroup
    (poissongroup template: rand() < rates*dt) drawing from the call site's
    stream: identical spikes in FP64."""
    dt = 1e-4
    rates = 5.0 + 80.0 * np.random.default_rng(2).random(400)
    m, names = cg.neuron_model("rates : Hz", "rand() < rates*dt", None, 0.0, None, {}, dt=dt, stream=3)
    code = nw.NetSpec("poisson_code", dt, 11, [nw.GroupSpec(0, 400, [], {0: rates}, code=m)], [])
    code.spike_monitors = [0]
    builtin = nw.NetSpec("poisson", dt, 11, [nw.GroupSpec(nw.HS_MODEL_POISSON, 400, [0.0, _rnd_stream(3, 0, 0)],
                                                          {0: rates})], [])
    builtin.spike_monitors = [0]
    a, sa, _ = oracle.build_ref(code, 64, oracle.SEQ)
    b, sb, _ = oracle.build_ref(builtin, 64, oracle.SEQ)
    a.run(2000)
    b.run(2000)
    ia, ta = a.spikes(sa[0])
    ib, tb = b.spikes(sb[0])
    assert len(ia) > 1000 and np.array_equal(ia, ib) and np.array_equal(ta, tb)


def test_rand_reference_test_rand():
    """brian2cuda's test_rand (tests/test_random_number_generation.py:194-204):
    dv/dt = rand() : second over 3 steps; the last values are not all 0 and
    differ from the step before."""
    m, _ = cg.neuron_model("dv/dt = rand() : second", None, None, 0.0, None, {}, dt=1e-4)
    spec = nw.NetSpec("test_rand", 1e-4, 1, [nw.GroupSpec(0, 1000, [], {}, code=m)], [])
    spec.state_monitors = [(0, 0, 0, 1000)]
    ref, _, st = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(3)
    v = ref.state_monitor(st[0], 1000)
    assert not np.all(v[-1] == 0) and not np.array_equal(v[-2], v[-1])
    # the second run of test_random_number_generation_with_multiple_runs continues the draws
    ref.run(2)
    v2 = ref.state_monitor(st[0], 1000)
    assert v2.shape == (5, 1000) and not np.array_equal(v2[-2], v2[-1])


@pytest.mark.parametrize("eqs,kw,msg", [
    ("dv/dt = rand()/ms : 1", {"method": "rk4"}, "need method 'euler'"),
    ("dv/dt = I/ms : 1\nI = rand() : 1", {"method": "euler"}, "subexpression"),
])
def test_rand_refusals(eqs, kw, msg):
    with pytest.raises(NotImplementedError, match=msg):
        cg.neuron_model(eqs, None, None, 0.0, kw["method"], {}, dt=1e-4)


# ---- stochastic methods and several noise variables ------------------------------------
def test_multiple_noise_variables_extended():
    """brian2cuda's test_multiple_noise_variables_extended (tests/test_stateupdaters.py:16-63):
    an oscillator with two noise variables scaled by noise_factor = 0, in both
    of the reference's forms, integrates exactly like the noise-free equations
    with euler and with heun."""
    dt = 1e-4
    base = "dx/dt = y : 1\ndy/dt = - 1*ms**-1*y - 40*ms**-2*x : Hz"
    forms = ["dx/dt = y : 1\ndy/dt = noise_factor*ms**-1.5*xi_1 + noise_factor*ms**-1.5*xi_2 "
             "- 1*ms**-1*y - 40*ms**-2*x : Hz",
             "dx/dt = y + noise_factor*ms**-0.5*xi_1: 1\ndy/dt = noise_factor*ms**-1.5*xi_2 "
             "- 1*ms**-1*y - 40*ms**-2*x : Hz"]
    init = {0: np.array([0.5, 1.0]), 1: np.array([0.0, 0.5])}

    def run(eqs, method):
        m, _ = cg.neuron_model(eqs, None, None, 0.0, method, {"noise_factor": 0.0}, dt=dt)
        spec = nw.NetSpec("noise", dt, 3, [nw.GroupSpec(0, 2, [], dict(init), code=m)], [])
        spec.state_monitors = [(0, 0, 0, 2), (0, 1, 0, 2)]
        ref, _, st = oracle.build_ref(spec, 64, oracle.SEQ)
        ref.run(100)
        return ref.state_monitor(st[0], 2), ref.state_monitor(st[1], 2)

    x0, y0 = run(base, "euler")
    assert np.abs(x0).max() > 0.1
    for eqs in forms:
        for method in ("euler", "heun"):
            x, y = run(eqs, method)
            assert np.array_equal(x, x0) and np.array_equal(y, y0), method


def _gbm(method, n=50, steps=200, dt=1e-4, seed=17, stream=6):
    """x' = mu x + s x xi / sqrt(ms) (geometric Brownian motion) on the oracle:
    (states, the primary noise draws of every step)."""
    mu, sg = 10.0, 0.3
    m, _ = cg.neuron_model("dx/dt = mu*x + s*x*xi/sqrt(ms) : 1", None, None, 0.0, method, {"mu": mu, "s": sg},
                           dt=dt, stream=stream)
    spec = nw.NetSpec("gbm", dt, seed, [nw.GroupSpec(0, n, [], {0: 1.0}, code=m)], [])
    spec.state_monitors = [(0, 0, 0, n)]
    ref, _, st = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(steps)
    z = np.array([[oracle.randn(seed, stream, i, t) for i in range(n)] for t in range(steps)])
    return ref.state_monitor(st[0], n), z, mu, sg * np.sqrt(1e3), dt


def test_heun_multiplicative_noise():
    """method='heun' (Stratonovich): the recurrence x_s = x + f dt + g dW,
    x' = x + f dt + dW (g(x) + g(x_s))/2 restated in numpy over the same draws;
    and strong agreement with the exact Stratonovich solution
    x0 exp(mu t + s W_t) of the same Brownian path."""
    rec, z, mu, s, dt = _gbm("heun")
    x = np.ones(z.shape[1])
    W = np.zeros_like(x)
    for t in range(z.shape[0]):
        # the monitor records each step's state before its update
        np.testing.assert_allclose(rec[t], x, rtol=1e-13)
        dW = np.sqrt(dt) * z[t]
        xs = x + mu * x * dt + s * x * dW
        x = x + mu * x * dt + dW * (s * x + s * xs) / 2
        W = W + dW
    exact = np.exp(mu * dt * z.shape[0] + s * W)
    assert np.max(np.abs(x / exact - 1)) < 5e-2  # strong order 1; s^2 t = 1.8 here


def test_milstein_multiplicative_noise():
    """method='milstein' (Ito, derivative-free, Kloeden and Platen 11.1.7):
    x_s = x + f dt + g sqrt(dt), x' = x + f dt + g dW + (g(x_s) - g(x))(dW^2 - dt)/(2 sqrt(dt))
    restated in numpy over the same draws; strong agreement with the exact Ito
    solution x0 exp((mu - s^2/2) t + s W_t)."""
    rec, z, mu, s, dt = _gbm("milstein")
    x = np.ones(z.shape[1])
    W = np.zeros_like(x)
    for t in range(z.shape[0]):
        np.testing.assert_allclose(rec[t], x, rtol=1e-13)
        dW = np.sqrt(dt) * z[t]
        xs = x + mu * x * dt + s * x * np.sqrt(dt)
        x = x + mu * x * dt + s * x * dW + (s * xs - s * x) * (dW * dW - dt) / (2 * np.sqrt(dt))
        W = W + dW
    exact = np.exp((mu - s * s / 2) * dt * z.shape[0] + s * W)
    assert np.max(np.abs(x / exact - 1)) < 5e-2  # strong order 1; s^2 t = 1.8 here


def test_second_noise_variable_draws():
    """dx/dt = xi_1/sqrt(ms), dy/dt = xi_2/sqrt(ms) (euler): x follows the group's
    own draws, y the draws of noise variable 1 (the stream's noise-variable
    class, hs_cg_rnd_stream(stream, 1, 2))."""
    dt, seed, stream = 1e-4, 23, 12
    m, names = cg.neuron_model("dx/dt = xi_1/sqrt(ms) : 1\ndy/dt = xi_2/sqrt(ms) : 1", None, None, 0.0, "euler", {},
                               dt=dt, stream=stream)
    spec = nw.NetSpec("xi2", dt, seed, [nw.GroupSpec(0, 8, [], {}, code=m)], [])
    ref, _, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(3)
    sq = np.sqrt(1e3)
    for var, st in ((0, stream), (1, _rnd_stream(stream, 1, 2))):
        want = np.zeros(8)
        for t in range(3):
            want = want + np.array([oracle.randn(seed, st, i, t) for i in range(8)]) * sq * np.sqrt(dt)
        np.testing.assert_allclose(ref.get_state(0, var), want, rtol=1e-13)


# ---- rand() / randn() in synaptic code ----------------------------------------------
def _syn_stream(pid, k, normal):
    """the stream of synaptic call site k (hs_cg_syn_stream in hs_cgfuncs.h)"""
    return (pid & 0xFFFFFF) | ((64 + k + 16 * normal) << 24)


def test_synaptic_rand_known_answer():
    """Three generator sources (spikes at steps 5, 5 and 40) onto four generator
    targets (spikes at step 20), all to all: on_pre 'x = rand(); y = randn()',
    on_post 'z = rand()' (the second rand() call site, k = 1).  Each value is the
    draw of its call site for (seed, pathway, i, j, step): Philox counter
    (i, step, j, stream) -- the same keys the device uses."""
    G = nw.GroupSpec(nw.HS_MODEL_SPIKEGEN, 3, [0.0], spikes=(np.array([0, 1, 2], np.int32), np.array([5, 5, 40])))
    H = nw.GroupSpec(nw.HS_MODEL_SPIKEGEN, 4, [0.0], spikes=(np.arange(4, dtype=np.int32), np.full(4, 20)))
    code = cg.synapse_code("x = rand()\ny = randn()", ["x", "y", "z"], ["_placeholder"], [], on_post="z = rand()")
    assert "hs_srand_u(0)" in code.on_pre and "hs_srand_n(0)" in code.on_pre and "hs_srand_u(1)" in code.on_post
    pre, post = np.repeat(np.arange(3), 4).astype(np.int32), np.tile(np.arange(4), 3).astype(np.int32)
    seed = 77
    spec = nw.NetSpec("syn_rand", DT, seed, [G, H], [nw.PathSpec(0, 1, 0, pre, post, code=code)])
    for prec in (64, 32):
        ref, _, _ = oracle.build_ref(spec, prec, oracle.SEQ)
        ref.run(60)
        got = {n: ref.synapse_var(0, nw.HS_SYN_CODE + k, 12) for k, n in enumerate(code.syn_names)}
        f = np.float64 if prec == 64 else np.float32
        for s, (i, j) in enumerate(zip(pre, post)):
            step = 5 if i < 2 else 40
            ctr = (int(j) << 32) | step
            assert got["x"][s] == f(oracle.rand(seed, _syn_stream(0, 0, 0), int(i), ctr))
            assert got["y"][s] == f(oracle.randn(seed, _syn_stream(0, 0, 1), int(i), ctr))
            assert got["z"][s] == f(oracle.rand(seed, _syn_stream(0, 1, 0), int(i), (int(j) << 32) | 20))


def test_stochastic_release_statistics():
    """code_synapses(release=True): 'released = rand() < 0.5' releases half the
    arrivals (within 4 sigma of the binomial count), every released arrival and
    only those adds to ge (the target's summed input equals 1.5 mV per release),
    the multiplicative randn() step keeps w in [0, 1], and two call sites of one
    statement list draw independently (released differs from a fresh draw)."""
    spec = nw.code_synapses(release=True)
    p = spec.paths[0]
    assert "hs_srand_u(0)" in p.code.on_pre and "hs_srand_n(0)" in p.code.on_pre
    ref, sm, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(2000)
    n = len(p.pre)
    hits = ref.synapse_var(0, nw.HS_SYN_CODE + 1, n)
    nrel = ref.synapse_var(0, nw.HS_SYN_CODE + 3, n)
    w = ref.synapse_var(0, nw.HS_SYN_CODE + 0, n)
    H, K = hits.sum(), nrel.sum()
    assert H == ref.events > 10000
    assert abs(K - 0.5 * H) < 4 * np.sqrt(0.25 * H)
    assert np.all(nrel <= hits) and np.all((w >= 0) & (w <= 1))
    assert len(ref.spikes(sm[0])[0]) > 20  # the released input drives the targets
    # the same network with every arrival released moves w the same way but
    # drives the targets harder
    code = cg.synapse_code("ge_post += 1.5*mV\nhits += 1", ["w", "hits"], ["v", "ge"], [])
    code.init = {"w": 0.5, "hits": 0.0}
    spec2 = nw.code_synapses(release=True)
    spec2.paths[0].code = code
    r2, sm2, _ = oracle.build_ref(spec2, 64, oracle.SEQ)
    r2.run(2000)
    assert len(r2.spikes(sm2[0])[0]) > len(ref.spikes(sm[0])[0])


def test_comparisons_as_numbers():
    """'a < b' in arithmetic is 1 or 0 (Brian2's booleans in expressions), a
    chain is the product of its links; in neuron models too."""
    code = cg.synapse_code("w += (i < j) + 2*(1 < j <= 3) + (i == j)", ["w"], ["v"], [])
    G = nw.GroupSpec(nw.HS_MODEL_SPIKEGEN, 5, [0.0], spikes=(np.arange(5, dtype=np.int32), np.full(5, 2)))
    H = nw.GroupSpec(nw.HS_MODEL_SPIKEGEN, 5, [0.0])
    pre, post = np.repeat(np.arange(5), 5).astype(np.int32), np.tile(np.arange(5), 5).astype(np.int32)
    code.post_names = []
    spec = nw.NetSpec("cmp", DT, 1, [G, H], [nw.PathSpec(0, 1, 0, pre, post, code=code)])
    ref, _, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(5)
    want = [(i < j) + 2 * (1 < j <= 3) + (i == j) for i, j in zip(pre, post)]
    assert np.array_equal(ref.synapse_var(0, nw.HS_SYN_CODE, 25), np.array(want, float))
    m, _ = cg.neuron_model("dv/dt = (v < 0.5)/(10*ms) : 1", None, None, 0.0, "euler", {}, dt=DT)
    assert "<" in m.update
    with pytest.raises(NotImplementedError, match="comparison"):
        cg.synapse_code("w += (i in j)", ["w"], ["v"], [])
