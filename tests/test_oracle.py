"""CPU tests that pin the oracle against facts it does not compute itself:
published known-answer vectors, closed-form solutions, textbook integrators,
the reference's own test expectations, and libm.  (No GPU.)"""
import math

import dataclasses

import numpy as np
import pytest
from scipy import linalg, stats

import oracle
from brian2cuda_amd import networks as nw
import refcases as rc

# ---------------------------------------------------------------------------
# Philox4x32-10: Random123 known-answer vectors (Salmon et al., SC'11)
# ---------------------------------------------------------------------------
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert oracle.philox4x32_10(ctr, key) == out


def test_randn_statistics():
    z = np.array([oracle.randn(1234, 0, i, s) for s in range(4) for i in range(50000)])
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # neighbouring counters are independent
    a = np.array([oracle.randn(7, 0, i, 5) for i in range(20000)])
    b = np.array([oracle.randn(7, 0, i, 6) for i in range(20000)])
    c = np.array([oracle.randn(7, 1, i, 5) for i in range(20000)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.03
    assert abs(np.corrcoef(a, c)[0, 1]) < 0.03


def test_randn_f32_statistics_and_relation_to_f64():
    """The FP32 form's draw: standard normal statistics; it uses the same Philox
    words as the FP64 draw with 24-bit uniforms, so the two agree to the uniforms'
    resolution except in the far tail."""
    z = np.array([oracle.randn(1234, 0, i, s, precision=32) for s in range(4) for i in range(50000)])
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-3
    d = np.array([oracle.randn(1234, 0, i, 3) for i in range(50000)])
    f = z[3 * 50000:]
    assert np.median(np.abs(f - d)) < 1e-6
    assert np.mean(np.abs(f - d) < 1e-4) > 0.999


def test_float_detmath_accuracy():
    L = oracle.lib()
    rng = np.random.default_rng(2)
    u = (rng.integers(1, 2**24, 20000) * 2.0 ** -24).astype(np.float32)
    got = np.array([L.ref_logf(float(v)) for v in u])
    want = np.log(u.astype(np.float64))
    assert np.all(np.abs(got - want) <= 3 * np.spacing(np.abs(want).astype(np.float32)).astype(np.float64) + 1e-45)
    x = rng.uniform(0, 80, 20000).astype(np.float32)
    assert np.array_equal(np.array([L.ref_sqrtf(float(v)) for v in x], np.float32), np.sqrt(x))
    c = np.array([L.ref_cos2pif(float(v)) for v in u])
    assert np.abs(c - np.cos(2 * np.pi * u.astype(np.float64))).max() < 3e-7


def test_randn_reproducible_and_seed_dependent():
    assert oracle.randn(99, 3, 17, 5) == oracle.randn(99, 3, 17, 5)
    assert oracle.randn(99, 3, 17, 5) != oracle.randn(100, 3, 17, 5)


# ---------------------------------------------------------------------------
# deterministic math against libm
# ---------------------------------------------------------------------------
def _ulps(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    sp = np.spacing(np.abs(b))
    return np.abs(a - b) / sp


def test_detmath_accuracy():
    L = oracle.lib()
    rng = np.random.default_rng(0)
    x = rng.uniform(-700, 700, 20000)
    assert _ulps([L.ref_exp(v) for v in x], np.exp(x)).max() <= 2
    y = 10.0 ** rng.uniform(-300, 300, 20000)
    assert _ulps([L.ref_log(v) for v in y], np.log(y)).max() <= 3
    assert _ulps([L.ref_sqrt(v) for v in y], np.sqrt(y)).max() == 0  # correctly rounded
    u = rng.random(20000)
    assert np.abs(np.array([L.ref_cos2pi(v) for v in u]) - np.cos(2 * np.pi * u)).max() < 2e-15
    xf = rng.uniform(-80, 80, 20000).astype(np.float32)
    got = np.array([L.ref_expf(float(v)) for v in xf])
    assert (np.abs(got - np.exp(xf.astype(np.float64))) / np.exp(xf.astype(np.float64))).max() < 2.5e-7
    assert L.ref_exp(0.0) == 1.0 and L.ref_log(1.0) == 0.0 and L.ref_sqrt(4.0) == 2.0


# ---------------------------------------------------------------------------
# reference conversions: delays and refractory steps
# ---------------------------------------------------------------------------
def test_delay_quantisation_matches_reference_rule():
    """(int)(d/dt + 0.5), synapses_push_spikes.cu:375."""
    L = oracle.lib()
    dt = 1e-4
    d = np.array([0.0, 0.04e-3, 0.05e-3, 0.06e-3, 0.1e-3, 0.15e-3, 0.5e-3, 4.96e-3, 5e-3])
    want = [L.ref_quantize_delay(float(v), dt) for v in d]
    assert list(nw.quantise_delay(d, dt)) == want
    assert want[:2] == [0, 0] and want[4] == 1 and want[-1] == 50


def test_timestep_function():
    """Brian2's timestep(t, dt) = (int64)((t + 1e-3 dt)/dt) (cuda_generator.py:997-1006)."""
    L = oracle.lib()
    assert L.ref_timestep(5e-3, 1e-4) == 50
    assert L.ref_timestep(2e-3, 1e-4) == 20
    assert L.ref_timestep(3e-3, 1e-4) == 30
    assert L.ref_timestep(0.3e-3, 1e-4) == 3  # 0.3e-3/1e-4 = 2.9999999999999996 without the guard
    for k in range(1, 5000, 7):
        assert L.ref_timestep(k * 1e-4, 1e-4) == k


# ---------------------------------------------------------------------------
# model expressions against closed forms / textbook integrators
# ---------------------------------------------------------------------------
def test_lif_cuba_exact_matches_matrix_exponential():
    """CUBA's 'exact' step vs expm of the linear system (v-El, ge, gi)."""
    p = nw.cuba_params()
    El, taum, taue, taui = p[0], p[3], p[4], p[5]
    dt = 1e-4
    A = np.array([[-1 / taum, 1 / taum, 1 / taum], [0, -1 / taue, 0], [0, 0, -1 / taui]])
    M = linalg.expm(A * dt)
    rng = np.random.default_rng(1)
    n = 200
    v = -0.06 + 0.01 * rng.random(n)
    ge = 0.004 * rng.random(n)
    gi = -0.004 * rng.random(n)
    last = np.full(n, -1073741824, np.int32)
    out = oracle.stateupdate(1, 64, [v, ge, gi], last, p, dt, 0, 10)
    x = M @ np.vstack([v - El, ge, gi])
    np.testing.assert_allclose(out[0], x[0] + El, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(out[1], x[1], rtol=1e-13)
    np.testing.assert_allclose(out[2], x[2], rtol=1e-13)


def test_lif_cuba_unless_refractory_holds_v():
    p = nw.cuba_params()  # refractory 5 ms = 50 steps
    v, ge, gi = np.array([-0.055, -0.055]), np.array([0.001, 0.001]), np.array([0.0, 0.0])
    last = np.array([90, -1073741824], np.int32)  # neuron 0 spiked 10 steps ago
    out = oracle.stateupdate(1, 64, [v, ge, gi], last, p, 1e-4, 0, 100)
    assert out[0][0] == -0.055 and out[0][1] != -0.055
    assert out[1][0] == out[1][1]  # ge decays in both


def test_lif_cuba_trajectory_vs_ode_solver():
    """10 ms of the oracle's exact integration vs a tight RK45 solution."""
    from scipy.integrate import solve_ivp
    p = nw.cuba_params()
    El, taum, taue, taui = p[0], p[3], p[4], p[5]
    spec = nw.NetSpec("one", 1e-4, 0, [nw.GroupSpec(1, 1, p, {0: -0.058, 1: 0.003, 2: -0.002})], [])
    spec.state_monitors = [(0, 0, 0, 1)]
    ref, _, st = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(101)
    v = ref.state_monitor(st[0], 1)[:, 0]

    def f(t, y):
        return [(y[1] + y[2] - (y[0] - El)) / taum, -y[1] / taue, -y[2] / taui]
    sol = solve_ivp(f, (0, 0.01), [-0.058, 0.003, -0.002], rtol=1e-12, atol=1e-15, t_eval=np.arange(101) * 1e-4)
    np.testing.assert_allclose(v, sol.y[0], rtol=1e-9)


def test_lif_noise_euler_maruyama_step():
    """V' = V + dt (mu - V)/tau + sqrt(dt) sigma sqrt(tau)/tau xi with the Philox draw."""
    p = nw.brunel_params()
    tau, mu, sigma = p[0], p[1], p[2]
    dt, seed, step = 1e-4, 42, 1234
    V = np.linspace(0.0, 0.02, 50)
    last = np.full(50, -1073741824, np.int32)
    out = oracle.stateupdate(2, 64, [V], last, p, dt, seed, step)[0]
    xi = np.array([oracle.randn(seed, 0, i, step) for i in range(50)])
    want = V + dt * (mu - V) / tau + math.sqrt(dt) * sigma * math.sqrt(tau) / tau * xi
    np.testing.assert_allclose(out, want, rtol=1e-13, atol=1e-18)


def test_lif_noise_fake_randn_half():
    """The reference's fake_randn check (tests/conftest.py:7-16, test_stateupdaters.py:118-134):
    with xi = 0.5 and no drift, one Euler-Maruyama step adds sqrt(dt) sigma sqrt(tau)/tau * 0.5."""
    tau, sigma, dt = 10e-3, 1.0, 1e-4
    k = oracle.model_constants(2, [tau, 0.0, sigma, 1.0, 0.0, 0.0, 0.0], dt)
    assert math.isclose(k[1] * 0.5, math.sqrt(dt) * sigma / math.sqrt(tau) * 0.5, rel_tol=1e-14)


def _hh_rates(v, VT):
    mV, ms = 1e-3, 1e-3
    am = 0.32 / mV * (13 * mV - v + VT) / (np.exp((13 * mV - v + VT) / (4 * mV)) - 1) / ms
    bm = 0.28 / mV * (v - VT - 40 * mV) / (np.exp((v - VT - 40 * mV) / (5 * mV)) - 1) / ms
    ah = 0.128 * np.exp((17 * mV - v + VT) / (18 * mV)) / ms
    bh = 4.0 / (1 + np.exp((40 * mV - v + VT) / (5 * mV))) / ms
    an = 0.032 / mV * (15 * mV - v + VT) / (np.exp((15 * mV - v + VT) / (5 * mV)) - 1) / ms
    bn = 0.5 * np.exp((10 * mV - v + VT) / (40 * mV)) / ms
    return am, bm, ah, bh, an, bn


def test_hh_exponential_euler_vs_rush_larsen():
    """COBAHH step vs the textbook form x' = x_inf + (x - x_inf) exp(-dt/tau_x),
    every right-hand side frozen at the start of the step (cobahh.py:160-185)."""
    p = nw.cobahh_params()
    Cm, gl, El, EK, ENa, gNa, gK, VT, Ee, Ei, taue, taui = p[:12]
    dt = 1e-4
    rng = np.random.default_rng(2)
    n = 300
    v = -0.07 + 0.06 * rng.random(n)
    m, h, nn = rng.random(n), rng.random(n), rng.random(n)
    ge, gi = 1e-8 * rng.random(n), 5e-8 * rng.random(n)
    last = np.full(n, -1073741824, np.int32)
    out = oracle.stateupdate(3, 64, [v, m, nn, h, ge, gi], last, p, dt, 0, 3)
    am, bm, ah, bh, an, bn = _hh_rates(v, VT)

    def rl(x, a, b):
        tau = 1 / (a + b)
        xinf = a * tau
        return xinf + (x - xinf) * np.exp(-dt / tau)
    gna, gkd = gNa * m ** 3 * h, gK * nn ** 4
    gtot = gl + ge + gi + gna + gkd
    vinf = (gl * El + ge * Ee + gi * Ei + gna * ENa + gkd * EK) / gtot
    v_want = vinf + (v - vinf) * np.exp(-dt * gtot / Cm)
    np.testing.assert_allclose(out[0], v_want, rtol=1e-11)
    np.testing.assert_allclose(out[1], rl(m, am, bm), rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(out[2], rl(nn, an, bn), rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(out[3], rl(h, ah, bh), rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(out[4], ge * np.exp(-dt / taue), rtol=1e-13)
    np.testing.assert_allclose(out[5], gi * np.exp(-dt / taui), rtol=1e-13)


def test_float32_form_rounds_like_float32():
    """The FP32 oracle differs from FP64 by float rounding only (one CUBA step)."""
    p = nw.cuba_params()
    v, ge, gi = np.array([-0.055]), np.array([0.001]), np.array([-0.002])
    last = np.full(1, -1073741824, np.int32)
    a = oracle.stateupdate(1, 64, [v, ge, gi], last, p, 1e-4, 0, 0)
    b = oracle.stateupdate(1, 32, [v, ge, gi], last, p, 1e-4, 0, 0)
    for x, y in zip(a, b):
        assert abs(x[0] - y[0]) <= 4 * np.spacing(np.float32(abs(x[0])))
        assert y[0] == float(np.float32(y[0]))


# ---------------------------------------------------------------------------
# the reference's spike-propagation expectations
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["push_outer_loop", "push_outer_loop2", "circular_eventspaces", "bundle_prefs"])
@pytest.mark.parametrize("sem", [oracle.SEQ, oracle.ACC])
def test_reference_propagation_cases(case, sem):
    spec, steps, expect = getattr(rc, case)()
    ref, sm, _ = oracle.build_ref(spec, 64, sem)
    ref.run(steps)
    got = rc.spikes_by_step(*ref.spikes(sm[0]))
    assert sorted(got) == sorted(expect)
    for k in expect:
        np.testing.assert_array_equal(got[k], expect[k])


@pytest.mark.parametrize("sem", [oracle.SEQ, oracle.ACC])
def test_reference_circular_eventspaces_spikegenerator(sem):
    spec, steps, expect = rc.circular_eventspaces_spikegenerator(sem)
    ref, sm, _ = oracle.build_ref(spec, 64, sem)
    ref.run(steps)
    got = rc.spikes_by_step(*ref.spikes(sm[0]))
    assert sorted(got) == sorted(expect)
    for k in expect:
        np.testing.assert_array_equal(got[k], expect[k])


@pytest.mark.parametrize("delays", [[1], [0, 1]])
@pytest.mark.parametrize("sem", [oracle.SEQ, oracle.ACC])
def test_reference_multiple_runs(delays, sem):
    """test_multiple_runs_with_scalar_delay ([[0, 0, 1, 2]]) and
    _with_heterogeneous_delay ([0, 1, 2, 3], [0, 0, 1, 2]): two runs of two steps."""
    spec, expect = rc.multiple_runs(delays)
    ref, _, st = oracle.build_ref(spec, 64, sem)
    ref.run(2)
    ref.run(2)
    np.testing.assert_array_equal(ref.state_monitor(st[0], len(delays)).T, expect)


def test_reference_effects_cancel():
    spec, steps, _ = rc.effects_cancel()
    for sem in (oracle.SEQ, oracle.ACC):
        ref, sm, _ = oracle.build_ref(spec, 64, sem)
        ref.run(steps)
        assert len(ref.spikes(sm[0])[0]) == 0
        assert np.all(ref.get_state(1, 0) == 0.0)
        assert ref.events == 2 * 15360 * (steps - 1)


def test_events_invariant_and_queue_vs_closed_form():
    """Delivered events = sum over spikes of their synapses whose arrival
    step (emission + delay) falls inside the run."""
    spec = nw.brunel(N=800, K=80, seed=5, monitors=True)
    ref, sm, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    steps = 700
    ref.run(steps)
    i, t = ref.spikes(sm[0])
    p = spec.paths[0]
    order = np.argsort(p.pre, kind="stable")
    starts = np.searchsorted(p.pre[order], np.arange(spec.groups[0].n + 1))
    expect = 0
    for ii, tt in zip(i, t):
        d = p.delay_steps[order[starts[ii]:starts[ii + 1]]]
        expect += int(np.sum(tt + d < steps))
    assert ref.events == expect


def test_refractoriness_in_network():
    """No neuron spikes twice within its refractory period (C3: 2 ms = 20 steps)."""
    spec = nw.brunel(N=2000, K=200, seed=6, monitors=True)
    ref, sm, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(3000)
    i, t = ref.spikes(sm[0])
    for n in np.unique(i)[:500]:
        tt = np.sort(t[i == n])
        if len(tt) > 1:
            assert np.diff(tt).min() >= 20


def test_seq_and_acc_agree_with_single_inputs():
    """With at most one input per target and step the two semantics are
    arithmetically identical (x + w == (0 + w) + x): the spike trains coincide."""
    spec, steps, expect = rc.circular_eventspaces()
    a, sa, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    b, sb, _ = oracle.build_ref(spec, 64, oracle.ACC)
    a.run(steps)
    b.run(steps)
    for x, y in zip(a.spikes(sa[0]), b.spikes(sb[0])):
        np.testing.assert_array_equal(x, y)


def test_cuba_network_rates_plausible():
    """C1 (4000 CUBA LIF, p=0.02): asynchronous irregular regime, a few Hz."""
    spec = nw.cuba(N=4000, seed=1234)
    ref, sm, st = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(5000)
    i, t = ref.spikes(sm[0])
    rate = len(i) / 4000 / 0.5
    assert 1.0 < rate < 50.0
    v = ref.state_monitor(st[0], 100)
    assert v.shape == (5000, 100)
    # sampled at the start of a step: after resets, so never above threshold
    assert v.max() <= -0.050 and v.min() > -0.1


@pytest.mark.parametrize("sem", [oracle.SEQ, oracle.ACC])
def test_reference_ring_uncoupled_closed_form(sem):
    """The ring without its synapses: after a spike a neuron is held 50 steps,
    then needs 7 exact steps from -60 mV to pass -50 mV (-40 - 20 e^{-0.7} mV),
    so it fires every 56 steps; every first spike comes by step 47, so each of
    the 1000 neurons fires 18 times in 1000 steps."""
    spec, steps = rc.ring(w=0.0)
    ref, sm, _ = oracle.build_ref(spec, 64, sem)
    ref.run(steps)
    i, t = ref.spikes(sm[0])
    assert np.array_equal(np.bincount(i, minlength=1000), np.full(1000, 18))
    assert t[0] == 0 and t.max() <= 999


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("sem", [oracle.SEQ, oracle.ACC])
def test_reference_ring(prec, sem):
    """test_cuda_standalone: the first spike at t = 0 and at least 17,000 spikes
    in 100 ms, as the reference expects; the count itself (18,068) equals an
    independent numpy restatement of the schedule.  The reference's upper bound
    of 18,000 is not met: the uncoupled ring already fires exactly 18,000 times
    under timestep-based refractoriness, and excitatory coupling can only add."""
    spec, steps = rc.ring()
    ref, sm, _ = oracle.build_ref(spec, prec, sem)
    ref.run(steps)
    i, t = ref.spikes(sm[0])
    assert len(i) >= 17000
    assert t[0] == 0
    if prec == 64:
        assert (len(i), int(t[0])) == rc.ring_numpy()


@pytest.mark.parametrize("prec,rtol", [(64, 1e-7), (32, 2e-6)])
def test_reference_example_run(prec, rtol):
    """example_run's closed form v = arange(N)/N exp(-1 ms/tau) after 1 ms (FP64 at
    the reference's rtol 1e-7; FP32 at 2e-6: ten roundings of a float product)."""
    spec, steps, expect = rc.example_run()
    ref, sm, _ = oracle.build_ref(spec, prec, oracle.SEQ)
    ref.run(steps)
    np.testing.assert_allclose(ref.get_state(0, 0), expect, rtol=rtol)
    assert len(ref.spikes(sm[0])[0]) == 0


def test_cuba_singular_time_constants_refused():
    """The shared closed-form constants reject taue or taui equal to taum."""
    spec, steps, _ = rc.example_run()
    g = spec.groups[0]
    bad = dataclasses.replace(spec, groups=[dataclasses.replace(g, params=g.params[:5] + [g.params[3]] + g.params[6:])])
    with pytest.raises(Exception):
        oracle.build_ref(bad, 64, oracle.SEQ)


@pytest.mark.parametrize("name,prec,sem,steps", [("cuba", 64, 0, 800), ("cuba", 32, 1, 800),
                                                 ("cobahh", 64, 0, 300), ("brunel", 32, 1, 400)])
def test_threaded_oracle_equals_serial(oracle_lib, name, prec, sem, steps):
    """The OpenMP step loop (ref_net_set_threads, the CPU baseline's form) gives
    the serial loop's results bit for bit: spikes, state and events."""
    from brian2cuda_amd import networks as nw
    spec = {"cuba": lambda: nw.cuba(N=3000, seed=3), "cobahh": lambda: nw.cobahh(N=3000, seed=3),
            "brunel": lambda: nw.brunel(N=12000, K=400, seed=3, monitors=True)}[name]()
    out = []
    for th in (1, 5):
        ref, sm, _ = oracle_lib.build_ref(spec, prec, sem, threads=th)
        ref.run(steps)
        out.append((*ref.spikes(sm[0]), ref.get_state(0, 0), ref.events))
        ref.close()
    (i1, t1, v1, e1), (i5, t5, v5, e5) = out
    assert len(i1) > 100
    assert np.array_equal(i1, i5) and np.array_equal(t1, t5)
    assert np.array_equal(v1, v5) and e1 == e5


def test_threaded_oracle_only_before_the_first_step(oracle_lib):
    from brian2cuda_amd import networks as nw
    ref, _, _ = oracle_lib.build_ref(nw.cuba(N=200, seed=1), 64, 0)
    ref.run(3)
    with pytest.raises(ValueError):
        ref.set_threads(4)
