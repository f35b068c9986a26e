"""GPU parity of spike generators and STDP pathways (plain form) against the oracle.

Bit-exact where the arithmetic is order-free: a generator's spikes, and a
feed-forward STDP network whose input fires at most once per step (each target
then receives at most one plastic effect per step, so no two floating-point
additions meet on one target).  The reference's recurrent STDP networks
(test_cpp_cuda_consistency.py) sum several effects into one conductance in the
same step in atomic order, as the reference's CUDA kernels do; they are compared
the way that test compares cpp_standalone with cuda_standalone: same spikes,
weights and voltages within a relative tolerance.
"""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw
from brian2cuda_amd import runtime as rt
from brian2cuda_amd._lib import HsError

pytestmark = pytest.mark.gpu

DT = 1e-4


def _bundle_order(spec, path):
    """permutation from synapse (creation) order to the device's bundle order"""
    d = path.delay_steps if path.delay_steps is not None else np.full(path.nsyn, path.scalar_delay)
    return np.lexsort((d, path.pre))


@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_spikegen_pattern(hs, oracle_lib, engine):
    rng = np.random.default_rng(3)
    n = 700
    idx = rng.integers(0, n, 4000)
    steps = rng.integers(0, 900, 4000)
    key = np.unique(steps * n + idx)
    idx, steps = (key % n).astype(np.int32), (key // n).astype(np.int64)
    for period in (0.0, 500 * DT):
        keep = steps < 500 if period > 0 else np.ones(steps.size, bool)  # a pattern lies inside its period
        spec = nw.NetSpec("gen", DT, 1, [nw.GroupSpec(rt.HS_MODEL_SPIKEGEN, n, [period],
                                                      spikes=(idx[keep], steps[keep]))], [], spike_monitors=[0])
        net, sm, _ = nw.build_hs(spec, 32, engine)
        ref, rsm, _ = oracle_lib.build_ref(spec, 32, oracle_lib.SEQ)
        net.run(1200)
        ref.run(1200)
        gi, gt = net.spikes(sm[0])
        ri, rt_ = ref.spikes(rsm[0])
        assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
        if period == 0.0:
            assert len(gi) == idx.size
        net.close()


def _feedforward(delay_steps, n_in=100, n_out=60, seed=5):
    """generator (one spike per step at most) -> LIF, all-to-all STDP"""
    rs = np.random.RandomState(seed)
    times = rs.choice(np.arange(3000), 1500, replace=False)
    src = rs.randint(0, n_in, times.size)
    G = nw.GroupSpec(rt.HS_MODEL_SPIKEGEN, n_in, [0.0], spikes=(src.astype(np.int32), times.astype(np.int64)))
    P = nw.GroupSpec(rt.HS_MODEL_LIF_CUBA, n_out, [-49e-3, -50e-3, -60e-3, 20e-3, 5e-3, 5e-3, 5e-3, 1.0],
                     {0: -60e-3 + 10e-3 * rs.rand(n_out), 1: 0.0})
    ii, jj = np.meshgrid(np.arange(n_in), np.arange(n_out), indexing="ij")
    S = nw.PathSpec(0, 1, 1, ii.ravel().astype(np.int32), jj.ravel().astype(np.int32), scalar_delay=delay_steps,
                    wmode=rt.HS_W_SYNAPSE, weights=1.5 + 2.0 * rs.rand(n_in * n_out),
                    stdp=(20e-3, 20e-3, 0.05, -0.055, 0.0, 4.0, 1e-3))
    return nw.NetSpec("stdp_ff", DT, seed, [G, P], [S], spike_monitors=[1], state_monitors=[(1, 0, 0, 20)])


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("delay", [0, 3])
def test_stdp_feedforward_bit_exact(hs, oracle_lib, prec, delay):
    spec = _feedforward(delay)
    net, sm, st = nw.build_hs(spec, prec, "plain")
    ref, rsm, rst = oracle_lib.build_ref(spec, prec, oracle_lib.SEQ)
    S = spec.paths[0]
    perm = _bundle_order(spec, S)
    for k in (700, 1100, 1300):
        net.run(k)
        ref.run(k)
        for var in (rt.HS_SYN_W, rt.HS_SYN_APRE, rt.HS_SYN_APOST, rt.HS_SYN_LASTUPDATE):
            np.testing.assert_array_equal(net.synapse_var(0, var), ref.synapse_var(0, var, S.nsyn)[perm],
                                          err_msg=f"synapse variable {var} after {net.stats()['steps']} steps")
    gi, gt = net.spikes(sm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(gi) > 50
    assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert np.array_equal(net.state_monitor(st[0], 20), ref.state_monitor(rst[0], 20))
    assert net.stats()["events"] == ref.events
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("heterog", [False, True])
def test_reference_stdp_examples(hs, oracle_lib, prec, heterog):
    """test_stdp_example / test_stdp_heterog_delays_example: w of synapses 0..9
    at t = 0 and 0.1 s, v of neurons 0..9 every step, spike count and rate."""
    spec = nw.stdp_example(heterog)
    net, sm, st = nw.build_hs(spec, prec, "plain")
    ref, rsm, rst = oracle_lib.build_ref(spec, prec, oracle_lib.SEQ)
    S = spec.paths[1]
    perm = _bundle_order(spec, S)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)  # creation index -> bundle position
    rtol = 1e-9
    for rec in range(2):  # StateMonitor(S, 'w', record=range(10), dt=0.1*second): t = 0, 0.1 s
        gw, rw = net.synapse_var(1, rt.HS_SYN_W)[inv[:10]], ref.synapse_var(1, 0, S.nsyn)[:10]
        if prec == 64 or rec == 0:
            np.testing.assert_allclose(gw, rw, rtol=rtol if prec == 64 else 1e-7)
        else:  # FP32: a spike moved by a step moves w by about one trace increment (0.03)
            assert np.mean(np.abs(gw - rw)) < 0.02
        net.run(1000)
        ref.run(1000)
    gi, gt = net.spikes(sm[0])
    ri, rt_ = ref.spikes(rsm[0])
    if prec == 64:
        assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    else:
        assert abs(len(gi) - len(ri)) <= 0.02 * len(ri)
    rate_g = np.bincount(gt, minlength=2000) / (100 * DT)
    rate_r = np.bincount(rt_, minlength=2000) / (100 * DT)
    if prec == 64:
        np.testing.assert_array_equal(rate_g, rate_r)
        # PopulationRateMonitor(P), as the reference's test records it
        np.testing.assert_array_equal(net.rate_monitor(net.rate_mids[0]), ref.rate_monitor(ref.rate_mids[0]))
        np.testing.assert_array_equal(net.rate_monitor(net.rate_mids[0]), rate_g)
        np.testing.assert_allclose(net.state_monitor(st[0], 10), ref.state_monitor(rst[0], 10), rtol=rtol,
                                   atol=1e-12)
        np.testing.assert_allclose(net.synapse_var(1, rt.HS_SYN_W), ref.synapse_var(1, 0, S.nsyn)[perm], rtol=rtol,
                                   atol=1e-12)
    net.close()


def test_stdp_rejects_unsupported(hs):
    spec = _feedforward(0)
    S = spec.paths[0]
    net = rt.Net(64, "plain", DT, 1)
    for g in spec.groups:
        net.add_group(g.model, g.n, g.params)
    pid = net.add_pathway(S.src, S.tgt, S.var, S.pre, S.post, None, 0, rt.HS_W_SCALAR, None, 1.0)
    with pytest.raises(HsError):
        net.pathway_stdp(pid, *S.stdp)  # scalar weight
    with pytest.raises(HsError):
        net.set_spikegen(0, [1, 1], [5, 5])  # twice in one step
    with pytest.raises(HsError):
        net.set_spikegen(1, [1], [5])  # not a generator
    net.close()


@pytest.mark.parametrize("case", ["extreme_period", "period_repeat", "rounding_first_bin", "rounding_separate"])
def test_reference_spikegenerator_cases_gpu(hs, case):
    """The expectations of the reference's test_spikegenerator.py on the device."""
    n, idx, times, period, steps = {
        "extreme_period": (5, [0, 1, 2], np.array([0, 1, 2]) * 1e-3, 1e6, 100),
        "period_repeat": (1, np.zeros(10, int), np.arange(0, 1, 0.1) * 1e-3, 1e-3, 50),
        "rounding_first_bin": (100, np.arange(100), np.linspace(0, 0.1e-3, 100, endpoint=False), 0.0, 1),
        "rounding_separate": (1, np.zeros(10000, int), np.arange(10000) * DT, 0.0, 10000),
    }[case]
    net = rt.Net(64, "plain", DT, 1)
    g = net.add_group(rt.HS_MODEL_SPIKEGEN, n, [period])
    net.set_spikegen(g, idx, rt.spikegen_steps(times, DT))
    m = net.add_spike_monitor(g)
    net.run(steps)
    i, t = net.spikes(m)
    if case == "extreme_period":
        assert i.tolist() == [0, 1, 2] and t.tolist() == [0, 10, 20]
    elif case == "period_repeat":
        assert np.array_equal(np.bincount(t // 10, minlength=5), [10] * 5)
    elif case == "rounding_first_bin":
        assert np.array_equal(np.bincount(i, minlength=100), np.ones(100)) and np.all(t == 0)
    else:
        assert np.array_equal(t, np.arange(10000))
    net.close()


def test_spikegen_period_validation_gpu(hs):
    net = rt.Net(64, "plain", DT, 1)
    g = net.add_group(rt.HS_MODEL_SPIKEGEN, 2, [10 * DT])
    with pytest.raises(HsError):
        net.set_spikegen(g, [0], [10])
    net.close()


@pytest.fixture(scope="module")
def stdp_poisson():
    return nw.stdp_poisson_example()


@pytest.mark.parametrize("prec", [64, 32])
def test_stdp_poisson_example_vs_oracle(hs, oracle_lib, stdp_poisson, prec):
    """examples/stdp.py at its size (10,000 Poisson inputs, 10 neurons, p = 0.1,
    STDP): spikes, weights and events bit for bit over 2000 steps."""
    spec = stdp_poisson
    net, sm, _ = nw.build_hs(spec, prec, "plain")
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, oracle_lib.SEQ)
    net.run(2000)
    ref.run(2000)
    gi, gt = net.spikes(sm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(ri) > 50
    assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    nsyn = net.pathway_info(0)["nsyn"]
    for var in range(4):
        assert np.array_equal(net.synapse_var(0, var), ref.synapse_var(0, var, nsyn)), var
    assert net.stats()["events"] == ref.events
    net.close()


def test_stdp_poisson_example_weights_split(hs, stdp_poisson):
    """What examples/stdp.py shows (Song, Miller & Abbott 2000): under STDP with
    depression slightly stronger than potentiation (dApost = -1.05 dApre) the
    uniform initial weights move toward both bounds and the output rate falls.
    Over 20 s: the share of weights within 10% of 0 or gmax grows from 0.20 to
    over 0.27, both tails grow, and the rate of the last 10 s is below the
    first's."""
    spec = stdp_poisson
    gmax = spec.info["gmax"]
    net, sm, _ = nw.build_hs(spec, 64, "plain")
    w0 = net.synapse_var(0, 0)
    assert abs(np.mean((w0 < 0.1 * gmax) | (w0 > 0.9 * gmax)) - 0.2) < 0.03
    net.run(100_000)
    n1 = len(net.spikes(sm[0])[0])
    net.run(100_000)
    n2 = len(net.spikes(sm[0])[0]) - n1
    w = net.synapse_var(0, 0)
    assert np.all((w >= 0) & (w <= gmax))
    assert np.mean((w < 0.1 * gmax) | (w > 0.9 * gmax)) > 0.27
    assert np.mean(w < 0.1 * gmax) > 0.15 and np.mean(w > 0.9 * gmax) > 0.1
    assert n2 < n1
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_synapse_monitor_vs_oracle(hs, oracle_lib, stdp_poisson, prec):
    """StateMonitor(S, 'w' / 'Apre' / 'Apost', record=...) of examples/stdp.py:182
    on the full-size example: every recorded row bit for bit against the oracle,
    and row k equals the variable read after k steps."""
    spec = stdp_poisson
    net, _, _ = nw.build_hs(spec, prec, "plain")
    ref, _, _ = oracle_lib.build_ref(spec, prec, oracle_lib.SEQ)
    nsyn = net.pathway_info(0)["nsyn"]
    sel = np.array([0, 1, 2, 3, 4, nsyn // 2, nsyn - 1])
    gm = [net.add_synapse_monitor(0, v, sel) for v in (rt.HS_SYN_W, rt.HS_SYN_APRE, rt.HS_SYN_APOST)]
    rm = [ref.add_synapse_monitor(0, v, sel) for v in (rt.HS_SYN_W, rt.HS_SYN_APRE, rt.HS_SYN_APOST)]
    net.run(700)
    ref.run(700)
    w700 = net.synapse_var(0, rt.HS_SYN_W)[sel]
    net.run(800)
    ref.run(800)
    for a, b in zip(gm, rm):
        g = net.state_monitor(a, sel.size)
        assert g.shape == (1500, sel.size)
        assert np.array_equal(g, ref.state_monitor(b, sel.size))
    np.testing.assert_array_equal(net.state_monitor(gm[0], sel.size)[700], w700)
    net.close()


def test_synapse_monitor_refusals(hs, stdp_poisson):
    spec = stdp_poisson
    net, _, _ = nw.build_hs(spec, 64, "plain")
    nsyn = net.pathway_info(0)["nsyn"]
    for args in ((0, rt.HS_SYN_W, [nsyn]), (0, rt.HS_SYN_W, [-1]), (0, rt.HS_SYN_LASTUPDATE, [0]), (1, 0, [0])):
        with pytest.raises(HsError):
            net.add_synapse_monitor(*args)
    net.close()
    fused = nw.cuba(N=300, seed=2)
    net, _, _ = nw.build_hs(fused, 64, "fused")
    net.run(10)  # built in the one-launch form: no per-step synapse monitor any more
    with pytest.raises(HsError):
        net.add_synapse_monitor(0, rt.HS_SYN_W, [0])
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_stdp_fused_form_vs_oracle(hs, oracle_lib, stdp_poisson, prec):
    """STDP in the fused (batched) form: one step per update launch, on_pre
    effects into the step's accumulator -- examples/stdp.py at full size against
    the oracle's accumulator semantics: spikes, weights, traces, last updates,
    events and a synapse StateMonitor with an interval, over 2000 steps."""
    spec = stdp_poisson
    net, sm, _ = nw.build_hs(spec, prec, "fused")
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, oracle_lib.ACC)
    sel = np.arange(0, 5000, 250)
    g = net.add_synapse_monitor(0, rt.HS_SYN_W, sel)
    r = ref.add_synapse_monitor(0, rt.HS_SYN_W, sel)
    net.set_monitor_interval(g, 10)
    ref.set_monitor_interval(r, 10)
    for k in (700, 1300):
        net.run(k)
        ref.run(k)
    gi, gt = net.spikes(sm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(ri) > 50
    assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    nsyn = net.pathway_info(0)["nsyn"]
    for var in range(4):
        assert np.array_equal(net.synapse_var(0, var), ref.synapse_var(0, var, nsyn)), var
    assert np.array_equal(net.state_monitor(g, sel.size), ref.state_monitor(r, sel.size))
    assert net.stats()["events"] == ref.events
    assert net.stats()["window"] == 1
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("delay", [0, 2])
def test_stdp_feedforward_fused_vs_oracle(hs, oracle_lib, prec, delay):
    spec = _feedforward(delay)
    net, sm, st = nw.build_hs(spec, prec, "fused")
    ref, rsm, rst = oracle_lib.build_ref(spec, prec, oracle_lib.ACC)
    net.run(1500)
    ref.run(1500)
    assert np.array_equal(net.spikes(sm[0])[0], ref.spikes(rsm[0])[0])
    nsyn = net.pathway_info(0)["nsyn"]
    for var in range(4):
        assert np.array_equal(net.synapse_var(0, var), ref.synapse_var(0, var, nsyn)), var
    assert np.array_equal(net.state_monitor(st[0], 20), ref.state_monitor(rst[0], 20))
    net.close()
