"""CPU tests that pin the oracle's plasticity, spike generator and unclamped
refractory LIF against closed forms (no GPU):

  - a spike generator fires exactly its pattern (and repeats it with a period);
  - STDP weight and trace trajectories for hand-placed pre/post spike times equal
    the closed forms with numpy's exp (the on_pre/on_post statement order, trace
    decay from the last update, pre pathway before post pathway in one step);
  - CUBA with free = 1 integrates v while refractory (only the threshold is
    blocked); free = 0 holds it.
"""
import math

import numpy as np
import pytest

import oracle
from brian2cuda_amd import networks as nw
from brian2cuda_amd.runtime import HS_MODEL_LIF_CUBA, HS_MODEL_SPIKEGEN, HS_SYN_APOST, HS_SYN_APRE, HS_SYN_LASTUPDATE, \
    HS_SYN_W, HS_W_SYNAPSE

DT = 1e-4


def test_spikegen_fires_its_pattern():
    net = oracle.RefNet(64, oracle.SEQ, DT, 1)
    g = net.add_group(HS_MODEL_SPIKEGEN, 10, [0.0])
    idx = np.array([3, 1, 3, 7, 0, 9])
    steps = np.array([5, 5, 6, 0, 12, 5])
    net.set_spikegen(g, idx, steps)
    m = net.add_spike_monitor(g)
    net.run(20)
    i, t = net.spikes(m)
    got = sorted(zip(t.tolist(), i.tolist()))
    assert got == sorted(zip(steps.tolist(), idx.tolist()))
    with pytest.raises(ValueError):
        net.set_spikegen(g, [2, 2], [4, 4])  # one spike per neuron and step


def test_spikegen_period_repeats():
    net = oracle.RefNet(64, oracle.SEQ, DT, 1)
    g = net.add_group(HS_MODEL_SPIKEGEN, 4, [7 * DT])  # period 7 steps
    net.set_spikegen(g, [0, 2], [1, 4])
    m = net.add_spike_monitor(g)
    net.run(30)
    i, t = net.spikes(m)
    exp = sorted([(s, 0) for s in range(1, 30, 7)] + [(s, 2) for s in range(4, 30, 7)])
    assert sorted(zip(t.tolist(), i.tolist())) == exp


def _pair_net(pre_steps, post_steps, w0, params, delay=0, prec=64):
    """generator A (1 neuron) -> generator B (1 neuron), one STDP synapse; B's
    placeholder variable receives the on_pre effects"""
    net = oracle.RefNet(prec, oracle.SEQ, DT, 1)
    a = net.add_group(HS_MODEL_SPIKEGEN, 1, [0.0])
    b = net.add_group(HS_MODEL_SPIKEGEN, 1, [0.0])
    net.set_spikegen(a, np.zeros(len(pre_steps), int), pre_steps)
    net.set_spikegen(b, np.zeros(len(post_steps), int), post_steps)
    pid = net.add_pathway(a, b, 0, [0], [0], [delay], 0, HS_W_SYNAPSE, [w0], 0.0)
    net.pathway_stdp(pid, params)
    return net, pid


PARAMS = (20e-3, 15e-3, 0.03, -0.035, -np.inf, np.inf, 1e-3)


def test_stdp_pre_post_pre_closed_form():
    taupre, taupost, dApre, dApost, _, _, scale = PARAMS
    w0 = 0.5
    net, pid = _pair_net([10, 45], [30], w0, PARAMS)
    net.run(31)
    a = dApre * math.exp(-20 * DT / taupre)  # Apre seen by the post spike at 30
    assert net.synapse_var(pid, HS_SYN_W, 1)[0] == pytest.approx(w0 + a, rel=1e-14)
    assert net.synapse_var(pid, HS_SYN_APOST, 1)[0] == pytest.approx(dApost, rel=1e-14)
    assert net.synapse_var(pid, HS_SYN_LASTUPDATE, 1)[0] == 30
    net.run(20)
    b = dApost * math.exp(-15 * DT / taupost)
    a2 = a * math.exp(-15 * DT / taupre) + dApre
    assert net.synapse_var(pid, HS_SYN_W, 1)[0] == pytest.approx(w0 + a + b, rel=1e-13)
    assert net.synapse_var(pid, HS_SYN_APRE, 1)[0] == pytest.approx(a2, rel=1e-13)
    assert net.synapse_var(pid, HS_SYN_APOST, 1)[0] == pytest.approx(b, rel=1e-13)
    # the effects on the target used the weight before each update: w0, then w0 + a
    assert net.get_state(1, 0)[0] == pytest.approx((w0 + (w0 + a)) * scale, rel=1e-13)


def test_stdp_same_step_runs_pre_before_post():
    """Pre and post spike in one step: on_pre (w += Apost, Apre += dApre), then
    on_post (Apost += dApost, w += Apre) -- the post pathway sees the new Apre."""
    taupre, taupost, dApre, dApost = PARAMS[:4]
    net, pid = _pair_net([12], [12], 1.0, PARAMS)
    net.run(13)
    assert net.synapse_var(pid, HS_SYN_W, 1)[0] == pytest.approx(1.0 + dApre, rel=1e-14)
    assert net.synapse_var(pid, HS_SYN_APOST, 1)[0] == pytest.approx(dApost, rel=1e-14)


def test_stdp_delay_and_clip():
    """A delayed pre spike arrives d steps later; weights are clipped to [wmin, wmax]."""
    taupre, taupost, dApre, dApost = PARAMS[:4]
    params = (taupre, taupost, dApre, dApost, 0.0, 0.52, 1e-3)
    net, pid = _pair_net([10], [20, 40], 0.5, params, delay=7)  # arrival at 17
    net.run(18)
    assert net.synapse_var(pid, HS_SYN_LASTUPDATE, 1)[0] == 17
    assert net.synapse_var(pid, HS_SYN_APRE, 1)[0] == dApre
    assert net.get_state(1, 0)[0] == pytest.approx(0.5e-3, rel=1e-15)
    net.run(3)
    a = dApre * math.exp(-3 * DT / taupre)
    assert net.synapse_var(pid, HS_SYN_APRE, 1)[0] == pytest.approx(a, rel=1e-14)
    assert 0.5 + a > 0.52 and net.synapse_var(pid, HS_SYN_W, 1)[0] == 0.52  # clipped at wmax
    net.run(30)
    assert net.synapse_var(pid, HS_SYN_W, 1)[0] == 0.52


def test_stdp_clip_bounds_hit():
    params = (20e-3, 20e-3, 0.3, -0.9, 0.0, 1.0, 1e-3)
    net, pid = _pair_net([5], [6, 8], 0.9, params)
    net.run(7)
    assert net.synapse_var(pid, HS_SYN_W, 1)[0] == 1.0  # 0.9 + 0.3 e^{-dt/tau} clipped
    net.run(2)
    w = net.synapse_var(pid, HS_SYN_W, 1)[0]
    assert w == pytest.approx(min(1.0, 1.0 + 0.3 * math.exp(-3 * DT / 20e-3)), rel=1e-14)


def test_cuba_free_integrates_while_refractory():
    El, Vt, Vr, taum, taue, taui, ref = -49e-3, -50e-3, -60e-3, 20e-3, 5e-3, 5e-3, 5e-3
    v0 = np.array([-55e-3])
    last = np.array([0], np.int32)  # spiked at step 0: refractory for 50 steps
    held = oracle.stateupdate(HS_MODEL_LIF_CUBA, 64, [v0, [0.0], [0.0]], last, [El, Vt, Vr, taum, taue, taui, ref],
                              DT, 0, 10)[0]
    free = oracle.stateupdate(HS_MODEL_LIF_CUBA, 64, [v0, [0.0], [0.0]], last,
                              [El, Vt, Vr, taum, taue, taui, ref, 1.0], DT, 0, 10)[0]
    assert held[0] == v0[0]
    assert free[0] == pytest.approx(El + (v0[0] - El) * math.exp(-DT / taum), rel=1e-14)


def test_stdp_example_network_runs_and_learns():
    """The reference's STDP networks on the oracle: both pathways deliver
    (T: one event per generator spike; S: 100 per P spike), weights move."""
    for het in (False, True):
        spec = nw.stdp_example(het)
        ref, sm, _ = oracle.build_ref(spec, 64, oracle.SEQ)
        w0 = ref.synapse_var(1, HS_SYN_W, 10000).copy()
        ref.run(2000)
        i, t = ref.spikes(sm[0])
        gen = spec.groups[1].spikes[1]
        S = spec.paths[1]
        d = np.zeros(S.nsyn, np.int64) if S.delay_steps is None else S.delay_steps.astype(np.int64)
        # S events: synapses of each spike whose arrival t + delay falls inside the run
        arrived = sum(int(np.sum(t_k + d[S.pre == i_k] < 2000)) for i_k, t_k in zip(i, t))
        assert ref.events == int(np.sum(gen < 2000)) + arrived
        assert het or arrived == 100 * len(i)
        assert len(i) > 100
        w1 = ref.synapse_var(1, HS_SYN_W, 10000)
        assert np.mean(np.abs(w1 - w0)) > 1e-3


# ---- the reference's SpikeGeneratorGroup expectations (test_spikegenerator.py) ----

def _gen_run(n, indices, times, period, steps, prec=64):
    from brian2cuda_amd.runtime import spikegen_steps
    net = oracle.RefNet(prec, oracle.SEQ, DT, 1)
    g = net.add_group(HS_MODEL_SPIKEGEN, n, [period])
    net.set_spikegen(g, indices, spikegen_steps(times, DT))
    m = net.add_spike_monitor(g)
    net.run(steps)
    return net.spikes(m)


def test_reference_spikegenerator_extreme_period():
    """test_spikegenerator_extreme_period: period 1e6 s, spikes at 0, 1, 2 ms."""
    i, t = _gen_run(5, [0, 1, 2], np.array([0, 1, 2]) * 1e-3, 1e6, 100)
    assert i.tolist() == [0, 1, 2]
    np.testing.assert_allclose(t * DT, [0, 1e-3, 2e-3])


def test_reference_spikegenerator_period_repeat():
    """test_spikegenerator_period_repeat: 10 spikes of neuron 0 every 0.1 ms,
    period 1 ms, five runs of 1 ms -> 10 spikes in each."""
    i, t = _gen_run(1, np.zeros(10, int), np.arange(0, 1, 0.1) * 1e-3, 1e-3, 50)
    assert np.array_equal(np.bincount(t // 10, minlength=5), [10] * 5)


def test_reference_spikegenerator_rounding():
    """test_spikegenerator_rounding: 100 times in [0, dt) all fall in the first
    bin; 10,000 times k*dt fall in 10,000 separate bins."""
    i, t = _gen_run(100, np.arange(100), np.linspace(0, 0.1e-3, 100, endpoint=False), 0.0, 1)
    assert np.array_equal(np.bincount(i, minlength=100), np.ones(100))
    i, t = _gen_run(1, np.zeros(10000, int), np.arange(10000) * DT, 0.0, 10000)
    assert np.array_equal(t, np.arange(10000))


def test_spikegen_rejects_times_outside_the_period():
    net = oracle.RefNet(64, oracle.SEQ, DT, 1)
    g = net.add_group(HS_MODEL_SPIKEGEN, 2, [10 * DT])
    with pytest.raises(ValueError):
        net.set_spikegen(g, [0], [10])


def test_synapse_monitor_records_at_step_start():
    """A synapse StateMonitor (when='start') records, at the start of step k, the
    values the synapses hold after k steps: row k equals the variable read after
    run(k), for w and both traces."""
    spec = nw.stdp_example(n_cells=40, seed=3)
    ref, _, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    pid = len(spec.paths) - 1  # the STDP pathway S
    nsyn = spec.paths[pid].pre.size
    sel = np.array([0, 7, nsyn // 2, nsyn - 1])
    mids = [ref.add_synapse_monitor(pid, v, sel) for v in (0, 1, 2)]
    snaps = []
    for k in (300, 450, 250):
        ref.run(k)
        snaps.append((ref.step, [ref.synapse_var(pid, v, nsyn)[sel] for v in (0, 1, 2)]))
    ref.run(1)
    changed = False
    for v, mid in enumerate(mids):
        rows = ref.state_monitor(mid, sel.size)
        assert rows.shape == (ref.step, sel.size)
        for step, vals in snaps:
            np.testing.assert_array_equal(rows[step], vals[v])
        changed |= bool(np.any(rows[0] != rows[-1]))
    assert changed
    with pytest.raises(ValueError):
        ref.add_synapse_monitor(pid, 0, [nsyn])
    ref.close()
