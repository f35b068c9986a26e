"""GPU network parity: the device step loop against the oracle on the same inputs.

Deterministic models (CUBA exact, COBAHH exponential Euler) and the noise model
(Philox draws are a pure function of (seed, neuron, step)) are compared bit for
bit: spike indices and steps, StateMonitor traces and the final state.
  plain engine  vs oracle SEQ (Brian2's in-place order)
  fused engine  vs oracle ACC (effects applied at the start of the next step)
"""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw

pytestmark = pytest.mark.gpu

ENGINES = [("plain", 0), ("fused", 1)]


def _compare(spec, prec, engine, sem, steps, oracle_lib, options=None, chunks=None):
    net, smon, stmon = nw.build_hs(spec, prec, engine, options=options)
    ref, rsmon, rstmon = oracle_lib.build_ref(spec, prec, sem)
    for k in (chunks or [steps]):
        net.run(k)
        ref.run(k)
    for a, b in zip(smon, rsmon):
        gi, gt = net.spikes(a)
        ri, rt = ref.spikes(b)
        assert len(gi) == len(ri), f"{len(gi)} spikes on the GPU vs {len(ri)} in the oracle"
        assert np.array_equal(gt, rt) and np.array_equal(gi, ri)
    for (a, b, m) in zip(stmon, rstmon, spec.state_monitors):
        w = nw.state_monitor_width(m)
        assert np.array_equal(net.state_monitor(a, w), ref.state_monitor(b, w), equal_nan=True)
    # equal_nan: the reference's HH rate alpha_n = 0.032(15 mV - v + VT)/(exp(...) - 1)
    # is 0/0 where v lands exactly on VT + 15 mV (it happens in FP32), and that
    # neuron's state is NaN on both sides
    for gid, g in enumerate(spec.groups):
        for v in range(len(g.init) if g.model != 3 else 6):
            assert np.array_equal(net.get_state(gid, v), ref.get_state(gid, v), equal_nan=True), f"group {gid} var {v}"
    assert net.stats()["events"] == ref.events
    return net, ref


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_cuba_parity(hs, oracle_lib, prec, engine, sem):
    """C1 at full size (4000 neurons, p=0.02, delay one dt), 0.3 s."""
    spec = nw.cuba(N=4000, seed=11)
    net, ref = _compare(spec, prec, engine, sem, 3000, oracle_lib)
    assert ref.events > 10000


@pytest.mark.parametrize("engine,sem", ENGINES)
def test_cuba_chunked_runs_and_options(hs, oracle_lib, engine, sem):
    """Several run() calls (graph chunks of odd sizes) and another block size give the same bits."""
    spec = nw.cuba(N=1500, seed=5)
    _compare(spec, 64, engine, sem, None, oracle_lib, options={0: 128, 2: 97}, chunks=[1, 250, 333, 416])


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_cobahh_parity(hs, oracle_lib, prec, engine, sem):
    """C2 at 4000 neurons (p=0.02, delay 0: two launches per step in both forms), 0.1 s."""
    spec = nw.cobahh(N=4000, seed=3)
    _compare(spec, prec, engine, sem, 1000, oracle_lib)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_brunel_parity_small(hs, oracle_lib, prec, engine, sem):
    """C3 scaled to N=4000, K=400 (heterogeneous delays 5..50 steps, noise), 0.2 s."""
    spec = nw.brunel(N=4000, K=400, seed=2, monitors=True)
    net, ref = _compare(spec, prec, engine, sem, 2000, oracle_lib)
    assert ref.events > 0


@pytest.mark.parametrize("pipelined", [1, 0])
@pytest.mark.parametrize("window", [1, 2, 3, 5])
@pytest.mark.parametrize("prec", [64, 32])
def test_brunel_windows(hs, oracle_lib, window, prec, pipelined):
    """The windowed fused forms are bit-identical to the oracle's ACC semantics for
    every window <= the minimum delay (5 steps), including runs whose length is
    not a multiple of the window:
      pipelined = 1: one launch per window (update + push of delays < 2W, pull of
                     the longer delays for the next window, next window's noise)
      pipelined = 0: delivery of w steps, then w update steps per tile"""
    spec = nw.brunel(N=3000, K=300, seed=4, monitors=True)
    net, ref = _compare(spec, prec, "fused", 1, None, oracle_lib, options={4: window, 2: 64, 9: pipelined},
                        chunks=[7, 600, 393, 1000])
    s = net.stats()
    assert s["window"] == window
    assert s["spikes"] == len(ref.spikes(0)[0])
    if window >= 2:
        assert s["launches_per_step"] == pytest.approx((1.0 if pipelined else 2.0) / window)


@pytest.mark.parametrize("dlo,dhi", [(0.2e-3, 0.3e-3), (0.2e-3, 0.7e-3), (0.4e-3, 2.5e-3)])
def test_pipelined_near_far_split(hs, oracle_lib, dlo, dhi):
    """Delay spans entirely below 2W (push only), straddling 2W, and mostly above
    it (pull only beyond 2W), with the window at the minimum delay."""
    spec = nw.brunel(N=2000, K=200, seed=12, monitors=True, dlo=dlo, dhi=dhi)
    for prec in (64, 32):
        net, ref = _compare(spec, prec, "fused", 1, None, oracle_lib, options={2: 50}, chunks=[3, 333, 164])
        assert net.stats()["window"] == int(round(dlo / 1e-4))


@pytest.mark.parametrize("delivery", [1, 2])
@pytest.mark.parametrize("engine,sem", ENGINES)
@pytest.mark.parametrize("prec", [64, 32])
def test_brunel_scatter_and_gather(hs, oracle_lib, delivery, engine, sem, prec):
    """Both delivery modes (per-(pre,delay) bundles with global atomics, and per
    target-tile gather with LDS accumulation) give the oracle's bits; the gather
    tile edge is exercised by N not a multiple of the 256-neuron tile."""
    spec = nw.brunel(N=2900, K=290, seed=8, monitors=True)
    _compare(spec, prec, engine, sem, None, oracle_lib, options={5: delivery}, chunks=[500, 700])


def test_cuba_gather_forced(hs, oracle_lib):
    """Gather on a homogeneous one-step delay (single-delay segments)."""
    spec = nw.cuba(N=1300, seed=21)
    _compare(spec, 64, "plain", 0, 1200, oracle_lib, options={5: 2})
    _compare(spec, 32, "fused", 1, 1200, oracle_lib, options={5: 2})


def test_window_is_capped_by_min_delay(hs):
    spec = nw.brunel(N=2000, K=100, seed=1, dlo=0.3e-3, dhi=1e-3)  # delays 3..10 steps
    net, _, _ = nw.build_hs(spec, 32, "fused", options={4: 16})
    net.run(10)
    assert net.stats()["window"] == 3


@pytest.mark.parametrize("engine,sem", ENGINES)
def test_batch_parity_small(hs, oracle_lib, engine, sem):
    """C4 scaled to 12 networks of 200..1500 neurons, FP32, 0.1 s."""
    spec = nw.batch_cuba(n_nets=12, nmin=200, nmax=1500, seed=9)
    _compare(spec, 32, engine, sem, 1000, oracle_lib)


@pytest.mark.parametrize("engine,sem,delivery,pipelined",
                         [("plain", 0, 1, 0), ("fused", 1, 2, 0), ("fused", 1, 1, 0), ("fused", 1, 1, 1)])
def test_kernel_timing_keeps_results(hs, oracle_lib, engine, sem, delivery, pipelined):
    """Kernel timing (no graphs, an event pair per launch) changes no bits and
    counts the launches of each kind: an update and a delivery launch per window,
    or one window launch (pipelined fused form)."""
    from brian2cuda_amd import runtime as rt
    spec = nw.brunel(N=2500, K=250, seed=6, monitors=True)
    net, _ = _compare(spec, 32, engine, sem, None, oracle_lib, options={5: delivery, 7: 1, 9: pipelined},
                      chunks=[300, 211])
    w = net.stats()["window"]
    chunk = net.get_option(rt.OPT_GRAPH_CHUNK)  # a run advances in chunks; a window never spans two
    pieces = [min(chunk, k - j) for k in (300, 211) for j in range(0, k, chunk)]
    launches = sum(-(-p // w) for p in pieces)
    counts = {k: net.kernel_time(k) for k in rt.KERNEL_NAMES}
    if pipelined:
        expect = {rt.KT_WINDOW: launches}
    else:
        expect = {rt.KT_UPDATE: launches, (rt.KT_SCATTER if delivery == 1 else rt.KT_GATHER): launches}
    for k, (ms, n) in counts.items():
        assert n == expect.get(k, 0), (rt.KERNEL_NAMES[k], n)
        assert (ms > 0) == (k in expect)


@pytest.mark.parametrize("case", ["push_outer_loop", "push_outer_loop2", "circular_eventspaces", "bundle_prefs"])
@pytest.mark.parametrize("engine", ["plain", "fused"])
@pytest.mark.parametrize("prec", [64, 32])
def test_reference_propagation_cases_gpu(hs, case, engine, prec):
    """The reference's spike-propagation expectations (test_synaptic_propagations.py),
    checked on the GPU directly against the reference's expected index sets."""
    import refcases as rc
    spec, steps, expect = getattr(rc, case)()
    net, sm, _ = nw.build_hs(spec, prec, engine)
    net.run(steps)
    got = rc.spikes_by_step(*net.spikes(sm[0]))
    assert sorted(got) == sorted(expect)
    for k in expect:
        np.testing.assert_array_equal(got[k], expect[k])


@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
@pytest.mark.parametrize("prec", [64, 32])
def test_reference_circular_eventspaces_spikegenerator_gpu(hs, engine, sem, prec):
    import refcases as rc
    spec, steps, expect = rc.circular_eventspaces_spikegenerator(sem)
    net, sm, _ = nw.build_hs(spec, prec, engine)
    net.run(steps)
    got = rc.spikes_by_step(*net.spikes(sm[0]))
    assert sorted(got) == sorted(expect)
    for k in expect:
        np.testing.assert_array_equal(got[k], expect[k])


@pytest.mark.parametrize("delays", [[1], [0, 1]])
@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_reference_multiple_runs_gpu(hs, delays, engine):
    import refcases as rc
    spec, expect = rc.multiple_runs(delays)
    net, _, st = nw.build_hs(spec, 64, engine)
    net.run(2)
    net.run(2)
    np.testing.assert_array_equal(net.state_monitor(st[0], len(delays)).T, expect)


@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_reference_effects_cancel_gpu(hs, engine):
    """test_synaptic_effect_modes: +1 and -1 onto every target each step cancel exactly."""
    import refcases as rc
    spec, steps, _ = rc.effects_cancel()
    net, sm, _ = nw.build_hs(spec, 64, engine)
    net.run(steps)
    assert len(net.spikes(sm[0])[0]) == 0
    assert np.all(net.get_state(1, 0) == 0.0)
    assert net.stats()["events"] == 2 * 15360 * (steps - 1)


@pytest.mark.parametrize("engine,sem,opts", [("plain", 0, {}), ("fused", 1, {}), ("fused", 1, {9: 1}),
                                             ("fused", 1, {5: 2}), ("plain", 0, {5: 2})])
def test_large_group_sub_lists(hs, oracle_lib, engine, sem, opts):
    """A group of 140,000 neurons (547 chunks of 256) splits each step's spike
    list into sub-lists; every delivery form still gives the oracle's bits."""
    spec = nw.brunel(N=140_000, K=20, seed=31, monitors=True)
    for prec in (32, 64):
        _compare(spec, prec, engine, sem, None, oracle_lib, options=opts, chunks=[150, 101])


def test_batch_sub_lists(hs, oracle_lib):
    """40 CUBA networks of up to 16,000 neurons in one index space (about 700
    chunks): sub-listed spike rings in the single-launch fused form."""
    spec = nw.batch_cuba(n_nets=40, seed=17)
    assert spec.n_neurons >= 512 * 256
    _compare(spec, 32, "fused", 1, 400, oracle_lib)
    _compare(spec, 64, "plain", 0, 400, oracle_lib)


@pytest.fixture(scope="module")
def batch_full():
    """The timed batch: 256 CUBA networks of 500-16,000 neurons (1,164,156 in
    all, 2.0e8 synapses), seed 1234."""
    return nw.batch_cuba(n_nets=256, seed=1234)


@pytest.mark.parametrize("prec,engine,sem", [(32, "fused", 1), (64, "plain", 0)])
def test_batch_full_size(hs, oracle_lib, batch_full, prec, engine, sem):
    _compare(batch_full, prec, engine, sem, None, oracle_lib, chunks=[120, 80])


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_cobahh_64k(hs, oracle_lib, prec, engine, sem):
    """COBAHH at 64,000 neurons (5.1e6 synapses), the largest size timed, over
    300 steps in two chunks.  In FP32 a few neurons reach v = VT + 15 mV
    exactly, where the reference's alpha_n is 0/0: their NaN state must match
    too."""
    spec = nw.cobahh(N=64000, seed=1234)
    net, ref = _compare(spec, prec, engine, sem, None, oracle_lib, chunks=[180, 120])
    if prec == 64:
        assert not np.isnan(ref.get_state(0, 0)).any()



@pytest.mark.parametrize("engine,options", [("plain", None), ("plain", {1: 1}), ("plain", {1: 4096}),
                                            ("plain", {5: 2}), ("fused", {0: 64}), ("fused", {1: 2048})])
def test_reference_bundle_prefs_launch_shapes(hs, engine, options):
    """test_threads_per_synapse_bundle_prefs varies the threads per bundle; here
    the delivery grid (1 block, 4096 blocks), gather delivery and update tiles
    vary, and neuron 3 still spikes at step 4, neuron 4 at step 5."""
    import refcases as rc
    spec, steps, expect = rc.bundle_prefs()
    net, sm, _ = nw.build_hs(spec, 64, engine, options=options)
    net.run(steps)
    got = rc.spikes_by_step(*net.spikes(sm[0]))
    assert sorted(got) == sorted(expect)
    for k in expect:
        np.testing.assert_array_equal(got[k], expect[k])


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_reference_ring_gpu(hs, oracle_lib, prec, engine, sem):
    """test_cuda_standalone's ring: bit for bit against the oracle, the first
    spike at t = 0 and at least 17,000 spikes (18,068 in FP64)."""
    import refcases as rc
    spec, steps = rc.ring()
    net, ref = _compare(spec, prec, engine, sem, steps, oracle_lib)
    i, t = net.spikes(0)
    assert len(i) >= 17000 and t[0] == 0
    if prec == 64:
        assert len(i) == 18068


@pytest.mark.parametrize("prec,rtol", [(64, 1e-7), (32, 2e-6)])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_reference_example_run_gpu(hs, oracle_lib, prec, rtol, engine, sem):
    """brian2cuda's example_run: v == arange(N)/N exp(-1 ms/tau) after 1 ms."""
    import refcases as rc
    spec, steps, expect = rc.example_run()
    net, ref = _compare(spec, prec, engine, sem, steps, oracle_lib)
    np.testing.assert_allclose(net.get_state(0, 0), expect, rtol=rtol)


def test_cuba_singular_time_constants_refused(hs):
    """taue or taui equal to taum: the closed form divides by their difference."""
    from brian2cuda_amd._lib import HsError
    net = hs.Net(64, "plain", 1e-4, 1)
    with pytest.raises(HsError):
        net.add_group(hs.HS_MODEL_LIF_CUBA, 10, nw.cuba_params(taum=10e-3, taui=10e-3))
    net.close()


@pytest.mark.parametrize("engine,options", [("plain", None), ("fused", None), ("fused", {9: 1}), ("fused", {2: 0})])
def test_prepare_then_run_same_bits(hs, engine, options):
    """hs_net_prepare instantiates the graphs a run launches without advancing
    the network: prepare + run gives the bits of run alone, and the step count
    is untouched by prepare."""
    spec = nw.brunel(N=3000, K=100, seed=9, monitors=True)
    out = []
    for prep in (False, True):
        net, smon, _ = nw.build_hs(spec, 32, engine, options=options)
        net.run(300)
        if prep:
            net.prepare(517)
            assert net.stats()["steps"] == 300
        net.run(517)
        out.append((*net.spikes(smon[0]), net.get_state(0, 0), net.stats()["events"]))
        net.close()
    (i0, t0, v0, e0), (i1, t1, v1, e1) = out
    assert len(i0) > 100
    assert np.array_equal(i0, i1) and np.array_equal(t0, t1) and np.array_equal(v0, v1) and e0 == e1


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("which", ["cuba", "cobahh", "cobahh_64k", "batch"])
def test_pipelined_single_step_windows(hs, oracle_lib, prec, which):
    """The pipelined form with windows of one step (option 9): one launch per
    step whose update blocks push the synapses of delays 0 and 1 of their own
    spikes -- for COBAHH (delay 0, otherwise two launches per step), CUBA and the
    batch (delay 1) -- equals the oracle's ACC semantics bit for bit, over
    chunked runs."""
    spec = {"cuba": lambda: nw.cuba(N=4000, seed=11), "cobahh": lambda: nw.cobahh(N=4000, seed=3),
            "cobahh_64k": lambda: nw.cobahh(N=64000, seed=1234),
            "batch": lambda: nw.batch_cuba(n_nets=24, seed=8)}[which]()
    net, ref = _compare(spec, prec, "fused", 1, None, oracle_lib, options={9: 1}, chunks=[1, 256, 343])
    st = net.stats()
    assert st["window"] == 1 and st["launches_per_step"] <= 1.0 + 1e-9
    assert ref.events > 1000


@pytest.mark.parametrize("opts", [{11: 1}, {11: 1, 9: 1}])
def test_cobahh_fast_math_error_vs_fp64(hs, oracle_lib, opts):
    """The FP32 fast form of the COBA HH update (device exponential, hardware
    reciprocal, FMA contraction; option 11) is not reproducible by the oracle
    bit for bit: its error is bounded against the FP64 oracle instead.  After 50
    steps almost every neuron is within 1 mV of FP64 (tolerance 1% of neurons
    beyond it), and over 1000 steps the spike count is within 3% of FP64's."""
    spec = nw.cobahh(N=4000, seed=3)
    net, sm, _ = nw.build_hs(spec, 32, "fused", options=opts)
    ref, rsm, _ = oracle_lib.build_ref(spec, 64, oracle_lib.SEQ)
    net.run(50)
    ref.run(50)
    dv = np.abs(net.get_state(0, 0) - ref.get_state(0, 0))
    assert np.isfinite(dv).mean() > 0.999
    assert np.mean(dv > 1e-3) < 0.01, f"{np.mean(dv > 1e-3):.4f} of neurons beyond 1 mV"
    net.run(950)
    ref.run(950)
    ng, nr = len(net.spikes(sm[0])[0]), len(ref.spikes(rsm[0])[0])
    assert nr > 1000 and abs(ng - nr) < 0.03 * nr, (ng, nr)
    net.close()


def test_fast_math_is_an_fp32_option(hs):
    from brian2cuda_amd._lib import HsError
    net = nw.build_hs(nw.cobahh(N=100, seed=1), 64, "fused")[0]
    with pytest.raises(HsError):
        net.set_option(11, 1)
    net.close()
