"""GPU parity at the edges of the input space, bit for bit against the oracle:
the largest delay range a pathway may have (4096 slots: delays of 1..4096
steps), a one-neuron group driving itself through an autapse, ragged group
sizes around the 256-neuron tile, pathways with no synapses, and a source
group that never spikes."""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw
from brian2cuda_amd import runtime as rt

pytestmark = pytest.mark.gpu

DT = 1e-4
ENGINES = [("plain", 0), ("fused", 1)]


def _compare(spec, prec, engine, sem, chunks, oracle_lib, options=None):
    net, sm, st = nw.build_hs(spec, prec, engine, options=options)
    ref, rsm, rst = oracle_lib.build_ref(spec, prec, sem)
    for k in chunks:
        net.run(k)
        ref.run(k)
    for a, b in zip(sm, rsm):
        gi, gt = net.spikes(a)
        ri, rt_ = ref.spikes(b)
        assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    for gid, g in enumerate(spec.groups):
        for v in range(rt.MODEL_NVARS[g.model]):
            assert np.array_equal(net.get_state(gid, v), ref.get_state(gid, v)), f"group {gid} var {v}"
    assert net.stats()["events"] == ref.events
    return net, ref


def _brunel_small(N, K, dlo_steps, dhi_steps, seed):
    rng = np.random.default_rng(seed)
    g = nw.GroupSpec(rt.HS_MODEL_LIF_NOISE, N, nw.brunel_params(), {0: 10e-3 + 5e-3 * rng.random(N)})
    pre, post = nw.bernoulli_connect(rng, N, N, K / N)
    d = rng.integers(dlo_steps, dhi_steps + 1, pre.size).astype(np.int32)
    d[0], d[-1] = dlo_steps, dhi_steps  # both ends of the range present
    ps = nw.PathSpec(0, 0, 0, pre, post, delay_steps=d, wmode=rt.HS_W_SCALAR, wscalar=-0.1e-3)
    return nw.NetSpec("brunel_edges", DT, seed, [g], [ps], spike_monitors=[0])


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_max_delay_range(hs, oracle_lib, prec, engine, sem):
    """Delays of 1..4096 steps: the largest bundle table a pathway accepts; the
    spike ring spans 4096 + 2 steps and the run crosses it."""
    spec = _brunel_small(500, 50, 1, 4096, seed=3)
    net, ref = _compare(spec, prec, engine, sem, [1000, 3500, 1], oracle_lib)
    assert net.pathway_info(0)["D"] == 4096
    assert ref.events > 0


def test_max_delay_range_gather(hs, oracle_lib):
    spec = _brunel_small(500, 50, 1, 4096, seed=4)
    _compare(spec, 64, "plain", 0, [2000, 2600], oracle_lib, options={5: 2})


def test_delay_range_over_the_limit_is_refused(hs):
    from brian2cuda_amd._lib import HsError
    spec = _brunel_small(100, 10, 1, 4097, seed=5)
    with pytest.raises(HsError):
        nw.build_hs(spec, 64, "plain")


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_single_neuron_autapse(hs, oracle_lib, prec, engine, sem):
    """One CUBA neuron (El above threshold: it fires periodically) exciting
    itself after 3 steps."""
    g = nw.GroupSpec(rt.HS_MODEL_LIF_CUBA, 1, nw.cuba_params(), {0: -55e-3})
    ps = nw.PathSpec(0, 0, 1, np.array([0], np.int32), np.array([0], np.int32), scalar_delay=3,
                     wmode=rt.HS_W_SCALAR, wscalar=1.62e-3)
    spec = nw.NetSpec("autapse", DT, 1, [g], [ps], spike_monitors=[0], state_monitors=[(0, 0, 0, 1)])
    net, ref = _compare(spec, prec, engine, sem, [7, 993, 1000], oracle_lib)
    i, t = net.spikes(0)
    assert len(i) >= 3 and np.all(i == 0)  # an ISI of about 45 ms


@pytest.mark.parametrize("n", [255, 256, 257, 513])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_ragged_group_sizes(hs, oracle_lib, n, engine, sem):
    spec = nw.cuba(N=n, seed=n, p=0.1)
    _compare(spec, 64, engine, sem, [300, 401], oracle_lib)


@pytest.mark.parametrize("engine,sem", ENGINES)
def test_empty_pathways_and_silent_source(hs, oracle_lib, engine, sem):
    """A pathway with no synapses, and one whose source group never spikes
    (v held far below threshold: El < Vt), beside a live network."""
    live = nw.cuba(N=800, seed=8)
    silent = nw.GroupSpec(rt.HS_MODEL_LIF_CUBA, 300,
                          nw.cuba_params(El=-70e-3, Vt=-50e-3, Vr=-60e-3), {0: -70e-3})
    groups = live.groups + [silent]
    rng = np.random.default_rng(2)
    pre, post = nw.bernoulli_connect(rng, 300, 800, 0.05)
    paths = live.paths + [
        nw.PathSpec(0, 1, 1, np.zeros(0, np.int32), np.zeros(0, np.int32), scalar_delay=1, wmode=rt.HS_W_SCALAR,
                    wscalar=1e-3),
        nw.PathSpec(1, 0, 2, pre, post, scalar_delay=2, wmode=rt.HS_W_SCALAR, wscalar=-9e-3),
    ]
    spec = nw.NetSpec("edges", DT, 8, groups, paths, spike_monitors=[0, 1])
    net, ref = _compare(spec, 64, engine, sem, [500, 500], oracle_lib)
    assert len(net.spikes(1)[0]) == 0 and len(net.spikes(0)[0]) > 0
