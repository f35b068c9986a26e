"""PoissonGroup (HS_MODEL_POISSON) on the GPU: bit for bit against the oracle,
alone and driving a CUBA group, in both forms and precisions; and every spike
decision against the standalone uniform draw (hs_rand) of the same
(seed, stream, neuron, step)."""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw
from brian2cuda_amd import runtime as rt

pytestmark = pytest.mark.gpu

DT = 1e-4


def _spec(with_target, seed=5):
    rng = np.random.default_rng(seed)
    groups = [nw.GroupSpec(rt.HS_MODEL_POISSON, 2000, [30.0, 1.0], {0: rng.uniform(0.0, 60.0, 2000)})]
    paths = []
    if with_target:
        groups.append(nw.GroupSpec(rt.HS_MODEL_LIF_CUBA, 500, nw.cuba_params(), {0: -60e-3, 1: 0.0, 2: 0.0}))
        pre, post = nw.bernoulli_connect(rng, 2000, 500, 0.05)
        d = rng.integers(1, 8, pre.size).astype(np.int32)
        paths.append(nw.PathSpec(0, 1, 1, pre, post, delay_steps=d, wmode=rt.HS_W_SCALAR, wscalar=1.5e-3))
    mons = [0, 1] if with_target else [0]
    return nw.NetSpec("poisson", DT, seed, groups, paths, spike_monitors=mons)


@pytest.mark.parametrize("with_target", [False, True])
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
def test_poisson_vs_oracle(hs, oracle_lib, with_target, prec, engine, sem):
    spec = _spec(with_target)
    net, sm, _ = nw.build_hs(spec, prec, engine)
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, sem)
    for k in (250, 351):
        net.run(k)
        ref.run(k)
    for a, b in zip(sm, rsm):
        gi, gt = net.spikes(a)
        ri, rt_ = ref.spikes(b)
        assert len(ri) > 0
        assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    for gid, g in enumerate(spec.groups):
        for v in range(rt.MODEL_NVARS[g.model]):
            assert np.array_equal(net.get_state(gid, v), ref.get_state(gid, v))
    assert net.stats()["events"] == ref.events


def test_poisson_decisions_match_the_uniform_draws(hs):
    """Spikes of step t are exactly {i : hs_rand(seed, stream, t)[i] < rates_i * dt}."""
    n, steps, seed, stream = 3000, 40, 99, 3
    rates = np.linspace(0.0, 2000.0, n)
    net = rt.Net(64, "plain", DT, seed)
    g = net.add_group(rt.HS_MODEL_POISSON, n, [0.0, float(stream)])
    net.set_state(g, 0, rates)
    m = net.add_spike_monitor(g)
    net.run(steps)
    i, t = net.spikes(m)
    for s in range(steps):
        want = np.nonzero(rt.rand(seed, stream, s, n) < rates * DT)[0]
        np.testing.assert_array_equal(i[t == s], want)
    assert i.size > 0
    net.close()


def test_poisson_parameters_refused(hs):
    from brian2cuda_amd._lib import HsError
    net = rt.Net(64, "plain", DT, 1)
    for params in ([-5.0], [1.0, 2.0, 3.0], []):
        with pytest.raises(HsError):
            net.add_group(rt.HS_MODEL_POISSON, 10, params)
    net.close()
