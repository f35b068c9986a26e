"""GPU: bundle tables built on the device from synapse arrays in HBM
(hs_net_add_pathway_device: device check, stable radix sort by (pre, delay))
equal the host-built tables entry for entry, and simulate to the oracle's bits."""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw
from brian2cuda_amd import runtime as rt
from brian2cuda_amd._lib import HsError

pytestmark = pytest.mark.gpu

DT = 1e-4


def _two_nets(prec, spec):
    out = []
    for dev in (False, True):
        net, _, _ = nw.build_hs(spec, prec, "plain", device_sort=dev)
        out.append(net)
    return out


def _shuffled(spec, seed):
    """the same synapses in a random creation order, with multapses added"""
    rng = np.random.default_rng(seed)
    for p in spec.paths:
        k = rng.permutation(p.nsyn)
        extra = rng.integers(0, p.nsyn, max(1, p.nsyn // 50))  # repeated (pre, post) pairs
        k = np.concatenate([k, extra])
        p.pre, p.post = p.pre[k], p.post[k]
        if p.delay_steps is not None:
            p.delay_steps = p.delay_steps[k]
        if p.wmode == rt.HS_W_SYNAPSE:
            p.weights = np.asarray(p.weights)[k] * (1 + 0.01 * rng.random(k.size))
    return spec


SPECS = {
    "brunel_delays": lambda: nw.brunel(N=3000, K=300, seed=4),
    "cuba_scalar_delay": lambda: nw.cuba(N=2000, seed=6),
    "batch_pre_weights": lambda: nw.batch_cuba(n_nets=10, nmin=200, nmax=1500, seed=3),
    "brunel_shuffled_multapses": lambda: _shuffled(nw.brunel(N=2500, K=200, seed=9), 1),
}


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("name", list(SPECS))
def test_device_table_equals_host_table(hs, prec, name):
    spec = SPECS[name]()
    host, dev = _two_nets(prec, spec)
    for pid in range(len(spec.paths)):
        assert dev.pathway_info(pid) == host.pathway_info(pid)
        for a, b in zip(dev.synapses(pid), host.synapses(pid)):
            np.testing.assert_array_equal(a, b)
    host.close()
    dev.close()


@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
def test_device_sorted_network_simulates_to_oracle_bits(hs, oracle_lib, engine, sem):
    spec = nw.brunel(N=3000, K=300, seed=12, monitors=True)
    net, sm, _ = nw.build_hs(spec, 32, engine, device_sort=True)
    ref, rsm, _ = oracle_lib.build_ref(spec, 32, sem)
    net.run(1200)
    ref.run(1200)
    gi, gt = net.spikes(sm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(gi) > 0 and np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert net.stats()["events"] == ref.events
    net.close()


def test_device_sort_empty_and_errors(hs):
    net = rt.Net(64, "plain", DT, 1)
    net.add_group(rt.HS_MODEL_LIF_CUBA, 100, nw.cuba_params())
    pid = net.add_pathway_device(0, 0, 1, np.zeros(0, np.int32), np.zeros(0, np.int32), None, 2)
    assert net.pathway_info(pid)["nsyn"] == 0
    for pre, post, d in [([0, 100], [1, 2], None), ([0, 1], [1, -1], None), ([0, 1], [1, 2], [3, -2]),
                         ([0, 1], [1, 2], [1, 5000])]:
        with pytest.raises(HsError):
            net.add_pathway_device(0, 0, 1, np.array(pre), np.array(post), None if d is None else np.array(d), 1)
    net.run(10)
    net.close()
