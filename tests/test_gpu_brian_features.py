"""GPU: brian2cuda's feature tests that the code generator can express, run
through the engine (code models compiled at run time) and checked against
their own expectations and against the oracle's bits:

  test_subexpressions       (brian2cuda/tests/test_stateupdaters.py:138-168)
  test_refractory           (test_stateupdaters.py:171-196)
  test_state_monitor_more_threads_than_single_block (test_monitor.py:9-20)
  test_spike_monitor_subgroups_all_attr (test_monitor.py:23-52): SpikeMonitors
      of a group and of three subgroups (recorded indices relative)
  test_group_variable_set_copy_to_host (test_neurongroup.py:8-20)
"""
import numpy as np
import pytest

from brian2cuda_amd import codegen as cg
from brian2cuda_amd import networks as nw
from brian2cuda_amd import runtime as rt

pytestmark = pytest.mark.gpu

DT = 1e-4


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("method", ["exponential_euler", "rk2", "rk4"])
def test_subexpressions(hs, oracle_lib, prec, method):
    m1, _ = cg.neuron_model("dv/dt = (-v + sin(2*pi*100*Hz*t)) / (10*ms) : 1", None, None, 0.0, method, {}, dt=DT)
    m2, _ = cg.neuron_model("dv/dt = I / (10*ms) : 1\nI = -v + sin(2*pi*100*Hz*t): 1", None, None, 0.0, method, {},
                            dt=DT)
    spec = nw.NetSpec("subexpr", DT, 1, [nw.GroupSpec(0, 1, [], {0: 1.0}, code=m1),
                                         nw.GroupSpec(0, 1, [], {0: 1.0}, code=m2)], [])
    spec.state_monitors = [(0, 0, 0, 1), (1, 0, 0, 1)]
    net, _, gst = nw.build_hs(spec, prec, "plain")
    net.run(100)
    a, b = net.state_monitor(gst[0], 1), net.state_monitor(gst[1], 1)
    assert a.shape == (100, 1) and np.array_equal(a, b)
    ref, _, rst = oracle_lib.build_ref(spec, prec, oracle_lib.SEQ)
    ref.run(100)
    assert np.array_equal(a, ref.state_monitor(rst[0], 1))
    net.close()


def _refractory_spec(method):
    """10 neurons of dv/dt = -v/(10 ms) from v = (i+1)/11, with and without
    '(unless refractory)' (refractory 1 ms, no threshold: never refractory)."""
    m_no, _ = cg.neuron_model("dv/dt = -v/(10*ms) : 1", None, None, 0.0, method, {}, dt=DT)
    m_ref, _ = cg.neuron_model("dv/dt = -v/(10*ms) : 1 (unless refractory)", None, None, 1e-3, method, {}, dt=DT)
    v0 = (np.arange(10) + 1) / 11.0
    return nw.NetSpec("refractory", DT, 1, [nw.GroupSpec(0, 10, [], {0: v0}, code=m_ref),
                                            nw.GroupSpec(0, 10, [], {0: v0}, code=m_no)], [])


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("method", ["linear", "exact", "independent", "euler", "exponential_euler", "rk2", "rk4"])
def test_refractory(hs, oracle_lib, prec, method):
    spec = _refractory_spec(method)
    net, _, _ = nw.build_hs(spec, prec, "plain")
    net.run(100)
    a, b = net.get_state(0, 0), net.get_state(1, 0)
    assert np.array_equal(a, b)
    ref, _, _ = oracle_lib.build_ref(spec, prec, oracle_lib.SEQ)
    ref.run(100)
    assert np.array_equal(a, ref.get_state(0, 0))
    net.close()


def test_state_monitor_more_threads_than_single_block(hs):
    m, _ = cg.neuron_model("v : 1", None, None, 0.0, None, {}, dt=DT)
    n = 2000
    spec = nw.NetSpec("statemon", DT, 1, [nw.GroupSpec(0, n, [], {0: np.arange(n, dtype=float)}, code=m)], [])
    spec.state_monitors = [(0, 0, 0, n)]
    net, _, gst = nw.build_hs(spec, 64, "plain")
    net.run(3)
    rows = net.state_monitor(gst[0], n)
    assert rows.shape == (3, n) and np.all(rows == np.arange(n))
    net.close()


@pytest.mark.parametrize("prec,engine", [(64, "plain"), (32, "fused")])
def test_spike_monitor_subgroups_all_attr(hs, prec, engine):
    m, _ = cg.neuron_model("do_spike : boolean", "do_spike", None, 0.0, None, {}, dt=DT)
    net = rt.Net(prec, engine, DT, 1)
    g = net.add_group_code(6, m)
    net.set_state(g, 0, np.array([1, 0, 0, 0, 1, 1], float))
    mons = [net.add_spike_monitor(g), net.add_spike_monitor(g, start=0, stop=2),
            net.add_spike_monitor(g, start=2, stop=4), net.add_spike_monitor(g, start=4, stop=6)]
    net.run(1)
    want = [([0, 4, 5], 6), ([0], 2), ([], 2), ([0, 1], 2)]  # recorded indices are relative to the subgroup
    for mid, (idx, size) in zip(mons, want):
        i, t = net.spikes(mid)
        assert list(i) == idx and np.all(t == 0)
        assert list(np.bincount(i, minlength=size)) == [1 if k in idx else 0 for k in range(size)]
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_group_variable_set_copy_to_host(hs, prec):
    m, names = cg.neuron_model("v : 1", None, None, 0.0, None, {}, dt=DT)
    c, _ = cg.connect_expression("v_pre == 50", {}, names, names)
    for v, nsyn in ((50.0, 1), (49.0, 0)):
        net = rt.Net(prec, "plain", DT, 1)
        g = net.add_group_code(1, m)
        net.set_state(g, 0, np.array([v]))
        pid = net.connect_code(g, g, 0, c, "", 1.0)
        assert net.pathway_info(pid)["nsyn"] == nsyn
        net.close()
