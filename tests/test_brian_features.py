"""CPU (oracle): brian2cuda's feature tests that the code generator can now
express, restated on the oracle with the generator's statements (the GPU runs
the same cases in tests/test_gpu_brian_features.py):

  test_subexpressions       (brian2cuda/tests/test_stateupdaters.py:138-168):
      dv/dt = (-v + sin(2 pi 100 Hz t))/(10 ms) written directly and through a
      subexpression I give equal v, for exponential_euler, rk2 and rk4;
  test_refractory           (test_stateupdaters.py:171-196): a neuron that never
      spikes integrates the same with and without '(unless refractory)', for
      linear, exact, independent, euler, exponential_euler, rk2 and rk4;
  test_state_monitor_more_threads_than_single_block (test_monitor.py:9-20):
      2000 neurons of 'v : 1' recorded for 3 steps keep v = arange(n);
  test_group_variable_set_copy_to_host (test_neurongroup.py:8-20): v set to 50,
      connect(condition='v_pre == 50') creates one synapse.
"""
import numpy as np
import pytest

import oracle
from brian2cuda_amd import codegen as cg
from brian2cuda_amd import networks as nw

DT = 1e-4


def _run(spec, steps, prec=64):
    ref, sm, st = oracle.build_ref(spec, prec, oracle.SEQ)
    ref.run(steps)
    return ref, sm, st


@pytest.mark.parametrize("method", ["exponential_euler", "rk2", "rk4"])
def test_subexpressions(method):
    m1, _ = cg.neuron_model("dv/dt = (-v + sin(2*pi*100*Hz*t)) / (10*ms) : 1", None, None, 0.0, method, {}, dt=DT)
    m2, _ = cg.neuron_model("dv/dt = I / (10*ms) : 1\nI = -v + sin(2*pi*100*Hz*t): 1", None, None, 0.0, method, {},
                            dt=DT)
    spec = nw.NetSpec("subexpr", DT, 1, [nw.GroupSpec(0, 1, [], {0: 1.0}, code=m1),
                                         nw.GroupSpec(0, 1, [], {0: 1.0}, code=m2)], [])
    spec.state_monitors = [(0, 0, 0, 1), (1, 0, 0, 1)]
    ref, _, st = _run(spec, 100)
    a, b = ref.state_monitor(st[0], 1), ref.state_monitor(st[1], 1)
    assert a.shape == (100, 1) and np.array_equal(a, b)
    assert np.all(np.abs(a) <= 1.0) and abs(a[-1, 0] - a[0, 0]) > 0.1  # it moved


METHODS = ["linear", "exact", "independent", "euler", "exponential_euler", "rk2", "rk4"]


def refractory_spec(method):
    m_no, _ = cg.neuron_model("dv/dt = -v/(10*ms) : 1", None, None, 0.0, method, {}, dt=DT)
    m_ref, _ = cg.neuron_model("dv/dt = -v/(10*ms) : 1 (unless refractory)", None, None, 1e-3, method, {}, dt=DT)
    v0 = (np.arange(10) + 1) / 11.0
    return nw.NetSpec("refractory", DT, 1, [nw.GroupSpec(0, 10, [], {0: v0}, code=m_ref),
                                            nw.GroupSpec(0, 10, [], {0: v0}, code=m_no)], [])


@pytest.mark.parametrize("method", METHODS)
def test_refractory(method):
    ref, _, _ = _run(refractory_spec(method), 100)
    a, b = ref.get_state(0, 0), ref.get_state(1, 0)
    assert np.array_equal(a, b)
    v0 = (np.arange(10) + 1) / 11.0
    assert np.allclose(a, v0 * np.exp(-1.0), rtol=1e-2)  # 10 ms of decay, tau = 10 ms (euler: 0.99**100)


def test_state_monitor_more_threads_than_single_block():
    m, _ = cg.neuron_model("v : 1", None, None, 0.0, None, {}, dt=DT)
    n = 2000
    spec = nw.NetSpec("statemon", DT, 1, [nw.GroupSpec(0, n, [], {0: np.arange(n, dtype=float)}, code=m)], [])
    spec.state_monitors = [(0, 0, 0, n)]
    ref, _, st = _run(spec, 3)
    rows = ref.state_monitor(st[0], n)
    assert rows.shape == (3, n) and np.all(rows == np.arange(n))


def test_group_variable_set_copy_to_host():
    m, names = cg.neuron_model("v : 1", None, None, 0.0, None, {}, dt=DT)
    c, const = cg.connect_expression("v_pre == 50", {}, names, names)
    out = oracle.connect_code((0, 1), (0, 1), c, "", 1.0, 0, False, 1, 0, (0.0, 0.0), DT, (1.0, 1.0),
                              [np.array([50.0])], [np.array([50.0])])
    assert out["pre"].size == 1
    out = oracle.connect_code((0, 1), (0, 1), c, "", 1.0, 0, False, 1, 0, (0.0, 0.0), DT, (1.0, 1.0),
                              [np.array([49.0])], [np.array([49.0])])
    assert out["pre"].size == 0
