"""GPU: StateMonitors with arbitrary record indices and their own dt
(StateMonitor(G, var, record=[...], dt=k*dt), statemonitor.cu:9-100 with
_indices; the form of the reference's consistency tests,
test_cpp_cuda_consistency.py:72,156), bit for bit against the oracle; the
interval of synapse monitors; and the result files of such monitors."""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw
from brian2cuda_amd import project
from brian2cuda_amd import runtime as rt

pytestmark = pytest.mark.gpu

IDX = np.array([3999, 0, 17, 2500, 1234, 3], np.int32)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem,opts", [("plain", 0, {}), ("fused", 1, {}), ("fused", 1, {9: 1})])
def test_indexed_interval_monitors_vs_oracle(hs, oracle_lib, prec, engine, sem, opts):
    spec = nw.cuba(N=4000, seed=11, record=0)
    spec.state_monitors = [nw.StateMon(0, 0, IDX, every=10), nw.StateMon(0, 1, np.arange(50)[::-1], every=1),
                           (0, 2, 100, 140)]
    net, _, gst = nw.build_hs(spec, prec, engine, options=opts)
    ref, _, rst = oracle_lib.build_ref(spec, prec, sem)
    for k in (37, 263, 700):
        net.run(k)
        ref.run(k)
    # added mid-run (step 1000): its clock ticks at multiples of 7 -> first row at 1001
    g7 = net.add_state_monitor_indices(0, 0, IDX[::-1], every=7)
    r7 = ref.add_state_monitor_indices(0, 0, IDX[::-1], every=7)
    net.run(321)
    ref.run(321)
    for a, b, m in zip(gst, rst, spec.state_monitors):
        w = nw.state_monitor_width(m)
        g, r = net.state_monitor(a, w), ref.state_monitor(b, w)
        assert g.shape == r.shape and g.shape[0] > 0
        assert np.array_equal(g, r)
    g, r = net.state_monitor(g7, IDX.size), ref.state_monitor(r7, IDX.size)
    assert g.shape == (len(range(1001, 1321, 7)), IDX.size)
    assert np.array_equal(g, r)
    net.close()


def test_indexed_monitor_files_and_rows(hs, tmp_path):
    """t = recorded step * dt, indices as given; a row equals the state read at
    that step; the interval of a monitor is fixed once it holds a row."""
    from brian2cuda_amd._lib import HsError
    spec = nw.cuba(N=2000, seed=4, record=0)
    net, _, _ = nw.build_hs(spec, 64, "plain")
    net.run(5)
    m = net.add_state_monitor_indices(0, 0, [7, 3, 1999], every=4)   # rows at 8, 12, ...
    net.run(3)                                                       # step 8: the first row is not taken yet
    v8 = net.get_state(0, 0)[[7, 3, 1999]]
    net.run(1)
    net.run(100)
    rows = net.state_monitor(m, 3)
    assert rows.shape == (len(range(8, 109, 4)), 3)
    np.testing.assert_array_equal(rows[0], v8)
    p = [tmp_path / k for k in ("t", "v", "idx", "N")]
    net.write_monitor(m, p)
    np.testing.assert_array_equal(np.fromfile(p[0], "<f8"), np.arange(8, 109, 4) * spec.dt)
    np.testing.assert_array_equal(np.fromfile(p[2], "<i4"), [7, 3, 1999])
    assert int(np.fromfile(p[3], "<i4")[0]) == rows.shape[0]
    with pytest.raises(HsError):
        net.set_monitor_interval(m, 2)
    for bad in ([5, 5], [2000], []):
        with pytest.raises(HsError):
            net.add_state_monitor_indices(0, 0, bad, every=1)
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_synapse_monitor_interval_vs_oracle(hs, oracle_lib, prec):
    """StateMonitor(S, 'w', record=range(n), dt=...) on the reference's STDP
    example network (test_cpp_cuda_consistency.py:72): rows at multiples of the
    interval, bit for bit."""
    spec = nw.stdp_poisson_example(N=2000, K=200)
    net, _, _ = nw.build_hs(spec, prec, "plain")
    ref, _, _ = oracle_lib.build_ref(spec, prec, oracle_lib.SEQ)
    sel = np.arange(20)
    g = net.add_synapse_monitor(0, rt.HS_SYN_W, sel)
    r = ref.add_synapse_monitor(0, rt.HS_SYN_W, sel)
    net.set_monitor_interval(g, 25)
    ref.set_monitor_interval(r, 25)
    net.run(1000)
    ref.run(1000)
    a, b = net.state_monitor(g, sel.size), ref.state_monitor(r, sel.size)
    assert a.shape == (40, sel.size) and np.array_equal(a, b)
    net.close()


def test_indexed_monitor_through_the_project(hs, oracle_lib, tmp_path):
    spec = nw.cuba(N=4000, seed=11, record=0)
    spec.state_monitors = [nw.StateMon(0, 0, IDX, every=10)]
    project.write_project(spec, tmp_path, 600)
    project.build_project(tmp_path)
    project.run_project(tmp_path, "results", timeout=120)
    res = project.read_results(tmp_path)
    ref, _, rst = oracle_lib.build_ref(spec, 64, oracle_lib.SEQ)
    ref.run(600)
    m = res["statemonitor"]
    np.testing.assert_array_equal(m["t"], np.arange(0, 600, 10) * spec.dt)
    np.testing.assert_array_equal(m["indices"], IDX)
    np.testing.assert_array_equal(m["v"], ref.state_monitor(rst[0], IDX.size))
