"""GPU: the result files of hs_net_write_monitor / hs_net_write_last_run_info
(the reference's raw little-endian arrays, objects.cu:550-595) hold the
oracle's records for C1 (CUBA, 4000 neurons: spikes of every neuron, v of
neurons 0..99 every step)."""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw
from brian2cuda_amd import results

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [64, 32])
def test_c1_result_files(hs, oracle_lib, tmp_path, prec):
    spec = nw.cuba(N=4000, seed=11)
    net, sm, st = nw.build_hs(spec, prec, "plain")
    ref, rsm, rst = oracle_lib.build_ref(spec, prec, oracle_lib.SEQ)
    net.run(1000)
    ref.run(1000)
    p = {k: tmp_path / k for k in ("i", "t", "count", "N", "st_t", "st_v", "st_idx", "st_N", "info")}
    net.write_monitor(sm[0], (p["i"], p["t"], p["count"], p["N"]))
    net.write_monitor(st[0], (p["st_t"], p["st_v"], p["st_idx"], p["st_N"]))
    net.write_last_run_info(p["info"])
    s = results.read_spike_monitor(p["i"], p["t"], p["count"], p["N"])
    ri, rstep = ref.spikes(rsm[0])
    assert s["N"] == len(ri) > 0
    np.testing.assert_array_equal(s["i"], ri)
    np.testing.assert_array_equal(s["t"], rstep.astype(np.float64) * spec.dt)
    np.testing.assert_array_equal(s["count"], np.bincount(ri, minlength=4000))
    v = results.read_state_monitor(p["st_t"], p["st_v"], 100, prec, p["st_idx"], p["st_N"])
    assert v["N"] == 1000 and v["values"].shape == (1000, 100)
    np.testing.assert_array_equal(v["t"], np.arange(1000) * spec.dt)
    np.testing.assert_array_equal(v["indices"], np.arange(100))
    np.testing.assert_array_equal(v["values"].astype(np.float64), ref.state_monitor(rst[0], 100))
    wall, frac = results.read_last_run_info(p["info"])
    assert wall > 0 and frac == 1.0
    net.close()


def test_partial_paths_and_errors(hs, tmp_path):
    from brian2cuda_amd._lib import HsError
    spec = nw.cuba(N=500, seed=2)
    net, sm, _ = nw.build_hs(spec, 64, "plain")
    net.run(50)
    net.write_monitor(sm[0], (None, tmp_path / "t", None, None))
    assert (tmp_path / "t").exists() and not (tmp_path / "i").exists()
    with pytest.raises(HsError):
        net.write_monitor(sm[0], (tmp_path / "no_such_dir" / "i", None, None, None))
    net.close()


@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
def test_rate_monitor(hs, oracle_lib, tmp_path, engine, sem):
    """PopulationRateMonitor over the whole group: the device counts give the
    oracle's rates, 1.0*count/dt/N, to the bit; added after 200 steps, it
    records from then on; its result files follow the format."""
    spec = nw.brunel(N=3000, K=300, seed=5, monitors=True)
    net, sm, _ = nw.build_hs(spec, 32, engine)
    ref, rsm, _ = oracle_lib.build_ref(spec, 32, sem)
    net.run(200)
    ref.run(200)
    a = net.add_rate_monitor(0)
    ra = ref.add_rate_monitor(0, 0, 3000)
    net.run(700)
    ref.run(700)
    ga, rra = net.rate_monitor(a), ref.rate_monitor(ra)
    assert ga.size == 700 and np.array_equal(ga, rra) and ga.sum() > 0
    gi, gt = net.spikes(sm[0])
    np.testing.assert_allclose(ga * 3000 * spec.dt, np.bincount(gt[gt >= 200] - 200, minlength=700)[:700],
                               rtol=1e-12)
    net.write_monitor(a, (tmp_path / "t", tmp_path / "rate", None, tmp_path / "N"))
    np.testing.assert_array_equal(np.fromfile(tmp_path / "rate", "<f8"), ga)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "t", "<f8"), (200 + np.arange(700)) * spec.dt)
    net.close()


def test_rate_monitor_subgroup(hs, oracle_lib):
    spec = nw.cuba(N=2000, seed=4)
    spec.rate_monitors = [(0, 100, 1600)]
    net, _, _ = nw.build_hs(spec, 64, "plain")
    ref, _, _ = oracle_lib.build_ref(spec, 64, oracle_lib.SEQ)
    net.run(1000)
    ref.run(1000)
    g, r = net.rate_monitor(net.rate_mids[0]), ref.rate_monitor(ref.rate_mids[0])
    assert np.array_equal(g, r) and g.sum() > 0
    net.close()
