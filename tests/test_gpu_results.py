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


@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_spike_monitor_subgroups_all_attr(hs, engine):
    """test_spike_monitor_subgroups_all_attr (test_monitor.py:23-52): neurons 0, 4
    and 5 of six spike at t = 0 (here a generator group); monitors of the whole
    group and of G[:2], G[2:4], G[4:] record relative indices."""
    from brian2cuda_amd import runtime as rt
    net = rt.Net(64, engine, 1e-4, 1)
    gid = net.add_group(rt.HS_MODEL_SPIKEGEN, 6, [0.0])
    net.set_spikegen(gid, np.array([0, 4, 5], np.int32), np.zeros(3, np.int64))
    m1 = net.add_spike_monitor(gid, start=0, stop=2)  # before the whole-group monitor
    m_all = net.add_spike_monitor(gid)
    m2 = net.add_spike_monitor(gid, start=2, stop=4)
    m3 = net.add_spike_monitor(gid, start=4, stop=6)
    net.run(1)
    want = {m_all: ([0, 4, 5], [1, 0, 0, 0, 1, 1]), m1: ([0], [1, 0]), m2: ([], [0, 0]), m3: ([0, 1], [1, 1])}
    for mid, (i_want, count_want) in want.items():
        i, t = net.spikes(mid)
        np.testing.assert_array_equal(i, i_want)
        np.testing.assert_array_equal(t, np.zeros(len(i_want)))
        np.testing.assert_array_equal(np.bincount(i, minlength=len(count_want)), count_want)
        assert net.mon_count(mid) == len(i_want)
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_range_monitors_vs_oracle(hs, oracle_lib, tmp_path, prec):
    """Overlapping range monitors on a live network (created before and after
    the whole-group one) hold the oracle's records of their ranges, and their
    result files count per range index."""
    spec = nw.cuba(N=800, seed=13)
    spec = nw.NetSpec(spec.name, spec.dt, spec.seed, spec.groups, spec.paths, spike_monitors=[0])
    net, sm, _ = nw.build_hs(spec, prec, "plain")
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, oracle_lib.SEQ)
    ranges = [(0, 100), (50, 300), (799, 800), (0, 800)]
    mids = [net.add_spike_monitor(0, start=a, stop=b) for a, b in ranges]
    net.run(600)
    ref.run(600)
    ri, rt_ = ref.spikes(rsm[0])
    assert len(ri) > 0
    for (a, b), mid in zip(ranges, mids):
        keep = (ri >= a) & (ri < b)
        i, t = net.spikes(mid)
        np.testing.assert_array_equal(i, ri[keep] - a)
        np.testing.assert_array_equal(t, rt_[keep])
    p = {k: tmp_path / k for k in ("i", "t", "count", "N")}
    net.write_monitor(mids[1], (p["i"], p["t"], p["count"], p["N"]))
    s = results.read_spike_monitor(p["i"], p["t"], p["count"], p["N"])
    keep = (ri >= 50) & (ri < 300)
    assert s["N"] == keep.sum()
    np.testing.assert_array_equal(s["count"], np.bincount(ri[keep] - 50, minlength=250))
    net.close()


def test_range_monitor_bad_range(hs):
    from brian2cuda_amd import runtime as rt
    from brian2cuda_amd._lib import HsError
    net = rt.Net(64, "plain", 1e-4, 1)
    gid = net.add_group(rt.HS_MODEL_SPIKEGEN, 6, [0.0])
    for a, b in [(-1, 3), (3, 3), (4, 2), (0, 7)]:
        with pytest.raises(HsError):
            net.add_spike_monitor(gid, start=a, stop=b)
    net.close()
