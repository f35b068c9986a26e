"""GPU parity of the device-built random connectivity (hs_net_connect_random,
and the fixed out-degree form hs_net_connect_fixed): the bundle table read back
through the C ABI equals the oracle's synapses (oracle.connect_random /
connect_fixed, creation order) stably sorted by (pre, delay), and a network
built this way simulates to the oracle's bits."""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw
from brian2cuda_amd import runtime as rt
from brian2cuda_amd._lib import HsError

pytestmark = pytest.mark.gpu

DT = 1e-4


def _bundle_order(s):
    order = np.lexsort((s["delay_steps"], s["pre"]))  # stable: creation order kept inside a bundle
    return {k: v[order] for k, v in s.items()}


def _net(prec, sizes):
    net = rt.Net(prec, "plain", DT, 1)
    for n in sizes:
        net.add_group(rt.HS_MODEL_LIF_CUBA, n, nw.cuba_params())
    return net


CASES = [
    # n_src, n_tgt, pre, post, p, autapses, delay (s), wmode, w
    (3000, 3000, None, None, 0.01, True, (0.5e-3, 5e-3), rt.HS_W_SYNAPSE, (-1e-4, -1e-4)),
    (2000, 1500, (100, 1900), (7, 1499), 0.03, True, (0.0, 2e-3), rt.HS_W_SYNAPSE, (-2.0, 3.0)),
    (1000, 1000, None, None, 0.2, False, (0.1e-3, 0.1e-3), rt.HS_W_SCALAR, (0.5, 0.5)),
    (700, 700, None, None, 0.6, False, (0.2e-3, 1.9e-3), rt.HS_W_SYNAPSE, (0.0, 1.0)),
    (500, 800, None, None, 1.0, True, (0.0, 0.0), rt.HS_W_SCALAR, (1.0, 1.0)),
    (1200, 1200, None, None, 0.05, True, (0.0, 25e-3), rt.HS_W_SYNAPSE, (1.0, 4.0)),   # D = 251: global tables
    (4000, 4000, (0, 3200), None, 0.02, True, (0.1e-3, 0.1e-3), rt.HS_W_SCALAR, (1.62e-3, 1.62e-3)),
    (900, 900, None, None, 0.0, True, (0.5e-3, 1e-3), rt.HS_W_SYNAPSE, (1.0, 2.0)),    # empty
    (900, 900, (300, 300), None, 0.1, True, (0.5e-3, 1e-3), rt.HS_W_SYNAPSE, (1.0, 2.0)),  # no rows
]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_device_connect_matches_oracle(hs, oracle_lib, prec, case):
    n_src, n_tgt, pre, post, p, autapses, delay, wmode, w = CASES[case]
    net = _net(prec, [n_src] if n_src == n_tgt else [n_src, n_tgt])
    tgt = 0 if n_src == n_tgt else 1
    seed, stream = 20240 + case, 3
    pid = net.connect_random(0, tgt, 1, p, pre, post, autapses, seed, stream, delay, wmode, w)
    ref = oracle_lib.connect_random(pre or (0, n_src), post or (0, n_tgt), p, (not autapses) and tgt == 0, seed,
                                    stream, delay, DT, w)
    ref = _bundle_order(ref)
    info = net.pathway_info(pid)
    assert info["nsyn"] == ref["pre"].size
    assert info["dmin"] == nw.quantise_delay(delay[0], DT)
    assert info["D"] == nw.quantise_delay(delay[1], DT) - info["dmin"] + 1
    gp, gq, gd, gw = net.synapses(pid)
    np.testing.assert_array_equal(gp, ref["pre"])
    np.testing.assert_array_equal(gq, ref["post"])
    np.testing.assert_array_equal(gd, ref["delay_steps"])
    rw = ref["w"] if wmode == rt.HS_W_SYNAPSE else np.full(ref["pre"].size, w[0])
    if prec == 32 and wmode == rt.HS_W_SYNAPSE:
        rw = rw.astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(gw, rw)
    net.close()


def test_device_connect_full_size_rows(hs, oracle_lib):
    """C3's connectivity (100,000 x 100,000, p = 0.01, delays U[0.5, 5] ms): the
    total is Binomial(10^10, 0.01), bundles are sorted by target, and three
    ranges of rows equal the oracle's rows exactly."""
    N, K = 100_000, 1000
    spec = nw.brunel(N=N, K=K, seed=77, device_connect=True)
    net, _, _ = nw.build_hs(spec, 32, "fused")
    info = net.pathway_info(0)
    mean = N * N * (K / N)
    assert abs(info["nsyn"] - mean) < 5 * np.sqrt(mean)
    gp, gq, gd, gw = net.synapses(0)
    assert np.all(np.diff(gp) >= 0)
    key = gp.astype(np.int64) * 64 + gd
    assert np.all(np.diff(key) >= 0)
    same = np.diff(key) == 0
    assert np.all(np.diff(gq)[same] > 0)  # creation (ascending target) order inside each bundle
    assert gd.min() == 5 and gd.max() == 50 and np.all(gw == np.float32(-1e-4))
    bounds = np.searchsorted(gp, np.arange(N + 1))
    for lo, hi in [(0, 400), (50_000, 50_300), (99_700, 100_000)]:
        ref = _bundle_order(oracle_lib.random_pathway(spec, spec.paths[0], rows=(lo, hi)))
        sl = slice(bounds[lo], bounds[hi])
        np.testing.assert_array_equal(gq[sl], ref["post"])
        np.testing.assert_array_equal(gd[sl], ref["delay_steps"])
        np.testing.assert_array_equal(gp[sl], ref["pre"])
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
def test_device_connected_brunel_simulates_to_oracle_bits(hs, oracle_lib, prec, engine, sem):
    spec = nw.brunel(N=4000, K=400, seed=5, monitors=True, device_connect=True)
    net, smon, _ = nw.build_hs(spec, prec, engine)
    ref, rsmon, _ = oracle_lib.build_ref(spec, prec, sem)
    net.run(1500)
    ref.run(1500)
    gi, gt = net.spikes(smon[0])
    ri, rt_ = ref.spikes(rsmon[0])
    assert len(gi) == len(ri) > 0
    assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert np.array_equal(net.get_state(0, 0), ref.get_state(0, 0))
    assert net.stats()["events"] == ref.events
    net.close()


def test_device_connect_rejects_bad_arguments(hs):
    net = _net(64, [100])
    for kw in [dict(p=-0.1), dict(p=1.5), dict(pre=(0, 101)), dict(post=(50, 20)), dict(delay=(2e-3, 1e-3)),
               dict(wmode=rt.HS_W_PRE)]:
        args = dict(p=0.1, pre=None, post=None, delay=(0.0, 1e-3), wmode=rt.HS_W_SCALAR)
        args.update(kw)
        with pytest.raises(HsError):
            net.connect_random(0, 0, 1, args["p"], args["pre"], args["post"], True, 1, 0, args["delay"],
                               args["wmode"], 1.0)
    net.close()
    net = _net(64, [100])
    net.set_option(5, 2)  # gather delivery: the device build writes scatter bundles only
    with pytest.raises(HsError):
        net.connect_random(0, 0, 1, 0.1)
    net.close()


FIXED = [
    # n_src, n_tgt, pre, post, K, autapses, delay (s), wmode, w
    (2000, 2000, None, None, 100, True, (0.5e-3, 5e-3), rt.HS_W_SYNAPSE, (-1.0, 1.0)),
    (2000, 2000, None, None, 100, False, (0.5e-3, 5e-3), rt.HS_W_SYNAPSE, (-1.0, 1.0)),   # j == i dropped
    (1500, 900, (10, 1400), (100, 800), 1, True, (0.0, 0.0), rt.HS_W_SCALAR, (0.5, 0.5)),
    (600, 600, None, None, 64, False, (0.1e-3, 0.1e-3), rt.HS_W_SYNAPSE, (1.0, 2.0)),    # one full batch
    (600, 600, None, None, 65, False, (0.1e-3, 2e-3), rt.HS_W_SYNAPSE, (1.0, 2.0)),      # a second batch of 1
    (300, 300, None, None, 300, False, (0.2e-3, 1e-3), rt.HS_W_SCALAR, (1.0, 1.0)),      # every target
    (400, 3000, None, None, 2048, True, (0.0, 25e-3), rt.HS_W_SYNAPSE, (0.0, 1.0)),      # largest K, D = 251
    (700, 700, None, None, 0, True, (0.5e-3, 1e-3), rt.HS_W_SYNAPSE, (1.0, 2.0)),        # K = 0: empty
]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("case", range(len(FIXED)))
def test_device_fixed_outdegree_matches_oracle(hs, oracle_lib, prec, case):
    n_src, n_tgt, pre, post, K, autapses, delay, wmode, w = FIXED[case]
    net = _net(prec, [n_src] if n_src == n_tgt else [n_src, n_tgt])
    tgt = 0 if n_src == n_tgt else 1
    seed, stream = 777 + case, 2
    pid = net.connect_fixed(0, tgt, 1, K, pre, post, autapses, seed, stream, delay, wmode, w)
    ref = oracle_lib.connect_fixed(pre or (0, n_src), post or (0, n_tgt), K, (not autapses) and tgt == 0, seed,
                                   stream, delay, DT, w)
    ref = _bundle_order(ref)
    info = net.pathway_info(pid)
    assert info["nsyn"] == ref["pre"].size
    gp, gq, gd, gw = net.synapses(pid)
    np.testing.assert_array_equal(gp, ref["pre"])
    np.testing.assert_array_equal(gq, ref["post"])
    np.testing.assert_array_equal(gd, ref["delay_steps"])
    rw = ref["w"] if wmode == rt.HS_W_SYNAPSE else np.full(ref["pre"].size, w[0])
    if prec == 32 and wmode == rt.HS_W_SYNAPSE:
        rw = rw.astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(gw, rw)
    net.close()


def test_device_fixed_outdegree_full_size(hs, oracle_lib):
    """C3 with a fixed out-degree (100,000 neurons, K = 1000 each, 1e8 synapses):
    every row holds exactly K distinct targets, and three ranges of rows equal
    the oracle's rows exactly."""
    N, K = 100_000, 1000
    spec = nw.brunel(N=N, K=K, seed=78, device_connect=True, fixed_outdegree=True)
    net, _, _ = nw.build_hs(spec, 32, "fused")
    assert net.pathway_info(0)["nsyn"] == N * K
    gp, gq, gd, gw = net.synapses(0)
    assert np.all(np.bincount(gp, minlength=N) == K)
    bounds = np.searchsorted(gp, np.arange(N + 1))
    for lo, hi in [(0, 300), (50_000, 50_200), (99_800, 100_000)]:
        ref = _bundle_order(oracle_lib.random_pathway(spec, spec.paths[0], rows=(lo, hi)))
        sl = slice(bounds[lo], bounds[hi])
        np.testing.assert_array_equal(gp[sl], ref["pre"])
        np.testing.assert_array_equal(gq[sl], ref["post"])
        np.testing.assert_array_equal(gd[sl], ref["delay_steps"])
    net.close()


@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
def test_fixed_outdegree_brunel_simulates_to_oracle_bits(hs, oracle_lib, engine, sem):
    spec = nw.brunel(N=4000, K=400, seed=6, monitors=True, device_connect=True, fixed_outdegree=True)
    net, smon, _ = nw.build_hs(spec, 32, engine)
    ref, rsmon, _ = oracle_lib.build_ref(spec, 32, sem)
    net.run(1200)
    ref.run(1200)
    gi, gt = net.spikes(smon[0])
    ri, rt_ = ref.spikes(rsmon[0])
    assert len(gi) == len(ri) > 0
    assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert net.stats()["events"] == ref.events
    net.close()


def test_device_fixed_outdegree_rejects_bad_arguments(hs):
    net = _net(64, [100])
    for K, post in [(-1, None), (101, None), (11, (0, 10))]:
        with pytest.raises(HsError):
            net.connect_fixed(0, 0, 1, K, None, post, True, 1, 0, 0.0)
    net.close()
    net = _net(64, [3000])
    with pytest.raises(HsError):
        net.connect_fixed(0, 0, 1, 2049)   # beyond the largest out-degree of the device sampler
    net.close()
