"""GPU parity of the individually testable kernels against the oracle (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F64, F32 = 64, 32


def test_detmath_bits_equal_host(hs, oracle_lib):
    """exp/log/sqrt/cos2pi and the float expf/logf/sqrtf/cos2pif on gfx950 give the host oracle's bits."""
    L = oracle_lib.lib()
    rng = np.random.default_rng(7)
    cases = {
        0: rng.uniform(-700, 700, 20000),
        1: np.concatenate([rng.random(10000) + 1e-300, 10.0 ** rng.uniform(-300, 300, 10000)]),
        2: np.concatenate([10.0 ** rng.uniform(-300, 300, 20000), rng.uniform(0, 80, 20000),
                           np.nextafter(np.arange(1.0, 2000.0) ** 2, 0), np.nextafter(np.arange(1.0, 2000.0) ** 2, 1e9),
                           rng.uniform(0, 1, 2000) * 2.0 ** -1060]),
        3: rng.random(20000),
        4: rng.uniform(-80, 80, 20000).astype(np.float32).astype(np.float64),
        5: np.concatenate([(np.arange(1, 20001) * 2.0 ** -24), rng.random(20000) * 1e30 + 1e-30]).astype(
            np.float32).astype(np.float64),
        6: np.concatenate([rng.uniform(0, 80, 20000), np.arange(1.0, 4000.0) ** 2]).astype(np.float32).astype(
            np.float64),
        7: (rng.integers(0, 2**24, 20000) * 2.0 ** -24),
    }
    host = {0: L.ref_exp, 1: L.ref_log, 2: L.ref_sqrt, 3: L.ref_cos2pi, 4: lambda x: float(L.ref_expf(x)),
            5: lambda x: float(L.ref_logf(x)), 6: lambda x: float(L.ref_sqrtf(x)), 7: lambda x: float(L.ref_cos2pif(x))}
    for fn, x in cases.items():
        y = hs.math_eval(fn, x)
        ref = np.array([host[fn](float(a)) for a in x])
        assert np.array_equal(y.view(np.uint64), ref.view(np.uint64)), f"fn {fn}: {np.sum(y != ref)} differ"


@pytest.mark.parametrize("prec", [F64, F32])
def test_philox_randn_matches_oracle(hs, oracle_lib, prec):
    for seed, stream, step in [(1234, 0, 0), (1234, 0, 7), (2**40 + 5, 3, 123456789)]:
        n = 5000
        z = hs.randn(seed, stream, step, n, prec)
        ref = np.array([oracle_lib.randn(seed, stream, i, step, precision=prec) for i in range(n)])
        if prec == F64:
            assert np.array_equal(z, ref)
        else:
            assert np.array_equal(z, ref.astype(np.float32))  # float draws, exactly representable


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 100_003])
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("sorted_", [True, False])
def test_threshold_compact_vs_brute_force(hs, n, prec, sorted_):
    dt = np.float64 if prec == F64 else np.float32
    rng = np.random.default_rng(n)
    v = rng.standard_normal(n).astype(dt)
    last = np.full(n, -1073741824, np.int32)
    last[rng.random(n) < 0.1] = 98  # refractory at step 100 with 5 steps
    vth, step, ref = 0.7, 100, 5
    expect = np.nonzero((v > dt(vth)) & ((step - last.astype(np.int64)) >= ref))[0].astype(np.int32)
    dv = hs.DeviceArray.from_numpy(v)
    dl = hs.DeviceArray.from_numpy(last)
    got = hs.threshold_compact(dv, vth, dl, step, ref, sorted_)
    if sorted_:
        assert np.array_equal(got, expect)
    else:
        assert np.array_equal(np.sort(got), expect)
        assert len(np.unique(got)) == len(got)
    new_last = dl.numpy()
    assert np.all(new_last[expect] == step)
    mask = np.ones(n, bool)
    mask[expect] = False
    assert np.array_equal(new_last[mask], last[mask])


@pytest.mark.parametrize("n_spike", ["all", "none"])
def test_threshold_compact_all_or_none(hs, n_spike):
    n = 4097
    v = np.full(n, 1.0 if n_spike == "all" else -1.0)
    dv = hs.DeviceArray.from_numpy(v)
    dl = hs.DeviceArray.from_numpy(np.full(n, -1073741824, np.int32))
    got = hs.threshold_compact(dv, 0.0, dl, 0, 0, True)
    assert len(got) == (n if n_spike == "all" else 0)
    if n_spike == "all":
        assert np.array_equal(got, np.arange(n))


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("model", [1, 2, 3])
def test_stateupdate_one_step_vs_oracle(hs, oracle_lib, prec, model):
    from brian2cuda_amd import networks as nw
    dtype = np.float64 if prec == F64 else np.float32
    n = 3000
    rng = np.random.default_rng(model)
    params = {1: nw.cuba_params(), 2: nw.brunel_params(), 3: nw.cobahh_params()}[model]
    nv = {1: 3, 2: 1, 3: 6}[model]
    if model == 1:
        init = [-0.06 + 0.012 * rng.random(n), 0.003 * rng.random(n), -0.003 * rng.random(n)]
    elif model == 2:
        init = [0.01 + 0.01 * rng.random(n)]
    else:
        init = [-0.065 + 0.05 * rng.random(n), rng.random(n), rng.random(n), rng.random(n),
                1e-8 * rng.random(n), 1e-8 * rng.random(n)]
    step = 37
    last = np.full(n, -1073741824, np.int32)
    last[::7] = step - 3  # some refractory neurons
    dvars = [hs.DeviceArray.from_numpy(np.asarray(a, dtype)) for a in init]
    dlast = hs.DeviceArray.from_numpy(last)
    hs.stateupdate(model, prec, dvars, dlast, params, 1e-4, 99, step)
    got = [d.numpy() for d in dvars]
    ref = oracle_lib.stateupdate(model, prec, [np.asarray(a, dtype).astype(np.float64) for a in init], last,
                                 params, 1e-4, 99, step)
    for v in range(nv):
        assert np.array_equal(got[v], ref[v].astype(dtype)), f"var {v}: {np.sum(got[v] != ref[v].astype(dtype))}"


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("n_events,n_target", [(0, 10), (1, 1), (1000, 7), (300_000, 100_000)])
def test_scatter_add_vs_serial_sum(hs, prec, n_events, n_target):
    """hs_scatter_add against numpy.add.at: every addend is the same weight, so
    the atomic sum is exact whatever the order."""
    dt = np.float64 if prec == 64 else np.float32
    rng = np.random.default_rng(n_events + n_target)
    idx = rng.integers(0, n_target, n_events).astype(np.int32)
    t0 = rng.standard_normal(n_target).astype(dt)
    w = dt(-1e-4)
    d_idx = hs.DeviceArray.from_numpy(idx if n_events else np.zeros(1, np.int32))
    d_idx.n = n_events
    d_t = hs.DeviceArray.from_numpy(t0)
    ms = hs.scatter_add(d_idx, float(w), d_t)
    want = t0.copy()
    np.add.at(want, idx, w)
    assert np.array_equal(d_t.numpy(), want)
    assert ms >= 0.0


def test_scatter_add_rejects_out_of_range(hs):
    from brian2cuda_amd._lib import HsError
    idx = np.array([0, 5, -1, 3, 9], np.int32)
    d_idx = hs.DeviceArray.from_numpy(idx)
    d_t = hs.DeviceArray.from_numpy(np.zeros(5, np.float64))
    with pytest.raises(HsError):
        hs.scatter_add(d_idx, 1.0, d_t)
    # the in-range entries were applied, the others skipped
    assert np.array_equal(d_t.numpy(), np.array([1.0, 0.0, 0.0, 1.0, 0.0]))
