"""CPU tests that pin the oracle's random connectivity (oracle.connect_random,
the restatement of Synapses.connect(p) of synapses_create_generator.cu:315-333
with per-synapse delay and weight draws) against facts it does not compute itself:

  - a vectorised numpy restatement of the draw layout (its own Philox, pinned by
    the Random123 vectors) and of the skip chain as a cumulative sum of geometric
    gaps, with numpy's log -- so an off-by-one in where a row starts, how a skip
    advances, which Philox half a draw takes or how a delay is quantised fails;
  - the distributions the algorithm must produce (Bernoulli counts, geometric
    gaps, quantised uniform delays, uniform weights), by chi-square tests;
  - special cases: p = 0, p = 1, dropping j == i, rows independent of the range.

The fixed out-degree form (oracle.connect_fixed) is pinned the same way: a
numpy restatement of its draw/dedupe/redraw rounds, K distinct sorted targets
per row, uniformity over the K-subsets (chi-square over all 56 subsets of a
small case), and K = 0, K = n, K > n.
"""
import numpy as np
import pytest
from scipy import stats

import oracle
from brian2cuda_amd import networks as nw

CTR = 0x434E0000
TAG_SKIP, TAG_SYN, TAG_PAIR = 0, 1, 2
M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox_np(c0, c1, c2, c3, seed):
    """Philox4x32-10 over arrays of counters (uint64 arithmetic, low 32 bits kept)."""
    c = [np.asarray(x, dtype=np.uint64) & MASK for x in np.broadcast_arrays(c0, c1, c2, c3)]
    k0, k1 = np.uint64(seed & 0xFFFFFFFF), np.uint64(seed >> 32)
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & MASK, p1 & MASK, ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & MASK,
             p0 & MASK]
        k0 = (k0 + np.uint64(W0)) & MASK
        k1 = (k1 + np.uint64(W1)) & MASK
    return c


def u53(a, b):
    return ((a >> np.uint64(5)) * np.uint64(1 << 26) + (b >> np.uint64(6))).astype(np.float64) * 2.0 ** -53


def test_numpy_philox_matches_known_answers():
    for ctr, key, out in [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
                          ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
                           (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]:
        w = philox_np(*ctr, key[0] | (key[1] << 32))
        assert tuple(int(x) for x in w) == out


def restate_rows(rows, post, p, seed, stream, delay, dt, weight, skip_self=False):
    """The same synapses from the draw layout, vectorised per row."""
    lo, hi = post
    out = {k: [] for k in ("pre", "post", "delay_steps", "w")}
    for i in range(*rows):
        if p < 0.25:
            nd = int(2 * (hi - lo) * p + 64)
            k = np.arange(nd)
            w = philox_np(i, k // 2, CTR | TAG_SKIP, stream, seed)
            r = np.where(k % 2 == 0, u53(w[0], w[1]), u53(w[2], w[3]))
            assert np.all(r > 0)
            gaps = np.floor(np.log(r) / np.log1p(-p)).astype(np.int64)
            js = lo + np.cumsum(gaps + 1) - 1
            assert js[-1] >= hi, "restatement drew too few skips"
            js = js[js < hi]
        else:
            m = np.arange(hi - lo)
            w = philox_np(i, m // 2, CTR | TAG_PAIR, stream, seed)
            r = np.where(m % 2 == 0, u53(w[0], w[1]), u53(w[2], w[3]))
            js = lo + m[r < p]
        if skip_self:
            js = js[js != i]
        w = philox_np(i, js, CTR | TAG_SYN, stream, seed)
        ud, uw = u53(w[0], w[1]), u53(w[2], w[3])
        out["pre"].append(np.full(js.size, i, np.int32))
        out["post"].append(js.astype(np.int32))
        out["delay_steps"].append(nw.quantise_delay(delay[0] + ud * (delay[1] - delay[0]), dt))
        out["w"].append(weight[0] + uw * (weight[1] - weight[0]))
    return {k: np.concatenate(v) for k, v in out.items()}


@pytest.mark.parametrize("p,post,skip_self", [(0.01, (0, 20000), False), (0.003, (500, 9000), True),
                                              (0.2, (0, 700), False), (0.3, (10, 610), True),
                                              (0.9, (0, 300), False)])
def test_rows_match_vectorised_restatement(p, post, skip_self):
    rows = (3, 40)
    a = oracle.connect_random(rows, post, p, skip_self, 987654321012345, 5, (0.2e-3, 3.3e-3), 1e-4, (-2.0, 3.0))
    b = restate_rows(rows, post, p, 987654321012345, 5, (0.2e-3, 3.3e-3), 1e-4, (-2.0, 3.0), skip_self)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_counts_and_gaps_follow_bernoulli():
    """Row counts ~ Binomial(n, p); gaps between consecutive targets ~ Geometric(p)."""
    n, p = 4000, 0.02
    s = oracle.connect_random((0, n), (0, n), p, False, 42, 0, (0, 0), 1e-4, (0, 0))
    cnt = np.bincount(s["pre"], minlength=n)
    mean, var = n * p, n * p * (1 - p)
    assert abs(cnt.mean() - mean) < 5 * np.sqrt(var / n)
    assert abs(cnt.var() / var - 1) < 0.1
    # gap from each start (row start, or one past a target) to the next target of
    # the row; only starts with 200 positions left in the row, so gaps >= 200 are
    # observed as such rather than cut off by the row end
    bounds = np.searchsorted(s["pre"], np.arange(n + 1))
    gaps = []
    for i in range(n):
        t = s["post"][bounds[i]:bounds[i + 1]].astype(np.int64)
        starts = np.r_[0, t + 1]
        nxt = np.r_[t, 10 ** 9]
        ok = starts + 200 <= n
        gaps.append((nxt - starts)[ok])
    gaps = np.concatenate(gaps)
    edges = np.arange(0, 200, 10)
    obs = np.array([np.sum((gaps >= a) & (gaps < a + 10)) for a in edges] + [np.sum(gaps >= 200)])
    cdf = lambda g: 1 - (1 - p) ** g
    prob = np.array([cdf(a + 10) - cdf(a) for a in edges] + [1 - cdf(200)])
    exp = prob / prob.sum() * obs.sum()
    assert stats.chisquare(obs, exp).pvalue > 1e-4


def test_pair_tests_follow_bernoulli():
    n, p = 600, 0.4
    s = oracle.connect_random((0, n), (0, n), p, False, 7, 3, (0, 0), 1e-4, (0, 0))
    m = np.zeros((n, n), bool)
    m[s["pre"], s["post"]] = True
    assert abs(m.mean() - p) < 5 * np.sqrt(p * (1 - p) / n ** 2)
    col = m.sum(0)
    assert stats.chisquare(col, np.full(n, col.mean())).pvalue > 1e-4  # no column preferred


def test_delays_and_weights_distribution():
    dt, dlo, dhi = 1e-4, 0.5e-3, 5e-3
    s = oracle.connect_random((0, 3000), (0, 3000), 0.05, False, 11, 2, (dlo, dhi), dt, (-1.0, 2.0))
    d = s["delay_steps"]
    assert d.min() == 5 and d.max() == 50
    ks = np.arange(5, 51)
    # P(q = k) = |[(k - 0.5) dt, (k + 0.5) dt) intersected with [dlo, dhi)| / (dhi - dlo)
    prob = np.clip(np.minimum((ks + 0.5) * dt, dhi) - np.maximum((ks - 0.5) * dt, dlo), 0, None) / (dhi - dlo)
    obs = np.bincount(d, minlength=51)[5:51]
    assert stats.chisquare(obs, prob / prob.sum() * obs.sum()).pvalue > 1e-4
    w = s["w"]
    assert w.min() >= -1.0 and w.max() < 2.0
    assert stats.kstest((w + 1.0) / 3.0, "uniform").pvalue > 1e-4
    # the delay and weight of a synapse come from separate words: uncorrelated
    assert abs(np.corrcoef(w, d)[0, 1]) < 0.02


def test_special_cases():
    n = 300
    assert oracle.connect_random((0, n), (0, n), 0.0, False, 1, 0, (0, 0), 1e-4, (0, 0))["pre"].size == 0
    full = oracle.connect_random((0, n), (0, n), 1.0, False, 1, 0, (0, 0), 1e-4, (0, 0))
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    np.testing.assert_array_equal(full["pre"], ii.ravel())
    np.testing.assert_array_equal(full["post"], jj.ravel())
    noself = oracle.connect_random((0, n), (0, n), 1.0, True, 1, 0, (0, 0), 1e-4, (0, 0))
    assert noself["pre"].size == n * n - n and np.all(noself["pre"] != noself["post"])
    for p in (0.05, 0.5):
        a = oracle.connect_random((0, n), (0, n), p, False, 9, 1, (0.1e-3, 0.9e-3), 1e-4, (0, 1))
        b = oracle.connect_random((0, n), (0, n), p, True, 9, 1, (0.1e-3, 0.9e-3), 1e-4, (0, 1))
        keep = a["pre"] != a["post"]
        for k in a:  # dropping j == i changes nothing else
            np.testing.assert_array_equal(a[k][keep], b[k])
        c = oracle.connect_random((100, 200), (0, n), p, False, 9, 1, (0.1e-3, 0.9e-3), 1e-4, (0, 1))
        sel = (a["pre"] >= 100) & (a["pre"] < 200)
        for k in a:  # a row does not depend on the other rows
            np.testing.assert_array_equal(a[k][sel], c[k])
    one = oracle.connect_random((0, n), (0, n), 0.1, False, 3, 0, (0.7e-3, 0.7e-3), 1e-4, (0.25, 0.25))
    assert np.all(one["delay_steps"] == 7) and np.all(one["w"] == 0.25)


def test_seed_and_stream_change_the_draws():
    a = oracle.connect_random((0, 50), (0, 2000), 0.01, False, 5, 0, (0, 0), 1e-4, (0, 0))
    b = oracle.connect_random((0, 50), (0, 2000), 0.01, False, 5, 1, (0, 0), 1e-4, (0, 0))
    c = oracle.connect_random((0, 50), (0, 2000), 0.01, False, 6, 0, (0, 0), 1e-4, (0, 0))
    assert not np.array_equal(a["post"][:50], b["post"][:50])
    assert not np.array_equal(a["post"][:50], c["post"][:50])


# --------------------------------------------------------------------------
# fixed out-degree: connect(j='k for k in sample(N_post, size=K)'), the
# reference's fixed-size sampling (synapses_create_generator.cu:250-312)
# --------------------------------------------------------------------------
TAG_FIX = 3


def restate_fixed(rows, post, K, seed, stream, delay, dt, weight, skip_self=False):
    """Draw K values, keep the distinct ones (np.unique sorts), draw the missing
    ones with the row's next counters; then the per-(i, j) delay and weight."""
    lo, hi = post
    n = hi - lo
    out = {k: [] for k in ("pre", "post", "delay_steps", "w")}
    for i in range(*rows):
        vals = np.empty(0, np.int64)
        c = 0
        while vals.size < K:
            need = K - vals.size
            w = philox_np(i, np.arange(c, c + need), CTR | TAG_FIX, stream, seed)
            c += need
            new = np.minimum(np.floor(u53(w[0], w[1]) * n).astype(np.int64), n - 1)
            vals = np.unique(np.r_[vals, new])
        js = lo + vals
        if skip_self:
            js = js[js != i]
        w = philox_np(i, js, CTR | TAG_SYN, stream, seed)
        out["pre"].append(np.full(js.size, i, np.int32))
        out["post"].append(js.astype(np.int32))
        out["delay_steps"].append(nw.quantise_delay(delay[0] + u53(w[0], w[1]) * (delay[1] - delay[0]), dt))
        out["w"].append(weight[0] + u53(w[2], w[3]) * (weight[1] - weight[0]))
    return {k: np.concatenate(v) for k, v in out.items()}


@pytest.mark.parametrize("K,post,skip_self", [(10, (0, 2000), False), (100, (0, 300), True), (250, (50, 350), False),
                                              (300, (0, 300), False), (1, (0, 7), True), (64, (0, 1000), True)])
def test_fixed_rows_match_vectorised_restatement(K, post, skip_self):
    rows = (0, 60)
    a = oracle.connect_fixed(rows, post, K, skip_self, 123456789, 7, (0.1e-3, 2.2e-3), 1e-4, (-1.0, 1.0))
    b = restate_fixed(rows, post, K, 123456789, 7, (0.1e-3, 2.2e-3), 1e-4, (-1.0, 1.0), skip_self)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_fixed_rows_hold_k_distinct_sorted_targets():
    n, K = 500, 40
    s = oracle.connect_fixed((0, n), (0, n), K, True, 3, 1, (0, 0), 1e-4, (0, 0))
    bounds = np.searchsorted(s["pre"], np.arange(n + 1))
    for i in range(n):
        t = s["post"][bounds[i]:bounds[i + 1]]
        assert np.all(np.diff(t) > 0) and t.min() >= 0 and t.max() < n and i not in t
        assert t.size in (K - 1, K)
    a = oracle.connect_fixed((0, n), (0, n), K, False, 3, 1, (0, 0), 1e-4, (0, 0))
    assert a["pre"].size == n * K
    keep = a["pre"] != a["post"]
    for k in a:  # dropping j == i changes nothing else
        np.testing.assert_array_equal(a[k][keep], s[k])
    # i is among the K sampled targets of its own row with probability K / n
    drops = n * K - s["pre"].size
    assert drops == np.sum(~keep) and abs(drops - K) < 5 * np.sqrt(K * (1 - K / n))


def test_fixed_subsets_are_uniform():
    """K = 3 of n = 8: all 56 subsets equally likely (chi-square over 28,000 rows);
    a sampler that favours low values, or repeats, fails it."""
    from itertools import combinations
    n, K, rows = 8, 3, 28000
    s = oracle.connect_fixed((0, rows), (0, n), K, False, 99, 4, (0, 0), 1e-4, (0, 0))
    t = s["post"].reshape(rows, K)
    code = (1 << t).sum(1)
    subsets = [sum(1 << j for j in c) for c in combinations(range(n), K)]
    obs = np.array([np.sum(code == c) for c in subsets])
    assert obs.sum() == rows
    assert stats.chisquare(obs).pvalue > 1e-4
    # large K / n (many redraw rounds): every target equally often
    s = oracle.connect_fixed((0, 4000), (0, 100), 90, False, 5, 0, (0, 0), 1e-4, (0, 0))
    col = np.bincount(s["post"], minlength=100)
    assert stats.chisquare(col).pvalue > 1e-4


def test_fixed_special_cases():
    n = 200
    full = oracle.connect_fixed((0, 50), (0, n), n, False, 1, 0, (0, 0), 1e-4, (0, 0))
    np.testing.assert_array_equal(full["post"], np.tile(np.arange(n), 50))
    noself = oracle.connect_fixed((0, n), (0, n), n, True, 1, 0, (0, 0), 1e-4, (0, 0))
    assert noself["pre"].size == n * n - n and np.all(noself["pre"] != noself["post"])
    assert oracle.connect_fixed((0, n), (0, n), 0, False, 1, 0, (0, 0), 1e-4, (0, 0))["pre"].size == 0
    with pytest.raises(ValueError):
        oracle.connect_fixed((0, n), (0, n), n + 1, False, 1, 0, (0, 0), 1e-4, (0, 0))
    a = oracle.connect_fixed((0, n), (10, 190), 30, False, 9, 2, (0.1e-3, 0.9e-3), 1e-4, (0, 1))
    c = oracle.connect_fixed((120, 170), (10, 190), 30, False, 9, 2, (0.1e-3, 0.9e-3), 1e-4, (0, 1))
    sel = (a["pre"] >= 120) & (a["pre"] < 170)
    for k in a:  # a row does not depend on the other rows
        np.testing.assert_array_equal(a[k][sel], c[k])
    assert a["post"].min() >= 10 and a["post"].max() < 190
    d = oracle.connect_fixed((0, n), (10, 190), 30, False, 9, 3, (0.1e-3, 0.9e-3), 1e-4, (0, 1))
    assert not np.array_equal(a["post"], d["post"])
