"""CPU: connect(condition=..., p=...) given as code -- the expression
translator (brian2cuda_amd.codegen.connect_expression) and the oracle's
restatement (oracle.connect_code, the expressions compiled for the host).

Pins:
  * deterministic sets (p = 1, or a probability that is 0 or 1 per pair) equal
    Python's own evaluation of the Brian2 expression over every pair: integer
    arithmetic, // and % (floor, divisor's sign), true division, powers,
    chained comparisons, and / or / not, functions, subgroup-relative i and j,
    N_pre / N_post, source and target variables;
  * with a constant p the synapses are those of connect_random (the same
    draws) kept where the condition holds, for both algorithms (skips below
    0.25, one draw per pair above); a fixed out-degree likewise filters
    connect_fixed; a probability expression that is constant per pair draws
    what connect_random draws for that constant;
  * rand() / randn() in a condition or probability are the draws of their call
    site for the pair (seed, stream, i, j): a known answer against the oracle's
    Philox words, and brian2cuda's
    test_random_values_init_synapses_fixed_and_random_seed
    (brian2cuda/tests/test_random_number_generation.py:563-600);
  * constants fold (a probability that reads no index is passed as a number),
    and what the translator does not know is refused with its name.
"""
import numpy as np
import pytest

import oracle
from brian2cuda_amd import codegen as cg

DT = 1e-4


def _pairs(out, pre_lo=0, post_lo=0):
    return set(zip((out["pre"] - pre_lo).tolist(), (out["post"] - post_lo).tolist()))


def _code(cond, p_expr="", p=1.0, K=0, pre=(0, 40), post=(0, 30), pre_vars=(), post_vars=(), skip_self=False,
          seed=11, delay=(0.0, 0.0), w=(1.0, 1.0), precision=64):
    return oracle.connect_code(pre, post, cond, p_expr, p, K, skip_self, seed, 0, delay, DT, w, pre_vars, post_vars,
                               precision)


EXPRS = [
    "i != j",
    "abs(i - j) < 3",
    "(i - j) // 3 == (i % 4) - 2 or (j - 7) % 5 == 1",
    "(i - 20) // 7 + (j - 11) % 4 > 0",
    "i / 3 > j / 2",
    "i * N_post / N_pre - j <= 1.5 and not (j > 2 * i)",
    "1 < j - i <= 5",
    "(i - 15) ** 3 > 40 * j",
    "2 ** 3 * i < j * 5",
    "clip(i - j, -2, 2) == 2",
    "int(i / 4) == int(j / 3)",
    "floor(i / 3.5) == ceil(j / 4.5)",
    "sign(i - j) == -1 and j - i < 6",
    "sqrt(i) + log(j + 1) < 5",
    "exp(-(i - j) ** 2 / 8.0) > 0.3",
]


@pytest.mark.parametrize("expr", EXPRS)
def test_condition_equals_python(expr):
    """p = 1: the pairs are exactly those where Python evaluates the expression
    true (i, j relative to the subgroup ranges)."""
    c, const = cg.connect_expression(expr)
    assert const is None
    pre, post = (5, 50), (3, 38)
    out = _code(c, pre=pre, post=post)
    ni, nj = pre[1] - pre[0], post[1] - post[0]
    env = {"N_pre": ni, "N_post": nj, "abs": abs, "sqrt": np.sqrt, "log": np.log, "exp": np.exp,
           "floor": np.floor, "ceil": np.ceil, "sign": np.sign, "clip": lambda x, a, b: min(max(x, a), b), "int": int}
    want = {(i, j) for i in range(ni) for j in range(nj) if eval(expr, dict(env, i=i, j=j))}
    assert _pairs(out, pre[0], post[0]) == want
    assert np.all(np.diff(out["pre"]) >= 0)  # creation order: rows ascending


def test_variables_of_source_and_target():
    rng = np.random.default_rng(3)
    x_pre, y_pre = rng.random(40), rng.random(40)
    x_post = rng.random(30)
    c, _ = cg.connect_expression("x_pre < x_post and y_pre > 0.3 or i == 2 * j", {}, ["x", "y"], ["x"])
    assert "hs_pre(0)" in c and "hs_pre(1)" in c and "hs_post(0)" in c
    out = _code(c, pre_vars=[x_pre, y_pre], post_vars=[x_post])
    want = {(i, j) for i in range(40) for j in range(30) if (x_pre[i] < x_post[j] and y_pre[i] > 0.3) or i == 2 * j}
    assert _pairs(out) == want
    # FP32 networks read the arrays in FP32
    x32 = np.float32(x_pre).astype(np.float64)
    c2, _ = cg.connect_expression("x_pre * 3 < 1.2", {}, ["x"], [])
    out = _code(c2, pre_vars=[x_pre], precision=32)
    assert _pairs(out) == {(i, j) for i in range(40) for j in range(30) if x32[i] * 3 < 1.2}


@pytest.mark.parametrize("p", [0.05, 0.2, 0.3, 0.8])
def test_constant_p_filters_connect_random(p):
    """connect(condition=c, p=p): connect(p)'s synapses where c holds, the same
    draws (geometric skips below 0.25, one per pair above), delays and weights
    included."""
    cond = "(i + j) % 3 != 0 and j < 2 * i + 5"
    c, _ = cg.connect_expression(cond)
    pre, post = (0, 300), (0, 200)
    kw = dict(pre=pre, post=post, delay=(1e-4, 3e-3), w=(0.5, 2.0))
    out = _code(c, p=p, **kw)
    ref = oracle.connect_random(pre, post, p, False, 11, 0, kw["delay"], DT, kw["w"])
    keep = np.array([(i + j) % 3 != 0 and j < 2 * i + 5 for i, j in zip(ref["pre"], ref["post"])], bool)
    assert keep.sum() == out["pre"].size > 100
    for k in ("pre", "post", "delay_steps", "w"):
        assert np.array_equal(out[k], ref[k][keep]), k


def test_fixed_out_degree_filtered():
    c, _ = cg.connect_expression("j % 2 == 0")
    out = _code(c, K=12, pre=(0, 100), post=(0, 80))
    ref = oracle.connect_fixed((0, 100), (0, 80), 12, False, 11, 0, (0.0, 0.0), DT, (1.0, 1.0))
    keep = ref["post"] % 2 == 0
    assert np.array_equal(out["pre"], ref["pre"][keep]) and np.array_equal(out["post"], ref["post"][keep])


def test_probability_expression():
    # constant per pair: the draws of connect_random for that constant (one per pair)
    pe, const = cg.connect_expression("0.5 + 0 * i")
    assert const is None
    out = _code("", pe, pre=(0, 200), post=(0, 150))
    ref = oracle.connect_random((0, 200), (0, 150), 0.5, False, 11, 0, (0.0, 0.0), DT, (1.0, 1.0))
    assert np.array_equal(out["pre"], ref["pre"]) and np.array_equal(out["post"], ref["post"])
    # 0 or 1 per pair: a closed form
    pe, _ = cg.connect_expression("1.0 * (j < 10) + 2.0 * (i == j)")
    out = _code("", pe, pre=(0, 40), post=(0, 30))
    assert _pairs(out) == {(i, j) for i in range(40) for j in range(30) if j < 10 or i == j}
    # distance-dependent with the condition: kept pairs satisfy the condition,
    # and each kept pair's own draw lies below its probability (Brian2's rand() < p)
    x = np.linspace(0.0, 1.0, 60)
    pe, _ = cg.connect_expression("exp(-(x_pre - x_post)**2 / (2 * 0.1**2))", {}, ["x"], ["x"])
    c, _ = cg.connect_expression("i != j")
    out = _code(c, pe, pre=(0, 60), post=(0, 60), pre_vars=[x], post_vars=[x])
    pairs = _pairs(out)
    assert all(i != j for i, j in pairs)
    near = sum(1 for i, j in pairs if abs(x[i] - x[j]) < 0.1)
    far = sum(1 for i, j in pairs if abs(x[i] - x[j]) > 0.4)
    assert near > 100 and far < 5


def test_constant_folding_and_skip_self():
    c, const = cg.connect_expression("2**3 * 0.01 + 20*Hz * 1*ms")
    assert const == pytest.approx(0.1) and c.startswith("(")
    c, const = cg.connect_expression("i != j and True")
    assert const is None
    out = _code(c, pre=(0, 20), post=(0, 20))
    assert _pairs(out) == {(i, j) for i in range(20) for j in range(20) if i != j}
    # skip_self drops j == i (absolute indices) whatever the condition says
    out = _code("", "", 1.0, pre=(0, 20), post=(0, 20), skip_self=True)
    assert _pairs(out) == {(i, j) for i in range(20) for j in range(20) if i != j}


@pytest.mark.parametrize("expr,msg", [("rand(1) < 0.1", "no arguments"), ("foo(i) > 1", "foo"), ("i < unknown", "unknown"),
                                      ("i in [1, 2]", "comparison"), ("[i for i in x]", "not supported"),
                                      ("i / (3 - 3) > 1", "division by zero")])
def test_refusals(expr, msg):
    with pytest.raises(NotImplementedError, match=msg):
        cg.connect_expression(expr)


def _conn_stream(stream, k, normal):
    """the stream of connect call site k (hs_cg_conn_stream in hs_cgfuncs.h)"""
    return (stream & 0xFFFFFF) | ((96 + k + 16 * normal) << 24)


def test_rand_in_condition_known_answer():
    """'rand() < 0.3 and randn() > -1' at p = 1: the pairs whose call-site
    draws, Philox counter (i, tag, j, stream), satisfy it -- about
    0.3 * 0.84 of the pairs -- and a probability expression drawing after the
    condition (first_draw) takes the next call site."""
    c, const = cg.connect_expression("rand() < 0.3 and randn() > -1")
    assert const is None and "hs_crand_u(0)" in c and "hs_crand_n(1)" in c
    seed, stream, tag = 11, 0, 0x43524E44
    out = _code(c)
    want = {(i, j) for i in range(40) for j in range(30)
            if oracle.rand(seed, _conn_stream(stream, 0, 0), i, (j << 32) | tag) < 0.3
            and oracle.randn(seed, _conn_stream(stream, 1, 1), i, (j << 32) | tag) > -1}
    assert _pairs(out) == want and 200 < len(want) < 400
    pe, _ = cg.connect_expression("0.5 * (rand() < 0.5) + 0.5", first_draw=2)
    assert "hs_crand_u(2)" in pe
    out = _code(c, pe)
    assert _pairs(out) <= want and len(_pairs(out)) > 0.5 * len(want)


def test_random_values_init_synapses_fixed_and_random_seed():
    """brian2cuda's test (test_random_number_generation.py:563-600): four
    connect('randn() + 0.5 < rand()') over 10 x 10, two with fresh seeds and two
    with seed 12345678: the two seeded calls give equal synapses, every other
    pair of calls different ones."""
    c, _ = cg.connect_expression("randn() + 0.5 < rand()")
    fresh = np.random.SeedSequence().generate_state(2, np.uint64)
    seeds = [int(fresh[0]), 12345678, int(fresh[1]), 12345678]
    idx = []
    for s in seeds:
        out = _code(c, pre=(0, 10), post=(0, 10), seed=s)
        idx.append(np.hstack([out["pre"], out["post"]]))
    same = lambda a, b: a.shape == b.shape and np.array_equal(a, b)  # noqa: E731
    assert same(idx[1], idx[3])
    assert not same(idx[0], idx[1]) and not same(idx[0], idx[2]) and not same(idx[1], idx[2])


@pytest.mark.parametrize("K", [0, 9])
def test_several_synapses_per_pair(K):
    """connect(n=3): each selected pair three times, consecutively in creation
    order; copy 0 carries the single synapse's delay and weight, the other
    copies their own draws (tag 1 | c << 4), so copies of one pair differ."""
    kw = dict(pre=(0, 60), post=(0, 50), delay=(1e-4, 5e-3), w=(0.5, 2.0), K=K)
    c, _ = cg.connect_expression("(i + j) % 4 != 1")
    one = _code(c, p=0.3 if not K else 1.0, **kw)
    three = oracle.connect_code(kw["pre"], kw["post"], c, "", 0.3 if not K else 1.0, K, False, 11, 0, kw["delay"], DT,
                                kw["w"], (), (), 64, 3)
    m = one["pre"].size
    assert m > 100 and three["pre"].size == 3 * m
    for k in ("pre", "post"):
        assert np.array_equal(three[k], np.repeat(one[k], 3)), k
    for k in ("delay_steps", "w"):
        assert np.array_equal(three[k][0::3], one[k]), k
    assert np.mean(three["w"][1::3] != three["w"][0::3]) > 0.99
    assert np.mean(three["delay_steps"][2::3] != three["delay_steps"][1::3]) > 0.8


def test_weight_expression():
    """S.w = '<expression>' after connect: each synapse's weight is the
    expression at its pair (a closed form of i, j and a source variable),
    draws in it are keyed by the pair and the copy (n=2 copies differ), and a
    uniform 'rand()' weight spreads over [0, 1)."""
    x = np.linspace(0.0, 1.0, 40)
    we, _ = cg.connect_expression("0.01*i + j + 2*x_pre", {}, ["x"], [])
    out = oracle.connect_code((0, 40), (0, 30), "", "", 0.3, 0, False, 11, 0, (0.0, 0.0), DT, (0.0, 0.0), [x], [],
                              64, 1, we)
    ref = oracle.connect_random((0, 40), (0, 30), 0.3, False, 11, 0, (0.0, 0.0), DT, (0.0, 0.0))
    assert np.array_equal(out["pre"], ref["pre"]) and np.array_equal(out["post"], ref["post"])
    assert np.array_equal(out["w"], 0.01 * out["pre"] + out["post"] + 2 * x[out["pre"]])
    wr, _ = cg.connect_expression("rand()")
    two = oracle.connect_code((0, 200), (0, 200), "", "", 0.2, 0, False, 11, 0, (0.0, 0.0), DT, (0.0, 0.0), (), (),
                              64, 2, wr)
    w = two["w"]
    assert np.all((w >= 0) & (w < 1)) and abs(w.mean() - 0.5) < 0.02 and np.all(w[0::2] != w[1::2])


def test_random_values_set_synapses_fixed_and_random_seed():
    """brian2cuda's test_random_values_set_synapses_random_seed / _fixed_seed
    (test_random_number_generation.py:480-511), with the synapse values set by
    a weight expression 'rand() + randn()' over connect() on a 10-neuron group
    and the connect call's seed as seed(): values vary; fresh seeds give
    different values, equal seeds equal ones."""
    we, _ = cg.connect_expression("rand() + randn()")
    fresh = np.random.SeedSequence().generate_state(2, np.uint64)
    v = [oracle.connect_code((0, 10), (0, 10), "", "", 1.0, 0, False, int(s), 0, (0.0, 0.0), DT, (0.0, 0.0), (), (),
                             64, 1, we)["w"] for s in (fresh[0], fresh[1], 12345678, 12345678)]
    assert all(np.var(x) > 0 for x in v) and v[0].size == 100
    assert np.var(v[0] - v[1]) > 0
    assert np.array_equal(v[2], v[3])
