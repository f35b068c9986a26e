"""GPU parity of connect(condition=..., p=...) given as code
(hs_net_connect_code: the connect passes compiled at run time with the
condition and the per-pair probability): the bundle table read back through
the C ABI equals the oracle's synapses (oracle.connect_code, the same
expressions compiled for the host) stably sorted by (pre, delay), in both
precisions; a network connected this way simulates to the oracle's bits; bad
expressions are refused without a launch."""
import numpy as np
import pytest

from brian2cuda_amd import codegen as cg
from brian2cuda_amd import networks as nw
from brian2cuda_amd import runtime as rt
from brian2cuda_amd._lib import HsError

pytestmark = pytest.mark.gpu

DT = 1e-4
NAMES = ["v", "ge", "gi"]  # the CUBA model's variables, in the runtime's order


def _bundle_order(s):
    order = np.lexsort((s["delay_steps"], s["pre"]))
    return {k: v[order] for k, v in s.items()}


CASES = [
    # n_src, n_tgt, pre, post, condition, p (number or expression), K, autapses, delay (s), wmode, w
    (1500, 1500, None, None, "abs(i - j) < 20", 1.0, 0, False, (0.1e-3, 0.1e-3), rt.HS_W_SCALAR, (0.5, 0.5)),
    (3000, 2000, (200, 2900), (5, 1990), "(i + j) % 3 != 0", 0.05, 0, True, (0.5e-3, 4e-3), rt.HS_W_SYNAPSE,
     (-1.0, 2.0)),
    (1200, 1200, None, None, "j < 2 * i and i - j != 7", 0.4, 0, True, (0.0, 2e-3), rt.HS_W_SYNAPSE, (0.0, 1.0)),
    (2000, 2000, None, None, "i != j", "exp(-(v_pre - v_post)**2 / (2 * 0.02**2))", 0, True, (0.2e-3, 1e-3),
     rt.HS_W_SYNAPSE, (1.0, 3.0)),
    (1000, 1500, None, None, "j % 2 == 0 and v_post > 0.25", 1.0, 60, True, (0.3e-3, 0.9e-3), rt.HS_W_SYNAPSE,
     (1.0, 2.0)),
    (900, 900, None, None, "", "0.3 * ((i // 7) % 2) + 0.1 * (j > N_post / 2)", 0, True, (0.0, 0.0),
     rt.HS_W_SCALAR, (1.0, 1.0)),
    (1600, 1400, None, (100, 1300), "rand() < 0.4 and randn() > -0.5", "0.5 + 0.5 * (rand() < 0.6)", 0, True,
     (0.1e-3, 2e-3), rt.HS_W_SYNAPSE, (0.0, 1.0)),  # draws in both expressions
    (800, 800, None, None, "i > N_pre", 0.5, 0, True, (0.0, 0.0), rt.HS_W_SCALAR, (1.0, 1.0)),  # empty
]


def _setup(prec, n_src, n_tgt):
    net = rt.Net(prec, "plain", DT, 1)
    sizes = [n_src] if n_src == n_tgt else [n_src, n_tgt]
    pos = []
    for g, n in enumerate(sizes):
        net.add_group(rt.HS_MODEL_LIF_CUBA, n, nw.cuba_params())
        x = np.random.default_rng(100 + g).random(n)  # positions in v, read by the expressions
        net.set_state(g, 0, x)
        pos.append(x)
    return net, pos


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_connect_code_matches_oracle(hs, oracle_lib, prec, case):
    n_src, n_tgt, pre, post, cond, p, K, autapses, delay, wmode, w = CASES[case]
    print(f"connect code case {case} prec {prec}: {cond!r} p={p!r} K={K}", flush=True)
    net, pos = _setup(prec, n_src, n_tgt)
    tgt = 0 if n_src == n_tgt else 1
    c, cconst = cg.connect_expression(cond, {}, NAMES, NAMES) if cond else ("", None)
    if isinstance(p, str):
        pe, pconst = cg.connect_expression(p, {}, NAMES, NAMES, first_draw=c.count("hs_crand_"))
        assert pconst is None
        pnum = 1.0
    else:
        pe, pnum = "", p
    seed, stream = 777 + case, 5
    pid = net.connect_code(0, tgt, 1, c, pe, pnum, K, pre, post, autapses, seed, stream, delay, wmode, w)
    ref = oracle_lib.connect_code(pre or (0, n_src), post or (0, n_tgt), c, pe, pnum, K,
                                  (not autapses) and tgt == 0, seed, stream, delay, DT, w, [pos[0], 0 * pos[0],
                                                                                            0 * pos[0]],
                                  [pos[tgt], 0 * pos[tgt], 0 * pos[tgt]], prec)
    ref = _bundle_order(ref)
    info = net.pathway_info(pid)
    assert info["nsyn"] == ref["pre"].size
    if case != len(CASES) - 1:
        assert ref["pre"].size > 500
    if info["nsyn"]:
        gp, gq, gd, gw = net.synapses(pid)
        np.testing.assert_array_equal(gp, ref["pre"])
        np.testing.assert_array_equal(gq, ref["post"])
        np.testing.assert_array_equal(gd, ref["delay_steps"])
        rw = ref["w"] if wmode == rt.HS_W_SYNAPSE else np.full(ref["pre"].size, w[0])
        if prec == 32 and wmode == rt.HS_W_SYNAPSE:
            rw = rw.astype(np.float32).astype(np.float64)
        np.testing.assert_array_equal(gw, rw)
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", [("plain", 0), ("fused", 1)])
def test_conditional_cuba_simulates_to_oracle_bits(hs, oracle_lib, prec, engine, sem):
    """C1's neurons connected by a probability expression and by a condition,
    both compiled for the device (networks.cuba_conditional), simulated in both
    engine forms and precisions: bits equal the oracle's."""
    spec = nw.cuba_conditional(N=4000, seed=9)
    print(f"cuba_conditional prec {prec} engine {engine}", flush=True)
    net, gsm, _ = nw.build_hs(spec, prec, engine)
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, sem)
    for p in range(2):
        assert net.pathway_info(p)["nsyn"] == oracle_lib.random_pathway(spec, spec.paths[p], precision=prec)["pre"].size
    for k in (100, 400):
        net.run(k)
        ref.run(k)
    gi, gt = net.spikes(gsm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(ri) > 100 and np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert np.array_equal(net.get_state(0, 0), ref.get_state(0, 0))
    assert net.stats()["events"] == ref.events
    net.close()


def test_connect_code_refusals(hs):
    net, _ = _setup(64, 100, 100)
    with pytest.raises(HsError, match="variables"):  # a variable the group does not have
        net.connect_code(0, 0, 1, "hs_pre(3) > 0.5", "", 1.0)
    with pytest.raises(HsError, match="expressions only"):
        net.connect_code(0, 0, 1, "i > 0; int x = 1", "", 1.0)
    with pytest.raises(HsError, match="do not compile"):
        net.connect_code(0, 0, 1, "i > undefined_name", "", 1.0)
    with pytest.raises(HsError, match="reserved"):  # the generated function's own names
        net.connect_code(0, 0, 1, "((const double *)hs_a.pre[5])[0] > 0.5", "", 1.0)
    with pytest.raises(HsError, match="fixed out-degree"):
        net.connect_code(0, 0, 1, "", "0.5", 1.0, K=5)
    # the network still works after the refusals
    pid = net.connect_code(0, 0, 1, "i == j", "", 1.0)
    assert net.pathway_info(pid)["nsyn"] == 100
    net.close()


CUBA_EQS = """
dv/dt  = (ge+gi-(v-El))/taum : volt (unless refractory)
dge/dt = -ge/taue : volt
dgi/dt = -gi/taui : volt
"""
CUBA_NS = {"taum": 20e-3, "taue": 5e-3, "taui": 10e-3, "Vt": -50e-3, "Vr": -60e-3, "El": -49e-3, "we": 1.62e-3,
           "wi": -9e-3}


@pytest.mark.parametrize("prec,engine,sem", [(64, "plain", 0), (32, "fused", 1)])
def test_generator_syntax_network(hs, oracle_lib, prec, engine, sem):
    """A CUBA network described with Brian2's generator syntax (device layer):
    excitatory 'k for k in range(i-60, i+61) if k != i and (k + i) % 3 == 0'
    and inhibitory sources from a subgroup 'k for k in sample(N_post, size=40)
    if k % 2 == 1', both generated on the device; spikes and state equal the
    oracle's."""
    from brian2cuda_amd import device as dv
    n = 4000
    v0 = -0.060 + 0.010 * np.random.default_rng(4).random(n)
    g = dv.GroupDesc("g", n, equations=CUBA_EQS, threshold="v>Vt", reset="v = Vr", refractory=5e-3,
                     namespace=CUBA_NS, init={"v": v0})
    syns = [dv.SynapsesDesc("se", "g", "g", "ge += we", namespace=CUBA_NS, source_range=(0, 3200),
                            j="k for k in range(i - 60, i + 61) if k != i and (k + i) % 3 == 0", delay=1e-4, seed=5),
            dv.SynapsesDesc("si", "g", "g", "gi += wi", namespace=CUBA_NS, source_range=(3200, n),
                            j="k for k in sample(N_post, size=40) if k % 2 == 1", delay=1e-4, seed=6)]
    mons = [dv.MonitorDesc("spikes", "spike", "g")]
    spec, _ = dv.network_spec([g], syns, mons, dt=1e-4, seed=8)
    print(f"generator-syntax network prec {prec} engine {engine}", flush=True)
    net, gsm, _ = nw.build_hs(spec, prec, engine)
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, sem)
    for k in (200, 300):
        net.run(k)
        ref.run(k)
    gi, gt = net.spikes(gsm[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(ri) > 100 and np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    assert np.array_equal(net.get_state(0, 0), ref.get_state(0, 0))
    assert net.stats()["events"] == ref.events
    net.close()


# brian2cuda's test_default_function_implementations (tests/test_cuda_generator.py:20-90):
# (function, zero point, offset)
DEFAULT_FUNCS = [("sin", 0, 0), ("cos", 1, 0), ("tan", 0, 0), ("sinh", 0, 0), ("cosh", 1, 0), ("tanh", 0, 0),
                 ("exp", 1, 0), ("sqrt", 0, 0), ("ceil", 0, 0), ("floor", 0, 0), ("arcsin", 0, 0),
                 ("arctan", 0, 0), ("int", 0, 0), ("arccos", 0, 1), ("log", 0, 1), ("log10", 0, 1),
                 ("abs", 1, -1), ("sign", 1, 1)]


@pytest.mark.parametrize("prec", [64, 32])
def test_default_function_implementations(hs, prec):
    """Each default function in a connect condition on a one-neuron group,
    'f(i + offset) == zero_point' with an integer and with a real argument
    (the reference's int and double overloads), and clip: exactly one synapse
    each, generated on the device."""
    conds = [c for f, z, o in DEFAULT_FUNCS for c in (f"{f}(i + {o}) == {z}", f"{f}(i * 1.0 + {o}) == {z}")]
    conds += ["clip(i+1,-1,0)==0", "clip(i+1,-1.0,0.0)==0"]
    net, _ = _setup(prec, 1, 1)
    for cond in conds:
        c, const = cg.connect_expression(cond)
        assert const is None
        pid = net.connect_code(0, 0, 1, c, "", 1.0)
        assert net.pathway_info(pid)["nsyn"] == 1, cond
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_random_values_init_synapses_fixed_and_random_seed(hs, prec):
    """brian2cuda's test (tests/test_random_number_generation.py:563-600) on the
    device: four connect('randn() + 0.5 < rand()') over a 10-neuron group, two
    with fresh seeds and two with seed 12345678: the seeded pair gives equal
    synapses, every other pair different ones."""
    c, _ = cg.connect_expression("randn() + 0.5 < rand()")
    fresh = np.random.SeedSequence().generate_state(2, np.uint64)
    net, _ = _setup(prec, 10, 10)
    idx = []
    for s in (int(fresh[0]), 12345678, int(fresh[1]), 12345678):
        pid = net.connect_code(0, 0, 1, c, "", 1.0, 0, None, None, True, s, 0)
        pre, post, _, _ = net.synapses(pid) if net.pathway_info(pid)["nsyn"] else (np.zeros(0),) * 4
        idx.append(np.hstack([pre, post]))
    same = lambda a, b: a.shape == b.shape and np.array_equal(a, b)  # noqa: E731
    assert same(idx[1], idx[3]) and idx[1].size > 0
    assert not same(idx[0], idx[1]) and not same(idx[0], idx[2]) and not same(idx[1], idx[2])
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("cond,p,K,n", [("", 0.1, 0, 2), ("i % 3 != 0", 0.5, 0, 3), ("", 1.0, 7, 16),
                                        ("j < i", "0.2", 0, 4)])
def test_several_synapses_per_pair_matches_oracle(hs, oracle_lib, prec, cond, p, K, n):
    """connect(n=...): the bundle table equals the oracle's synapses (each
    selected pair n times, each copy its own delay and weight) stably sorted
    by (pre, delay), so copies of one pair keep their order within a bundle."""
    net, pos = _setup(prec, 1200, 900)
    c = cg.connect_expression(cond)[0] if cond else ""
    pe, pnum = (cg.connect_expression(p)[0], 1.0) if isinstance(p, str) else ("", p)
    print(f"connect n={n} prec {prec}: {cond!r} p={p!r} K={K}", flush=True)
    pid = net.connect_code(0, 1, 1, c, pe, pnum, K, None, None, True, 31, 2, (0.1e-3, 3e-3), rt.HS_W_SYNAPSE,
                           (0.0, 1.0), n)
    ref = _bundle_order(oracle_lib.connect_code((0, 1200), (0, 900), c, pe, pnum, K, False, 31, 2, (0.1e-3, 3e-3), DT,
                                                (0.0, 1.0), (), (), prec, n))
    assert net.pathway_info(pid)["nsyn"] == ref["pre"].size > 5000
    gp, gq, gd, gw = net.synapses(pid)
    np.testing.assert_array_equal(gp, ref["pre"])
    np.testing.assert_array_equal(gq, ref["post"])
    np.testing.assert_array_equal(gd, ref["delay_steps"])
    rw = ref["w"].astype(np.float32).astype(np.float64) if prec == 32 else ref["w"]
    np.testing.assert_array_equal(gw, rw)
    with pytest.raises(HsError, match="per pair"):
        net.connect_code(0, 1, 1, "", "", 0.1, 0, None, None, True, 31, 2, 0.0, rt.HS_W_SCALAR, 1.0, 17)
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("wexpr,cond,p,n", [("0.01*i + j + 2*v_pre - v_post", "", 0.1, 1),
                                            ("rand()*0.5 + 0.1*randn()", "j % 2 == 0", 0.3, 2),
                                            ("exp(-abs(i - j)/50.0)", "", "0.4*(rand() < 0.5)", 3)])
def test_weight_expression_matches_oracle(hs, oracle_lib, prec, wexpr, cond, p, n):
    """S.w = '<expression>' after connect, evaluated by the device generator per
    synapse (draws keyed by pair and copy): the bundle table's weights equal the
    oracle's (the same expression compiled for the host), in the network's
    precision."""
    net, pos = _setup(prec, 1500, 1100)
    c = cg.connect_expression(cond)[0] if cond else ""
    pe, pnum = (cg.connect_expression(p)[0], 1.0) if isinstance(p, str) else ("", p)
    we, wconst = cg.connect_expression(wexpr, {}, NAMES, NAMES, first_draw=(c + pe).count("hs_crand_"))
    assert wconst is None
    print(f"weight expression {wexpr!r} prec {prec} n={n}", flush=True)
    pid = net.connect_code(0, 1, 1, c, pe, pnum, 0, None, None, True, 41, 3, (0.1e-3, 2e-3), rt.HS_W_SYNAPSE,
                           (0.0, 0.0), n, we)
    z = [0 * pos[0], 0 * pos[0]]
    ref = _bundle_order(oracle_lib.connect_code((0, 1500), (0, 1100), c, pe, pnum, 0, False, 41, 3, (0.1e-3, 2e-3),
                                                DT, (0.0, 0.0), [pos[0]] + z, [pos[1], 0 * pos[1], 0 * pos[1]], prec,
                                                n, we))
    assert net.pathway_info(pid)["nsyn"] == ref["pre"].size > 5000
    gp, gq, gd, gw = net.synapses(pid)
    np.testing.assert_array_equal(gp, ref["pre"])
    np.testing.assert_array_equal(gd, ref["delay_steps"])
    rw = ref["w"].astype(np.float32).astype(np.float64) if prec == 32 else ref["w"]
    np.testing.assert_array_equal(gw, rw)
    with pytest.raises(HsError, match="per-synapse weights"):
        net.connect_code(0, 1, 1, "", "", 0.1, 0, None, None, True, 41, 3, 0.0, rt.HS_W_SCALAR, 1.0, 1, we)
    net.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_random_values_set_synapses_fixed_and_random_seed(hs, prec):
    """brian2cuda's test_random_values_set_synapses_random_seed / _fixed_seed
    (tests/test_random_number_generation.py:480-511) on the device: the values
    'rand() + randn()' set per synapse by the generator over connect() on a
    10-neuron group, the connect call's seed as seed(): values vary, fresh seeds
    differ, equal seeds agree."""
    we, _ = cg.connect_expression("rand() + randn()")
    fresh = np.random.SeedSequence().generate_state(2, np.uint64)
    net, _ = _setup(prec, 10, 10)
    v = []
    for s in (int(fresh[0]), int(fresh[1]), 12345678, 12345678):
        pid = net.connect_code(0, 0, 1, "", "", 1.0, 0, None, None, True, s, 0, 0.0, rt.HS_W_SYNAPSE, (0.0, 0.0), 1,
                               we)
        v.append(net.synapses(pid)[3])
    assert all(np.var(x) > 0 for x in v) and v[0].size == 100
    assert np.var(v[0] - v[1]) > 0
    assert np.array_equal(v[2], v[3])
    net.close()
