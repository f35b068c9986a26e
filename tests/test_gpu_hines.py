"""Multicompartment trees on the GPU (hs_hines_*) against the oracle's serial
restatement of the reference's spatial state updater: bit-exact v, gates and
soma spike counts, FP64 and FP32 channel arithmetic."""
import numpy as np
import pytest

import oracle
from brian2cuda_amd import morphology as mo

pytestmark = pytest.mark.gpu


def _run_both(tb, prec, steps, chunks=None, options=None):
    g = mo.build_hs(tb, prec, options)
    r = oracle.RefMC(tb, prec)
    for k in (chunks or [steps]):
        g.run(k)
        r.run(k)
    return g, r


def _assert_same(g, r):
    for var in (0, 1, 2, 3):
        a, b = g.get_state(var), r.get_state(var)
        np.testing.assert_array_equal(a, b, err_msg=f"var {var}")
    np.testing.assert_array_equal(g.spike_counts(), r.spike_counts())


@pytest.mark.parametrize("prec", [64, 32])
def test_hines_one_step(hs, prec):
    tb = mo.random_trees(n_trees=16, seed=5)
    g, r = _run_both(tb, prec, 1)
    _assert_same(g, r)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("seed", [1, 2])
def test_hines_spiking_trees(hs, prec, seed):
    tb = mo.random_trees(n_trees=64, seed=seed)
    g, r = _run_both(tb, prec, 2000)
    assert r.spike_counts().sum() > 0
    _assert_same(g, r)


@pytest.mark.parametrize("spl", [1, 7, 1000])
def test_hines_steps_per_launch(hs, spl):
    tb = mo.random_trees(n_trees=8, seed=3)
    g, r = _run_both(tb, 64, 300, chunks=[100, 37, 163], options={0: spl})
    _assert_same(g, r)


@pytest.mark.parametrize("cmin,cmax", [(1, 12), (60, 64), (65, 130), (700, 990)])
def test_hines_tree_sizes(hs, cmin, cmax):
    tb = mo.random_trees(n_trees=12, seed=9, cmin=cmin, cmax=cmax)
    g, r = _run_both(tb, 64, 200)
    _assert_same(g, r)


def test_hines_cable_and_state_roundtrip(hs):
    tb = mo.cable(n_comp=100, I=0.05e-9)
    g, r = _run_both(tb, 64, 500)
    _assert_same(g, r)
    np.testing.assert_array_equal(g.get_state(4), tb.init[4])
    v = np.linspace(-0.01, 0.01, tb.n_comp)
    g.set_state(0, v)
    np.testing.assert_array_equal(g.get_state(0), v)


@pytest.mark.parametrize("prec", [64, 32])
def test_hines_full_batch(hs, prec):
    """The timed batch itself (2048 trees, 528,986 compartments, seed 1234) over
    40 steps in two launches: every compartment's state and every soma's spike
    count bit for bit."""
    tb = mo.random_trees(n_trees=2048, seed=1234)
    g, r = _run_both(tb, prec, 40, chunks=[25, 15])
    assert r.spike_counts().sum() > 0
    _assert_same(g, r)
