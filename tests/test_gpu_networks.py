"""GPU network parity: the device step loop against the oracle on the same inputs.

Deterministic models (CUBA exact, COBAHH exponential Euler) and the noise model
(Philox draws are a pure function of (seed, neuron, step)) are compared bit for
bit: spike indices and steps, StateMonitor traces and the final state.
  plain engine  vs oracle SEQ (Brian2's in-place order)
  fused engine  vs oracle ACC (effects applied at the start of the next step)
"""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw

pytestmark = pytest.mark.gpu

ENGINES = [("plain", 0), ("fused", 1)]


def _compare(spec, prec, engine, sem, steps, oracle_lib, options=None, chunks=None):
    net, smon, stmon = nw.build_hs(spec, prec, engine, options=options)
    ref, rsmon, rstmon = oracle_lib.build_ref(spec, prec, sem)
    for k in (chunks or [steps]):
        net.run(k)
        ref.run(k)
    for a, b in zip(smon, rsmon):
        gi, gt = net.spikes(a)
        ri, rt = ref.spikes(b)
        assert len(gi) == len(ri), f"{len(gi)} spikes on the GPU vs {len(ri)} in the oracle"
        assert np.array_equal(gt, rt) and np.array_equal(gi, ri)
    for (a, b, m) in zip(stmon, rstmon, spec.state_monitors):
        w = m[3] - m[2]
        assert np.array_equal(net.state_monitor(a, w), ref.state_monitor(b, w))
    for gid, g in enumerate(spec.groups):
        for v in range(len(g.init) if g.model != 3 else 6):
            assert np.array_equal(net.get_state(gid, v), ref.get_state(gid, v)), f"group {gid} var {v}"
    assert net.stats()["events"] == ref.events
    return net, ref


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_cuba_parity(hs, oracle_lib, prec, engine, sem):
    """C1 at full size (4000 neurons, p=0.02, delay one dt), 0.3 s."""
    spec = nw.cuba(N=4000, seed=11)
    net, ref = _compare(spec, prec, engine, sem, 3000, oracle_lib)
    assert ref.events > 10000


@pytest.mark.parametrize("engine,sem", ENGINES)
def test_cuba_chunked_runs_and_options(hs, oracle_lib, engine, sem):
    """Several run() calls (graph chunks of odd sizes) and another block size give the same bits."""
    spec = nw.cuba(N=1500, seed=5)
    _compare(spec, 64, engine, sem, None, oracle_lib, options={0: 128, 2: 97}, chunks=[1, 250, 333, 416])


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_cobahh_parity(hs, oracle_lib, prec, engine, sem):
    """C2 at 4000 neurons (p=0.02, delay 0: two launches per step in both forms), 0.1 s."""
    spec = nw.cobahh(N=4000, seed=3)
    _compare(spec, prec, engine, sem, 1000, oracle_lib)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("engine,sem", ENGINES)
def test_brunel_parity_small(hs, oracle_lib, prec, engine, sem):
    """C3 scaled to N=4000, K=400 (heterogeneous delays 5..50 steps, noise), 0.2 s."""
    spec = nw.brunel(N=4000, K=400, seed=2, monitors=True)
    net, ref = _compare(spec, prec, engine, sem, 2000, oracle_lib)
    assert ref.events > 0


@pytest.mark.parametrize("engine,sem", ENGINES)
def test_batch_parity_small(hs, oracle_lib, engine, sem):
    """C4 scaled to 12 networks of 200..1500 neurons, FP32, 0.1 s."""
    spec = nw.batch_cuba(n_nets=12, nmin=200, nmax=1500, seed=9)
    _compare(spec, 32, engine, sem, 1000, oracle_lib)
