"""COBAHH at its largest configured size: N = 512,000 Hodgkin-Huxley neurons,
80/20 exc/inh, p = 80/N (4.1e7 synapses), seed 1234, exponential Euler, delay 0.

  parity     : both timed forms (FP32 batched, FP64 plain), the other two
               (FP64 fused, FP32 plain) and the FP32 pipelined form (built in
               and as code) bit for bit against the oracle in their
               own precision and semantics over 60 steps in two chunks: every
               spike, the final state of all six variables and the event count.
               NaN counts as equal to NaN: the reference's alpha_n is 0/0 where v
               lands exactly on VT + 15 mV (examples/cobahh.py:174), which FP32
               reaches for about one neuron in 30,000.
  invariants : over 1000 steps (100 ms), which the oracle is not run for, the
               delivered-event count equals the count derived from the recorded
               spikes and the connectivity (delay 0: each spike of neuron i
               delivers all out-degree(i) synapses in its own step, the last
               step included), and two spikes of one neuron are at least the
               refractory 3 ms (30 steps) apart.
The oracle runs its OpenMP step loop (bit-identical to its serial loop).
"""
import os

import numpy as np
import pytest

from brian2cuda_amd import networks as nw

pytestmark = pytest.mark.gpu

N = 512_000


def _threads():
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(omp) if omp and omp.isdigit() else n, 16))


@pytest.fixture(scope="module")
def c2_512k():
    spec = nw.cobahh(N=N, seed=1234)
    out_degree = sum(np.bincount(p.pre, minlength=N) for p in spec.paths)
    return spec, out_degree


def _parity(net, smon, ref, rsmon, nvars=6):
    gi, gt = net.spikes(smon[0])
    ri, rt = ref.spikes(rsmon[0])
    assert len(ri) > 1000, "the network must fire within the compared steps"
    assert len(gi) == len(ri), f"{len(gi)} spikes on the GPU vs {len(ri)} in the oracle"
    assert np.array_equal(gt, rt) and np.array_equal(gi, ri)
    for v in range(nvars):
        assert np.array_equal(net.get_state(0, v), ref.get_state(0, v), equal_nan=True), f"var {v}"
    assert net.stats()["events"] == ref.events


@pytest.mark.parametrize("prec,engine,sem", [(32, "fused", 1), (64, "plain", 0), (64, "fused", 1), (32, "plain", 0)])
def test_cobahh_512k_bit_exact(hs, oracle_lib, c2_512k, prec, engine, sem):
    spec = c2_512k[0]
    print(f"cobahh 512k parity: prec {prec} engine {engine}", flush=True)
    net, smon, _ = nw.build_hs(spec, prec, engine, spike_capacity=1 << 24)
    ref, rsmon, _ = oracle_lib.build_ref(spec, prec, sem, threads=_threads())
    for k in (35, 25):
        net.run(k)
        ref.run(k)
    _parity(net, smon, ref, rsmon)
    if prec == 64:
        assert not np.isnan(ref.get_state(0, 0)).any()
    net.close()
    ref.close()


@pytest.mark.parametrize("code", [False, True])
def test_cobahh_512k_pipelined_bit_exact(hs, oracle_lib, c2_512k, code):
    """The timed FP32 form, one pipelined launch per step (option 9): its
    8,000 update waves take the 8-wave-per-SIMD window kernel, built in and,
    with the model given as examples/cobahh.py's equations, compiled at run
    time."""
    spec = nw.code_cobahh(N=N, seed=1234) if code else c2_512k[0]
    print(f"cobahh 512k pipelined parity: code {code}", flush=True)
    net, smon, _ = nw.build_hs(spec, 32, "fused", spike_capacity=1 << 24, options={9: 1})
    ref, rsmon, _ = oracle_lib.build_ref(spec, 32, 1, threads=_threads())
    for k in (35, 25):
        net.run(k)
        ref.run(k)
    _parity(net, smon, ref, rsmon)
    net.close()
    ref.close()


@pytest.mark.parametrize("prec,engine", [(32, "fused"), (64, "plain")])
def test_cobahh_512k_invariants_over_1000_steps(hs, c2_512k, prec, engine):
    spec, out_degree = c2_512k
    steps = 1000
    print(f"cobahh 512k invariants: prec {prec} engine {engine}", flush=True)
    net, smon, _ = nw.build_hs(spec, prec, engine, spike_capacity=1 << 24)
    net.run(steps)
    i, t = net.spikes(smon[0])
    ev = net.stats()["events"]
    net.close()
    assert len(i) > 50_000
    assert t.min() >= 0 and t.max() < steps and i.min() >= 0 and i.max() < N
    assert ev == int(out_degree[i].sum())
    o = np.lexsort((t, i))
    i, t = i[o], t[o]
    same = i[1:] == i[:-1]
    assert np.all(np.diff(t)[same] >= 30)
