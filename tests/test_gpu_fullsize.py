"""Full-size Brunel-Hakim network (N = 100,000, K = 1000, 1e8 synapses, 46
delay slots): too large for the serial oracle, so the GPU run is checked through
properties that do not depend on size.

  events   : the device's delivered-event count equals the count derived from
             the recorded spikes and the connectivity (each spike (i, s) delivers
             the synapses of i whose arrival step s + d falls inside the run)
  refractoriness : two spikes of one neuron are >= 2 ms apart
  forms    : FP32 batched and FP64 plain fire at the same rate (within 3 %)
and, over the first 1000 steps (100 ms: the oracle finishes them in seconds),
bit for bit against the oracle in the form's own precision and semantics: every
spike, the final state and the event count, with default and with tuned
launch shapes.
"""
import dataclasses

import numpy as np
import pytest

from brian2cuda_amd import networks as nw

pytestmark = pytest.mark.gpu

STEPS = 1200


@pytest.fixture(scope="module")
def c3():
    spec = nw.brunel(N=100_000, K=1000, seed=1234, monitors=False)
    p = spec.paths[0]
    dmin = int(p.delay_steps.min())
    D = int(p.delay_steps.max()) - dmin + 1
    key = p.pre.astype(np.int64) * D + (p.delay_steps - dmin)
    per = np.bincount(key, minlength=spec.n_neurons * D).reshape(spec.n_neurons, D)
    return dataclasses.replace(spec, spike_monitors=[0]), np.cumsum(per, axis=1), dmin, D


def _run(spec, prec, engine, steps):
    net, smon, _ = nw.build_hs(spec, prec, engine, spike_capacity=1 << 22)
    net.run(steps)
    i, t = net.spikes(smon[0])
    ev = net.stats()["events"]
    net.close()
    return i, t, ev


def _expected_events(i, t, cum, dmin, D, steps):
    last_slot = np.minimum(D - 1, steps - 1 - t - dmin)  # largest delay slot arriving by step steps-1
    ok = last_slot >= 0
    return int(cum[i[ok], last_slot[ok]].sum())


@pytest.mark.parametrize("prec,engine", [(32, "fused"), (64, "plain")])
def test_fullsize_event_count_and_refractoriness(hs, c3, prec, engine):
    spec, cum, dmin, D = c3
    i, t, ev = _run(spec, prec, engine, STEPS)
    assert len(i) > 10_000
    assert ev == _expected_events(i, t, cum, dmin, D, STEPS)
    o = np.lexsort((t, i))
    i, t = i[o], t[o]
    same = i[1:] == i[:-1]
    assert np.all(np.diff(t)[same] >= 20)


def test_fullsize_forms_fire_alike(hs, c3):
    spec = c3[0]
    n32 = len(_run(spec, 32, "fused", STEPS)[0])
    n64 = len(_run(spec, 64, "plain", STEPS)[0])
    assert abs(n32 - n64) <= 0.03 * n64


# the launch shapes the autotuner picks for the FP32 form on this network
# (pipelined windows, 64-thread update tiles, 768 noise and 768 pull blocks)
TUNED_FP32 = {9: 1, 0: 64, 8: 768, 1: 768}


@pytest.mark.parametrize("prec,engine,sem,options", [(64, "plain", 0, None), (32, "fused", 1, None),
                                                     (32, "fused", 1, TUNED_FP32), (64, "plain", 0, {0: 64})])
def test_fullsize_bit_exact_vs_oracle(hs, oracle_lib, c3, prec, engine, sem, options):
    spec = c3[0]
    steps = 1000
    net, smon, _ = nw.build_hs(spec, prec, engine, spike_capacity=1 << 22, options=options)
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, sem)
    net.run(steps)
    ref.run(steps)
    gi, gt = net.spikes(smon[0])
    ri, rt_ = ref.spikes(rsm[0])
    assert len(ri) > 10_000
    assert np.array_equal(gi, ri) and np.array_equal(gt, rt_)
    for v in range(hs.MODEL_NVARS[spec.groups[0].model]):
        assert np.array_equal(net.get_state(0, v), ref.get_state(0, v)), f"var {v}"
    assert net.stats()["events"] == ref.events
    net.close()
