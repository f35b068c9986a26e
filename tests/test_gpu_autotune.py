"""GPU: every value the autotuner may pick leaves the results bit-identical to
the oracle, and the autotuner returns valid options and caches them."""
import time

import numpy as np
import pytest

from brian2cuda_amd import _lib, autotune as at
from brian2cuda_amd import networks as nw

pytestmark = pytest.mark.gpu


def _same_as_oracle(spec, prec, engine, sem, options, oracle_lib, steps=400):
    net, sm, _ = nw.build_hs(spec, prec, engine, options=options)
    ref, rsm, _ = oracle_lib.build_ref(spec, prec, sem)
    net.run(steps)
    ref.run(steps)
    gi, gt = net.spikes(sm[0])
    ri, rt_ = ref.spikes(rsm[0])
    ok = np.array_equal(gi, ri) and np.array_equal(gt, rt_) and net.stats()["events"] == ref.events
    ok = ok and np.array_equal(net.get_state(0, 0), ref.get_state(0, 0))
    net.close()
    return ok


@pytest.mark.parametrize("engine,sem", [("fused", 1), ("plain", 0)])
def test_every_candidate_keeps_the_bits(hs, oracle_lib, engine, sem):
    n_cu = _lib.device_info()["compute_units"]
    spec = nw.brunel(N=2500, K=250, seed=3, monitors=True)
    for key, values in at.candidate_space(n_cu, engine).items():
        for v in values[1:]:
            assert _same_as_oracle(spec, 32, engine, sem, {key: v}, oracle_lib), (key, v)


def test_autotune_returns_valid_options_and_caches(hs, oracle_lib, tmp_path):
    spec = nw.brunel(N=3000, K=300, seed=2, monitors=True)
    cache = str(tmp_path / "tune.json")
    n_cu = _lib.device_info()["compute_units"]
    space = at.candidate_space(n_cu, "fused")
    opts = at.autotune(spec, 32, "fused", steps=400, warmup=100, cache=cache, rounds=1)
    for k, v in opts.items():
        assert k in space and v in space[k] and v != space[k][0]
    t0 = time.time()
    assert at.autotune(spec, 32, "fused", cache=cache) == opts  # from the cache
    assert time.time() - t0 < 0.5
    assert _same_as_oracle(spec, 32, "fused", 1, opts, oracle_lib)


@pytest.mark.parametrize("prec", [32, 64])
def test_tuned_c3_shape_keeps_the_bits(hs, oracle_lib, prec):
    """The launch shapes the autotuner picks for the 100k-neuron network
    (pipelined windows, 64-thread tiles, 3 pull and 3 noise blocks per CU),
    on a smaller network with the same delay span."""
    n_cu = _lib.device_info()["compute_units"]
    opts = {9: 1, 0: 64, 8: 3 * n_cu, 1: 3 * n_cu}
    spec = nw.brunel(N=4000, K=400, seed=7, monitors=True)
    assert _same_as_oracle(spec, prec, "fused", 1, opts, oracle_lib, steps=1500)
