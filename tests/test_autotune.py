"""CPU: the autotuner's search and cache logic with a synthetic cost (no GPU)."""
import json

from brian2cuda_amd import autotune as at
from brian2cuda_amd import networks as nw


def test_coordinate_descent_finds_the_separable_minimum():
    space = {0: [256, 128, 64], 10: [0, 1024, 2048], 4: [0, 4, 3]}
    cost = {0: {256: 3.0, 128: 1.0, 64: 2.0}, 10: {0: 5.0, 1024: 4.0, 2048: 6.0}, 4: {0: 1.0, 4: 2.0, 3: 3.0}}
    best, t, trials = at.coordinate_descent(space, lambda o: sum(cost[k][v] for k, v in o.items()))
    assert best == {0: 128, 10: 1024, 4: 0} and t == 6.0
    assert trials[0][0] == {0: 256, 10: 0, 4: 0}  # starts from the defaults


def test_coordinate_descent_keeps_defaults_within_the_margin():
    space = {0: [256, 128], 1: [0, 512]}
    best, t, _ = at.coordinate_descent(space, lambda o: 100.0 - (0.5 if o[0] == 128 else 0.0), margin=0.02)
    assert best == {0: 256, 1: 0}


def test_signature_separates_networks_and_forms():
    a = nw.cuba(N=500, seed=1)
    b = nw.cuba(N=600, seed=1)
    sa = at.signature(a, 32, "fused", "gfx950")
    assert sa != at.signature(b, 32, "fused", "gfx950")
    assert sa != at.signature(a, 64, "fused", "gfx950")
    assert sa != at.signature(a, 32, "plain", "gfx950")
    json.dumps({sa: 1})  # usable as a cache key


def test_tune_forms_finds_a_form_that_only_wins_with_its_own_shapes():
    """Form 1 is slower with the default shapes but fastest with its own: a
    single descent from the defaults misses it, one descent per form does not."""
    space = {9: [0, 1], 8: [0, 1024]}

    def cost(o):
        if o[9] == 0:
            return 5.0 if o[8] == 0 else 5.05
        return 5.1 if o[8] == 0 else 4.7

    best, t, _ = at.coordinate_descent(space, cost)
    assert best == {9: 0, 8: 0}
    best, t, trials = at.tune_forms(space, cost, 9)
    assert best == {9: 1, 8: 1024} and t == 4.7
    assert len(trials) >= 4


def test_cache_write_failure_is_not_fatal(tmp_path, monkeypatch):
    """An unwritable cache location only loses the cache: the tuned options are
    still returned."""
    from brian2cuda_amd import autotune as at
    from brian2cuda_amd import _lib
    monkeypatch.setattr(_lib, "device_info", lambda: {"arch": "test", "compute_units": 4})
    monkeypatch.setattr(at, "measure_us_per_step", lambda *a, **k: 1.0)
    blocked = tmp_path / "file"
    blocked.write_text("not a directory")
    logs = []
    opts = at.autotune(nw.cuba(N=200, seed=1), 64, "plain", cache=str(blocked / "sub" / "autotune.json"),
                       log=logs.append)
    assert opts == {}
    assert any("cache not written" in m for m in logs)
