"""CPU pins of the oracle's StateMonitor forms: a monitor of arbitrary indices
with its own dt holds exactly the columns and rows of a full every-step monitor
of the same run (record=[...], dt=k*dt as in test_cpp_cuda_consistency.py:72);
rows follow the monitor's clock (steps that are multiples of k) also when the
monitor is added mid-run; synapse monitors take an interval the same way."""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw


def test_indexed_interval_monitor_is_a_subsample(oracle_lib):
    spec = nw.cuba(N=600, seed=3, record=600)
    ref, _, rst = oracle_lib.build_ref(spec, 64, oracle_lib.SEQ)
    idx = np.array([599, 0, 17, 311, 42], np.int32)
    ref.run(37)                                    # added mid-run: first row at step 40
    m = ref.add_state_monitor_indices(0, 0, idx, every=10)
    m1 = ref.add_state_monitor_indices(0, 2, idx[::-1], every=1)
    ref.run(263)                                   # steps 37..299
    full = ref.state_monitor(rst[0], 600)          # rows = steps 0..299
    got = ref.state_monitor(m, idx.size)
    steps = np.arange(40, 300, 10)
    assert got.shape == (steps.size, idx.size)
    np.testing.assert_array_equal(got, full[steps][:, idx])
    gi = ref.state_monitor(m1, idx.size)
    assert gi.shape == (263, idx.size) and np.any(gi != 0)


def test_interval_is_refused_after_the_first_row(oracle_lib):
    spec = nw.cuba(N=100, seed=1, record=10)
    ref, _, rst = oracle_lib.build_ref(spec, 64, oracle_lib.SEQ)
    ref.run(5)
    with pytest.raises(ValueError):
        ref.set_monitor_interval(rst[0], 10)
    with pytest.raises(ValueError):
        ref.add_state_monitor_indices(0, 0, [100], every=2)   # index outside the group
