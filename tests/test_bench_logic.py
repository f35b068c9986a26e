"""Host logic of bench.py (no GPU): the drift matcher and the byte model."""
import numpy as np

import bench


def test_match_spikes_identical_trains():
    i = np.array([0, 0, 1, 2, 2, 2], np.int32)
    t = np.array([5, 40, 7, 1, 100, 300], np.int64)
    r = bench.match_spikes(i, t, i.copy(), t.copy(), 3, 50)
    assert r["matched_pairs"] == 6 and r["rms_dt_steps"] == 0.0 and r["unmatched_frac"] == 0.0


def test_match_spikes_shift_and_extra():
    i64 = np.array([0, 0, 1], np.int32)
    t64 = np.array([10, 200, 50], np.int64)
    i32 = np.array([0, 0, 1, 1], np.int32)     # one extra spike, one shifted by 3, one by -4
    t32 = np.array([13, 196, 50, 900], np.int64)
    r = bench.match_spikes(i64, t64, i32, t32, 2, 50)
    assert r["matched_pairs"] == 3
    assert np.isclose(r["rms_dt_steps"], np.sqrt((9 + 16 + 0) / 3))
    assert np.isclose(r["unmatched_frac"], 1 / 7)


def test_match_spikes_outside_window_is_unmatched():
    r = bench.match_spikes(np.array([0]), np.array([0]), np.array([0]), np.array([100]), 1, 50)
    assert r["matched_pairs"] == 0 and r["unmatched_frac"] == 1.0 and r["rms_dt_steps"] is None


def test_state_bytes_follow_brian2_arrays():
    assert bench.state_bytes("brunel", 64) == (17, 9)
    assert bench.state_bytes("brunel", 32) == (9, 5)
