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


def test_chunk_launches_counts_windows_per_graph_chunk():
    # 10,000 steps in chunks of 256 steps, windows of 5: 39 chunks of 52 launches + one of 16 steps (4 launches)
    assert bench.chunk_launches(10000, 256, 5) == 39 * 52 + 4
    assert bench.chunk_launches(10, 256, 5) == 2
    assert bench.chunk_launches(12, 0, 5) == 3  # no graphs: the whole run is one chunk
    assert bench.chunk_launches(7, 4, 1) == 7


def test_drift_horizons_longest_prefix_below_one_step():
    # neuron 0 agrees exactly until step 1000, then a spike moves by 3 steps
    i = np.zeros(4, np.int32)
    t64 = np.array([50, 400, 900, 1200], np.int64)
    t32 = np.array([50, 400, 900, 1203], np.int64)
    rows, below = bench.drift_horizons(i, t64, i, t32, 1, 2000)
    assert below == 1000
    assert rows["1000"]["rms_dt_steps"] == 0.0 and rows["1500"]["rms_dt_steps"] > 1.0
    assert "3000" not in rows  # past the run


def test_hines_valu_roofline_from_committed_counters():
    out = {"compartments": 528986, "fp64": {"us_per_step": 31.9}, "fp32": {"us_per_step": 24.0}}
    r = bench.hines_valu_roofline(out, "f64", 2048, 1234)
    assert r["bound"] == "valu"
    # 4,380 wave instructions per tree-step (the committed counters) x 2048 trees / 31.9 us
    assert abs(r["valu_instr_per_tree_step"] - 4380) < 1
    assert abs(r["achieved"] - 4380.4 * 2048 / 31.9e-6 / 1e9) < 1
    assert 0.3 < r["frac"] < 0.6
    assert bench.hines_valu_roofline(out, "f64", 100, 1234)["frac"] is None  # another batch: no counters


def test_cpu_info_reports_cores():
    ci = bench.cpu_info()
    assert ci["nproc"] >= 1
