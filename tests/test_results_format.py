"""CPU: the result-file readers against files laid out as the reference's
_write_arrays writes them (objects.cu:550-595): raw little-endian arrays,
2-D monitor values time-major (one row of recorded indices per step)."""
import numpy as np

from brian2cuda_amd import results


def test_spike_monitor_files(tmp_path):
    i = np.array([3, 1, 4], "<i4")
    t = np.array([0.1e-3, 0.2e-3, 0.2e-3], "<f8")
    i.tofile(tmp_path / "i")
    t.tofile(tmp_path / "t")
    np.array([0, 1, 0, 1, 1], "<i4").tofile(tmp_path / "c")
    np.array([3], "<i4").tofile(tmp_path / "n")
    s = results.read_spike_monitor(tmp_path / "i", tmp_path / "t", tmp_path / "c", tmp_path / "n")
    assert s["N"] == 3 and s["i"].tolist() == [3, 1, 4] and s["t"].tolist() == t.tolist()
    assert s["count"].tolist() == [0, 1, 0, 1, 1]


def test_state_monitor_time_major(tmp_path):
    rows, w = 4, 3
    vals = np.arange(rows * w, dtype="<f4").reshape(rows, w)  # row r: step r, all recorded indices
    np.arange(rows, dtype="<f8").tofile(tmp_path / "t")
    vals.tofile(tmp_path / "v")
    out = results.read_state_monitor(tmp_path / "t", tmp_path / "v", w, precision=32)
    assert out["values"].shape == (rows, w)
    assert out["values"][2].tolist() == [6.0, 7.0, 8.0]


def test_last_run_info(tmp_path):
    (tmp_path / "info").write_text("0.25 1\n")
    assert results.read_last_run_info(tmp_path / "info") == (0.25, 1.0)
