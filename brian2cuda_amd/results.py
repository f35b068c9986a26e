"""Readers of the result files that hs_net_write_monitor / hs_net_write_last_run_info
write: raw little-endian arrays, one per file, as the reference's _write_arrays
writes them (brian2cuda/templates/objects.cu:550-563, 2-D monitors time-major
:589-595) and as Brian2's device reads them back (numpy.fromfile with the
variable's dtype)."""
import numpy as np


def read_spike_monitor(path_i, path_t, path_count=None, path_N=None):
    """dict of i (int32), t (float64 seconds), and count / N when given"""
    out = {"i": np.fromfile(path_i, dtype="<i4"), "t": np.fromfile(path_t, dtype="<f8")}
    if path_count is not None:
        out["count"] = np.fromfile(path_count, dtype="<i4")
    if path_N is not None:
        out["N"] = int(np.fromfile(path_N, dtype="<i4")[0])
    return out


def read_state_monitor(path_t, path_values, n_indices, precision=64, path_indices=None, path_N=None):
    """dict of t (float64), values ([rows, n_indices], time-major) and indices / N when given"""
    t = np.fromfile(path_t, dtype="<f8")
    v = np.fromfile(path_values, dtype="<f8" if precision == 64 else "<f4").reshape(t.size, n_indices)
    out = {"t": t, "values": v}
    if path_indices is not None:
        out["indices"] = np.fromfile(path_indices, dtype="<i4")
    if path_N is not None:
        out["N"] = int(np.fromfile(path_N, dtype="<i4")[0])
    return out


def read_last_run_info(path):
    """(wall seconds of the last run, completed fraction)"""
    with open(path) as f:
        a, b = f.read().split()
    return float(a), float(b)
