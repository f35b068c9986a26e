"""The allocation-free threshold entry point (hs_threshold_compact_noalloc)
against hs_threshold_compact, through the C ABI: same spike list, count and
last-spike steps, bit for bit, on two calls in a row (the second smaller, so
tile counts left by the first call sit in the shared scratch), in both
precisions and both orders; sizes above the maximum are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F64, F32 = 64, 32


def _case(n, prec, seed):
    dt = np.float64 if prec == F64 else np.float32
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n).astype(dt)
    last = np.full(n, -1073741824, np.int32)
    last[rng.random(n) < 0.1] = 98
    return v, last


def _run(hs, entry, v, last, sorted_, prec):
    dv = hs.DeviceArray.from_numpy(v)
    dl = hs.DeviceArray.from_numpy(last)
    spikes = hs.DeviceArray(max(len(v), 1), np.int32)
    count = hs.DeviceArray(1, np.int32)
    count.upload(np.array([-7], np.int32))
    hs.call(entry, prec, dv.ptr, len(v), 0.7, dl.ptr, 100, 5, 1 if sorted_ else 0, spikes.ptr, count.ptr)
    k = int(count.numpy()[0])
    return spikes.numpy()[:k], dl.numpy()


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("sorted_", [True, False])
def test_noalloc_equals_allocating_entry(hs, prec, sorted_):
    for n, seed in [(1_000_003, 11), (4_097, 12)]:
        v, last = _case(n, prec, seed)
        print(f"noalloc case prec={prec} sorted={sorted_} n={n}", flush=True)
        ref_spk, ref_last = _run(hs, "hs_threshold_compact", v, last, sorted_, prec)
        got_spk, got_last = _run(hs, "hs_threshold_compact_noalloc", v, last, sorted_, prec)
        if sorted_:
            assert np.array_equal(got_spk, ref_spk)
        else:
            assert np.array_equal(np.sort(got_spk), np.sort(ref_spk))
        assert np.array_equal(got_last, ref_last)
        expect = np.nonzero((v > v.dtype.type(0.7)) & ((100 - last.astype(np.int64)) >= 5))[0]
        assert np.array_equal(np.sort(got_spk), expect)


def test_noalloc_refuses_sizes_above_its_maximum(hs):
    from brian2cuda_amd._lib import HsError
    dv = hs.DeviceArray(1, np.float32)
    dl = hs.DeviceArray(1, np.int32)
    sp = hs.DeviceArray(1, np.int32)
    cnt = hs.DeviceArray(1, np.int32)
    with pytest.raises(HsError, match="exceeds the maximum"):
        hs.call("hs_threshold_compact_noalloc", F32, dv.ptr, (1 << 25) + 1, 0.0, dl.ptr, 0, 0, 1, sp.ptr, cnt.ptr)
