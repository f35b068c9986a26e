"""hip_standalone -- an MI355X-native (gfx950) standalone device for Brian2.

The per-timestep engine (threshold/spike compaction, heterogeneous-delay spike
ring, synaptic effect application, state updates, Philox noise, Hines tree
solve) lives in the HIP library libhs.so behind the C ABI of include/hs.h.
This package holds the Python host side:

  runtime    ctypes-level objects (device arrays, standalone kernels, Net)
  networks   seeded synthetic workloads (CUBA, COBAHH, Brunel-Hakim, batches)
  build      builds libhs.so for gfx950

Importing the package never needs Brian2 and never touches the GPU.
"""
from ._lib import HsError, device_info, init, load_library  # noqa: F401

__all__ = ["HsError", "device_info", "init", "load_library"]
