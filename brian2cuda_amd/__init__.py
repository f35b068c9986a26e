"""hip_standalone -- an MI355X-native (gfx950) standalone device for Brian2.

The per-timestep engine (threshold/spike compaction, heterogeneous-delay spike
ring, synaptic effect application, state updates, Philox noise, Hines tree
solve) lives in the HIP library libhs.so behind the C ABI of include/hs.h.
This package holds the Python host side:

  runtime    ctypes-level objects (device arrays, standalone kernels, Net)
  networks   seeded synthetic workloads (CUBA, COBAHH, Brunel-Hakim, batches)
  device     the Brian2 Device plugin (registered when brian2 is importable)
  build      builds libhs.so for gfx950
"""
from ._lib import HsError, device_info, init, load_library  # noqa: F401

__all__ = ["HsError", "device_info", "init", "load_library"]

try:  # register set_device("hip_standalone") when Brian2 is present
    import brian2  # noqa: F401
except ImportError:  # Brian2 is optional: the runtime works without it
    brian2 = None
else:
    from . import device  # noqa: F401
