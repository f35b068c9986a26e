"""CPU checks of the drop-in boundary: libhs.so loads and exports every entry
point include/*.h declares; the Python binding declares each of them; calls
that need a device fail with a clear error instead of a crash.  (No GPU.)"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "brian2cuda_amd", "libhs.so")


def declared_entry_points():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(hs_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_the_engine():
    names = declared_entry_points()
    for n in ["hs_init", "hs_last_error", "hs_threshold_compact", "hs_randn", "hs_stateupdate", "hs_deliver",
              "hs_net_create", "hs_net_add_pathway", "hs_net_run", "hs_net_get_spikes", "hs_hines_create",
              "hs_hines_run"]:
        assert n in names, n


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhs.so not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared_entry_points()) if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhs.so not built")
def test_python_binding_covers_the_header():
    from brian2cuda_amd import _lib
    assert declared_entry_points() <= set(_lib.SIGNATURES), declared_entry_points() - set(_lib.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhs.so not built")
def test_calls_before_init_fail_cleanly():
    """No exit(), no crash: a negative code and a message (cuda_utils.h:77-104 exits instead)."""
    lib = ctypes.CDLL(LIB)
    lib.hs_last_error.restype = ctypes.c_char_p
    p = ctypes.c_void_p()
    assert lib.hs_alloc(ctypes.c_size_t(16), ctypes.byref(p)) == -4  # HS_ERR_STATE
    assert b"hs_init" in lib.hs_last_error()
    assert lib.hs_version() >= 1
