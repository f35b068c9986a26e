"""Builds libhs.so (the HIP runtime of the hip_standalone device) in-tree for gfx950.

Usage: python -m brian2cuda_amd.build   (or brian2cuda_amd.build.build())

Flags that carry the numerical contract:
  -ffp-contract=off       no fused multiply-add: the expression text fixes every rounding
  -Xclang -target-feature -Xclang -packed-fp32-ops
                          no packed FP32 instructions (v_pk_*_f32) anywhere
  -munsafe-fp-atomics     hardware global FP atomics (effects are additions)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhs.so")
SOURCES = ["hs_runtime.hip", "hs_engine.hip", "hs_ops.hip", "hs_hines.hip", "hs_connect.hip", "hs_plastic.hip", "hs_sort.hip"]
ARCH = os.environ.get("HS_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def hipcc_path():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found: the HIP runtime cannot be built")
    return p


def build(verbose=False):
    """Compile every source to an object, then link libhs.so. Always recompiles."""
    hipcc = hipcc_path()
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(build_dir, src.replace(".hip", ".o"))
        cmd = [hipcc, *HIPCC_FLAGS, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
