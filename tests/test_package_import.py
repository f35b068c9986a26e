"""The package imports wherever it is installed: with Brian2 absent, and with a
Brian2 present (here a stub module put into sys.modules).  Each case runs in a
fresh interpreter so that neither the stub nor the import leaks into the other
tests.  (No GPU.)"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(code):
    return subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)


def test_import_without_brian2():
    r = _run("import sys; sys.modules['brian2'] = None\n"  # any `import brian2` raises ImportError
             "import brian2cuda_amd, brian2cuda_amd.runtime, brian2cuda_amd.networks\n"
             "print('ok')")
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr


def test_import_with_a_stub_brian2():
    r = _run("import sys, types\n"
             "sys.modules['brian2'] = types.ModuleType('brian2')\n"
             "import brian2cuda_amd, brian2cuda_amd.runtime, brian2cuda_amd.networks\n"
             "print('ok')")
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr


def test_import_does_not_load_the_gpu_library():
    """Importing is free of side effects: libhs.so is loaded on first use only."""
    r = _run("import brian2cuda_amd, brian2cuda_amd._lib as L\n"
             "print('loaded' if L._lib is not None else 'not loaded')")
    assert r.returncode == 0 and r.stdout.strip() == "not loaded", r.stderr
