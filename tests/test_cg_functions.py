"""CPU: the functions generated code may call beyond exp / log / sqrt -- sin,
cos, tan, sinh, cosh, tanh, arcsin, arccos, arctan, expm1, log1p, exprel
(hs_detmath.h, shared by the device and the oracle):

  * accuracy against mpmath at 120 bits over wide ranges (at most 4 ulp);
  * the special values the reference's function test relies on are exact;
  * brian2cuda's test_default_function_implementations
    (brian2cuda/tests/test_cuda_generator.py:20-90): for every default function
    the condition 'f(i + offset) == zero_point' on a one-neuron group creates
    exactly one synapse, with an integer and with a real argument, and likewise
    clip -- here through the oracle's connect restatement (the GPU runs the
    same cases in tests/test_gpu_connect_code.py).
"""
import numpy as np
import pytest

import oracle
from brian2cuda_amd import codegen as cg

mpmath = pytest.importorskip("mpmath")

FUNCS = ["sin", "cos", "tan", "sinh", "cosh", "tanh", "arcsin", "arccos", "arctan", "expm1", "log1p", "exprel"]

_SRC = '#include "hs_cgfuncs.h"\nextern "C" {\n' + "".join(
    f"double f_{f}(double x) {{ return hs_cg_{f}(x); }}\nfloat f_{f}_f(float x) {{ return hs_cg_{f}(x); }}\n"
    for f in FUNCS) + "}\n"


@pytest.fixture(scope="module")
def lib():
    import ctypes as C
    L = oracle._host_lib(_SRC)
    for f in FUNCS:
        getattr(L, f"f_{f}").restype = C.c_double
        getattr(L, f"f_{f}").argtypes = [C.c_double]
        getattr(L, f"f_{f}_f").restype = C.c_float
        getattr(L, f"f_{f}_f").argtypes = [C.c_float]
    return L


def _ref(f):
    return {"sin": mpmath.sin, "cos": mpmath.cos, "tan": mpmath.tan, "sinh": mpmath.sinh, "cosh": mpmath.cosh,
            "tanh": mpmath.tanh, "arcsin": mpmath.asin, "arccos": mpmath.acos, "arctan": mpmath.atan,
            "expm1": mpmath.expm1, "log1p": mpmath.log1p,
            "exprel": lambda x: mpmath.expm1(x) / x if x != 0 else mpmath.mpf(1)}[f]


def _args(f, rng):
    if f in ("sin", "cos", "tan"):
        return np.concatenate([rng.uniform(-10, 10, 600), rng.uniform(-1e5, 1e5, 200), [1e-300, -1e-20, 1e-8]])
    if f in ("arcsin", "arccos"):
        return np.concatenate([rng.uniform(-1, 1, 600), [1e-300, 0.999999, -0.9999999999]])
    if f == "arctan":
        return np.concatenate([rng.uniform(-50, 50, 600), rng.uniform(-0.3, 0.3, 300), [1e-300, 1e300, 1.0]])
    if f == "log1p":
        return np.concatenate([rng.uniform(-0.99, 10, 600), rng.uniform(-1e-6, 1e-6, 200)])
    return np.concatenate([rng.uniform(-20, 20, 600), rng.uniform(-0.6, 0.6, 300), [1e-300, 1e-12, -0.4999, 0.5]])


@pytest.mark.parametrize("f", FUNCS)
def test_accuracy_against_mpmath(lib, f):
    mpmath.mp.prec = 120
    rng = np.random.default_rng(FUNCS.index(f))
    fn = getattr(lib, f"f_{f}")
    worst = 0.0
    for x in _args(f, rng):
        got = fn(float(x))
        want = _ref(f)(mpmath.mpf(float(x)))
        if want == 0:
            assert got == 0.0
            continue
        worst = max(worst, float(abs(mpmath.mpf(got) - want) / abs(want)) / 2.0 ** -52)
    assert worst <= 4.0, f"{f}: {worst:.2f} ulp"
    # the FP32 form is the double result rounded once
    for x in _args(f, rng)[:50]:
        x32 = float(np.float32(x))
        assert getattr(lib, f"f_{f}_f")(x32) == np.float32(fn(x32)) or np.isnan(fn(x32))


def test_special_values_and_domains(lib):
    exact = {"sin": (0.0, 0.0), "cos": (0.0, 1.0), "tan": (0.0, 0.0), "sinh": (0.0, 0.0), "cosh": (0.0, 1.0),
             "tanh": (0.0, 0.0), "arcsin": (0.0, 0.0), "arccos": (1.0, 0.0), "arctan": (0.0, 0.0),
             "expm1": (0.0, 0.0), "log1p": (0.0, 0.0), "exprel": (0.0, 1.0)}
    for f, (x, y) in exact.items():
        assert getattr(lib, f"f_{f}")(x) == y, f
    assert np.isnan(lib.f_arcsin(1.5)) and np.isnan(lib.f_arccos(-1.01)) and np.isnan(lib.f_sin(np.inf))
    assert lib.f_arccos(-1.0) == np.pi and lib.f_arcsin(1.0) == np.pi / 2 and lib.f_arctan(np.inf) == np.pi / 2
    assert lib.f_tanh(400.0) == 1.0 and lib.f_tanh(-400.0) == -1.0


# (function, zero point, offset) of the reference's test_default_function_implementations
DEFAULT_FUNCS = [("sin", 0, 0), ("cos", 1, 0), ("tan", 0, 0), ("sinh", 0, 0), ("cosh", 1, 0), ("tanh", 0, 0),
                 ("exp", 1, 0), ("sqrt", 0, 0), ("ceil", 0, 0), ("floor", 0, 0), ("arcsin", 0, 0),
                 ("arctan", 0, 0), ("int", 0, 0), ("arccos", 0, 1), ("log", 0, 1), ("log10", 0, 1),
                 ("abs", 1, -1), ("sign", 1, 1)]


def default_function_conditions():
    """The reference test's conditions: (integer form, real form) per function, and clip's two."""
    out = [(f"{f}(i + {o}) == {z}", f"{f}(i * 1.0 + {o}) == {z}") for f, z, o in DEFAULT_FUNCS]
    return out + [("clip(i+1,-1,0)==0", "clip(i+1,-1.0,0.0)==0")]


@pytest.mark.parametrize("conds", default_function_conditions(), ids=lambda c: c[0].split("(")[0])
def test_default_function_implementations_oracle(conds):
    for cond in conds:
        c, const = cg.connect_expression(cond)
        assert const is None
        out = oracle.connect_code((0, 1), (0, 1), c, "", 1.0, 0, False, 1, 0, (0.0, 0.0), 1e-4, (1.0, 1.0))
        assert out["pre"].size == 1, cond
