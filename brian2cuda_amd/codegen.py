"""Code models from Brian2 model descriptions: a NeuronGroup's equations,
threshold, reset and refractory period turned into the C statements of
hs_net_add_group_code (update, threshold, reset), which the runtime compiles
into the engine's kernels for the device.

The reference hands the equations to Brian2's state-update methods and
translates the resulting abstract code statement by statement into CUDA
(cuda_generator.py:393-461, translate_one_statement_sequence; the kernels of
stateupdate.cu, threshold.cu and reset.cu).  This module does both steps for
the methods Brian2 offers for point neurons:

  euler              x += dt*f(x)  (+ sqrt(dt)*sum_k g_k*xi_k for additive noise: Euler-Maruyama)
  heun               the stochastic Heun method (Stratonovich): x_s = x + f dt + sum_k g_k dW_k,
                     x += f dt + sum_k dW_k (g_k(x) + g_k(x_s, t+dt))/2; multiplicative noise
  milstein           Kloeden and Platen's derivative-free Milstein scheme (Ito, one noise
                     variable): x_s = x + f dt + g sqrt(dt),
                     x += f dt + g dW + (g(x_s) - g(x))(dW^2 - dt)/(2 sqrt(dt))
  rk2                midpoint: k = dt*f(x, t); x += dt*f(x + k/2, t + dt/2)
  rk4                the classical four-stage Runge-Kutta
  exponential_euler  dx/dt = A x + B with A, B free of x: x = -B/A + (x + B/A) exp(A dt)
  exact              linear systems with constant coefficients: X' = M X + c solved
                     over dt with the matrix exponential (numerically, at code
                     generation), X(t+dt) = E X + b

Equations are Brian2's text form: 'dx/dt = expr : unit [(flags)]',
'name = expr : unit' (subexpressions, substituted where used) and
'name : unit [(constant)]' (per-neuron parameters, state variables the update
leaves alone).  Names resolve to the group's variables, its subexpressions,
the namespace (SI values), Brian2's units, and t, dt, xi, i.  Functions: exp,
log, log10, sqrt, abs, clip, floor, ceil, sign, sin, cos, tan, sinh, cosh,
tanh, arcsin, arccos, arctan, expm1, log1p, exprel (evaluated by the library's
deterministic implementations on the device and in the oracle alike), and
rand() / randn() in euler updates, thresholds and resets (each call site its
own Philox draw per neuron and step, hs_rand_u(k) / hs_rand_n(k)) and in
synaptic code (per synapse pair (i, j) and step, hs_srand_u(k) / hs_srand_n(k)).
A comparison inside arithmetic is the number 1 or 0 (w*(rand() < p)).
'(unless refractory)' holds
a variable while the neuron is refractory.  Anything else raises
NotImplementedError naming it.

Constants and units are folded in at generation time; every literal is cast to
the network's precision R, so an FP32 network computes in FP32 throughout.
"""
import ast
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import sympy as sp


@dataclass
class CodeModel:
    """The arguments of hs_net_add_group_code for one group."""
    names: List[str]
    update: str
    threshold: str = ""
    reset: str = ""
    refractory: float = 0.0
    noise: bool = False
    stream: int = 0


# ---- units ------------------------------------------------------------------------
def _unit_table():
    base = {"volt": 1.0, "V": 1.0, "amp": 1.0, "ampere": 1.0, "A": 1.0, "second": 1.0, "s": 1.0,
            "siemens": 1.0, "S": 1.0, "farad": 1.0, "F": 1.0, "ohm": 1.0, "hertz": 1.0, "Hz": 1.0,
            "metre": 1.0, "meter": 1.0, "m": 1.0, "gram": 1e-3, "gramme": 1e-3, "g": 1e-3, "mole": 1.0, "mol": 1.0,
            "molar": 1e3, "M": 1e3, "coulomb": 1.0, "C": 1.0, "newton": 1.0, "N": 1.0, "joule": 1.0, "J": 1.0,
            "watt": 1.0, "W": 1.0, "liter": 1e-3, "litre": 1e-3, "l": 1e-3, "L": 1e-3}
    prefixes = {"y": 1e-24, "z": 1e-21, "a": 1e-18, "f": 1e-15, "p": 1e-12, "n": 1e-9, "u": 1e-6, "m": 1e-3,
                "c": 1e-2, "d": 1e-1, "da": 1e1, "h": 1e2, "k": 1e3, "M": 1e6, "G": 1e9, "T": 1e12}
    out = dict(base)
    for b, v in base.items():
        for p, f in prefixes.items():
            out.setdefault(p + b, f * v)
    for b in ("metre", "meter", "m"):  # area and volume forms (cm2, um**2 spelled as words)
        for p, f in list(prefixes.items()) + [("", 1.0)]:
            out.setdefault(p + b + "2", (f * base[b]) ** 2)
            out.setdefault(p + b + "3", (f * base[b]) ** 3)
    return out


UNITS = _unit_table()
# Brian2's default constants, where the namespace does not name them
CONSTANTS = {"pi": float(np.pi), "e": float(np.e), "inf": float("inf")}
RESERVED = {"t", "dt", "xi", "i", "not_refractory"}


class clip(sp.Function):  # noqa: N801 -- printed as the runtime's clip()
    nargs = 3


class log10(sp.Function):  # noqa: N801
    nargs = 1


class expm1(sp.Function):  # noqa: N801 -- Brian2's expm1, log1p, exprel, printed by name
    nargs = 1


class log1p(sp.Function):  # noqa: N801
    nargs = 1


class exprel(sp.Function):  # noqa: N801
    nargs = 1


class cmp_(sp.Function):  # noqa: N801 -- a comparison as a number (1 or 0): (a, b, operator index)
    nargs = 3


_CMP_OPS = [(ast.Lt, "<"), (ast.LtE, "<="), (ast.Gt, ">"), (ast.GtE, ">="), (ast.Eq, "=="), (ast.NotEq, "!=")]


_FUNCS = {"exp": sp.exp, "log": sp.log, "sqrt": sp.sqrt, "abs": sp.Abs, "floor": sp.floor, "ceil": sp.ceiling,
          "sign": sp.sign, "clip": clip, "log10": log10, "sin": sp.sin, "cos": sp.cos, "tan": sp.tan,
          "sinh": sp.sinh, "cosh": sp.cosh, "tanh": sp.tanh, "arcsin": sp.asin, "arccos": sp.acos,
          "arctan": sp.atan, "expm1": expm1, "log1p": log1p, "exprel": exprel}
# sympy function class -> the name the generated code calls
_FUNC_NAMES = {sp.sin: "sin", sp.cos: "cos", sp.tan: "tan", sp.sinh: "sinh", sp.cosh: "cosh", sp.tanh: "tanh",
               sp.asin: "arcsin", sp.acos: "arccos", sp.atan: "arctan", expm1: "expm1", log1p: "log1p",
               exprel: "exprel"}


# ---- parsing ------------------------------------------------------------------------
@dataclass
class _Eq:
    kind: str          # "diff" | "sub" | "param"
    name: str
    expr: Optional[str]
    flags: List[str]


def parse_equations(text):
    """Brian2 equation text -> list of _Eq (continuation lines joined, comments dropped)."""
    start = re.compile(r"^\s*(d[A-Za-z_]\w*\s*/\s*dt\s*=|[A-Za-z_]\w*\s*=|[A-Za-z_]\w*\s*:)")
    stmts, cur = [], ""
    for raw in text.replace("\r", "").split("\n"):
        line = raw.split("#")[0].rstrip()
        if not line.strip():
            continue
        if start.match(line) and cur.strip() and ":" in cur:
            stmts.append(cur)
            cur = line
        else:
            cur += " " + line
    if cur.strip():
        stmts.append(cur)
    eqs = []
    for st in stmts:
        st = st.strip()
        flags = []
        m = re.match(r"^(.*):([^:]*?)(\(([^()]*)\))?\s*$", st)
        if not m:
            raise NotImplementedError(f"equation {st!r}: no ': unit'")
        body = m.group(1).strip()
        if m.group(4):
            flags = [f.strip() for f in m.group(4).split(",") if f.strip()]
        d = re.match(r"^d([A-Za-z_]\w*)\s*/\s*dt\s*=\s*(.+)$", body)
        s = re.match(r"^([A-Za-z_]\w*)\s*=\s*(.+)$", body)
        p = re.match(r"^([A-Za-z_]\w*)$", body)
        if d:
            eqs.append(_Eq("diff", d.group(1), d.group(2), flags))
        elif s:
            eqs.append(_Eq("sub", s.group(1), s.group(2), flags))
        elif p:
            eqs.append(_Eq("param", p.group(1), None, flags))
        else:
            raise NotImplementedError(f"equation {st!r} is not 'dx/dt = ...', 'x = ...' or 'x : unit'")
        for f in flags:
            if f not in ("unless refractory", "constant", "event-driven"):
                raise NotImplementedError(f"equation {st!r}: flag {f!r} is not supported")
    names = [e.name for e in eqs]
    dup = {n for n in names if names.count(n) > 1}
    if dup:
        raise NotImplementedError(f"equations define {sorted(dup)} more than once")
    return eqs


class _Conv(ast.NodeVisitor):
    """Python/Brian2 expression AST -> sympy, names through `sym`."""

    def __init__(self, sym, what):
        self.sym, self.what = sym, what

    def visit_Expression(self, n):
        return self.visit(n.body)

    def visit_Constant(self, n):
        if isinstance(n.value, bool):
            return sp.Integer(int(n.value))
        if isinstance(n.value, int):
            return sp.Integer(n.value)
        if isinstance(n.value, float):
            return sp.Float(n.value)
        raise NotImplementedError(f"{self.what}: constant {n.value!r}")

    def visit_Name(self, n):
        return self.sym(n.id)

    def visit_UnaryOp(self, n):
        x = self.visit(n.operand)
        if isinstance(n.op, ast.USub):
            return -x
        if isinstance(n.op, ast.UAdd):
            return x
        raise NotImplementedError(f"{self.what}: operator {type(n.op).__name__}")

    def visit_BinOp(self, n):
        a, b = self.visit(n.left), self.visit(n.right)
        ops = {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b, ast.Div: lambda: a / b,
               ast.Pow: lambda: a ** b}
        f = ops.get(type(n.op))
        if f is None:
            raise NotImplementedError(f"{self.what}: operator {type(n.op).__name__}")
        return f()

    def visit_Compare(self, n):
        """a < b as a number, 1 or 0 (Brian2's boolean in arithmetic: w*(rand() < p));
        a chain a < b < c is the product of its links"""
        out, left = sp.Integer(1), self.visit(n.left)
        for o, right in zip(n.ops, n.comparators):
            k = next((k for k, (t, _) in enumerate(_CMP_OPS) if isinstance(o, t)), None)
            if k is None:
                raise NotImplementedError(f"{self.what}: comparison {type(o).__name__}")
            r = self.visit(right)
            out, left = out * cmp_(left, r, sp.Integer(k)), r
        return out

    def visit_Call(self, n):
        if isinstance(n.func, ast.Name) and n.func.id in ("rand", "randn") and not n.args and not n.keywords:
            return self.sym(n.func.id + "()")  # a draw: the caller's name resolution decides
        if not isinstance(n.func, ast.Name) or n.func.id not in _FUNCS or n.keywords:
            raise NotImplementedError(f"{self.what}: function {ast.unparse(n.func)!r} is not supported")
        return _FUNCS[n.func.id](*[self.visit(a) for a in n.args])

    def generic_visit(self, n):
        raise NotImplementedError(f"{self.what}: {type(n).__name__} is not supported here")


def _parse(expr, sym, what):
    try:
        tree = ast.parse(expr.strip(), mode="eval")
    except SyntaxError as e:
        raise NotImplementedError(f"{what}: cannot parse {expr!r} ({e.msg})")
    return _Conv(sym, what).visit(tree)


# ---- printing ---------------------------------------------------------------------
def _lit(x):
    x = float(x)
    if not np.isfinite(x):
        raise NotImplementedError(f"non-finite constant {x}")
    return f"((R){x!r})"


class _Printer:
    """sympy -> C over the code model's names (R: the network's precision)."""

    def __init__(self, local):
        self.local = local  # Symbol -> C expression

    def __call__(self, e):
        return self.p(e)

    def p(self, e):
        if e in self.local:
            return self.local[e]
        if e.is_Symbol:
            raise NotImplementedError(f"unresolved name {e.name!r}")
        if e.is_Integer:
            return f"((R){int(e)})"
        if e.is_Rational:
            return f"(((R){int(e.p)})/((R){int(e.q)}))"
        if e.is_Number:
            return _lit(e)
        if e.is_Add:
            return "(" + " + ".join(self.p(a) for a in e.args) + ")"
        if e.is_Mul:
            num, den, sign = [], [], 1
            for a in e.args:
                if a.is_Number and a == -1:
                    sign = -sign
                elif a.is_Pow and a.exp.is_Integer and a.exp < 0:
                    den.append(self.pow(a.base, -a.exp))
                elif a.is_Rational and not a.is_Integer and a.p in (1, -1):
                    sign *= int(a.p)
                    den.append(f"((R){int(a.q)})")
                else:
                    num.append(self.p(a))
            s = "*".join(num) if num else "((R)1)"
            if den:
                s = f"({s})/({'*'.join(den)})"
            return f"(-({s}))" if sign < 0 else f"({s})"
        if e.is_Pow:
            return self.pow(e.base, e.exp)
        if isinstance(e, sp.exp):
            return f"exp({self.p(e.args[0])})"
        if isinstance(e, sp.log):
            return f"log({self.p(e.args[0])})"
        if isinstance(e, log10):
            return f"log10({self.p(e.args[0])})"
        if isinstance(e, sp.Abs):
            return f"abs({self.p(e.args[0])})"
        if isinstance(e, sp.floor):
            return f"floor({self.p(e.args[0])})"
        if isinstance(e, sp.ceiling):
            return f"ceil({self.p(e.args[0])})"
        if isinstance(e, sp.sign):
            return f"sign({self.p(e.args[0])})"
        if isinstance(e, clip):
            return "clip(" + ", ".join(self.p(a) for a in e.args) + ")"
        if isinstance(e, cmp_):
            return f"((R)({self.p(e.args[0])} {_CMP_OPS[int(e.args[2])][1]} {self.p(e.args[1])}))"
        if type(e) in _FUNC_NAMES:
            return f"{_FUNC_NAMES[type(e)]}({self.p(e.args[0])})"
        raise NotImplementedError(f"cannot print {e}")

    def pow(self, b, x):
        bs = self.p(b)
        if x.is_Integer:
            n = int(x)
            if n == 0:
                return "((R)1)"
            if 0 < n <= 8:
                return "(" + "*".join([bs] * n) + ")"
            if -8 <= n < 0:
                return "(((R)1)/(" + "*".join([bs] * -n) + "))"
        if x == sp.Rational(1, 2):
            return f"sqrt({bs})"
        if x == sp.Rational(-1, 2):
            return f"(((R)1)/sqrt({bs}))"
        return f"exp({self.p(x)}*log({bs}))"


# ---- the model ----------------------------------------------------------------------
def neuron_model(equations, threshold=None, reset=None, refractory=0.0, method=None, namespace=None, dt=None,
                 stream=0, what="NeuronGroup"):
    """CodeModel of a NeuronGroup description (method None: 'exact' where the
    system is linear with constant coefficients, else 'euler').  dt (seconds)
    is needed by 'exact' only."""
    ns = dict(namespace or {})
    eqs = parse_equations(equations)
    for e in eqs:
        if "event-driven" in e.flags:
            raise NotImplementedError(f"{what}: '(event-driven)' belongs to Synapses equations")
    diff = [e for e in eqs if e.kind == "diff"]
    params = [e for e in eqs if e.kind == "param"]
    subs = {e.name: e for e in eqs if e.kind == "sub"}
    names = [e.name for e in eqs if e.kind != "sub"] or ["_placeholder"]  # model='': one unused variable
    if len(names) > 8:
        raise NotImplementedError(f"{what}: {len(names)} state variables (at most 8)")
    for n in names + list(subs):
        if n in RESERVED or n in _FUNCS:
            raise NotImplementedError(f"{what}: variable name {n!r} is reserved")
    state = {n: sp.Symbol(n, real=True) for n in names}
    T, DT, XI, I = (sp.Symbol(n, real=True) for n in ("t", "dt", "xi", "i"))
    xi_names = set()
    xi_syms = {}  # each noise variable (xi, xi_1, ...) its own symbol and draw
    sub_expr = {}
    draws = []  # rand() / randn() call sites in order: their symbols and kinds ("u", "n")

    def sym(name, stack=()):
        if name in ("rand()", "randn()"):
            if stack:
                raise NotImplementedError(f"{what}: {name} inside a subexpression (its uses would share one draw)")
            kind = "u" if name == "rand()" else "n"
            k = sum(1 for _, q in draws if q == kind)
            if k >= 16:
                raise NotImplementedError(f"{what}: more than 16 {name} calls")
            d = sp.Symbol(f"_hs_rnd_{kind}{k}", real=True)
            draws.append((d, kind))
            return d
        if name in state:
            return state[name]
        if name in subs:
            if name in stack:
                raise NotImplementedError(f"{what}: subexpression {name!r} refers to itself")
            if name not in sub_expr:
                sub_expr[name] = _parse(subs[name].expr, lambda n: sym(n, stack + (name,)),
                                        f"{what}: {name} = {subs[name].expr}")
            return sub_expr[name]
        if name == "t":
            return T
        if name == "dt":
            return DT
        if name == "i":
            return I
        if name == "xi" or name.startswith("xi_"):
            xi_names.add(name)
            if name not in xi_syms:
                xi_syms[name] = sp.Symbol(f"_hs_noise_{name}", real=True)
            return xi_syms[name]
        if name in ns:
            return sp.Float(float(ns[name]))
        if name in UNITS:
            return sp.Float(UNITS[name])
        if name in CONSTANTS and np.isfinite(CONSTANTS[name]):
            return sp.Float(CONSTANTS[name])
        raise NotImplementedError(f"{what}: no value for {name!r} (not a variable, subexpression, constant or unit)")

    rhs = {e.name: _parse(e.expr, sym, f"{what}: d{e.name}/dt = {e.expr}") for e in diff}
    if len(xi_names) > 16:
        raise NotImplementedError(f"{what}: more than 16 noise variables")
    # the first noise variable (by name) is the engine's per-neuron draw xi; the
    # others are normal draws of their own (hs_rand_xi(k), k = 1, 2, ...)
    noise = [xi_syms[n] for n in sorted(xi_names)]
    noisy = bool(noise)
    ndraw_update = len(draws)
    unless = {e.name for e in diff if "unless refractory" in e.flags}
    if not diff:
        method = "none"  # nothing to integrate
    elif method is None:
        method = "euler" if noisy else ("exact" if _linear(rhs, state, T) else "euler")
    if ndraw_update and method != "euler" and method != "none":
        raise NotImplementedError(f"{what}: rand() / randn() in the equations need method 'euler' "
                                  f"(not {method!r})")
    if noisy and method not in ("euler", "heun", "milstein"):
        raise NotImplementedError(f"{what}: method {method!r} takes no noise (use 'euler', 'heun' or 'milstein')")
    local = {s: n for n, s in state.items()}
    local.update({T: "((R)t)", DT: "dt", XI: "xi", I: "((R)i)"})
    for k, z in enumerate(noise):
        local[z] = "xi" if k == 0 else f"((R)hs_rand_xi({k}))"

    class _DrawLocal(dict):  # draw symbols print as the draw of their call site
        def __missing__(self, key):
            raise KeyError(key)

        def __contains__(self, key):
            return dict.__contains__(self, key) or (isinstance(key, sp.Symbol) and
                                                    key.name.startswith("_hs_rnd_"))

        def __getitem__(self, key):
            if dict.__contains__(self, key):
                return dict.__getitem__(self, key)
            kind, k = key.name[len("_hs_rnd_")], key.name[len("_hs_rnd_") + 1:]
            return f"((R)hs_rand_{kind}({k}))"

    local = _DrawLocal(local)
    pr = _Printer(local)
    out = []
    order = [e.name for e in diff]
    if method == "none":
        pass
    elif method in ("euler", "heun", "milstein"):
        # dx = f dt + sum_k g_k dW_k, dW_k = sqrt(dt) xi_k: f and the g_k from the
        # equations (the noise must enter linearly)
        drift, diffusion = {}, {}
        for x in order:
            f = rhs[x]
            gs = [sp.diff(f, z) for z in noise]
            if any(g.has(z) for g in gs for z in noise):
                raise NotImplementedError(f"{what}: the noise of d{x}/dt is not linear in the noise variables")
            drift[x] = f.subs({z: 0 for z in noise}, simultaneous=True)
            diffusion[x] = gs
        additive = all(not g.has(y) for x in order for g in diffusion[x] for y in state.values())
        SQ = sp.sqrt(DT)
        if method == "euler" or (method == "milstein" and additive):
            # Euler-Maruyama (milstein's correction vanishes for additive noise)
            if not additive:
                raise NotImplementedError(f"{what}: the noise is not additive (euler); use 'heun' (Stratonovich) "
                                          "or 'milstein' (Ito)")
            for x in order:
                new = state[x] + DT * drift[x] + SQ * sum(g * z for g, z in zip(diffusion[x], noise))
                out.append(f"const R _hs_{x} = {pr(new)};")
        elif method == "heun":
            # the stochastic Heun method (Stratonovich):
            #   x_s = x + f(x,t) dt + sum_k g_k(x,t) dW_k
            #   x'  = x + f(x,t) dt + sum_k dW_k (g_k(x,t) + g_k(x_s, t+dt)) / 2
            sup = {state[x]: sp.Symbol(f"_hs_s_{x}", real=True) for x in order}
            for x in order:
                local[sup[state[x]]] = f"_hs_s_{x}"
                new = state[x] + DT * drift[x] + SQ * sum(g * z for g, z in zip(diffusion[x], noise))
                out.append(f"const R _hs_s_{x} = {pr(new)};")
            for x in order:
                corr = sum(z * (g + g.subs({**sup, T: T + DT}, simultaneous=True))
                           for g, z in zip(diffusion[x], noise))
                new = state[x] + DT * drift[x] + SQ * corr / 2
                out.append(f"const R _hs_{x} = {pr(new)};")
        else:
            # milstein with multiplicative noise: Kloeden & Platen's derivative-free
            # Milstein scheme for Ito equations (one noise variable):
            #   x_s = x + f dt + g sqrt(dt)
            #   x'  = x + f dt + g dW + (g(x_s) - g(x)) (dW^2 - dt) / (2 sqrt(dt))
            if len(noise) != 1:
                raise NotImplementedError(f"{what}: 'milstein' takes one noise variable (use 'heun')")
            z = noise[0]
            sup = {state[x]: sp.Symbol(f"_hs_s_{x}", real=True) for x in order}
            for x in order:
                local[sup[state[x]]] = f"_hs_s_{x}"
                out.append(f"const R _hs_s_{x} = {pr(state[x] + DT * drift[x] + SQ * diffusion[x][0])};")
            for x in order:
                g = diffusion[x][0]
                gsup = g.subs(sup, simultaneous=True)
                new = (state[x] + DT * drift[x] + g * SQ * z + (gsup - g) * (DT * z * z - DT) / (2 * SQ))
                out.append(f"const R _hs_{x} = {pr(new)};")
    elif method in ("rk2", "rk4"):
        if noisy:
            raise NotImplementedError(f"{what}: method {method!r} takes no noise (use 'euler')")
        def stage(k, shift, tshift):
            vals = {state[x]: state[x] + shift * sp.Symbol(f"_hs_k{k - 1}_{x}", real=True) for x in order} \
                if k > 1 else {}
            for x in order:
                ex = DT * rhs[x].subs({**vals, T: T + tshift * DT}, simultaneous=True)
                out.append(f"const R _hs_k{k}_{x} = {pr(ex)};")

        for x in order:
            for k in (1, 2, 3, 4):
                local[sp.Symbol(f"_hs_k{k}_{x}", real=True)] = f"_hs_k{k}_{x}"
        if method == "rk2":
            stage(1, 0, 0)
            stage(2, sp.Rational(1, 2), sp.Rational(1, 2))
            for x in order:
                out.append(f"const R _hs_{x} = {x} + _hs_k2_{x};")
        else:
            stage(1, 0, 0)
            stage(2, sp.Rational(1, 2), sp.Rational(1, 2))
            stage(3, sp.Rational(1, 2), sp.Rational(1, 2))
            stage(4, 1, 1)
            for x in order:
                out.append(f"const R _hs_{x} = {x} + (_hs_k1_{x} + ((R)2)*_hs_k2_{x} + ((R)2)*_hs_k3_{x} + "
                           f"_hs_k4_{x})/((R)6);")
    elif method == "exponential_euler":
        if noisy:
            raise NotImplementedError(f"{what}: method 'exponential_euler' takes no noise (use 'euler')")
        for x in order:
            f = rhs[x]
            A = sp.diff(f, state[x])
            if A.has(state[x]):
                raise NotImplementedError(f"{what}: d{x}/dt is not linear in {x} (exponential_euler)")
            if A == 0:
                out.append(f"const R _hs_{x} = {pr(state[x] + DT * f)};")
                continue
            B = f.subs(state[x], 0)
            out.append(f"const R _hs_BA_{x} = {pr(B / A)};")
            out.append(f"const R _hs_{x} = -_hs_BA_{x} + (_hs_BA_{x} + {x})*exp({pr(A * DT)});")
    elif method in ("exact", "linear", "independent"):  # Brian2's older names of the linear solution
        if dt is None:
            raise NotImplementedError(f"{what}: 'exact' needs the time step")
        if noisy or not _linear(rhs, state, T):
            raise NotImplementedError(f"{what}: 'exact' needs a linear system with constant coefficients")
        E, b = _exact_update(rhs, state, float(dt), set())
        lines = _exact_lines(E, b, list(state), order)
        if unless:
            E2, b2 = _exact_update(rhs, state, float(dt), unless)
            # variables not held keep the free system's coefficients where the two
            # agree (their dynamics do not involve a held variable): the same
            # numbers whether or not the neuron is refractory
            for r in range(E.shape[0]):
                if np.allclose(E2[r], E[r], rtol=1e-12, atol=0) and np.isclose(b2[r], b[r], rtol=1e-12, atol=0):
                    E2[r], b2[r] = E[r], b[r]
            lines2 = _exact_lines(E2, b2, list(state), order)
            out.append("R " + ", ".join("_hs_" + x for x in order) + ";")
            out.append("if (not_refractory) {")
            out += ["    " + ln.replace("const R ", "") for ln in lines]
            out.append("} else {")
            out += ["    " + ln.replace("const R ", "") for ln in lines2]
            out.append("}")
            unless = set()  # handled by the two systems
        else:
            out += lines
    else:
        raise NotImplementedError(f"{what}: method {method!r} is not supported "
                                  "(euler, heun, milstein, rk2, rk4, exponential_euler, exact)")
    for x in order:
        if x in unless:
            out.append(f"{x} = not_refractory ? _hs_{x} : {x};")
        else:
            out.append(f"{x} = _hs_{x};")
    update = "\n".join(out)
    thr = _condition(threshold, sym, pr, f"{what}: threshold") if threshold else ""
    rst = _statements(reset, state, sym, pr, f"{what}: reset") if reset else ""
    return CodeModel(names, update, thr, rst, float(refractory), noisy, int(stream)), names


def _linear(rhs, state, T):
    """X' = M X + c with M and c numbers: free of the state, of t, dt, i and
    of noise (every constant is folded in by then)."""
    xs = list(state.values())
    for f in rhs.values():
        if any(not sp.diff(f, x).is_Number for x in xs):
            return False
        if not f.subs({x: 0 for x in xs}).is_Number:
            return False
    return True


def _exact_update(rhs, state, dt, frozen):
    """(E, b) with X(t+dt) = E X(t) + b for X' = M X + c, X every state
    variable (parameters have zero rows: they hold); rows of frozen variables
    zero too (they hold while refractory).  The matrix exponential of the
    augmented system [[M, c], [0, 0]] * dt gives both."""
    n = len(state)
    keys = list(state)
    M = np.zeros((n + 1, n + 1))
    for r, x in enumerate(keys):
        if x not in rhs or x in frozen:
            continue
        f = rhs[x]
        for c_, y in enumerate(keys):
            M[r, c_] = float(sp.diff(f, state[y]))
        M[r, n] = float(f.subs({s: 0 for s in state.values()}))
    from scipy.linalg import expm
    P = expm(M * dt)
    return P[:n, :n], P[:n, n]


def _exact_lines(E, b, keys, order):
    lines = []
    for x in order:
        r = keys.index(x)
        terms = [y if E[r, c] == 1.0 else f"{_lit(E[r, c])}*{y}" for c, y in enumerate(keys) if E[r, c] != 0.0]
        if b[r] != 0.0:
            terms.append(_lit(b[r]))
        lines.append(f"const R _hs_{x} = {' + '.join(terms) if terms else '((R)0)'};")
    return lines


def _condition(text, sym, pr, what):
    """A Brian2 boolean expression -> C."""
    try:
        tree = ast.parse(text.strip(), mode="eval").body
    except SyntaxError as e:
        raise NotImplementedError(f"{what}: cannot parse {text!r} ({e.msg})")

    def c(n):
        if isinstance(n, ast.BoolOp):
            op = " && " if isinstance(n.op, ast.And) else " || "
            return "(" + op.join(c(v) for v in n.values) + ")"
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.Not):
            return f"(!{c(n.operand)})"
        if isinstance(n, ast.Compare):
            ops = {ast.Gt: ">", ast.Lt: "<", ast.GtE: ">=", ast.LtE: "<=", ast.Eq: "==", ast.NotEq: "!="}
            parts, left = [], n.left
            for o, right in zip(n.ops, n.comparators):
                if type(o) not in ops:
                    raise NotImplementedError(f"{what}: comparison {type(o).__name__}")
                parts.append(f"({pr(_Conv(sym, what).visit(left))} {ops[type(o)]} "
                             f"{pr(_Conv(sym, what).visit(right))})")
                left = right
            return "(" + " && ".join(parts) + ")"
        if isinstance(n, ast.Name) and n.id == "not_refractory":
            return "not_refractory"
        if isinstance(n, ast.Constant) and isinstance(n.value, bool):
            return "true" if n.value else "false"
        if isinstance(n, (ast.Name, ast.Call, ast.BinOp, ast.UnaryOp, ast.Constant)):
            # a number as a condition (a boolean variable 'x : boolean' holds 0 or 1)
            return f"({pr(_Conv(sym, what).visit(n))} != ((R)0))"
        raise NotImplementedError(f"{what}: {ast.unparse(n)!r} is not a condition")
    return c(tree)


def _statements(text, state, sym, pr, what):
    """Brian2 reset statements ('x = e', 'x += e', ...; newline- or ';'-separated) -> C."""
    out = []
    for st in re.split(r"[;\n]", text):
        st = st.split("#")[0].strip()
        if not st:
            continue
        m = re.match(r"^([A-Za-z_]\w*)\s*(=|\+=|-=|\*=|/=)\s*(.+)$", st)
        if not m:
            raise NotImplementedError(f"{what}: {st!r} is not an assignment")
        x, op, e = m.groups()
        if x not in state:
            raise NotImplementedError(f"{what}: {x!r} is not a state variable")
        out.append(f"{x} {op} {pr(_parse(e, sym, what))};")
    return "\n".join(out)


# ---- synaptic code ---------------------------------------------------------------------
@dataclass
class SynapseCode:
    """The arguments of hs_net_pathway_code (and hs_net_pathway_code_post) for one pathway."""
    syn_names: List[str]
    on_pre: str
    post_names: List[str]          # target group variables on_pre adds to
    pre_names: List[str]           # source group variables on_pre and on_post read
    init: Optional[Dict[str, object]] = None  # synapse variable -> scalar or array (synapse creation order)
    on_post: str = ""              # statements run for each synapse of a spiking target neuron
    post_read_names: List[str] = field(default_factory=list)  # target variables on_post reads


def synapse_model(model):
    """(variable names, event-driven equations [(name, rhs)]) of a Synapses model:
    'name : unit' variables and 'dx/dt = f : unit (event-driven)' traces."""
    names, ed = [], []
    for e in parse_equations(model or ""):
        if e.kind == "param":
            names.append(e.name)
        elif e.kind == "diff" and "event-driven" in e.flags:
            names.append(e.name)
            ed.append((e.name, e.expr))
        else:
            raise NotImplementedError(f"Synapses model line {e.name!r}: only 'name : unit' variables and "
                                      "'(event-driven)' equations are supported")
    return names, ed


def synapse_variables(model):
    """The variables of a Synapses model ('name : unit' lines and event-driven traces)."""
    return synapse_model(model)[0]


def _advance_code(ed, ns, what):
    """Statements that bring the event-driven variables from lastupdate to t, each
    solved exactly: dx/dt = a x + b (a, b constants) over the elapsed time."""
    if not ed:
        return []
    out = ["const R _hs_el = (R)(t - lastupdate);"]
    EL = sp.Symbol("_hs_el", real=True)
    for name, rhs in ed:
        x = sp.Symbol(name, real=True)

        def sym(n):
            if n == name:
                return x
            if n in ns:
                return sp.Float(float(ns[n]))
            if n in UNITS:
                return sp.Float(UNITS[n])
            raise NotImplementedError(f"{what}: event-driven d{name}/dt may use only {name} and constants ({n!r})")

        f = _parse(rhs, sym, f"{what}: d{name}/dt = {rhs}")
        a, b = sp.diff(f, x), f.subs(x, 0)
        if not (a.is_Number and b.is_Number):
            raise NotImplementedError(f"{what}: event-driven d{name}/dt is not linear with constant coefficients")
        pr = _Printer({x: name, EL: "_hs_el"})
        if a == 0:
            out.append(f"{name} = {pr(x + b * EL)};")
        elif b == 0:
            out.append(f"{name} = {name}*exp({pr(a * EL)});")
        else:
            out.append(f"{name} = {pr(-b / a)} + ({name} + {pr(b / a)})*exp({pr(a * EL)});")
    return out


def synapse_code(on_pre, syn_vars, post_vars, pre_vars, namespace=None, what="Synapses", on_post=None, model=None):
    """SynapseCode of a Synapses object: on_pre statements ('x = e', 'x += e', ...
    on synapse variables; 'y += e' / 'y -= e' on the target's variables, y or
    y_post) and on_post statements (synapse variables only).  Expressions read
    synapse variables, the source's variables (y_pre), the namespace, units, t,
    dt, i, j and lastupdate; on_post may also read the target's variables
    (y_post); on_pre may not (other synapses of the step add to them).
    model: the Synapses model text; its event-driven equations advance their
    variables from lastupdate to t before either statement list, exactly.
    rand() / randn(): each call site (numbered through on_pre, then on_post) a
    Philox draw keyed by the network seed, the pathway, i, j and the step."""
    ns = dict(namespace or {})
    syn_vars, post_vars, pre_vars = list(syn_vars), list(post_vars), list(pre_vars)
    ed = []
    if model:
        mnames, ed = synapse_model(model)
        syn_vars += [n for n in mnames if n not in syn_vars]
    advance = _advance_code(ed, ns, what)
    used_post, used_pre, read_post = [], [], []
    ndraw = {"u": 0, "n": 0}  # rand() / randn() call sites so far (on_pre, then on_post)

    def translate(text, post_reads):
        syms = {}

        def S(name):
            if name not in syms:
                syms[name] = sp.Symbol(name, real=True)
            return syms[name]

        def sym(name):
            if name in ("rand()", "randn()"):
                kind = "u" if name == "rand()" else "n"
                if ndraw[kind] >= 16:
                    raise NotImplementedError(f"{what}: more than 16 {name} calls")
                ndraw[kind] += 1
                return S(f"_hs_rnd_{kind}{ndraw[kind] - 1}")
            if name in syn_vars:
                return S(name)
            if name.endswith("_pre") and name[:-4] in pre_vars:
                if name[:-4] not in used_pre:
                    used_pre.append(name[:-4])
                return S(name)
            if name in ("t", "dt", "i", "j", "lastupdate"):
                return S("_hs_" + name)
            if (name.endswith("_post") and name[:-5] in post_vars) or name in post_vars:
                if not post_reads:
                    raise NotImplementedError(f"{what}: on_pre reads the target variable {name!r} (not supported)")
                base = name[:-5] if name.endswith("_post") else name
                if base not in read_post:
                    read_post.append(base)
                return S(base + "_post")
            if name in ns:
                return sp.Float(float(ns[name]))
            if name in UNITS:
                return sp.Float(UNITS[name])
            raise NotImplementedError(f"{what}: no value for {name!r}")

        out = []
        for st in re.split(r"[;\n]", text or ""):
            st = st.split("#")[0].strip()
            if not st:
                continue
            m = re.match(r"^([A-Za-z_]\w*)\s*(=|\+=|-=|\*=|/=)\s*(.+)$", st)
            if not m:
                raise NotImplementedError(f"{what}: {st!r} is not an assignment")
            x, op, e = m.groups()
            expr = _parse(e, sym, f"{what}: {st!r}")
            if x in syn_vars:
                target = x
            elif post_reads:
                raise NotImplementedError(f"{what}: on_post assigns to {x!r}; it may assign synapse variables only")
            else:
                base = x[:-5] if x.endswith("_post") else x
                if base not in post_vars:
                    raise NotImplementedError(f"{what}: on_pre assigns to {x!r}, not a synapse or target variable")
                if op not in ("+=", "-="):
                    raise NotImplementedError(f"{what}: on_pre {st!r}: target variables take additions only")
                if base not in used_post:
                    used_post.append(base)
                target = base + "_post"
            local = {s_: n for n, s_ in syms.items()}
            local.update({syms[k]: v for k, v in (("_hs_t", "((R)t)"), ("_hs_dt", "dt"), ("_hs_i", "((R)i)"),
                                                      ("_hs_j", "((R)j)"), ("_hs_lastupdate", "((R)lastupdate)"))
                          if k in syms})
            local.update({s_: f"((R)hs_srand_{n[8]}({n[9:]}))" for n, s_ in syms.items() if n.startswith("_hs_rnd_")})
            out.append(f"{target} {op} {_Printer(local)(expr)};")
        return out

    pre_code = translate(on_pre, False)
    post_code = translate(on_post, True) if on_post else []
    return SynapseCode(syn_vars, "\n".join(advance + pre_code), used_post, used_pre, None,
                       "\n".join(advance + post_code) if on_post else "", read_post)


# ---- connect conditions and probabilities --------------------------------------------
_CONN_FUNCS = {"exp": "float", "log": "float", "log10": "float", "sqrt": "float", "abs": None, "floor": "float",
               "ceil": "float", "sign": None, "clip": "float", "int": "int", "float": "float", "sin": "float",
               "cos": "float", "tan": "float", "sinh": "float", "cosh": "float", "tanh": "float", "arcsin": "float",
               "arccos": "float", "arctan": "float", "expm1": "float", "log1p": "float", "exprel": "float"}
_CONN_PY = {"exp": np.exp, "log": np.log, "log10": np.log10, "sqrt": np.sqrt, "abs": abs, "floor": np.floor,
            "ceil": np.ceil, "sign": np.sign, "sin": np.sin, "cos": np.cos, "tan": np.tan, "sinh": np.sinh,
            "cosh": np.cosh, "tanh": np.tanh, "arcsin": np.arcsin, "arccos": np.arccos, "arctan": np.arctan,
            "expm1": np.expm1, "log1p": np.log1p, "exprel": lambda x: np.expm1(x) / x if x != 0 else 1.0}


def connect_expression(expr, namespace=None, pre_names=(), post_names=(), what="connect", first_draw=0):
    """A Brian2 connect condition or probability -> (C expression, constant).

    The expression is Brian2's (Python's) over i and j (indices within the
    source and target of the Synapses), N_pre, N_post, source / target
    variables as name_pre / name_post (pre_names / post_names: the groups'
    variables in the runtime's order), the namespace (SI values) and Brian2's
    units; operators + - * / // % **, comparisons (chained), and / or / not,
    and the functions exp, log, log10, sqrt, abs, floor, ceil, sign, clip, int,
    float.  Python's numbers are kept: integers are long long in C, / is true
    division, // and % floor and take the divisor's sign (hs_cg_floordiv,
    hs_cg_pymod), a non-literal power is hs_cg_pow.  The C text is that of
    hs_net_connect_code (variables as hs_pre(k) / hs_post(k)).  constant is the
    expression's value when it reads neither an index nor a variable (else
    None), so a constant probability can take the geometric skips.  rand() and
    randn() are draws keyed by the connect call's seed and stream and the pair
    (hs_crand_u(k) / hs_crand_n(k)), call sites numbered from first_draw (the
    draws of a condition, when the probability expression of the same call
    also draws); anything else raises NotImplementedError naming it.
    """
    ns = dict(namespace or {})
    ndraw = [int(first_draw)]
    pre_idx = {n: k for k, n in enumerate(pre_names)}
    post_idx = {n: k for k, n in enumerate(post_names)}
    try:
        tree = ast.parse(str(expr).strip(), mode="eval").body
    except SyntaxError as e:
        raise NotImplementedError(f"{what}: cannot parse {expr!r} ({e.msg})")

    def lit(v):
        if isinstance(v, (bool, np.bool_)):
            return ("true" if v else "false"), "bool", bool(v)
        if isinstance(v, (int, np.integer)):
            return f"{int(v)}LL", "int", int(v)
        v = float(v)
        if not np.isfinite(v):
            raise NotImplementedError(f"{what}: non-finite constant {v}")
        return f"({v!r})", "float", v

    def as_num(t):  # bool operands of arithmetic are 0 / 1
        c, k, v = t
        return (f"((long long)({c}))", "int", None if v is None else int(v)) if k == "bool" else t

    def num(kind_a, kind_b):
        return "float" if "float" in (kind_a, kind_b) else "int"

    def fold(v, f):
        return None if v is None else f()

    def c(n):  # subtrees that read no index and no variable become literals
        t = c0(n)
        return lit(t[2]) if t[2] is not None else t

    def c0(n):
        if isinstance(n, ast.Constant):
            if isinstance(n.value, (bool, int, float)):
                return lit(n.value)
            raise NotImplementedError(f"{what}: constant {n.value!r}")
        if isinstance(n, ast.Name):
            name = n.id
            if name in ("i", "j", "N_pre", "N_post"):
                return name, "int", None
            if name.endswith("_pre") and name[:-4] in pre_idx:
                return f"hs_pre({pre_idx[name[:-4]]})", "float", None
            if name.endswith("_post") and name[:-5] in post_idx:
                return f"hs_post({post_idx[name[:-5]]})", "float", None
            if name in ns:
                return lit(ns[name])
            if name in UNITS:
                return lit(UNITS[name])
            if name in CONSTANTS and np.isfinite(CONSTANTS[name]):
                return lit(CONSTANTS[name])
            if name in ("True", "False"):
                return lit(name == "True")
            raise NotImplementedError(f"{what}: unknown name {name!r} in {expr!r}")
        if isinstance(n, ast.UnaryOp):
            a = c(n.operand)
            if isinstance(n.op, ast.Not):
                return f"(!{a[0]})", "bool", fold(a[2], lambda: not a[2])
            a = as_num(a)
            if isinstance(n.op, ast.USub):
                return f"(-{a[0]})", a[1], fold(a[2], lambda: -a[2])
            if isinstance(n.op, ast.UAdd):
                return a
            raise NotImplementedError(f"{what}: operator {type(n.op).__name__}")
        if isinstance(n, ast.BinOp):
            a, b = as_num(c(n.left)), as_num(c(n.right))
            both = None if a[2] is None or b[2] is None else (a[2], b[2])
            k = num(a[1], b[1])
            if isinstance(n.op, (ast.Div, ast.FloorDiv, ast.Mod)) and b[2] is not None and b[2] == 0:
                raise NotImplementedError(f"{what}: division by zero in {expr!r}")
            try:
                if isinstance(n.op, (ast.Add, ast.Sub, ast.Mult)):
                    sym = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*"}[type(n.op)]
                    f = {ast.Add: lambda x, y: x + y, ast.Sub: lambda x, y: x - y,
                         ast.Mult: lambda x, y: x * y}[type(n.op)]
                    return f"({a[0]} {sym} {b[0]})", k, fold(both, lambda: f(*both))
                if isinstance(n.op, ast.Div):
                    return (f"((double)({a[0]}) / (double)({b[0]}))", "float",
                            fold(both, lambda: both[0] / both[1]))
                if isinstance(n.op, ast.FloorDiv):
                    return f"hs_cg_floordiv({a[0]}, {b[0]})", k, fold(both, lambda: both[0] // both[1])
                if isinstance(n.op, ast.Mod):
                    return f"hs_cg_pymod({a[0]}, {b[0]})", k, fold(both, lambda: both[0] % both[1])
                if isinstance(n.op, ast.Pow):
                    value = fold(both, lambda: both[0] ** both[1])
                    e = b[2]
                    if b[1] == "int" and e is not None and -8 <= e <= 8:
                        if e == 0:
                            return ("1LL" if a[1] == "int" else "(1.0)"), a[1], value
                        prod = "(" + " * ".join([a[0]] * abs(e)) + ")"
                        if e > 0:
                            return prod, a[1], value
                        return f"(1.0 / (double){prod})", "float", value
                    return f"hs_cg_pow((double)({a[0]}), (double)({b[0]}))", "float", value
            except ZeroDivisionError:
                raise NotImplementedError(f"{what}: division by zero in {expr!r}")
            raise NotImplementedError(f"{what}: operator {type(n.op).__name__}")
        if isinstance(n, ast.Compare):
            ops = {ast.Gt: ">", ast.Lt: "<", ast.GtE: ">=", ast.LtE: "<=", ast.Eq: "==", ast.NotEq: "!="}
            pyops = {ast.Gt: lambda x, y: x > y, ast.Lt: lambda x, y: x < y, ast.GtE: lambda x, y: x >= y,
                     ast.LtE: lambda x, y: x <= y, ast.Eq: lambda x, y: x == y, ast.NotEq: lambda x, y: x != y}
            parts, vals, left = [], [], as_num(c(n.left))
            for o, rn in zip(n.ops, n.comparators):
                if type(o) not in ops:
                    raise NotImplementedError(f"{what}: comparison {type(o).__name__}")
                right = as_num(c(rn))
                parts.append(f"({left[0]} {ops[type(o)]} {right[0]})")
                vals.append(None if left[2] is None or right[2] is None else pyops[type(o)](left[2], right[2]))
                left = right
            value = None if any(v is None for v in vals) else all(vals)
            return "(" + " && ".join(parts) + ")", "bool", value
        if isinstance(n, ast.BoolOp):
            items = [c(v) for v in n.values]
            op = " && " if isinstance(n.op, ast.And) else " || "
            vals = [t[2] for t in items]
            value = None
            if all(v is not None for v in vals):
                value = all(vals) if isinstance(n.op, ast.And) else any(vals)
            return "(" + op.join(f"(bool)({t[0]})" for t in items) + ")", "bool", value
        if isinstance(n, ast.Call):
            if not isinstance(n.func, ast.Name) or n.keywords:
                raise NotImplementedError(f"{what}: call {ast.unparse(n)!r} is not supported")
            fn = n.func.id
            if fn in ("rand", "randn"):
                if n.args:
                    raise NotImplementedError(f"{what}: {fn}() takes no arguments")
                if ndraw[0] >= 16:
                    raise NotImplementedError(f"{what}: more than 16 rand() / randn() calls")
                ndraw[0] += 1
                return f"((double)hs_crand_{'u' if fn == 'rand' else 'n'}({ndraw[0] - 1}))", "float", None
            if fn not in _CONN_FUNCS:
                raise NotImplementedError(f"{what}: function {fn!r} is not supported")
            args = [as_num(c(x)) for x in n.args]
            nargs = 3 if fn == "clip" else 1
            if len(args) != nargs:
                raise NotImplementedError(f"{what}: {fn}() takes {nargs} argument(s)")
            vals = [t[2] for t in args]
            const = None if any(v is None for v in vals) else vals
            if fn == "int":
                return f"((long long)({args[0][0]}))", "int", fold(const, lambda: int(const[0]))
            if fn == "float":
                return f"((double)({args[0][0]}))", "float", fold(const, lambda: float(const[0]))
            if fn == "clip":
                k = "float" if any(t[1] == "float" for t in args) else "int"
                return ("clip(" + ", ".join(t[0] for t in args) + ")", k,
                        fold(const, lambda: min(max(const[0], const[1]), const[2])))
            kind = _CONN_FUNCS[fn] or args[0][1]
            arg = args[0][0] if kind == "int" or fn in ("abs", "sign") else f"(double)({args[0][0]})"
            return f"{fn}({arg})", kind, fold(const, lambda: _CONN_PY[fn](const[0]))
        raise NotImplementedError(f"{what}: {ast.unparse(n)!r} is not supported in a connect expression")

    text, kind, value = c(tree)
    if value is not None and not isinstance(value, bool):
        value = float(value)
    return text, value
