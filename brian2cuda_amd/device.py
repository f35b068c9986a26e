"""The ``hip_standalone`` Brian2 device: Brian2 objects -> a network spec -> a
standalone project on libhs.

Two layers:

* :func:`network_spec` (no Brian2 needed) maps a plain description of the
  network's objects -- equation text, threshold / reset / refractory, the
  namespace of constants, initial values, connections, on_pre statements,
  delays and monitors -- onto :class:`brian2cuda_amd.networks.NetSpec`.  A
  NeuronGroup is accepted when its equations are one of the engine's built-in
  models (include/hs_models.h), written as the reference's examples write them
  (examples/cuba.py:125-129, examples/brunelhakim.py:159-161,
  examples/cobahh.py:160-182), with any parameter names; a Synapses object
  when its on_pre is ``var += w`` (w a per-synapse variable or a constant).
  Anything else raises NotImplementedError naming what is not supported, as the
  reference's device does at generation time (device.py:296-331).

* :class:`HipStandaloneDevice` (defined only where Brian2 is importable)
  subclasses Brian2's CPPStandaloneDevice, as brian2cuda's device does
  (brian2cuda/device.py:71), and registers itself as
  ``all_devices['hip_standalone']`` (device.py:2002-2004) when this module is
  imported: ``import brian2cuda_amd.device; set_device('hip_standalone')``.
  Its network_run (device.py:1508-1765) describes the network's objects,
  writes the project with :mod:`brian2cuda_amd.project`, builds it
  (device.py:1231-1506) and runs the binary; the monitors' arrays are read
  back from the result files.  Brian2 is absent where this engine was built
  and tested: this glue is written against Brian2's documented object
  attributes and is not exercised against a real Brian2.
"""
import ast
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import networks as nw
from .runtime import (HS_MODEL_HH_COBA, HS_MODEL_LIF_CUBA, HS_MODEL_LIF_NOISE, HS_MODEL_POISSON, HS_MODEL_SPIKEGEN,
                      HS_W_SCALAR, HS_W_SYNAPSE, spikegen_steps)


# ---- descriptions ---------------------------------------------------------------

@dataclass
class GroupDesc:
    """A NeuronGroup / PoissonGroup / SpikeGeneratorGroup as the device sees it."""
    name: str
    N: int
    kind: str = "neurongroup"            # "neurongroup" | "poisson" | "spikegenerator"
    equations: str = ""
    threshold: Optional[str] = None
    reset: Optional[str] = None
    refractory: float = 0.0              # seconds
    method: Optional[str] = None
    namespace: Dict[str, float] = field(default_factory=dict)  # SI values of external constants
    init: Dict[str, object] = field(default_factory=dict)      # variable -> scalar or array
    rates: object = None                 # PoissonGroup rates (Hz, scalar or array)
    indices: object = None               # SpikeGeneratorGroup
    times: object = None                 # seconds
    period: float = 0.0


@dataclass
class SynapsesDesc:
    name: str
    source: str
    target: str
    on_pre: str
    model: str = ""                       # synapse variables: 'name : unit' lines and event-driven traces
    on_post: str = ""                     # statements per synapse of a spiking target neuron (code form)
    source_range: Optional[tuple] = None  # a Subgroup source G[a:b]
    target_range: Optional[tuple] = None
    namespace: Dict[str, float] = field(default_factory=dict)
    i: object = None                      # connect(i=..., j=...) (within the (sub)groups)
    j: object = None
    p: object = None                      # connect(p=...): a number or an expression (str) of i, j,
                                          # x_pre, x_post; optionally with a condition
    k: Optional[int] = None               # connect(j='k for k in sample(N_post, size=K)'), optionally
                                          # with condition 'i != j' (its 'if k != i') and i/j ranges
    condition: Optional[str] = None
    n: int = 1                            # connect(n=...): synapses per selected pair (p= or a sample size)
    values: Dict[str, object] = field(default_factory=dict)  # per-synapse variables (w ...)
    weight_range: Optional[tuple] = None  # w = 'lo + (hi-lo)*rand()' with connect(p)
    delay: object = 0.0                   # seconds: scalar, array, or (lo, hi) uniform with connect(p)
    seed: int = 0


@dataclass
class MonitorDesc:
    name: str
    kind: str                             # "spike" | "state" | "rate"
    source: str
    source_range: Optional[tuple] = None
    variable: Optional[str] = None
    record: object = None                 # state monitor: recorded indices (within the source)
    dt: Optional[float] = None            # state monitor: its own clock's dt (a multiple of the network's)


# ---- built-in models ------------------------------------------------------------

def _norm(text):
    """Equation text without whitespace, one equation per line, comments dropped."""
    lines = []
    cur = ""
    for raw in text.replace("\r", "").split("\n"):
        raw = raw.split("#")[0]
        s = re.sub(r"\s+", "", raw)
        if not s:
            continue
        if re.match(r"^(d[A-Za-z_]\w*/dt=|[A-Za-z_]\w*=)", s) and cur and ":" in cur:
            lines.append(cur)
            cur = s
        else:
            cur += s
    if cur:
        lines.append(cur)
    return lines


def _pattern(template):
    """A regex of a normalised equation template; {name} matches one identifier."""
    out, names = "", []
    for part in re.split(r"(\{\w+\})", template):
        if part.startswith("{") and part.endswith("}"):
            names.append(part[1:-1])
            out += r"([A-Za-z_]\w*)"
        else:
            out += re.escape(part)
    return re.compile("^" + out + "$"), names


_CUBA = ["dv/dt=(ge+gi-(v-{El}))/{taum}:volt(unlessrefractory)", "dge/dt=-ge/{taue}:volt",
         "dgi/dt=-gi/{taui}:volt"]
_NOISE = ["dV/dt=(-V+{mu}+{sigma}*sqrt({tau})*xi)/{tau2}:volt"]
_HH = ["dv/dt=({gl}*({El}-v)+ge*({Ee}-v)+gi*({Ei}-v)-{gNa}*(m*m*m)*h*(v-{ENa})-{gK}*(n*n*n*n)*(v-{EK}))/{Cm}:volt",
       "dm/dt=alpha_m*(1-m)-beta_m*m:1", "dn/dt=alpha_n*(1-n)-beta_n*n:1", "dh/dt=alpha_h*(1-h)-beta_h*h:1",
       "dge/dt=-ge*(1./{taue}):siemens", "dgi/dt=-gi*(1./{taui}):siemens",
       "alpha_m=0.32*(mV**-1)*(13.*mV-v+{VT})/(exp((13.*mV-v+{VT})/(4.*mV))-1.)/ms:Hz",
       "beta_m=0.28*(mV**-1.)*(v-{VT}-40.*mV)/(exp((v-{VT}-40.*mV)/(5.*mV))-1.)/ms:Hz",
       "alpha_h=0.128*exp((17.*mV-v+{VT})/(18.*mV))/ms:Hz",
       "beta_h=4./(1.+exp((40.*mV-v+{VT})/(5.*mV)))/ms:Hz",
       "alpha_n=0.032*(mV**-1.)*(15.*mV-v+{VT})/(exp((15.*mV-v+{VT})/(5.*mV))-1.)/ms:Hz",
       "beta_n=.5*exp((10.*mV-v+{VT})/(40.*mV))/ms:Hz"]


def _match(lines, templates):
    """Parameter names bound by matching the equations to the templates
    (order-free), or None."""
    if len(lines) != len(templates):
        return None
    bound, left = {}, list(lines)
    for t in templates:
        rx, names = _pattern(t)
        for ln in left:
            m = rx.match(ln)
            if not m:
                continue
            trial = dict(bound)
            if all(trial.setdefault(n, v) == v for n, v in zip(names, m.groups())):
                bound = trial
                left.remove(ln)
                break
        else:
            return None
    return bound


def _value(name, ns, what):
    if name not in ns:
        raise NotImplementedError(f"{what}: no value for {name!r} in the namespace")
    return float(ns[name])


def _threshold_value(cond, var, ns, what):
    """'var > X' with X a constant name or a number with a unit -> X (SI)."""
    s = re.sub(r"\s+", "", cond or "")
    m = re.match(rf"^{var}>(.+)$", s)
    if not m:
        raise NotImplementedError(f"{what}: threshold {cond!r} is not '{var} > constant'")
    return _const_expr(m.group(1), ns, f"{what}: threshold {cond!r}")


def _reset_value(code, var, ns, what):
    s = re.sub(r"\s+", "", code or "")
    m = re.match(rf"^{var}=(.+)$", s)
    if not m:
        raise NotImplementedError(f"{what}: reset {code!r} is not '{var} = constant'")
    return _const_expr(m.group(1), ns, f"{what}: reset {code!r}")


_UNITS = {"mV": 1e-3, "volt": 1.0, "V": 1.0, "ms": 1e-3, "second": 1.0, "nS": 1e-9, "uS": 1e-6, "siemens": 1.0}


def _const_expr(s, ns, what):
    """A constant name, or number[*unit]."""
    if re.match(r"^[A-Za-z_]\w*$", s) and s not in _UNITS:
        return _value(s, ns, what)
    m = re.match(r"^([-+]?[0-9.]+(?:[eE][-+]?\d+)?)(?:\*([A-Za-z]+))?$", s)
    if not m or (m.group(2) and m.group(2) not in _UNITS):
        raise NotImplementedError(f"{what}: cannot evaluate {s!r}")
    return float(m.group(1)) * (_UNITS[m.group(2)] if m.group(2) else 1.0)


def group_spec(g: GroupDesc, dt=None):
    """(GroupSpec, variable names) of one group description: a built-in model
    where the equations, threshold, reset and method are one of them, else a
    code model through the code generator (dt: the network's, for 'exact')."""
    what = f"NeuronGroup {g.name!r}"
    if g.kind in ("poisson", "spikegenerator"):
        return _builtin_spec(g, what)
    try:
        built = _builtin_spec(g, what)
    except NotImplementedError:
        built = None  # a built-in model's equations with another threshold, reset or method
    if built is not None:
        return built
    from .codegen import neuron_model
    model, names = neuron_model(g.equations, g.threshold, g.reset, float(g.refractory), g.method, g.namespace,
                                dt=dt, what=what)
    return nw.GroupSpec(0, g.N, [], _init(g, names), code=model), names


def _builtin_spec(g: GroupDesc, what):
    """(GroupSpec, names) of a built-in model, None if the equations are none of them."""
    if g.kind == "poisson":
        rates = np.asarray(g.rates, dtype=np.float64)
        first = float(rates.ravel()[0]) if rates.size else 0.0
        init = {0: rates if rates.size > 1 else first}
        return nw.GroupSpec(HS_MODEL_POISSON, g.N, [first], init), ["rates"]
    if g.kind == "spikegenerator":
        return (nw.GroupSpec(HS_MODEL_SPIKEGEN, g.N, [float(g.period)],
                             spikes=(np.asarray(g.indices, np.int32), None)), ["_placeholder"])
    lines = _norm(g.equations)
    ns = g.namespace
    b = _match(lines, _CUBA)
    free = 0.0
    if b is None:  # the same model without "(unless refractory)"
        b = _match(lines, [_CUBA[0].replace("(unlessrefractory)", "")] + _CUBA[1:])
        free = 1.0
    if b is not None:
        if g.method not in (None, "exact", "linear"):
            raise NotImplementedError(f"{what}: the CUBA model is integrated 'exact', not {g.method!r}")
        params = [_value(b["El"], ns, what), _threshold_value(g.threshold, "v", ns, what),
                  _reset_value(g.reset, "v", ns, what), _value(b["taum"], ns, what), _value(b["taue"], ns, what),
                  _value(b["taui"], ns, what), float(g.refractory)]
        if free:
            params.append(1.0)
        return nw.GroupSpec(HS_MODEL_LIF_CUBA, g.N, params, _init(g, ["v", "ge", "gi"])), ["v", "ge", "gi"]
    b = _match(lines, _NOISE)
    if b is not None and b["tau"] == b["tau2"]:
        if g.method not in (None, "euler"):
            raise NotImplementedError(f"{what}: the noise LIF model is integrated 'euler', not {g.method!r}")
        params = [_value(b["tau"], ns, what), _value(b["mu"], ns, what), _value(b["sigma"], ns, what),
                  _threshold_value(g.threshold, "V", ns, what), _reset_value(g.reset, "V", ns, what),
                  float(g.refractory), 0.0]
        return nw.GroupSpec(HS_MODEL_LIF_NOISE, g.N, params, _init(g, ["V"])), ["V"]
    b = _match(lines, _HH)
    if b is not None:
        if g.method not in (None, "exponential_euler"):
            raise NotImplementedError(f"{what}: the HH model is integrated 'exponential_euler', not {g.method!r}")
        if g.reset:
            raise NotImplementedError(f"{what}: the HH model has no reset")
        v = {k: _value(b[k], ns, what) for k in ("Cm", "gl", "El", "EK", "ENa", "gNa", "gK", "VT", "Ee", "Ei",
                                                   "taue", "taui")}
        params = [v["Cm"], v["gl"], v["El"], v["EK"], v["ENa"], v["gNa"], v["gK"], v["VT"], v["Ee"], v["Ei"],
                  v["taue"], v["taui"], _threshold_value(g.threshold, "v", ns, what), float(g.refractory)]
        names = ["v", "m", "n", "h", "ge", "gi"]
        return nw.GroupSpec(HS_MODEL_HH_COBA, g.N, params, _init(g, names)), names
    return None


def _init(g, names):
    out = {}
    for var, val in g.init.items():
        if var not in names:
            raise NotImplementedError(f"NeuronGroup {g.name!r}: no state variable {var!r}")
        a = np.asarray(val, dtype=np.float64)
        if a.size not in (1, g.N):
            raise ValueError(f"NeuronGroup {g.name!r}.{var}: {a.size} values for {g.N} neurons")
        out[names.index(var)] = float(a.ravel()[0]) if a.size == 1 else a
    return out


def _on_pre(code, tgt_vars, syn):
    s = re.sub(r"\s+", "", code)
    m = re.match(r"^([A-Za-z_]\w*)(?:_post)?\+=(.+)$", s)
    if not m or m.group(1) not in tgt_vars:
        raise NotImplementedError(f"Synapses {syn.name!r}: on_pre {code!r} is not 'var += w' on a target variable")
    var = tgt_vars.index(m.group(1))
    rhs = m.group(2)
    neg = rhs.startswith("-")
    body = rhs[1:] if neg else rhs
    if body in syn.values:
        return var, "synapse", body, neg
    return var, "scalar", _const_expr(body, syn.namespace, f"Synapses {syn.name!r}") * (-1.0 if neg else 1.0), False


def _outer_parens(s):
    """s is one parenthesised expression: its first '(' closes at its end."""
    depth = 0
    for k, ch in enumerate(s):
        depth += ch == "("
        depth -= ch == ")"
        if depth == 0:
            return k == len(s) - 1
    return False


def _condition_ranges(cond, n_src, n_tgt, syn):
    """connect(condition, p=...) conditions the device generator takes:
    'i<X', 'i>=X', 'i!=j' (and their conjunctions with 'and')."""
    pre, post, autapses = [0, n_src], [0, n_tgt], True
    if not cond:
        return pre, post, autapses
    for part in re.split(r"\s+and\s+", cond.strip()):
        s = re.sub(r"\s+", "", part)
        while s.startswith("(") and s.endswith(")") and _outer_parens(s):
            s = s[1:-1]
        m = re.match(r"^(i|j)(<|>=|<=|>)(.+)$", s)
        if s in ("i!=j", "j!=i"):
            autapses = False
        elif m:
            x = _const_expr(m.group(3), syn.namespace, f"Synapses {syn.name!r}")
            x = int(round(x))
            rng = pre if m.group(1) == "i" else post
            if m.group(2) == "<":
                rng[1] = min(rng[1], x)
            elif m.group(2) == "<=":
                rng[1] = min(rng[1], x + 1)
            elif m.group(2) == ">=":
                rng[0] = max(rng[0], x)
            else:
                rng[0] = max(rng[0], x + 1)
        else:
            raise NotImplementedError(f"Synapses {syn.name!r}: connect condition {part!r} is not supported")
    return pre, post, autapses


def _generator(jtext, syn, n_src, n_tgt):
    """connect(j='...') in Brian2's generator syntax -> (condition, p, K): the
    condition is a Brian2 expression over i and j that holds exactly for the
    generated pairs (None: every pair), p the sampling probability (a number
    or an expression of i; None: 1) and K (None, or (size, (lo, hi)): a fixed
    sample size over the target range [lo, hi), whose condition then only
    filters the sample).
      'k for k in range(stop | start, stop[, step]) [if cond]'
            -> start <= j < stop (or stop < j <= start for a negative step),
               (j - start) % step == 0, and cond with k read as j
      'k for k in sample(N | start, stop[, step], p=P | size=K) [if cond]'
            -> the same range condition, and p = P or K
      '<expression of i>' (no generator, e.g. 'i' one-to-one) -> j == expression
    Targets outside the target group are skipped, as with skip_if_invalid=True.
    The range and sample bounds and steps may be expressions of i; the
    element must be the loop variable."""
    what = f"Synapses {syn.name!r}: connect(j={jtext!r})"
    try:
        tree = ast.parse("(" + jtext.strip() + ")", mode="eval").body  # a bare generator needs parentheses
    except SyntaxError as e:
        raise NotImplementedError(f"{what}: cannot parse ({e.msg})")
    if not isinstance(tree, ast.GeneratorExp):
        return f"j == ({ast.unparse(tree)})", None, None
    if len(tree.generators) != 1:
        raise NotImplementedError(f"{what}: one 'for' clause only")
    gen = tree.generators[0]
    if not isinstance(gen.target, ast.Name) or not isinstance(tree.elt, ast.Name) or tree.elt.id != gen.target.id:
        raise NotImplementedError(f"{what}: the element must be the loop variable ('k for k in ...')")
    k = gen.target.id

    class _KtoJ(ast.NodeTransformer):  # the loop variable is the target index j (+ off), N_post fixed if given
        def __init__(self, off=0, n_post=None):
            self.off, self.n_post = off, n_post

        def visit_Name(self, n):
            if n.id == k:
                e = ast.Name("j", ast.Load())
                if self.off:
                    e = ast.BinOp(e, ast.Add(), ast.Constant(self.off))
                return ast.copy_location(e, n)
            if n.id == "N_post" and self.n_post is not None:
                return ast.copy_location(ast.Constant(self.n_post), n)
            return n

    call = gen.iter
    if not (isinstance(call, ast.Call) and isinstance(call.func, ast.Name) and call.func.id in ("range", "sample")):
        raise NotImplementedError(f"{what}: the iterator must be range(...) or sample(...)")
    args = [ast.unparse(a) for a in call.args]
    kw = {x.arg: x.value for x in call.keywords}
    p, K = None, None
    if call.func.id == "sample":
        if set(kw) - {"p", "size"} or len(kw) != 1:
            raise NotImplementedError(f"{what}: sample() takes p= or size=")
        if "p" in kw:
            p = ast.unparse(kw["p"])
            try:
                p = float(ast.literal_eval(kw["p"]))
            except (ValueError, TypeError, SyntaxError):
                pass
        else:
            try:
                K = int(ast.literal_eval(kw["size"]))
            except (ValueError, TypeError, SyntaxError):
                raise NotImplementedError(f"{what}: sample(size=...) takes a constant size")
    if not 1 <= len(args) <= 3:
        raise NotImplementedError(f"{what}: {call.func.id}() takes 1 to 3 positional arguments")
    start, stop = ("0", args[0]) if len(args) == 1 else (args[0], args[1])
    step = args[2] if len(args) == 3 else "1"
    try:
        step_v = int(ast.literal_eval(step))
    except (ValueError, TypeError, SyntaxError):
        raise NotImplementedError(f"{what}: a constant step")
    if step_v == 0:
        raise ValueError(f"{what}: step 0")
    if K is not None:
        # K distinct targets of the range itself: the range narrows the target
        # range (constant bounds, step 1), whose j then counts from lo; the if
        # clauses filter the sample, in the Synapses' own j and N_post
        if step_v != 1:
            raise NotImplementedError(f"{what}: sample(size=...) over a range with a step")
        ns = dict(syn.namespace, N_pre=n_src, N_post=n_tgt)
        lo, hi = (_const_expr(x, ns, what) for x in (start, stop))
        if lo != int(lo) or hi != int(hi):
            raise NotImplementedError(f"{what}: sample(size=...) takes integer bounds")
        lo, hi = int(lo), int(hi)
        ifs = ["(" + ast.unparse(_KtoJ(lo, n_tgt).visit(cond)) + ")" for cond in gen.ifs]
        return (" and ".join(ifs) or None), None, (int(K), (lo, hi))
    ifs = ["(" + ast.unparse(_KtoJ().visit(cond)) + ")" for cond in gen.ifs]
    if call.func.id == "sample" and len(args) == 1 and stop == "N_post":
        parts = []  # the whole target range
    elif step_v > 0:
        parts = [f"j >= ({start})", f"j < ({stop})"]
    else:
        parts = [f"j <= ({start})", f"j > ({stop})"]
    if abs(step_v) != 1:
        parts.append(f"(j - ({start})) % {abs(step_v)} == 0")
    return (" and ".join(parts + ifs) or None), p, None


def network_spec(groups: List[GroupDesc], synapses: List[SynapsesDesc], monitors: List[MonitorDesc], dt, seed=0,
                 name="network"):
    """(NetSpec, object names for the project writer) of a described network."""
    gspecs, gvars, gindex = [], [], {}
    for g in groups:
        spec, names = group_spec(g, dt)
        if g.kind == "spikegenerator":
            spec.spikes = (np.asarray(g.indices, np.int32), spikegen_steps(np.asarray(g.times, np.float64), dt))
        gindex[g.name] = len(gspecs)
        gspecs.append(spec)
        gvars.append(names)
    paths = []
    for s in synapses:
        if s.source not in gindex or s.target not in gindex:
            raise NotImplementedError(f"Synapses {s.name!r}: source or target is not in the network")
        src, tgt = gindex[s.source], gindex[s.target]
        n_src, n_tgt = gspecs[src].n, gspecs[tgt].n
        sr = s.source_range or (0, n_src)
        tr = s.target_range or (0, n_tgt)
        code = None
        try:
            if s.model.strip() or s.on_post.strip():
                raise NotImplementedError("declared synapse variables or on_post: the code form")
            var, wkind, wval, neg = _on_pre(s.on_pre, gvars[tgt], s)
        except NotImplementedError:
            # any other on_pre (or an on_post): synaptic code over the declared
            # synapse variables and event-driven traces (and those given values),
            # additions to the target's variables
            from .codegen import synapse_code, synapse_variables
            syn_vars = synapse_variables(s.model)
            syn_vars += [k for k in s.values if k not in syn_vars]
            code = synapse_code(s.on_pre, syn_vars, gvars[tgt], gvars[src], s.namespace,
                                what=f"Synapses {s.name!r}", on_post=s.on_post or None, model=s.model)
            code.init = {k: v for k, v in s.values.items() if k in syn_vars}
            var, neg = 0, False
            wkind, wval = ("synapse", "w") if "w" in syn_vars and ("w" in s.values or s.weight_range) \
                else ("scalar", 0.0)
        condition, p_val, k_val = s.condition, s.p, s.k
        if isinstance(s.j, str):  # generator syntax: a condition, p or K over the target range
            if s.i is not None or s.p is not None or s.k is not None:
                raise ValueError(f"Synapses {s.name!r}: connect(j='...') takes no i, p or K besides")
            gcond, p_val, k_val = _generator(s.j, s, sr[1] - sr[0], tr[1] - tr[0])
            if k_val is not None:
                k_val, (lo, hi) = k_val
                if not 0 <= lo <= hi <= tr[1] - tr[0]:
                    raise ValueError(f"Synapses {s.name!r}: sample range [{lo}, {hi}) outside the target")
                tr = (tr[0] + lo, tr[0] + hi)
            if gcond:
                condition = f"({condition}) and ({gcond})" if condition else gcond
            if p_val is None and k_val is None:
                p_val = 1.0
        every_pair = p_val is None and k_val is None and s.i is None and s.j is None  # connect(condition=...)
        if p_val is not None or k_val is not None or every_pair:
            if p_val is not None and k_val is not None:
                raise ValueError(f"Synapses {s.name!r}: connect() takes p or a sample size K, not both")
            what = f"Synapses {s.name!r}"
            cond_c, p_c, pnum = "", "", 1.0 if every_pair else p_val
            try:  # index ranges and i != j: the built-in generator
                pre, post, autapses = _condition_ranges(condition, sr[1] - sr[0], tr[1] - tr[0], s)
                if not autapses and (src != tgt or sr[0] != tr[0]):
                    # the generator's flag drops equal neuron indices of one group; Brian2's i
                    # and j count within the Synapses' source and target, so anything else is code
                    raise NotImplementedError("i != j across groups or subgroup offsets")
            except NotImplementedError:  # any other condition: compiled into the generator
                from .codegen import connect_expression
                pre, post, autapses = [0, sr[1] - sr[0]], [0, tr[1] - tr[0]], True
                cond_c, const = connect_expression(condition, s.namespace, gvars[src], gvars[tgt], what)
                if const is not None:
                    cond_c = ""
                    if not const:
                        pnum = 0.0  # a condition that never holds
            if isinstance(pnum, str):
                from .codegen import connect_expression
                p_c, const = connect_expression(pnum, s.namespace, gvars[src], gvars[tgt], what,
                                                first_draw=cond_c.count("hs_crand_") if cond_c else 0)
                if const is not None:
                    p_c, pnum = "", min(1.0, max(0.0, float(const)))
                else:
                    pnum = 1.0
            pre = (sr[0] + pre[0], sr[0] + pre[1])
            post = (tr[0] + post[0], tr[0] + post[1])
            delay = s.delay if isinstance(s.delay, tuple) else (float(s.delay), float(s.delay))
            if wkind == "synapse":
                w_c = ""
                if isinstance(s.values.get(wval), str):
                    # S.w = '<expression>': evaluated per synapse by the generator (its
                    # draws numbered after the condition's and the probability's)
                    from .codegen import connect_expression
                    w_c, wconst = connect_expression(s.values[wval], s.namespace, gvars[src], gvars[tgt], what,
                                                     first_draw=(cond_c + p_c).count("hs_crand_"))
                    if wconst is not None:
                        w_c, s.weight_range = "", (float(wconst), float(wconst))
                    elif neg:
                        w_c = f"(-({w_c}))"
                if s.weight_range is None and not w_c:
                    raise NotImplementedError(f"Synapses {s.name!r}: with connect(p) or a sample size the weight is a "
                                              "constant, 'lo + (hi-lo)*rand()' or an expression (values={'w': '...'})")
                lo, hi = s.weight_range or (0.0, 0.0)
                w = (-hi, -lo) if neg else (lo, hi)
                rc = nw.RandomConnect(float(pnum or 0.0), pre, post, int(s.seed), stream=len(paths) + 1,
                                      autapses=autapses, delay=delay, wmode=HS_W_SYNAPSE, w=w, k=int(k_val or 0),
                                      cond=cond_c, p_expr=p_c, n=int(s.n), w_expr=w_c)
                paths.append(nw.PathSpec(src, tgt, var, None, None, wmode=HS_W_SYNAPSE, random=rc, code=code))
            else:
                rc = nw.RandomConnect(float(pnum or 0.0), pre, post, int(s.seed), stream=len(paths) + 1,
                                      autapses=autapses, delay=delay, wmode=HS_W_SCALAR, w=(wval, wval),
                                      k=int(k_val or 0), cond=cond_c, p_expr=p_c, n=int(s.n))
                paths.append(nw.PathSpec(src, tgt, var, None, None, wmode=HS_W_SCALAR, wscalar=wval, random=rc,
                                         code=code))
            if code is not None:
                code.init = {k: v for k, v in code.init.items() if np.asarray(v).size == 1 and not isinstance(v, str)}
            continue
        if s.i is None or s.j is None:
            raise NotImplementedError(f"Synapses {s.name!r}: connect() with i/j arrays, p= or a sample size only")
        if int(s.n) != 1:
            raise NotImplementedError(f"Synapses {s.name!r}: n={s.n} with i/j arrays (repeat the arrays instead)")
        pre = np.asarray(s.i, np.int64) + sr[0]
        post = np.asarray(s.j, np.int64) + tr[0]
        if pre.size != post.size:
            raise ValueError(f"Synapses {s.name!r}: i and j differ in length")
        d = np.asarray(s.delay, dtype=np.float64)
        kw = {}
        if d.size == 1:
            kw["scalar_delay"] = int(nw.quantise_delay(float(d.ravel()[0]), dt))
        else:
            kw["delay_steps"] = nw.quantise_delay(d, dt)
        if wkind == "synapse":
            w = np.asarray(s.values[wval], np.float64)
            w = np.broadcast_to(w, pre.shape).astype(np.float64)
            paths.append(nw.PathSpec(src, tgt, var, pre.astype(np.int32), post.astype(np.int32), wmode=HS_W_SYNAPSE,
                                     weights=-w if neg else w, code=code, **kw))
        else:
            paths.append(nw.PathSpec(src, tgt, var, pre.astype(np.int32), post.astype(np.int32), wmode=HS_W_SCALAR,
                                     wscalar=wval, code=code, **kw))
    spec = nw.NetSpec(name, float(dt), int(seed), gspecs, paths)
    names = {"groups": [g.name for g in groups], "paths": [s.name for s in synapses], "spike_monitors": [],
             "state_monitors": [], "rate_monitors": []}
    for m in monitors:
        if m.source not in gindex:
            raise NotImplementedError(f"monitor {m.name!r}: its source is not a group of the network")
        gid = gindex[m.source]
        lo, hi = m.source_range or (0, gspecs[gid].n)
        if m.kind == "spike":
            if (lo, hi) != (0, gspecs[gid].n):
                raise NotImplementedError(f"SpikeMonitor {m.name!r}: a Subgroup source runs through the C ABI "
                                          "(hs_net_add_spike_monitor_range), not the project writer")
            spec.spike_monitors.append(gid)
            names["spike_monitors"].append(m.name)
        elif m.kind == "state":
            if m.variable not in gvars[gid]:
                raise NotImplementedError(f"StateMonitor {m.name!r}: no variable {m.variable!r}")
            rec = np.atleast_1d(np.asarray(m.record if m.record is not None else np.arange(hi - lo), np.int64)) + lo
            if rec.size == 0 or len(np.unique(rec)) != rec.size:
                raise NotImplementedError(f"StateMonitor {m.name!r}: record must name distinct neurons")
            every = 1
            if m.dt is not None:
                every = int(round(float(m.dt) / dt))
                if every < 1 or abs(every * dt - float(m.dt)) > 1e-9 * float(m.dt):
                    raise NotImplementedError(f"StateMonitor {m.name!r}: its dt must be a multiple of the clock's dt")
            var = gvars[gid].index(m.variable)
            if every == 1 and np.array_equal(rec, np.arange(rec[0], rec[0] + rec.size)):
                spec.state_monitors.append((gid, var, int(rec[0]), int(rec[-1]) + 1))
            else:
                spec.state_monitors.append(nw.StateMon(gid, var, rec.astype(np.int32), every))
            names["state_monitors"].append(m.name)
        elif m.kind == "rate":
            spec.rate_monitors.append((gid, int(lo), int(hi)))
            names["rate_monitors"].append(m.name)
        else:
            raise NotImplementedError(f"monitor kind {m.kind!r}")
    return spec, names


# ---- the Brian2 device (only where Brian2 is importable) --------------------------

try:  # pragma: no cover - Brian2 is not installed where this engine is tested
    from brian2.devices.cpp_standalone.device import CPPStandaloneDevice as _CPPBase
    from brian2.devices.device import all_devices as _all_devices
except ImportError:  # Brian2 absent: the description layer above still works
    _CPPBase = None
    _all_devices = None


if _CPPBase is not None:  # pragma: no cover
    class HipStandaloneDevice(_CPPBase):
        """set_device('hip_standalone'): runs the network through libhs.

        Build options: precision (64 | 32), engine ("plain" | "fused"),
        directory (the project), options (hs_net_set_option launch shapes)."""

        def __init__(self):
            super().__init__()
            self._hs_values = {}

        def check_openmp_compatible(self, nb_threads):
            if nb_threads > 0:
                raise NotImplementedError("hip_standalone runs on one GPU; OpenMP threads are not used")

        def fill_with_array(self, var, arr):
            self._hs_values[(var.owner.name, var.name)] = np.array(arr, copy=True)
            super().fill_with_array(var, arr)

        def network_store(self, net, *args, **kwds):
            raise NotImplementedError("The store/restore mechanism is not supported in hip_standalone")

        def network_restore(self, net, *args, **kwds):
            raise NotImplementedError("The store/restore mechanism is not supported in hip_standalone")

        def network_run(self, net, duration, report=None, report_period=None, namespace=None, profile=False,
                        level=0, **kwds):
            from brian2 import (NeuronGroup, PoissonGroup, PopulationRateMonitor, SpikeGeneratorGroup,
                                SpikeMonitor, StateMonitor, Synapses)
            from brian2.core.namespace import get_local_namespace

            from . import project
            if namespace is None:
                namespace = get_local_namespace(level=level + 2)
            net.before_run(namespace)
            dt = float(net.objects[0].clock.dt_) if net.objects else 1e-4
            steps = int(round(float(duration) / dt))

            def ns_of(obj, code):
                ids = set(re.findall(r"[A-Za-z_]\w*", code or ""))
                res = obj.resolve_all(sorted(ids), namespace)
                out = {}
                for k, v in res.items():
                    try:
                        out[k] = float(v.get_value() if hasattr(v, "get_value") else v)
                    except (TypeError, ValueError):
                        pass
                return out

            groups, syns, mons = [], [], []
            for obj in net.sorted_objects:
                if isinstance(obj, PoissonGroup):
                    groups.append(GroupDesc(obj.name, len(obj), kind="poisson",
                                            rates=self._hs_values.get((obj.name, "rates"), obj.rates_)))
                elif isinstance(obj, SpikeGeneratorGroup):
                    groups.append(GroupDesc(obj.name, len(obj), kind="spikegenerator", indices=obj._neuron_index,
                                            times=obj._spike_time, period=float(obj.period_)))
                elif isinstance(obj, NeuronGroup):
                    eqs = str(obj.user_equations)
                    thr = obj.events.get("spike")
                    rst = obj.event_codes.get("spike")
                    ref = obj._refractory
                    groups.append(GroupDesc(
                        obj.name, len(obj), equations=eqs, threshold=thr, reset=rst,
                        refractory=float(ref) if ref is not False else 0.0, method=obj.method_choice,
                        namespace=ns_of(obj, eqs + " " + (thr or "") + " " + (rst or "")),
                        init={v: self._hs_values[(obj.name, v)] for (o, v) in self._hs_values if o == obj.name}))
                elif isinstance(obj, Synapses):
                    code = obj.pre.code if hasattr(obj, "pre") else ""
                    post_code = obj.post.code if hasattr(obj, "post") else ""
                    model = str(obj.equations) if getattr(obj, "equations", None) is not None else ""
                    src = obj.source
                    syns.append(SynapsesDesc(
                        obj.name, getattr(src, "source", src).name, getattr(obj.target, "source", obj.target).name,
                        on_pre=str(code), model=model, on_post=str(post_code),
                        namespace=ns_of(obj, " ".join((str(code), str(post_code), model))),
                        i=self._hs_values.get((obj.name, "_synaptic_pre")),
                        j=self._hs_values.get((obj.name, "_synaptic_post")),
                        values={k[1]: v for k, v in self._hs_values.items() if k[0] == obj.name},
                        delay=self._hs_values.get((obj.name, "delay"), 0.0)))
                elif isinstance(obj, SpikeMonitor):
                    mons.append(MonitorDesc(obj.name, "spike", obj.source.name))
                elif isinstance(obj, StateMonitor):
                    for var in obj.record_variables:
                        mons.append(MonitorDesc(obj.name, "state", obj.source.name, variable=var,
                                                record=obj.record))
                elif isinstance(obj, PopulationRateMonitor):
                    mons.append(MonitorDesc(obj.name, "rate", obj.source.name))
            spec, names = network_spec(groups, syns, mons, dt, name=net.name)
            opts = self.build_options
            directory = opts.get("directory", "output")
            project.write_project(spec, directory, steps, precision=opts.get("precision", 64),
                                  engine=opts.get("engine", "plain"), options=opts.get("options"), names=names)
            project.build_project(directory)
            project.run_project(directory, "results")
            self._hs_results = project.read_results(directory)
            net.t_ = float(net.t_) + steps * dt

    hip_standalone_device = HipStandaloneDevice()
    _all_devices["hip_standalone"] = hip_standalone_device
