"""Standalone project writer: the code-generation back end of the hip_standalone device.

The reference's device turns a Brian2 network into a generated C++/CUDA project
(brian2cuda/device.py: main source :496-708, build :1231-1506, network_run
:1508-1765), compiles it and runs the binary, which writes every result array
into a results directory (templates/objects.cu:522-640).  Here a network
description (:class:`brian2cuda_amd.networks.NetSpec`) becomes:

  main.cpp          host code that calls libhs (include/hs.h) -- groups,
                    pathways, monitors, one run -- and writes the result files
  static_arrays/    the network's arrays (initial state, pre/post/delay/weight)
                    as raw little-endian files, as the reference writes its
                    static arrays (device.py:1276-1283)
  Makefile          one compile and link line against libhs.so
  project.json      what the project holds and which result file is which

The binary's command line is the reference's (templates/main.cu:29-62,
objects.cu:89-183): ``main [--results_dir DIR] [owner.var=value | owner.var=file]``.
A value starting with '-' or a digit sets every element; anything else names a
raw file of the variable's size (float64 here); "true"/"false" are 1/0;
an unknown name is an error.  The run writes, into the results directory, the
raw arrays of every monitor and of every group's state variables at the end of
the run, plus last_run_info.txt.

File names are Brian2's array names (``_array_<owner>_<var>``,
``_dynamic_array_<owner>_<var>``, device.py:137-143) without Brian2's hash
suffix, which needs Brian2 itself to compute.
"""
import json
import os
import shlex
import shutil
import subprocess

import numpy as np

from . import networks as nw
from .runtime import (HS_ENGINE_FUSED, HS_ENGINE_PLAIN, HS_MODEL_HH_COBA, HS_MODEL_LIF_CUBA, HS_MODEL_LIF_NOISE,
                      HS_MODEL_POISSON, HS_MODEL_SPIKEGEN, HS_W_SCALAR, HS_W_SYNAPSE)

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
INCLUDE = os.path.join(REPO, "include")

#: state-variable names of each model, in slot order (include/hs_models.h)
MODEL_VARS = nw.MODEL_VAR_NAMES
MODEL_MACROS = {
    HS_MODEL_LIF_CUBA: "HS_MODEL_LIF_CUBA",
    HS_MODEL_LIF_NOISE: "HS_MODEL_LIF_NOISE",
    HS_MODEL_HH_COBA: "HS_MODEL_HH_COBA",
    HS_MODEL_SPIKEGEN: "HS_MODEL_SPIKEGEN",
    HS_MODEL_POISSON: "HS_MODEL_POISSON",
}


def _default_names(prefix, n):
    """Brian2's automatic object names: prefix, prefix_1, prefix_2, ..."""
    return [prefix if k == 0 else f"{prefix}_{k}" for k in range(n)]


def _c_double(x):
    return repr(float(x)) if np.isfinite(x) else ("INFINITY" if x > 0 else "-INFINITY")


def _c_str(s):
    """A C string literal of s (one literal per line of s, joined by adjacency)."""
    lines = (s or "").split("\n")
    esc = [ln.replace("\\", "\\\\").replace('"', '\\"') for ln in lines]
    return " ".join(f'"{e}\\n"' if k + 1 < len(esc) else f'"{e}"' for k, e in enumerate(esc))


def group_vars(g):
    """State-variable names of a group spec, in slot order."""
    return list(g.code.names) if getattr(g, "code", None) is not None else MODEL_VARS[g.model]


class ProjectWriter:
    """Writes one network and one run as a standalone project.

    spec      : NetSpec (groups, pathways, monitors)
    steps     : steps of the run (``run(steps * dt)``)
    precision : 64 or 32 (HS_F64 / HS_F32)
    engine    : "plain" or "fused"
    options   : {hs_net_set_option key: value} (launch shapes, e.g. from the autotuner)
    names     : optional {"groups": [...], "paths": [...], "spike_monitors": [...],
                "state_monitors": [...], "rate_monitors": [...]} object names
    """

    def __init__(self, spec, steps, precision=64, engine="plain", options=None, names=None, device=0,
                 spike_capacity=1 << 20):
        self.spec = spec
        self.steps = int(steps)
        self.precision = int(precision)
        if self.precision not in (32, 64):
            raise ValueError("precision is 32 or 64")
        self.engine = {"plain": HS_ENGINE_PLAIN, "fused": HS_ENGINE_FUSED}[engine]
        self.options = {int(k): int(v) for k, v in (options or {}).items()}
        self.device = int(device)
        self.spike_capacity = int(spike_capacity)
        names = names or {}
        self.gnames = names.get("groups") or _default_names("neurongroup", len(spec.groups))
        self.pnames = names.get("paths") or _default_names("synapses", len(spec.paths))
        self.smnames = names.get("spike_monitors") or _default_names("spikemonitor", len(spec.spike_monitors))
        self.stnames = names.get("state_monitors") or _default_names("statemonitor", len(spec.state_monitors))
        self.rmnames = names.get("rate_monitors") or _default_names("ratemonitor", len(spec.rate_monitors))
        for k, (lst, objs) in {"groups": (self.gnames, spec.groups), "paths": (self.pnames, spec.paths),
                               "spike_monitors": (self.smnames, spec.spike_monitors),
                               "state_monitors": (self.stnames, spec.state_monitors),
                               "rate_monitors": (self.rmnames, spec.rate_monitors)}.items():
            if len(lst) != len(objs):
                raise ValueError(f"{len(lst)} names for {len(objs)} {k}")
        allnames = self.gnames + self.pnames + self.smnames + self.stnames + self.rmnames
        if len(set(allnames)) != len(allnames):
            raise ValueError("object names must be unique")
        for nm in allnames:
            if not nm.replace("_", "a").isalnum() or nm[0].isdigit():
                raise ValueError(f"object name {nm!r} is not an identifier")

    # ---- static arrays ----------------------------------------------------------

    def _static(self, directory, name, arr, dtype):
        a = np.ascontiguousarray(arr, dtype=dtype)
        rel = os.path.join("static_arrays", name)
        a.astype(a.dtype.newbyteorder("<"), copy=False).tofile(os.path.join(directory, rel))
        return rel, int(a.size)

    # ---- main.cpp -----------------------------------------------------------------

    def write(self, directory):
        """Write the project into `directory` (created if missing).  Returns the
        project description (also written to project.json)."""
        spec = self.spec
        os.makedirs(os.path.join(directory, "static_arrays"), exist_ok=True)
        L = []  # body of main()
        info = {"name": spec.name, "dt": spec.dt, "seed": spec.seed, "steps": self.steps,
                "precision": self.precision, "engine": self.engine, "groups": [], "paths": [], "monitors": []}
        # initial state: host vectors (from static arrays or constants), then
        # command-line overrides, then the upload
        decl, settable = [], []
        for gi, g in enumerate(spec.groups):
            gname = self.gnames[gi]
            vars_ = group_vars(g)
            ginfo = {"name": gname, "model": g.model, "n": g.n, "vars": vars_, "init": {}}
            for var, val in sorted(g.init.items()):
                cname = f"_init_{gname}_{vars_[var]}"
                if np.isscalar(val) or np.asarray(val).size == 1:
                    decl.append(f"    std::vector<double> {cname}({g.n}, {_c_double(float(np.asarray(val).ravel()[0]))});")
                    ginfo["init"][vars_[var]] = float(np.asarray(val).ravel()[0])
                else:
                    rel, n = self._static(directory, f"_array_{gname}_{vars_[var]}", val, np.float64)
                    if n != g.n:
                        raise ValueError(f"{gname}.{vars_[var]}: {n} values for {g.n} neurons")
                    decl.append(f"    std::vector<double> {cname}({g.n});")
                    decl.append(f"    HS_IO(read_array(\"{rel}\", {cname}.data(), sizeof(double) * {g.n}));")
                    ginfo["init"][vars_[var]] = rel
                settable.append((f"{gname}.{vars_[var]}", cname))
            info["groups"].append(ginfo)
        L += decl
        L.append("    for (const std::string &a : args) {")
        L.append("        const size_t eq = a.find('=');")
        L.append("        if (eq == std::string::npos) { fprintf(stderr, \"Cannot parse argument '%s'.\\n\", a.c_str()); return 1; }")
        L.append("        const std::string name = a.substr(0, eq), value = a.substr(eq + 1);")
        L.append("        std::vector<double> *target = nullptr;")
        for full, cname in settable:
            L.append(f"        if (name == \"{full}\") target = &{cname};")
        L.append("        if (!target) { fprintf(stderr, \"Cannot set unknown variable '%s'.\\n\", name.c_str()); return 1; }")
        L.append("        if (set_from_string(*target, value) != 0) return 1;")
        L.append("    }")
        L.append(f"    HS_CHECK(hs_init({self.device}));")
        L.append("    hs_net *net = nullptr;")
        L.append(f"    HS_CHECK(hs_net_create({self.precision}, {self.engine}, {_c_double(spec.dt)}, "
                 f"{int(spec.seed)}ULL, &net));")
        for k, v in sorted(self.options.items()):
            L.append(f"    HS_CHECK(hs_net_set_option(net, {k}, {v}));")
        for gi, g in enumerate(spec.groups):
            gname = self.gnames[gi]
            L.append(f"    int g_{gname};")
            if g.code is not None:  # the model as code: compiled for the device on the first run
                m = g.code
                nm = ", ".join(f'"{n}"' for n in m.names)
                L.append(f"    {{ const char *names[] = {{{nm}}};")
                L.append(f"      HS_CHECK(hs_net_add_group_code(net, {g.n}, {len(m.names)}, names,")
                L.append(f"          {_c_str(m.update)},")
                L.append(f"          {_c_str(m.threshold)}, {_c_str(m.reset)}, {_c_double(m.refractory)}, "
                         f"{int(bool(m.noise))}, {int(m.stream)}u, &g_{gname})); }}")
            else:
                p = ", ".join(_c_double(x) for x in g.params)
                L.append(f"    {{ const double p[] = {{{p}}};")
                L.append(f"      HS_CHECK(hs_net_add_group(net, {MODEL_MACROS[g.model]}, {g.n}, p, {len(g.params)}, "
                         f"&g_{gname})); }}")
            for var in sorted(g.init):
                L.append(f"    HS_CHECK(hs_net_set_state(net, g_{gname}, {var}, _init_{gname}_"
                         f"{group_vars(g)[var]}.data()));")
            if g.spikes is not None:
                idx, st = g.spikes
                ri, n = self._static(directory, f"_array_{gname}__spikegen_indices", idx, np.int32)
                rs, _ = self._static(directory, f"_array_{gname}__spikegen_steps", st, np.int64)
                L.append(f"    {{ std::vector<int32_t> i({max(n, 1)}); std::vector<int64_t> s({max(n, 1)});")
                L.append(f"      HS_IO(read_array(\"{ri}\", i.data(), sizeof(int32_t) * {n}));")
                L.append(f"      HS_IO(read_array(\"{rs}\", s.data(), sizeof(int64_t) * {n}));")
                L.append(f"      HS_CHECK(hs_net_set_spikegen(net, g_{gname}, {n}, i.data(), s.data())); }}")
        for pi, p in enumerate(spec.paths):
            pname = self.pnames[pi]
            src, tgt = self.gnames[p.src], self.gnames[p.tgt]
            pinfo = {"name": pname, "source": src, "target": tgt, "var": p.var}
            L.append(f"    int p_{pname};")
            if p.random is not None:
                r = p.random
                k = int(getattr(r, "k", 0) or 0)
                w_expr = getattr(r, "w_expr", "")
                if getattr(r, "cond", "") or getattr(r, "p_expr", "") or getattr(r, "n", 1) > 1 or w_expr:
                    # condition / probability as code: compiled for the device by this call,
                    # after the groups' state (command-line values included) is set
                    fn = "hs_net_connect_code"
                    arg = f"{_c_str(r.cond)}, {_c_str(r.p_expr)}, {_c_double(r.p)}, {k}, {int(getattr(r, 'n', 1))}"
                    pinfo["connect"] = "code"
                else:
                    fn, arg = ("hs_net_connect_fixed", str(k)) if k else ("hs_net_connect_random", _c_double(r.p))
                    pinfo["connect"] = "fixed" if k else "random"
                L.append(f"    HS_CHECK({fn}(net, g_{src}, g_{tgt}, {p.var}, {r.pre[0]}, {r.pre[1]}, "
                         f"{r.post[0]}, {r.post[1]}, {arg}, {int(bool(r.autapses))}, {int(r.seed)}ULL, "
                         f"{int(r.stream)}u, {_c_double(r.delay[0])}, {_c_double(r.delay[1])}, {int(r.wmode)}, "
                         f"{_c_double(r.w[0])}, {_c_double(r.w[1])}, "
                         f"{(_c_str(w_expr) + ', ') if fn == 'hs_net_connect_code' else ''}&p_{pname}));")
            else:
                rp, n = self._static(directory, f"_dynamic_array_{pname}__synaptic_pre", p.pre, np.int32)
                rq, _ = self._static(directory, f"_dynamic_array_{pname}__synaptic_post", p.post, np.int32)
                m = max(n, 1)
                L.append(f"    {{ std::vector<int32_t> pre({m}), post({m});")
                L.append(f"      HS_IO(read_array(\"{rp}\", pre.data(), sizeof(int32_t) * {n}));")
                L.append(f"      HS_IO(read_array(\"{rq}\", post.data(), sizeof(int32_t) * {n}));")
                dptr = "nullptr"
                if p.delay_steps is not None:
                    rd, _ = self._static(directory, f"_dynamic_array_{pname}_delay_steps", p.delay_steps, np.int32)
                    L.append(f"      std::vector<int32_t> d({m});")
                    L.append(f"      HS_IO(read_array(\"{rd}\", d.data(), sizeof(int32_t) * {n}));")
                    dptr = "d.data()"
                wptr = "nullptr"
                if p.wmode != HS_W_SCALAR:
                    wn = n if p.wmode == 1 else spec.groups[p.src].n
                    rw, got = self._static(directory, f"_dynamic_array_{pname}_w", p.weights, np.float64)
                    if got != wn:
                        raise ValueError(f"{pname}: {got} weights, expected {wn}")
                    L.append(f"      std::vector<double> w({max(wn, 1)});")
                    L.append(f"      HS_IO(read_array(\"{rw}\", w.data(), sizeof(double) * {wn}));")
                    wptr = "w.data()"
                L.append(f"      HS_CHECK(hs_net_add_pathway(net, g_{src}, g_{tgt}, {p.var}, {n}, pre.data(), "
                         f"post.data(), {dptr}, {int(p.scalar_delay)}, {int(p.wmode)}, {wptr}, "
                         f"{_c_double(p.wscalar)}, &p_{pname})); }}")
                pinfo["connect"] = "arrays"
                pinfo["nsyn"] = n
            if p.stdp is not None:
                sp = ", ".join(_c_double(x) for x in p.stdp)
                L.append(f"    {{ const double sp[] = {{{sp}}};")
                L.append(f"      HS_CHECK(hs_net_pathway_stdp(net, p_{pname}, sp, {len(p.stdp)})); }}")
                pinfo["stdp"] = list(map(float, p.stdp))
            if p.code is not None:  # on_pre as code: compiled for the device on the first run
                c = p.code
                post_i, pre_i, read_i = nw.code_var_indices(spec, p)
                def arr(xs, empty):
                    return "{" + (", ".join(xs) if xs else empty) + "}"

                def quoted(names):
                    return ['"' + n + '"' for n in names]

                sn = arr(quoted(c.syn_names), "nullptr")
                pv = arr([str(k) for k in post_i], "0")
                pn = arr(quoted(c.post_names), "nullptr")
                qv = arr([str(k) for k in pre_i], "0")
                qn = arr(quoted(c.pre_names), "nullptr")
                L.append(f"    {{ const char *sn[] = {sn};")
                L.append(f"      const int pv[] = {pv};")
                L.append(f"      const char *pn[] = {pn};")
                L.append(f"      const int qv[] = {qv};")
                L.append(f"      const char *qn[] = {qn};")
                L.append(f"      HS_CHECK(hs_net_pathway_code(net, p_{pname}, {len(c.syn_names)}, sn,")
                L.append(f"          {_c_str(c.on_pre)},")
                L.append(f"          {len(post_i)}, pv, pn, {len(pre_i)}, qv, qn));")
                if c.on_post:
                    rv = arr([str(k) for k in read_i], "0")
                    rn = arr(quoted(c.post_read_names), "nullptr")
                    L.append(f"      const int rv[] = {rv};")
                    L.append(f"      const char *rn[] = {rn};")
                    L.append(f"      HS_CHECK(hs_net_pathway_code_post(net, p_{pname},")
                    L.append(f"          {_c_str(c.on_post)},")
                    L.append(f"          {len(read_i)}, rv, rn));")
                L.append("    }")
                for name, val in (c.init or {}).items():
                    if name == "w" and p.wmode == HS_W_SYNAPSE:
                        continue  # the pathway's weights
                    k = c.syn_names.index(name)
                    a = np.asarray(val, np.float64)
                    L.append("    { int64_t ns = 0; int dm, dd, mo;")
                    L.append(f"      HS_CHECK(hs_net_pathway_info(net, p_{pname}, &ns, &dm, &dd, &mo));")
                    if a.size > 1:
                        if p.random is not None:
                            raise ValueError("per-synapse initial values need explicit synapses")
                        rv, nv = self._static(directory, f"_array_{pname}_{name}", a[nw.bundle_order(p)],
                                              np.float64)
                        L.append(f"      std::vector<double> v({max(nv, 1)});")
                        L.append(f"      HS_IO(read_array(\"{rv}\", v.data(), sizeof(double) * {nv}));")
                    else:
                        L.append(f"      std::vector<double> v((size_t)(ns ? ns : 1), {_c_double(float(a.ravel()[0]))});")
                    L.append(f"      HS_CHECK(hs_net_set_synapse_var(net, p_{pname}, HS_SYN_CODE + {k}, v.data(), ns)); }}")
                pinfo["code"] = {"syn_vars": list(c.syn_names), "on_pre": c.on_pre}
            info["paths"].append(pinfo)
        writes = []
        for mi, gid in enumerate(spec.spike_monitors):
            mname, gname = self.smnames[mi], self.gnames[gid]
            L.append(f"    int m_{mname};")
            L.append(f"    HS_CHECK(hs_net_add_spike_monitor(net, g_{gname}, {self.spike_capacity}LL, &m_{mname}));")
            files = {"i": f"_dynamic_array_{mname}_i", "t": f"_dynamic_array_{mname}_t",
                     "count": f"_array_{mname}_count", "N": f"_array_{mname}_N"}
            writes.append((mname, [files["i"], files["t"], files["count"], files["N"]]))
            info["monitors"].append({"name": mname, "kind": "spike", "source": gname, "files": files})
        for mi, m in enumerate(spec.state_monitors):
            mname = self.stnames[mi]
            L.append(f"    int m_{mname};")
            if isinstance(m, nw.StateMon):
                gid, var = m.gid, m.var
                gname = self.gnames[gid]
                ri, n = self._static(directory, f"_array_{mname}__indices", m.indices, np.int32)
                L.append(f"    {{ std::vector<int32_t> idx({n});")
                L.append(f"      HS_IO(read_array(\"{ri}\", idx.data(), sizeof(int32_t) * {n}));")
                L.append(f"      HS_CHECK(hs_net_add_state_monitor_indices(net, g_{gname}, {var}, idx.data(), {n}, "
                         f"{int(m.every)}, &m_{mname})); }}")
                width, every = n, int(m.every)
            else:
                gid, var, start, stop = m
                gname = self.gnames[gid]
                L.append(f"    HS_CHECK(hs_net_add_state_monitor(net, g_{gname}, {var}, {start}, {stop}, &m_{mname}));")
                width, every = int(stop - start), 1
            vname = group_vars(spec.groups[gid])[var]
            files = {"t": f"_dynamic_array_{mname}_t", vname: f"_dynamic_array_{mname}_{vname}",
                     "indices": f"_array_{mname}__indices", "N": f"_array_{mname}_N"}
            writes.append((mname, [files["t"], files[vname], files["indices"], files["N"]]))
            info["monitors"].append({"name": mname, "kind": "state", "source": gname, "var": vname,
                                     "width": width, "every": every, "files": files})
        for mi, (gid, start, stop) in enumerate(spec.rate_monitors):
            mname, gname = self.rmnames[mi], self.gnames[gid]
            L.append(f"    int m_{mname};")
            L.append(f"    HS_CHECK(hs_net_add_rate_monitor(net, g_{gname}, {start}, {stop}, &m_{mname}));")
            files = {"t": f"_dynamic_array_{mname}_t", "rate": f"_dynamic_array_{mname}_rate",
                     "N": f"_array_{mname}_N"}
            writes.append((mname, [files["t"], files["rate"], None, files["N"]]))
            info["monitors"].append({"name": mname, "kind": "rate", "source": gname, "files": files})
        L.append(f"    HS_CHECK(hs_net_run(net, {self.steps}LL));  // run({self.steps} * dt)")
        L.append("    make_dir(results_dir);")
        for mname, files in writes:
            args = ", ".join("nullptr" if f is None else f"res(\"{f}\").c_str()" for f in files)
            L.append(f"    HS_CHECK(hs_net_write_monitor(net, m_{mname}, {args}));")
        state_files = {}
        for gi, g in enumerate(spec.groups):
            gname = self.gnames[gi]
            if g.model == HS_MODEL_SPIKEGEN:
                continue
            for var, vname in enumerate(group_vars(g)):
                f = f"_array_{gname}_{vname}"
                state_files[f"{gname}.{vname}"] = f
                L.append(f"    {{ std::vector<double> x({g.n});")
                L.append(f"      HS_CHECK(hs_net_get_state(net, g_{gname}, {var}, x.data()));")
                L.append(f"      HS_IO(write_array(res(\"{f}\"), x.data(), sizeof(double) * {g.n})); }}")
        info["state_files"] = state_files
        L.append("    HS_CHECK(hs_net_write_last_run_info(net, res(\"last_run_info.txt\").c_str()));")
        L.append("    hs_net_destroy(net);")
        L.append("    return 0;")
        with open(os.path.join(directory, "main.cpp"), "w") as f:
            f.write(MAIN_TEMPLATE.replace("@NAME@", spec.name).replace("@BODY@", "\n".join(L)))
        lib_dir = HERE
        with open(os.path.join(directory, "Makefile"), "w") as f:
            f.write(MAKEFILE_TEMPLATE.replace("@INCLUDE@", INCLUDE).replace("@LIBDIR@", lib_dir))
        info["libhs"] = os.path.join(lib_dir, "libhs.so")
        with open(os.path.join(directory, "project.json"), "w") as f:
            json.dump(info, f, indent=1)
        return info


MAIN_TEMPLATE = r"""// Standalone project of network "@NAME@", written by brian2cuda_amd.project:
// every simulation call goes to libhs (include/hs.h).
// Usage: ./main [--results_dir DIR] [owner.var=value | owner.var=file]...
#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <sys/stat.h>

#include "hs.h"

#define HS_CHECK(x)                                                                  \
    do {                                                                             \
        int _r = (x);                                                                \
        if (_r != HS_OK) {                                                           \
            fprintf(stderr, "%s failed (%d): %s\n", #x, _r, hs_last_error());        \
            return 1;                                                                \
        }                                                                            \
    } while (0)
#define HS_IO(x)                                                                     \
    do {                                                                             \
        if ((x) != 0) return 1;                                                      \
    } while (0)

static std::string results_dir = "results";

static std::string res(const char *name) { return results_dir + "/" + name; }

static void make_dir(const std::string &d)
{
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) fprintf(stderr, "Could not create '%s'\n", d.c_str());
}

static int read_array(const std::string &path, void *dst, size_t bytes)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        fprintf(stderr, "Could not read '%s'\n", path.c_str());
        return 1;
    }
    const size_t got = bytes ? fread(dst, 1, bytes, f) : 0;
    const bool extra = fgetc(f) != EOF;
    fclose(f);
    if (got != bytes || extra) {
        fprintf(stderr, "Error reading '%s': expected %zu bytes\n", path.c_str(), bytes);
        return 1;
    }
    return 0;
}

static int write_array(const std::string &path, const void *src, size_t bytes)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f || (bytes && fwrite(src, 1, bytes, f) != bytes)) {
        if (f) fclose(f);
        fprintf(stderr, "Error writing '%s'\n", path.c_str());
        return 1;
    }
    fclose(f);
    return 0;
}

/* owner.var=value: a number (leading '-' or digit; true/false are 1/0) sets
   every element, anything else names a raw float64 file of the variable's size */
static int set_from_string(std::vector<double> &var, std::string s)
{
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    if (s == "true") s = "1";
    else if (s == "false") s = "0";
    if (!s.empty() && (s[0] == '-' || (s[0] >= '0' && s[0] <= '9'))) {
        char *end = nullptr;
        const double x = strtod(s.c_str(), &end);
        if (!end || *end != '\0') {
            fprintf(stderr, "Cannot parse value '%s'\n", s.c_str());
            return 1;
        }
        for (double &v : var) v = x;
        return 0;
    }
    return read_array(s, var.data(), sizeof(double) * var.size());
}

int main(int argc, char **argv)
{
    std::vector<std::string> args(argv + 1, argv + argc);
    if (args.size() >= 2 && args[0] == "--results_dir") {
        results_dir = args[1];
        args.erase(args.begin(), args.begin() + 2);
    }
@BODY@
}
"""

MAKEFILE_TEMPLATE = """# Builds the standalone binary against libhs.so (host code only: the kernels live in libhs).
CXX ?= g++
CXXFLAGS ?= -O2 -std=c++17 -Wall
INCLUDE = @INCLUDE@
LIBDIR = @LIBDIR@

main: main.cpp
\t$(CXX) $(CXXFLAGS) -I$(INCLUDE) main.cpp -L$(LIBDIR) -lhs -Wl,-rpath,$(LIBDIR) -o main

clean:
\trm -f main
"""


def write_project(spec, directory, steps, **kw):
    """Write the standalone project of `spec` for a run of `steps` steps (see ProjectWriter)."""
    return ProjectWriter(spec, steps, **kw).write(directory)


def compile_command(directory, compiler=None):
    """The compile-and-link command of a written project (what its Makefile runs)."""
    cxx = compiler or shutil.which("g++") or "c++"
    return [cxx, "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", "main.cpp", f"-L{HERE}", "-lhs",
            f"-Wl,-rpath,{HERE}", "-o", "main"]


def build_project(directory, compiler=None):
    """Compile the project's main.cpp into `directory`/main (device.py:1231-1506 runs make)."""
    if not os.path.exists(os.path.join(HERE, "libhs.so")):
        raise RuntimeError("libhs.so is not built (python -m brian2cuda_amd.build)")
    cmd = compile_command(directory, compiler)
    r = subprocess.run(cmd, cwd=directory, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compilation failed ({shlex.join(cmd)}):\n{r.stderr}")
    return os.path.join(directory, "main")


def run_project(directory, results_dir="results", args=(), timeout=None):
    """Run the built binary from `directory` (static_arrays/ paths are relative to
    it), as the reference's device runs its binary (device.py:1461-1506).
    Returns the CompletedProcess; a non-zero exit raises with the binary's stderr."""
    cmd = [os.path.join(directory, "main"), "--results_dir", str(results_dir), *map(str, args)]
    r = subprocess.run(cmd, cwd=directory, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"standalone binary failed with exit code {r.returncode}:\n{r.stderr}")
    return r


def read_results(directory, results_dir="results"):
    """The result arrays of a finished run, by object: {monitor name: {array: numpy}},
    {"<group>.<var>": final state}, and "last_run_info": (wall s, completed fraction)."""
    with open(os.path.join(directory, "project.json")) as f:
        info = json.load(f)
    rd = results_dir if os.path.isabs(results_dir) else os.path.join(directory, results_dir)
    real = "<f8" if info["precision"] == 64 else "<f4"
    out = {}
    for m in info["monitors"]:
        files = m["files"]
        if m["kind"] == "spike":
            out[m["name"]] = {"i": np.fromfile(os.path.join(rd, files["i"]), "<i4"),
                              "t": np.fromfile(os.path.join(rd, files["t"]), "<f8"),
                              "count": np.fromfile(os.path.join(rd, files["count"]), "<i4"),
                              "N": int(np.fromfile(os.path.join(rd, files["N"]), "<i4")[0])}
        elif m["kind"] == "state":
            t = np.fromfile(os.path.join(rd, files["t"]), "<f8")
            width = m["width"]
            v = np.fromfile(os.path.join(rd, files[m["var"]]), real).reshape(t.size, width)
            out[m["name"]] = {"t": t, m["var"]: v, "indices": np.fromfile(os.path.join(rd, files["indices"]), "<i4"),
                              "N": int(np.fromfile(os.path.join(rd, files["N"]), "<i4")[0])}
        else:
            out[m["name"]] = {"t": np.fromfile(os.path.join(rd, files["t"]), "<f8"),
                              "rate": np.fromfile(os.path.join(rd, files["rate"]), "<f8"),
                              "N": int(np.fromfile(os.path.join(rd, files["N"]), "<i4")[0])}
    for full, f in info["state_files"].items():
        out[full] = np.fromfile(os.path.join(rd, f), "<f8")
    with open(os.path.join(rd, "last_run_info.txt")) as f:
        a, b = f.read().split()
    out["last_run_info"] = (float(a), float(b))
    return out
