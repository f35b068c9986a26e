"""Launch-shape autotuner for the per-timestep engine.

Picks the engine's tuning options (hs_net_set_option keys) for one network,
precision and engine form by measurement: coordinate descent from the
defaults, one option at a time over its candidate values, each candidate timed
on a fresh network built from the same spec (so tuning never touches the
network the caller simulates), keeping a value only when it beats the best
time by a margin.  Results are cached in a JSON file keyed by the device and a
signature of the network, so a second call costs nothing.

The options only change launch shapes and the order of floating-point
additions inside atomics; bit-exact tests run with tuned options too.

Replaces the fixed launch policies of the reference: threads per bundle and
blocks chosen from the bundle sizes (brian2cuda/templates/synapses.cu:171-322),
and block sizes from cudaOccupancyMaxPotentialBlockSize in the generated
kernels (common_group.cu:207, threshold.cu:123).
"""
import json
import os
import time

from . import runtime as rt

def candidate_space(n_cu, engine):
    """candidate values per option key; the first of each list is the engine's
    default.  OPT_PIPELINED selects the form of the windowed engine: the launch
    shapes that suit one form do not suit the other, so autotune() descends
    once per form (tune_forms) over the other keys."""
    space = {}
    if engine == "fused":
        space[rt.OPT_PIPELINED] = [0, 1]
        space[rt.OPT_WINDOW] = [0, 4, 3, 2]
    space[rt.OPT_UPDATE_BLOCK] = [256, 128, 64]
    if engine == "fused":
        space[rt.OPT_NOISE_BLOCKS] = [0, 1 * n_cu, 3 * n_cu, 4 * n_cu, 8 * n_cu]
    space[rt.OPT_DELIVER_GRID] = [0, 2 * n_cu, 3 * n_cu, 6 * n_cu, 8 * n_cu]
    if engine == "fused":
        space[rt.OPT_FUSED_GRID] = [0, 1 * n_cu, 4 * n_cu]
    space[rt.OPT_UPDATE_GRID] = [0, 4 * n_cu, 8 * n_cu]
    return space


def signature(spec, precision, engine, device_name):
    groups = ";".join(f"{g.model}:{g.n}" for g in spec.groups)
    paths = ";".join(f"{p.src}>{p.tgt}:{p.nsyn}:{p.scalar_delay if p.delay_steps is None else 'd'}"
                     for p in spec.paths)
    return f"{device_name}|{precision}|{engine}|{spec.name}|{groups}|{paths}"


def tune_forms(space, measure, form_key, rounds=3, margin=0.005, log=None):
    """coordinate_descent over space without form_key, once for each value of
    form_key (held fixed); the best of the descents.  Returns (options, time,
    trials) like coordinate_descent."""
    if form_key not in space:
        return coordinate_descent(space, measure, rounds, margin, log)
    rest = {k: v for k, v in space.items() if k != form_key}
    best = None
    trials = []
    for f in space[form_key]:
        opts, t, tr = coordinate_descent(rest, lambda o: measure({form_key: f, **o}), rounds, margin, log)
        trials += [({form_key: f, **o}, tt) for o, tt in tr]
        if best is None or t < best[1] * (1.0 - margin):
            best = ({form_key: f, **opts}, t)
    return best[0], best[1], trials


def coordinate_descent(space, measure, rounds=3, margin=0.005, log=None):
    """Minimise measure(options) over the product of space's candidate lists by
    changing one key at a time.  A candidate replaces the current value only if
    it is faster by more than `margin` (relative).  Returns (options, time,
    trials) where trials lists every (options, time) measured."""
    best = {k: v[0] for k, v in space.items()}
    trials = []

    def run(opts):
        t = measure(dict(opts))
        trials.append((dict(opts), t))
        if log:
            log(f"autotune {opts} -> {t:.3f}")
        return t

    best_t = run(best)
    for _ in range(rounds):
        improved = False
        for key, values in space.items():
            for v in values:
                if v == best[key]:
                    continue
                cand = dict(best)
                cand[key] = v
                t = run(cand)
                if t < best_t * (1.0 - margin):
                    best, best_t, improved = cand, t, True
        if not improved:
            break
    return best, best_t, trials


def measure_us_per_step(spec, precision, engine, options, steps, warmup, build=None):
    """µs per step of a fresh network with `options`, timed by the runtime's
    own clock around hs_net_run (warm-up steps first)."""
    from . import networks as nw
    net, _, _ = (build or nw.build_hs)(spec, precision, engine, options=options)
    try:
        net.run(warmup)
        net.run(steps)
        return net.stats()["last_run_wall_s"] * 1e6 / steps
    finally:
        net.close()


def autotune(spec, precision, engine, steps=4000, warmup=300, cache=None, rounds=3, log=None, build=None):
    """Options (key -> value, defaults omitted) for this network, precision and
    form; cache: JSON path (None: $BRIAN2CUDA_AMD_CACHE or ~/.cache/brian2cuda_amd)."""
    from . import _lib
    info = _lib.device_info()
    key = signature(spec, precision, engine, info["arch"])
    path = cache or os.path.join(os.environ.get("BRIAN2CUDA_AMD_CACHE",
                                                os.path.join(os.path.expanduser("~"), ".cache", "brian2cuda_amd")),
                                 "autotune.json")
    table = {}
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = {}  # no cache yet, or unreadable: tune afresh
    if key in table:
        return {int(k): v for k, v in table[key]["options"].items()}
    space = candidate_space(info["compute_units"], engine)
    t0 = time.time()
    measure = lambda o: measure_us_per_step(spec, precision, engine, o, steps, warmup, build)
    best, best_t, trials = tune_forms(space, measure, rt.OPT_PIPELINED, rounds=rounds, log=log)
    opts = {k: v for k, v in best.items() if v != space[k][0]}
    if opts:
        # confirm against the defaults with fresh, interleaved measurements: a
        # winner picked on the noise of single short runs is dropped
        d = [measure({}) for _ in range(2)]
        b = [measure(dict(opts)) for _ in range(2)]
        if log:
            log(f"autotune confirm: defaults {d} vs {opts} {b}")
        if min(b) >= min(d) * (1.0 - 0.005):
            opts = {}
        best_t = min(b) if opts else min(d)
    table[key] = {"options": {str(k): v for k, v in opts.items()}, "us_per_step": best_t,
                  "default_us_per_step": trials[0][1], "trials": len(trials), "seconds": time.time() - t0}
    try:  # the cache is an optimisation: a read-only home must not fail the caller
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1)
        os.replace(tmp, path)
    except OSError as e:
        if log:
            log(f"autotune: cache not written ({e})")
    return opts
