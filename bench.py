"""Benchmark of the per-timestep engine on one MI355X.

Metric: synaptic events/s and sim-s/wall-s, FP32 batched form vs FP64 plain form.
Workload (default): the Brunel-Hakim network -- N=100,000 noise-driven LIF
neurons, connect(p=1000/N) (about 1e8 synapses, float32 weights J=-0.1 mV),
delays U[0.5, 5] ms quantised to dt=0.1 ms (5..50 steps), sigma=1 mV, mu=25 mV,
Philox noise; 1 s of simulated time (10,000 steps).

  fp32 batched : FP32 state, weights and arithmetic; effects accumulated per
                 step; windows of W = min(8, min delay) steps (autotuned: one
                 pipelined launch per window, or a delivery launch and an
                 update launch); K steps per captured HIP graph
  fp64 plain   : FP64 state, weights and arithmetic; update kernel + delivery
                 kernel per step, effects applied in place (Brian2's order)

A "step" is one simulated time step of the whole network (dt = 0.1 ms).  The
network starts quiescent (V0 = Vr: about 220 steps to the first spike), so every
form first runs at least 1000 untimed steps; then every HIP graph the timed run
replays is instantiated (hs_net_prepare), and the timed region is the
configuration's whole 1 s (10,000 steps) -- more if --steps asks for more --
with inputs already in HBM: the step loop only, like Network::_last_run_time.
`value` is the FP32 batched form's synaptic events per wall-clock second over
the timed steps.  --exact-steps times exactly --steps after --warmup (profiling).

Roofline (dominant kernel of the FP32 form, the one with the largest device time
in a kernel-timed run of the same network): algorithmic bytes per launch are the
terms of the per-step byte model that the kernel touches, summed over the steps
of a launch:
  update  : N*(r_s + w_s) + n_spk*4          r_s = V + lastspike + not_ref =
            (FP32) 4+4+1, w_s = V + not_ref = 4+1 (halved FP fields of 17/9)
  delivery: n_spk*(8 + 12*b) + E*(4 + w_b)   b = delay slots per spike (46),
            w_b = 4 (FP32 per-synapse weight)
  window  : update + delivery (the pipelined form does both in one launch)
Bytes and launch time come from one run: when the dominant kernel is the timed
form's only kernel (the pipelined windows), the timed run itself -- n_spk and E
its own means, the launch time its HIP events over its launches; otherwise the
kernel-timed run (an event pair around every launch), with its own n_spk and E.
roofline.atomics sets the launch's synaptic events per second (one accumulator
atomic each) against hs_scatter_add, random FP32 atomics alone into an array of
N values.

other_configs: CUBA (4000), COBAHH (4k, 64k, 512k) and the 256-network batch at
full size, both forms, default launch shapes.

Drift: both forms are re-run with spike monitors over `--drift-steps` steps;
spikes of each neuron are matched greedily in time order (pairs within 5 ms) and
the RMS of the matched time differences is reported with the unmatched fraction,
by horizon, and the longest horizon over which it stays below one dt -- for the
Brunel-Hakim network (with controls: FP64 fused vs FP64 plain, and FP64 plain
with neuron 0's initial V moved by one FP32 ulp, one synaptic event or ten),
and for the noise-free CUBA and COBAHH networks of 4000 neurons.

Multicompartment: `--hines-trees` random branched HH trees (100-400
compartments) advanced `--hines-steps` steps, FP64 and FP32 channel arithmetic;
its roofline is VALU issue (committed SQ counters per tree-step over the
measured step time), since the trees stay on chip.

CPU baseline: the oracle (C, cpp_standalone semantics, FP64) on the same
network -- its OpenMP step loop (bit-identical to the serial one) over
min(nproc, OMP_NUM_THREADS) host threads on the GPU run's timed window, and the
serial loop on one thread over a bounded window.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_spec(workload, args):
    from brian2cuda_amd import networks as nw
    if workload == "brunel":
        return nw.brunel(N=args.N, K=args.K, seed=args.seed, monitors=False)
    if workload == "brunel_code":  # the same network with its model as generated code
        return nw.code_brunel(N=args.N, K=args.K, seed=args.seed, monitors=False)
    if workload == "cuba":
        return nw.cuba(N=args.N, seed=args.seed, monitors=False)
    if workload == "batch":
        return nw.batch_cuba(n_nets=args.nets, seed=args.seed, monitors=False)
    if workload == "cobahh":
        return nw.cobahh(N=args.N, seed=args.seed, monitors=False)
    raise ValueError(workload)


def state_bytes(workload, prec):
    """(r_s, w_s): bytes of per-neuron state read and written per step (Brian2's arrays)."""
    f = 8 if prec == 64 else 4
    if workload in ("brunel", "brunel_code"):  # V, lastspike, not_refractory
        return 2 * f + 1, f + 1
    if workload in ("cuba", "batch"):  # v, ge, gi, lastspike, not_refractory
        return 4 * f + 1, 3 * f + 1
    return 7 * f + 1, 6 * f + 1    # COBAHH: v, m, n, h, ge, gi, lastspike, not_refractory


def weight_bytes(spec, prec):
    from brian2cuda_amd.runtime import HS_W_SYNAPSE
    return (8 if prec == 64 else 4) if any(p.wmode == HS_W_SYNAPSE for p in spec.paths) else 0


def chunk_launches(steps, chunk, window):
    """Launches of the window kernel in a run of `steps` steps replayed as graph
    chunks of `chunk` steps, each chunk in windows of `window` steps."""
    if chunk <= 0:
        chunk = steps
    sizes = [chunk] * (steps // chunk) + ([steps % chunk] if steps % chunk else [])
    return sum(-(-k // window) for k in sizes)


def run_form(spec, prec, engine, steps, settle, options=None):
    """Build, settle `settle` untimed steps (the network leaves its quiescent
    start), instantiate every graph the timed run launches (hs_net_prepare), then
    time `steps` steps: the step loop only, as Network::_last_run_time."""
    from brian2cuda_amd import networks as nw
    from brian2cuda_amd import runtime as rt
    t0 = time.time()
    net, _, _ = nw.build_hs(spec, prec, engine, options=options)
    build_s = time.time() - t0
    net.run(settle)
    net.prepare(steps)
    s0 = net.stats()
    t1 = time.perf_counter()
    net.run(steps)
    wall_py = time.perf_counter() - t1
    s1 = net.stats()
    window = int(s1["window"])
    res = {
        "engine": engine,
        "precision": prec,
        "settle_steps": settle,
        "steps": steps,
        "wall_s": s1["last_run_wall_s"],
        "wall_s_python": wall_py,
        "gpu_ms": s1["last_run_gpu_ms"],
        "events": int(s1["events"] - s0["events"]),
        "spikes": int(s1["spikes"] - s0["spikes"]),
        "launches_per_step": float(s1["launches_per_step"]),
        "window": window,
        "window_launches": chunk_launches(steps, net.get_option(rt.OPT_GRAPH_CHUNK), window),
        "build_s": build_s,
    }
    res["events_per_s"] = res["events"] / res["wall_s"]
    res["sim_s_per_wall_s"] = steps * spec.dt / res["wall_s"]
    res["ms_per_step"] = 1e3 * res["wall_s"] / steps
    res["spikes_per_step"] = res["spikes"] / steps
    res["events_per_step"] = res["events"] / steps
    net.close()
    return res


def kernel_times(spec, prec, engine, steps, warmup, options=None):
    """Per-kernel device time with an event pair around every launch."""
    from brian2cuda_amd import networks as nw
    from brian2cuda_amd import runtime as rt
    net, _, _ = nw.build_hs(spec, prec, engine, options=options)
    net.run(warmup)
    net.set_option(rt.OPT_KERNEL_TIMING, 1)
    net.prepare(steps)
    s0 = net.stats()
    net.run(steps)
    s1 = net.stats()
    out = {}
    for kind, name in rt.KERNEL_NAMES.items():
        ms, n = net.kernel_time(kind)
        if n:
            out[name] = {"total_ms": ms, "launches": n, "avg_us": 1e3 * ms / n}
    window = int(s1["window"])
    spikes = s1["spikes"] - s0["spikes"]
    events = s1["events"] - s0["events"]
    net.close()
    return out, window, spikes / steps, events / steps


def spike_trains(spec, prec, engine, steps, v_at=None, options=None):
    """(i, step) of every spike over `steps` steps; with v_at, also a dict
    {h: V of every neuron after h steps} for each h in v_at (h < steps)."""
    from brian2cuda_amd import networks as nw
    sp = dataclasses.replace(spec, spike_monitors=[0])
    net, smon, _ = nw.build_hs(sp, prec, engine, spike_capacity=1 << 24, options=options)
    vs, done = {}, 0
    for h in sorted(v_at or ()):
        if 0 < h < steps:
            net.run(h - done)
            done = h
            vs[h] = net.get_state(0, 0)
    net.run(steps - done)
    if v_at is not None:
        vs[steps] = net.get_state(0, 0)
    i, t = net.spikes(smon[0])
    net.close()
    return (i, t, vs) if v_at is not None else (i, t)


def train_stats(i, t, n, steps, dt):
    """Population rate (Hz) and mean coefficient of variation of the inter-spike intervals."""
    o = np.lexsort((t, i))
    i, t = i[o], t[o]
    isi = np.diff(t).astype(np.float64)
    cvs = []
    b = np.searchsorted(i, np.arange(n + 1))
    for k in np.nonzero(np.diff(b) >= 3)[0]:
        x = isi[b[k]:b[k + 1] - 1]
        cvs.append(x.std() / x.mean())
    return {"rate_hz": len(t) / (n * steps * dt), "cv_isi": float(np.mean(cvs)) if cvs else None}


def drift(spec, steps, tol_steps=50):
    """Greedy time-ordered matching of FP32-batched to FP64-plain spikes per neuron,
    and the same for FP64 fused vs FP64 plain (same precision, effects applied one
    step later), and for FP64 plain with one initial value moved by one FP32 ulp,
    one synaptic event or ten vs FP64 plain: the last two separate the divergence that the network's own
    sensitivity produces from any change of arithmetic.  by_horizon: the first
    and last comparison over the first h steps."""
    horizons = (500, 1000, 2000, 5000)
    i64, t64, v64 = spike_trains(spec, 64, "plain", steps, v_at=horizons)
    i32, t32, v32 = spike_trains(spec, 32, "fused", steps, v_at=horizons)
    res = match_spikes(i64, t64, i32, t32, spec.n_neurons, tol_steps)
    res["steps"] = steps
    # RMS of V (FP32 batched - FP64 plain) over all neurons after h steps, in mV
    res["v_rms_mV"] = {str(h): float(np.sqrt(np.mean((v32[h] - v64[h]) ** 2))) * 1e3 for h in sorted(v64)}
    res["fp64_plain"] = train_stats(i64, t64, spec.n_neurons, steps, spec.dt)
    res["fp32_batched"] = train_stats(i32, t32, spec.n_neurons, steps, spec.dt)
    res["fp32_by_horizon"], res["rms_below_1dt_through_steps"] = drift_horizons(i64, t64, i32, t32,
                                                                                 spec.n_neurons, steps, tol_steps)
    f64, g64 = spike_trains(spec, 64, "fused", steps)
    res["fp64_fused_vs_fp64_plain"] = match_spikes(i64, t64, f64, g64, spec.n_neurons, tol_steps)
    # FP64 plain against itself with neuron 0's initial V raised by (a) one FP32
    # ulp, (b) |J| = 0.1 mV (one synaptic event) and (c) 1 mV, which moves that
    # neuron's first spike: how far the network alone carries each
    g0 = spec.groups[0]
    base = np.full(g0.n, g0.init.get(0, 0.0), dtype=np.float64)
    pert = {}
    for name, eps in (("fp32_ulp", float(np.spacing(np.float32(base[0])))), ("one_event", 1e-4),
                      ("ten_events", 1e-3)):
        v0 = base.copy()
        v0[0] += eps
        sp1 = dataclasses.replace(spec, groups=[dataclasses.replace(g0, init={**g0.init, 0: v0})]
                                  + list(spec.groups[1:]))
        u64, w64 = spike_trains(sp1, 64, "plain", steps)
        pert[name] = (u64, w64)
        r = match_spikes(i64, t64, u64, w64, spec.n_neurons, tol_steps)
        r["perturbation"] = {"neuron": 0, "var": "V", "delta": eps}
        res[f"fp64_plain_{name}_vs_fp64_plain"] = r
    # the comparisons over the first h steps only
    res["by_horizon"] = {}
    for h in horizons:
        if h >= steps:
            continue
        a, b = t64 < h, t32 < h
        r32 = match_spikes(i64[a], t64[a], i32[b], t32[b], spec.n_neurons, tol_steps)
        row = {"fp32_batched": {"rms_dt_steps": r32["rms_dt_steps"], "unmatched_frac": r32["unmatched_frac"]}}
        for name, (u64, w64) in pert.items():
            c = w64 < h
            r1 = match_spikes(i64[a], t64[a], u64[c], w64[c], spec.n_neurons, tol_steps)
            row[f"fp64_plain_{name}"] = {"rms_dt_steps": r1["rms_dt_steps"], "unmatched_frac": r1["unmatched_frac"]}
        res["by_horizon"][str(h)] = row
    return res


def atomic_ceiling(n_target, ev_launch, launch_us, seed):
    """The dominant kernel's synaptic events per second (one accumulator atomic
    each) against random FP32 atomics alone (hs_scatter_add) into an array of the
    same size, both timed with an event pair around the launch: at the launch's
    own event count (one launch, its fixed cost included) and at 64 times that
    count (the sustained rate)."""
    from brian2cuda_amd import runtime as rt
    rng = np.random.default_rng(seed)
    out = {"events_per_launch": int(ev_launch), "achieved_per_s": ev_launch / (launch_us * 1e-6)}
    for key, n in (("same_count", int(ev_launch)), ("sustained", 64 * int(ev_launch))):
        idx = rt.DeviceArray.from_numpy(rng.integers(0, n_target, max(n, 1)).astype(np.int32))
        acc = rt.DeviceArray(n_target, np.float32)
        acc.zero()
        rt.scatter_add(idx, 1e-4, acc)
        ms = min(rt.scatter_add(idx, 1e-4, acc) for _ in range(5))
        out[f"scatter_{key}_us"] = round(ms * 1e3, 3)
        out[f"scatter_{key}_per_s"] = n / (ms * 1e-3)
        idx.free()
        acc.free()
    out["frac_of_sustained"] = round(out["achieved_per_s"] / out["scatter_sustained_per_s"], 4)
    return out


def match_spikes(i64, t64, i32, t32, n, tol_steps):
    """Per neuron, walk both time-sorted trains: a pair within tol_steps matches,
    otherwise the earlier spike is unmatched."""
    o64 = np.lexsort((t64, i64))
    o32 = np.lexsort((t32, i32))
    i64, t64, i32, t32 = i64[o64], t64[o64], i32[o32], t32[o32]
    b64 = np.searchsorted(i64, np.arange(n + 1))
    b32 = np.searchsorted(i32, np.arange(n + 1))
    d = []
    unmatched = 0
    for k in range(n):
        a = t64[b64[k]:b64[k + 1]]
        b = t32[b32[k]:b32[k + 1]]
        p = q = 0
        while p < len(a) and q < len(b):
            if abs(int(a[p]) - int(b[q])) <= tol_steps:
                d.append(int(b[q]) - int(a[p]))
                p += 1
                q += 1
            elif a[p] < b[q]:
                p += 1
                unmatched += 1
            else:
                q += 1
                unmatched += 1
        unmatched += (len(a) - p) + (len(b) - q)
    d = np.asarray(d, dtype=np.float64)
    total = len(t64) + len(t32)
    return {"spikes_fp64": int(len(t64)), "spikes_fp32": int(len(t32)),
            "matched_pairs": int(d.size), "rms_dt_steps": float(np.sqrt(np.mean(d * d))) if d.size else None,
            "unmatched_frac": unmatched / total if total else 0.0, "match_window_steps": tol_steps}


DRIFT_GRID = (100, 200, 300, 500, 700, 1000, 1500, 2000, 3000, 5000, 7000, 10000)


def drift_horizons(i64, t64, i32, t32, n, steps, tol_steps=50):
    """The FP32-vs-FP64 matching over the first h steps for h on a grid, and the
    longest prefix of the grid over which the RMS spike-time drift stays below
    one step (dt)."""
    rows, below, ok = {}, 0, True
    for h in DRIFT_GRID:
        if h > steps:
            break
        a, b = t64 < h, t32 < h
        r = match_spikes(i64[a], t64[a], i32[b], t32[b], n, tol_steps)
        rows[str(h)] = {"rms_dt_steps": r["rms_dt_steps"], "unmatched_frac": r["unmatched_frac"],
                        "spikes_fp64": r["spikes_fp64"]}
        ok = ok and (r["rms_dt_steps"] or 0.0) < 1.0
        if ok:
            below = h
    return rows, below


def drift_config(spec, steps, tol_steps=50, options=None):
    """FP32 batched (with `options`, e.g. the fast HH form) against FP64 plain
    for a configuration without noise: spike matching over the whole run and by
    horizon, the V RMS after 100 steps and after the run, and each form's rate
    and CV."""
    i64, t64, v64 = spike_trains(spec, 64, "plain", steps, v_at=(100,))
    i32, t32, v32 = spike_trains(spec, 32, "fused", steps, v_at=(100,), options=options)
    res = match_spikes(i64, t64, i32, t32, spec.n_neurons, tol_steps)
    res["steps"] = steps
    if options:
        res["fp32_options"] = {str(k): v for k, v in options.items()}
    res["v_rms_mV_after_100_steps"] = float(np.sqrt(np.nanmean((v32[100] - v64[100]) ** 2))) * 1e3
    res["v_rms_mV_at_end"] = float(np.sqrt(np.nanmean((v32[steps] - v64[steps]) ** 2))) * 1e3
    res["by_horizon"], res["rms_below_1dt_through_steps"] = drift_horizons(i64, t64, i32, t32, spec.n_neurons,
                                                                            steps, tol_steps)
    res["fp64_plain"] = train_stats(i64, t64, spec.n_neurons, steps, spec.dt)
    res["fp32_batched"] = train_stats(i32, t32, spec.n_neurons, steps, spec.dt)
    return res


HINES_BYTES_PER_COMP_STEP = 68  # a, d, v, rhs r (32) + v w (8) + parent index (4) + m, h, n r/w (24)


def hines_bench(n_trees, steps, seed):
    """Multicompartment batch (random branched HH trees, 100-400 compartments,
    dt = 0.025 ms): device time per step for FP64 and FP32 channel arithmetic."""
    from brian2cuda_amd import morphology as mo
    tb = mo.random_trees(n_trees=n_trees, seed=seed)
    out = {"trees": n_trees, "compartments": tb.n_comp, "sections": tb.n_sections, "steps": steps}
    for prec in (64, 32):
        h = mo.build_hs(tb, prec)
        h.run(min(200, steps))
        h.run(steps)
        s = h.stats()
        step_s = s["last_run_gpu_ms"] * 1e-3 / steps
        out[f"fp{prec}"] = {"us_per_step": round(step_s * 1e6, 3),
                            "compartment_steps_per_s": tb.n_comp / step_s,
                            "sim_s_per_wall_s": steps * tb.dt / s["last_run_wall_s"],
                            "soma_spikes": int(h.spike_counts().sum())}
        h.close()
    out["roofline_fp64"] = hines_valu_roofline(out, "f64", n_trees, seed)
    out["roofline_fp32"] = hines_valu_roofline(out, "f32", n_trees, seed)
    return out


# MI355X: 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock; a wave64 VALU
# instruction (FP64, or FP32 unpacked) occupies its SIMD for 4 cycles
SIMDS, CLOCK_HZ, VALU_CYCLES = 1024, 2.4e9, 4
VALU_PEAK_GINSTR = SIMDS * CLOCK_HZ / VALU_CYCLES / 1e9


def hines_valu_roofline(out, prec, n_trees, seed):
    """hs_k_hines is bound by instruction issue, not by HBM (the trees stay on
    chip: measured traffic is under 1 B per compartment-step).  achieved = wave
    VALU instructions issued per second: the committed SQ_INSTS_VALU per
    tree-step (rocprofv3, profiles/hines_c5_counters.json, the same 2048 trees)
    x trees / this run's step time; peak = what the SIMDs can issue.  None for
    another batch.  The HBM figure (68 B per compartment-step, the algorithmic
    byte model) is kept beside it for reference."""
    key = "fp64" if prec == "f64" else "fp32"
    step_s = out[key]["us_per_step"] * 1e-6
    hbm = HINES_BYTES_PER_COMP_STEP * out["compartments"] / step_s / 1e9
    base = {"kernel": "hs_k_hines", "precision_channels": key, "hbm_model_gbs": round(hbm, 3),
            "hbm_model_frac": round(hbm / HBM_PEAK_GBS, 6), "bytes_per_compartment_step_model": HINES_BYTES_PER_COMP_STEP,
            "traffic": hines_traffic(n_trees, seed)}
    path = os.path.join(ROOT, "profiles", "hines_c5_counters.json")
    if n_trees != 2048 or seed != 1234 or not os.path.exists(path):
        return {**base, "bound": "valu", "achieved": None, "peak": round(VALU_PEAK_GINSTR, 3),
                "unit": "G wave-instr/s", "frac": None}
    with open(path) as f:
        c = json.load(f)
    k = c["kernels"][f"hs_k_hines_{prec}"]
    per_tree_step = k["SQ_INSTS_VALU"] / k["SQ_WAVES"] / (c["steps_counted"] / 2)
    ach = per_tree_step * n_trees / step_s / 1e9
    return {**base, "bound": "valu", "achieved": round(ach, 3), "peak": round(VALU_PEAK_GINSTR, 3),
            "unit": "G wave-instr/s", "frac": round(ach / VALU_PEAK_GINSTR, 4),
            "valu_instr_per_tree_step": round(per_tree_step, 1),
            "valu_active_frac_per_wave": round(k["SQ_ACTIVE_INST_VALU"] / k["SQ_WAVE_CYCLES"], 4),
            "wait_frac_per_wave": round(k["SQ_WAIT_INST_ANY"] / k["SQ_WAVE_CYCLES"], 4)}


def hines_traffic(n_trees, seed):
    """Measured HBM bytes of one hs_k_hines launch (FP64 and FP32, rocprofv3
    FETCH_SIZE / WRITE_SIZE scaled by a copy of known size) for the batch they
    were measured on (profiles/hines_c5_traffic.json); None for another batch.
    A launch reads the trees once and writes them back once, whatever its step
    count."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hines_c5_traffic.json")
    if n_trees != 2048 or seed != 1234 or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    out = {}
    for prec in ("f64", "f32"):
        k = t["kernels"][f"hs_k_hines_{prec}"]
        out[f"{prec}_bytes_per_launch"] = int(round(k["fetch_size_kib_per_dispatch"] * 1024 * t["fetch_scale"]
                                                  + k["write_size_kib_per_dispatch"] * 1024 * t["write_scale"]))
        out[f"{prec}_steps_per_launch_measured"] = k["steps_per_dispatch"]
    return out


def measured_traffic(args, kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, each scaled by what the counter
    reads for a plain copy of known size), for the workload it was measured on;
    None otherwise."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "brunel_c3_traffic.json")
    if args.workload != "brunel" or args.N != 100_000 or args.K != 1000 or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t["kernels"].get(kernel)
    if k is None:
        return None
    return int(1024 * (k["fetch_size_kib_per_dispatch"] * t["fetch_scale"] +
                       k["write_size_kib_per_dispatch"] * t["write_scale"]))


def connect_bench(args, repeats=3):
    """Device build of the same connectivity (connect(p=K/N), delays U[0.5, 5] ms,
    per-synapse FP32 weights): wall time of hs_net_connect_random (count pass,
    prefix sum, write pass, two scalar readbacks), best of `repeats` fresh nets."""
    from brian2cuda_amd import networks as nw
    from brian2cuda_amd import runtime as rt
    spec = nw.brunel(N=args.N, K=args.K, seed=args.seed, device_connect=True)
    r, g = spec.paths[0].random, spec.groups[0]
    best, nsyn, nslots = None, 0, 0
    for _ in range(repeats):
        net = rt.Net(32, "fused", spec.dt, spec.seed)
        net.add_group(g.model, g.n, g.params)
        t0 = time.perf_counter()
        pid = net.connect_random(0, 0, 0, r.p, r.pre, r.post, r.autapses, r.seed, r.stream, r.delay, r.wmode, r.w)
        dt_s = time.perf_counter() - t0
        pi = net.pathway_info(pid)
        nsyn, nslots = pi["nsyn"], pi["D"]
        net.close()
        best = dt_s if best is None else min(best, dt_s)
    return {"device_s": round(best, 5), "synapses": int(nsyn), "synapses_per_s": round(nsyn / best, 1),
            "bytes_written": int(nsyn * 8 + 4 * (args.N * nslots + 1))}


def cpu_info():
    """host CPU model, nproc (CPUs this process may run on) and OMP_NUM_THREADS"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    return {"model": model, "nproc": nproc, "omp_num_threads": int(omp) if omp and omp.isdigit() else None}


def cpu_baseline(spec, settle, steps, threads, serial_skip, serial_steps):
    """The oracle (cpp_standalone semantics, FP64, SEQ order) on the same network:
    its OpenMP step loop over `threads` threads on the timed window of the GPU
    run (settle, then `steps` steps), and the serial loop (one thread, the
    cpp_standalone default) on a bounded window.  Both give the same bits."""
    import oracle
    out = {}
    if threads > 1:
        t0 = time.time()
        ref, _, _ = oracle.build_ref(spec, 64, oracle.SEQ, threads=threads)
        build_s = time.time() - t0
        ref.run(settle)
        e0 = ref.events
        t1 = time.perf_counter()
        ref.run(steps)
        wall = time.perf_counter() - t1
        ev = ref.events - e0
        ref.close()
        out["threaded"] = {"threads": threads, "value": ev / wall, "wall_s": wall, "events": int(ev), "steps": steps,
                           "skip": settle, "sim_s_per_wall_s": steps * spec.dt / wall, "build_s": build_s}
    ref, _, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    ref.run(serial_skip)
    e0 = ref.events
    t1 = time.perf_counter()
    ref.run(serial_steps)
    wall = time.perf_counter() - t1
    ev = ref.events - e0
    ref.close()
    out["serial"] = {"threads": 1, "value": ev / wall, "wall_s": wall, "events": int(ev), "steps": serial_steps,
                     "skip": serial_skip, "sim_s_per_wall_s": serial_steps * spec.dt / wall}
    return out


def other_spec(name, seed):
    """the other configurations at full size (no monitors)"""
    from brian2cuda_amd import networks as nw
    if name == "cuba_4k":
        return nw.cuba(N=4000, seed=seed, monitors=False)
    if name.startswith("cobahh_"):
        return nw.cobahh(N={"cobahh_4k": 4000, "cobahh_64k": 64000, "cobahh_512k": 512000}[name], seed=seed,
                         monitors=False)
    if name == "batch_256_cuba":
        return nw.batch_cuba(n_nets=256, seed=seed, monitors=False)
    raise ValueError(name)


OTHER_CONFIGS = ("cuba_4k", "cobahh_4k", "cobahh_64k", "cobahh_512k", "batch_256_cuba")


def other_configs(args):
    """The other configurations at their full sizes: FP32 batched and FP64 plain,
    `--other-steps` timed steps after `--other-settle` (default launch shapes),
    with the per-step byte model's HBM fraction of the FP32 form's step."""
    out = {}
    for name in OTHER_CONFIGS:
        t0 = time.time()
        spec = other_spec(name, args.seed)
        gen_s = time.time() - t0
        row = {"neurons": spec.n_neurons, "synapses": spec.n_synapses, "spec_s": round(gen_s, 2)}
        rnd = lambda r: {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()}
        # FP32 batched: the faster of the two exact forms (one launch per step
        # with update and delivery blocks, or the pipelined form whose update
        # blocks push their spikes' synapses); COBAHH also in the fast form
        forms = {"default": {}, "pipelined": {9: 1}}
        if name.startswith("cobahh"):
            forms.update({"fast": {11: 1}, "fast_pipelined": {9: 1, 11: 1}})
        res = {fn: rnd(run_form(spec, 32, "fused", args.other_steps, args.other_settle, options=o))
               for fn, o in forms.items()}
        row["fp32_forms_ms_per_step"] = {fn: r["ms_per_step"] for fn, r in res.items()}
        exact = min(("default", "pipelined"), key=lambda fn: res[fn]["ms_per_step"])
        row["fp32_batched"] = dict(res[exact], options={str(k): v for k, v in forms[exact].items()})
        row["fp64_plain"] = rnd(run_form(spec, 64, "plain", args.other_steps, args.other_settle))
        f, p = row["fp32_batched"], row["fp64_plain"]
        row["speedup_fp32_batched_vs_fp64_plain"] = round(p["ms_per_step"] / f["ms_per_step"], 4)
        if "fast" in res:
            fast = min(("fast", "fast_pipelined"), key=lambda fn: res[fn]["ms_per_step"])
            row["fp32_fast"] = dict(res[fast], options={str(k): v for k, v in forms[fast].items()})
            row["speedup_fp32_fast_vs_fp64_plain"] = round(p["ms_per_step"] / res[fast]["ms_per_step"], 4)
        wl = "batch" if name.startswith("batch") else ("cuba" if name.startswith("cuba") else "cobahh")
        r_s, w_s = state_bytes(wl, 32)
        bstep = spec.n_neurons * (r_s + w_s) + f["spikes_per_step"] * 12 + f["events_per_step"] * 4
        row["fp32_step_bytes_model"] = int(bstep)
        row["fp32_step_hbm_frac"] = round(bstep / (f["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
        if name in ("cuba_4k", "cobahh_512k") and not args.no_code:
            # the same network with its model as generated code (compiled at run
            # time into the engine kernels), in the FP32 form chosen above
            cspec = code_spec(name, args.seed)
            c = rnd(run_form(cspec, 32, "fused", args.other_steps, args.other_settle, options=forms[exact]))
            row["fp32_code_model"] = c
            row["code_model_vs_builtin_ms_per_step"] = round(c["ms_per_step"] / f["ms_per_step"], 4)
        out[name] = row
        log(name, row)
    if not args.no_code:
        # C3 with its model as generated code, in the headline's FP32 pipelined form
        cspec = code_spec("brunel_100k", args.seed)
        opts = {9: 1, 0: 128, 1: 512}
        c = rnd(run_form(cspec, 32, "fused", args.other_steps, args.other_settle, options=opts))
        b = rnd(run_form(other_spec_brunel(args.seed), 32, "fused", args.other_steps, args.other_settle,
                         options=opts))
        out["brunel_100k_code_model"] = {"options": {str(k): v for k, v in opts.items()}, "fp32_code_model": c,
                                         "fp32_builtin": b,
                                         "code_model_vs_builtin_ms_per_step": round(c["ms_per_step"] /
                                                                                    b["ms_per_step"], 4)}
        log("brunel_100k_code_model", out["brunel_100k_code_model"])
    return out


def code_spec(name, seed):
    """the code-model form of a configuration (brian2cuda_amd.networks.code_*)"""
    from brian2cuda_amd import networks as nw
    if name == "cuba_4k":
        return nw.code_cuba(N=4000, seed=seed, monitors=False)
    if name == "cobahh_512k":
        return nw.code_cobahh(N=512000, seed=seed, monitors=False)
    if name == "brunel_100k":
        return nw.code_brunel(N=100_000, K=1000, seed=seed)
    raise ValueError(name)


def other_spec_brunel(seed):
    from brian2cuda_amd import networks as nw
    return nw.brunel(N=100_000, K=1000, seed=seed)


CONFIG_STEPS = 10000  # every configuration simulates 1 s at dt = 0.1 ms
SETTLE_MIN = 1000     # the networks start quiescent (C3: V0 = Vr, about 220 steps to the first spike)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=CONFIG_STEPS,
                    help="timed simulation steps requested (dt=0.1 ms); fewer than the configuration's "
                         f"{CONFIG_STEPS} (1 s) time the {CONFIG_STEPS} anyway")
    ap.add_argument("--warmup", type=int, default=SETTLE_MIN,
                    help=f"untimed steps before timing requested; fewer than {SETTLE_MIN} settle {SETTLE_MIN}")
    ap.add_argument("--workload", default="brunel", choices=["brunel", "brunel_code", "cuba", "batch", "cobahh"])
    ap.add_argument("--N", type=int, default=None)
    ap.add_argument("--K", type=int, default=1000)
    ap.add_argument("--nets", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--exact-steps", action="store_true",
                    help="time exactly --steps after exactly --warmup steps (profiling runs)")
    ap.add_argument("--cpu-steps", type=int, default=2000,
                    help="serial oracle timed steps (0: skip every CPU leg); 2000 is about 10 s of one core")
    ap.add_argument("--cpu-skip", type=int, default=300, help="serial oracle untimed steps before its window")
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="threads of the OpenMP oracle run (default: min(nproc, OMP_NUM_THREADS))")
    ap.add_argument("--kt-steps", type=int, default=2000, help="steps of the kernel-timed run (0: skip)")
    ap.add_argument("--drift-steps", type=int, default=CONFIG_STEPS, help="steps of the drift comparisons (0: skip)")
    ap.add_argument("--hines-trees", type=int, default=2048, help="trees of the multicompartment batch")
    ap.add_argument("--no-code", action="store_true",
                    help="skip the code-model forms of the other configurations")
    ap.add_argument("--hines-steps", type=int, default=2000, help="steps of the multicompartment run (0: skip)")
    ap.add_argument("--other-steps", type=int, default=3000, help="timed steps of the other configurations (0: skip)")
    ap.add_argument("--other-settle", type=int, default=500, help="untimed steps of the other configurations")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-connect", action="store_true", help="skip the device connectivity-build timing")
    ap.add_argument("--options", default=None,
                    help='JSON {"key": value} of engine options for the FP32 batched form instead of autotuning '
                         '(e.g. what a previous run reported as tuned_options.fp32_batched)')
    ap.add_argument("--fp64-options", default=None, help="the same for the FP64 plain form")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time both forms with the engine's default launch shapes instead of autotuned ones")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 256 MiB device-to-device copies (the reference for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 3 x 256 MiB")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the engine has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.N is None:
        args.N = {"brunel": 100_000, "brunel_code": 100_000, "cuba": 4000, "cobahh": 4000, "batch": 0}[args.workload]
    steps = args.steps if args.exact_steps else max(args.steps, CONFIG_STEPS)
    settle = args.warmup if args.exact_steps else max(args.warmup, SETTLE_MIN)
    log(f"timing {steps} steps after {settle} settle steps (asked: --steps {args.steps} --warmup {args.warmup})")

    import brian2cuda_amd
    brian2cuda_amd.init(0)
    info = brian2cuda_amd.device_info()
    log("device", info)
    t0 = time.time()
    spec = make_spec(args.workload, args)
    spec_s = time.time() - t0
    log(f"built {spec.name}: N={spec.n_neurons} S={spec.n_synapses} in {spec_s:.1f}s")
    D = max((p.delay_steps.max() - p.delay_steps.min() + 1) if p.delay_steps is not None else 1 for p in spec.paths)

    tuned = {"fp32_batched": {}, "fp64_plain": {}}
    if args.options or args.fp64_options:
        if args.options:
            tuned["fp32_batched"] = {int(k): v for k, v in json.loads(args.options).items()}
        if args.fp64_options:
            tuned["fp64_plain"] = {int(k): v for k, v in json.loads(args.fp64_options).items()}
    elif not args.no_autotune:
        from brian2cuda_amd import autotune as at
        from brian2cuda_amd import networks as nw

        def build(sp, prec, engine, options=None):  # same tables as the host build, bundled on the device
            return nw.build_hs(sp, prec, engine, options=options, device_sort=True)

        t0 = time.time()
        tuned["fp32_batched"] = at.autotune(spec, 32, "fused", build=build, log=log)
        if not args.no_fp64:
            tuned["fp64_plain"] = at.autotune(spec, 64, "plain", build=build, log=log)
        log(f"autotuned in {time.time() - t0:.1f}s: {tuned}")
    fp32 = run_form(spec, 32, "fused", steps, settle, options=tuned["fp32_batched"])
    log("fp32 batched", fp32)
    fp64 = None
    if not args.no_fp64:
        fp64 = run_form(spec, 64, "plain", steps, settle, options=tuned["fp64_plain"])
        log("fp64 plain", fp64)

    N = spec.n_neurons
    roofline = None
    if args.kt_steps > 0:
        kt, W, kn_spk, kE = kernel_times(spec, 32, "fused", args.kt_steps, settle, options=tuned["fp32_batched"])
        log("kernel times (fp32 batched)", kt)
        r_s, w_s = state_bytes(args.workload, 32)
        w_b = weight_bytes(spec, 32)

        def per_step(name, n_spk, E):
            return {
                "hs_k_update": N * (r_s + w_s) + n_spk * 4,
                "hs_k_deliver": n_spk * (8 + 12 * D) + E * (4 + w_b),
                "hs_k_gather": n_spk * (8 + 12 * D) + E * (4 + w_b),
                "hs_k_fused": N * (r_s + w_s) + n_spk * (12 + 12 * D) + E * (4 + w_b),
                "hs_k_window": N * (r_s + w_s) + n_spk * (12 + 12 * D) + E * (4 + w_b),
            }[name]

        name = max(kt, key=lambda k: kt[k]["total_ms"])
        if len(kt) == 1 and name == "hs_k_window":
            # the dominant kernel is the timed form's only kernel: its bytes and
            # its launch time both from the timed run itself (HIP events around
            # its steps on the runtime's stream, over its launches)
            launches = fp32["window_launches"]
            steps_per_launch = steps / launches
            bpl = per_step(name, fp32["spikes_per_step"], fp32["events_per_step"]) * steps_per_launch
            avg_us = fp32["gpu_ms"] * 1e3 / launches
            src = "timed run: its HIP events over its window launches; bytes from its own spike and event counts"
            ev_launch = fp32["events_per_step"] * steps_per_launch
        else:
            # several kernels per step: the kernel-timed run (an event pair per
            # launch), bytes from that run's own counts
            steps_per_launch = args.kt_steps / kt[name]["launches"]
            bpl = per_step(name, kn_spk, kE) * steps_per_launch
            avg_us = kt[name]["avg_us"]
            src = "kernel-timed run: an event pair per launch; bytes from its own spike and event counts"
            ev_launch = kE * steps_per_launch
        achieved = bpl / (avg_us * 1e-6) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(args, name),
                    "kernel": name, "bytes_per_launch": int(bpl), "steps_per_launch": round(steps_per_launch, 3),
                    "avg_launch_us": round(avg_us, 3), "avg_launch_from": src,
                    "avg_launch_us_event_pairs": round(kt[name]["avg_us"], 3),
                    "kernels": {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in kt.items()}}
        roofline["atomics"] = atomic_ceiling(N, ev_launch, avg_us, args.seed)
        log("atomics", roofline["atomics"])

    dr = None
    if args.drift_steps > 0 and not args.no_fp64 and args.workload == "brunel":
        dr = {"brunel_hakim_c3": drift(spec, args.drift_steps)}
        for name in ("cuba_4k", "cobahh_4k"):
            dr[name] = drift_config(other_spec(name, args.seed), args.drift_steps)
        # the FP32 fast form of the HH update (option 11): its error against FP64
        dr["cobahh_4k_fp32_fast"] = drift_config(other_spec("cobahh_4k", args.seed), args.drift_steps,
                                                 options={9: 1, 11: 1})
        log("drift", dr)

    others = None
    if args.other_steps > 0 and not args.no_fp64 and args.workload == "brunel":
        others = other_configs(args)

    conn = None
    if args.workload == "brunel" and not args.no_connect:
        conn = connect_bench(args)
        conn["host_numpy_generate_s"] = round(spec_s, 4)
        conn["host_bundle_upload_s"] = round(fp32["build_s"], 4)
        log("connectivity build", conn)

    hines = None
    if args.hines_steps > 0:
        hines = hines_bench(args.hines_trees, args.hines_steps, args.seed)
        log("hines", hines)

    cpu = None
    if args.cpu_steps > 0:
        ci = cpu_info()
        threads = args.cpu_threads or min(ci["nproc"], ci["omp_num_threads"] or ci["nproc"])
        cb = cpu_baseline(spec, settle, steps, threads, args.cpu_skip, args.cpu_steps)
        log("cpu baseline", cb)
        main_leg = cb.get("threaded", cb["serial"])
        ser = cb["serial"]
        cpu = {"value": round(main_leg["value"], 1), "unit": "synaptic events/s", "cores": main_leg["threads"],
               "kind": "port",
               "sample": f"oracle (C, FP64, cpp_standalone semantics and order) on the same network: OpenMP step "
                         f"loop over {main_leg['threads']} threads, steps {main_leg['skip']}.."
                         f"{main_leg['skip'] + main_leg['steps']} (the GPU run's timed window; "
                         f"{main_leg['events']} events in {main_leg['wall_s']:.2f} s, "
                         f"{main_leg['sim_s_per_wall_s']:.4f} sim-s/wall-s); serial loop on one thread, steps "
                         f"{ser['skip']}..{ser['skip'] + ser['steps']} ({ser['value']:.4g} events/s, "
                         f"{ser['sim_s_per_wall_s']:.4f} sim-s/wall-s)",
               "serial_1_thread": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in ser.items()},
               "cpu": ci}

    out = {
        "metric": "synaptic events/s and sim-s/wall-s: FP32 batched vs FP64 plain, 1 MI355X",
        "value": round(fp32["events_per_s"], 1),
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "steps_timed": steps,
        "settle_steps_run": settle,
        "step_accounting": (f"a step is one 0.1 ms simulation step of the whole network; the timed region is "
                            f"{steps} steps ({steps * spec.dt:g} s simulated, the configuration's duration) after "
                            f"{settle} untimed steps, whatever smaller --steps/--warmup ask for "
                            f"(--exact-steps times exactly what is asked)"),
        "ms_per_step": round(fp32["ms_per_step"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator; inputs resident in HBM before timing)",
        "config": {"workload": {"brunel": "brunel_hakim_heterog_delays", "cuba": "cuba", "batch": "batch_cuba",
                                "cobahh": "cobahh",
                                "brunel_code": "brunel_hakim_heterog_delays_code_model"}[args.workload],
                   "neurons": N, "synapses": spec.n_synapses, "dt_ms": spec.dt * 1e3,
                   "sim_s": steps * spec.dt, "delay_slots": int(D), "seed": args.seed},
        "sim_s_per_wall_s": round(fp32["sim_s_per_wall_s"], 4),
        "fp32_batched": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in fp32.items()},
        "fp64_plain": ({k: (round(v, 6) if isinstance(v, float) else v) for k, v in fp64.items()} if fp64 else None),
        "speedup_fp32_batched_vs_fp64_plain": (round(fp32["events_per_s"] / fp64["events_per_s"], 4)
                                               if fp64 and fp64["events_per_s"] > 0 else None),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "other_configs": others,
        "drift_fp32_batched_vs_fp64_plain": dr,
        "tuned_options": {k: {str(kk): vv for kk, vv in v.items()} for k, v in tuned.items()},
        "multicompartment": hines,
        "connectivity_build": conn,
        "device": info,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
