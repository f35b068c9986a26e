"""Benchmark of the per-timestep engine on one MI355X.

Metric: synaptic events/s and sim-s/wall-s, FP32 batched form vs FP64 plain form.
Workload (default): the Brunel-Hakim network -- N=100,000 noise-driven LIF
neurons, connect(p=1000/N) (about 1e8 synapses, float32 weights J=-0.1 mV),
delays U[0.5, 5] ms quantised to dt=0.1 ms (5..50 steps), sigma=1 mV, mu=25 mV,
Philox noise; 1 s of simulated time (10,000 steps) after a warm-up.

  fp32 batched : FP32 state, weights and arithmetic; one launch per step (update
                 of step t and delivery of step t in one kernel), K steps per
                 captured HIP graph
  fp64 plain   : FP64 state, weights and arithmetic; update kernel + delivery
                 kernel per step, effects applied in place (Brian2's order)

A "step" is one simulated time step of the whole network (dt = 0.1 ms).  Inputs
are built on the host from a seed and copied to HBM before timing; the timed
region is the step loop only (like Network::_last_run_time).  `value` is the
FP32 batched form's synaptic events per wall-clock second over the K timed steps.

Roofline: the dominant kernel is the fused step kernel of the FP32 form.  Its
algorithmic bytes per launch (one step) are
   N*20 + n_spk*4 + n_spk*D*12 + E*8
(state: V r/w 4+4, last-spike r 4, accumulator r/w 4+4; spike list 4 per spike;
per (spike, delay) bundle: spike index 4 + two bundle offsets 8; per event: post
index 4 + weight 4), with n_spk and E the measured per-step means.  The launch
duration is the device time of the timed chunks (HIP events on the runtime's
stream) divided by the number of launches.

CPU baseline: the oracle (serial C, cpp_standalone semantics, FP64) on the same
network, timed over a bounded window of steps on one host core.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_step(workload, N, n_spk, E, D, prec, engine):
    """Algorithmic HBM bytes of one step of the headline kernel(s)."""
    r = 4 if prec == 32 else 8
    if workload == "brunel":
        state = N * (r + 4 + r)          # V read, last read, V write
        if engine == "fused":
            state += N * 2 * r           # accumulator read + zero
        wbytes = r                       # per-synapse weight
    else:  # CUBA-type (cuba, batch): v, ge, gi r/w
        state = N * (3 * r + 4 + 3 * r)
        if engine == "fused":
            state += N * 2 * 2 * r       # two accumulators (ge, gi)
        wbytes = 0
    return state + n_spk * 4 + n_spk * D * 12 + E * (4 + wbytes)


def make_spec(workload, args):
    from brian2cuda_amd import networks as nw
    if workload == "brunel":
        return nw.brunel(N=args.N, K=args.K, seed=args.seed, monitors=False)
    if workload == "cuba":
        return nw.cuba(N=args.N, seed=args.seed, monitors=False)
    if workload == "batch":
        return nw.batch_cuba(n_nets=args.nets, seed=args.seed, monitors=False)
    if workload == "cobahh":
        return nw.cobahh(N=args.N, seed=args.seed, monitors=False)
    raise ValueError(workload)


def run_form(spec, prec, engine, steps, warmup, options=None):
    from brian2cuda_amd import networks as nw
    t0 = time.time()
    net, _, _ = nw.build_hs(spec, prec, engine, options=options)
    build_s = time.time() - t0
    net.run(warmup)
    s0 = net.stats()
    t1 = time.perf_counter()
    net.run(steps)
    wall_py = time.perf_counter() - t1
    s1 = net.stats()
    res = {
        "engine": engine,
        "precision": prec,
        "steps": steps,
        "wall_s": s1["last_run_wall_s"],
        "wall_s_python": wall_py,
        "gpu_ms": s1["last_run_gpu_ms"],
        "events": int(s1["events"] - s0["events"]),
        "spikes": int(s1["spikes"] - s0["spikes"]),
        "launches_per_step": float(s1["launches_per_step"]),
        "window": int(s1["window"]),
        "build_s": build_s,
    }
    res["events_per_s"] = res["events"] / res["wall_s"]
    res["sim_s_per_wall_s"] = steps * spec.dt / res["wall_s"]
    res["ms_per_step"] = 1e3 * res["wall_s"] / steps
    net.close()
    return res


def cpu_baseline(spec, skip_steps, steps):
    """Oracle (cpp_standalone semantics, FP64, 1 core) on a window of steps."""
    import oracle
    t0 = time.time()
    ref, _, _ = oracle.build_ref(spec, 64, oracle.SEQ)
    build_s = time.time() - t0
    ref.run(skip_steps)
    e0 = ref.events
    t1 = time.perf_counter()
    ref.run(steps)
    wall = time.perf_counter() - t1
    ev = ref.events - e0
    ref.close()
    return {"value": ev / wall, "wall_s": wall, "events": int(ev), "steps": steps, "skip": skip_steps,
            "sim_s_per_wall_s": steps * spec.dt / wall, "build_s": build_s}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000, help="timed simulation steps (dt=0.1 ms)")
    ap.add_argument("--warmup", type=int, default=1000, help="untimed steps before timing")
    ap.add_argument("--workload", default="brunel", choices=["brunel", "cuba", "batch", "cobahh"])
    ap.add_argument("--N", type=int, default=None)
    ap.add_argument("--K", type=int, default=1000)
    ap.add_argument("--nets", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--cpu-steps", type=int, default=300, help="oracle timed steps (0: skip)")
    ap.add_argument("--cpu-skip", type=int, default=300, help="oracle untimed steps before its window")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the engine has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.N is None:
        args.N = {"brunel": 100_000, "cuba": 4000, "cobahh": 4000, "batch": 0}[args.workload]

    import brian2cuda_amd
    brian2cuda_amd.init(0)
    info = brian2cuda_amd.device_info()
    log("device", info)
    t0 = time.time()
    spec = make_spec(args.workload, args)
    log(f"built {spec.name}: N={spec.n_neurons} S={spec.n_synapses} in {time.time() - t0:.1f}s")
    D = max((p.delay_steps.max() - p.delay_steps.min() + 1) if p.delay_steps is not None else 1 for p in spec.paths)

    fp32 = run_form(spec, 32, "fused", args.steps, args.warmup)
    log("fp32 batched", fp32)
    fp64 = None
    if not args.no_fp64:
        fp64 = run_form(spec, 64, "plain", args.steps, args.warmup)
        log("fp64 plain", fp64)

    N = spec.n_neurons
    n_spk = fp32["spikes"] / args.steps
    E = fp32["events"] / args.steps
    bpl = bytes_per_step(args.workload, N, n_spk, E, D, 32, "fused")
    # device time per simulated step over the timed region (all kernels of the step)
    step_s = fp32["gpu_ms"] / 1e3 / args.steps
    achieved = bpl / step_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                "kernel": "fused step kernels (FP32 form)", "bytes_per_step": int(bpl),
                "step_us": round(step_s * 1e6, 3)}

    cpu = None
    if args.cpu_steps > 0:
        cb = cpu_baseline(spec, args.cpu_skip, args.cpu_steps)
        log("cpu baseline", cb)
        cpu = {"value": round(cb["value"], 1), "unit": "synaptic events/s", "cores": 1, "kind": "port",
               "sample": f"oracle (serial C, FP64, cpp_standalone semantics) on the same network, steps "
                         f"{args.cpu_skip}..{args.cpu_skip + args.cpu_steps} ({cb['events']} events in "
                         f"{cb['wall_s']:.2f} s; {cb['sim_s_per_wall_s']:.4f} sim-s/wall-s)"}

    out = {
        "metric": "synaptic events/s and sim-s/wall-s: FP32 batched vs FP64 plain, 1 MI355X",
        "value": round(fp32["events_per_s"], 1),
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp32["ms_per_step"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator; inputs resident in HBM before timing)",
        "config": {"workload": {"brunel": "brunel_hakim_heterog_delays", "cuba": "cuba", "batch": "batch_cuba",
                                "cobahh": "cobahh"}[args.workload],
                   "neurons": N, "synapses": spec.n_synapses, "dt_ms": spec.dt * 1e3,
                   "sim_s": args.steps * spec.dt, "delay_slots": int(D), "seed": args.seed},
        "sim_s_per_wall_s": round(fp32["sim_s_per_wall_s"], 4),
        "fp32_batched": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in fp32.items()},
        "fp64_plain": ({k: (round(v, 6) if isinstance(v, float) else v) for k, v in fp64.items()} if fp64 else None),
        "speedup_fp32_batched_vs_fp64_plain": (round(fp32["events_per_s"] / fp64["events_per_s"], 4)
                                               if fp64 and fp64["events_per_s"] > 0 else None),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device": info,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
