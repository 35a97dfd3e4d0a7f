#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the DSMC hot path on one MI355X.

Workload: 2D supersonic argon flow over a cut-cell cylinder -- Mach 10,
T_inf = 200 K, Kn = 0.01 on the diameter, R = 0.1 m in a 0.6 m x 0.3 m
domain, 400x200 cells, diffuse wall at 500 K, freestream inflow/outflow,
30 particles per fluid cell (about 2.0M particles, rising in the shock
layer), FP32 state, 20,000 steps with sampling from step 10,000.

A step is one pass of the hot path (move + walls + inflow + bin, scan,
scatter, stable cell order + reorder + NTC collisions + sampling) over all
particles, captured in a hipGraph.  value = sum over the K timed steps of the
particles moved, divided by the wall time of those K steps, state resident in
HBM.  The W warm-up steps and initialisation are untimed.

Per-kernel durations come from HIP events on the library's stream over an
eager window after the timed run; roofline.achieved = algorithmic bytes per
launch of the dominant kernel / its mean duration.  cpu_baseline times the
CPU oracle (C, OpenMP) from the GPU's final state on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s; time-to-steady warm-start vs plain at <=2% field error"


def cpu_baseline(run, info, seconds, threads):
    """Times the oracle (same algorithm, same Philox streams) on this host,
    starting from the GPU's current state, for about `seconds`."""
    import numpy as np
    from oracle import oracle as O

    p_gpu = run.download()
    sm, mr, _ = run.cell_state()
    st0 = run.status()
    oc, _, _ = O.case_from_params(run.params, run._body_xy)
    st = O.OracleState(oc, p_gpu, int(info.capacity), info.sigcmax0, step=st0.step, next_pid=st0.next_pid)
    st.sigcmax[:] = sm
    st.mrem[:] = mr
    psteps = 0
    steps = 0
    t0 = time.perf_counter()
    while True:
        n = st.st.n
        st.step(sample=True, nthreads=threads)
        psteps += n
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return dict(value=psteps / dt, unit="particle-steps/s", cores=threads, kind="port",
                sample=(f"cylinder Mach 10 400x200 from the GPU state at step {st0.step}: "
                        f"{steps} oracle steps, {psteps} particle-steps in {dt:.1f} s"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=19990)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--timing-steps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--eager", action="store_true", help="launch steps eagerly instead of hipGraph replay")
    ap.add_argument("--save-state", default=None,
                    help="write a checkpoint (.npz) of the state the timing window starts from")
    ap.add_argument("--workload", default="cylinder_mach10_400x200",
                    choices=["cylinder_mach10_400x200", "sweep64"],
                    help="sweep64: the 64-instance batched sweep (1000 timed steps unless --steps)")
    ap.add_argument("--tune", action="store_true",
                    help="run the launch-settings autotuner in the untimed warm-up (always on for sweep64)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path runs on a single MI355X "
              "(sharding across GPUs is designed, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "sweep64":
        return bench_sweep(args)

    from cudadsmc_amd import cases
    from cudadsmc_amd.run import Run

    cfg = cases.cylinder_mach10()
    p = cfg["params"]
    run = Run(p, cfg["body"])
    info = run.info
    run.initialize()
    n_init = run.status().n_live
    graph = not args.eager
    run.solver(args.warmup, graph=graph)
    run.sync()
    tuned = None
    if args.tune:
        from cudadsmc_amd import autotune
        tuned, _ = autotune.tune(autotune.RunTarget(run), autotune.device_cus(), steps=10,
                                 log=lambda m: print(m, file=sys.stderr, flush=True))
    s0 = run.status()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:     # replays are throttled in the library, so this tracks the device
        k = min(4000, args.steps - done)
        run.solver(k, graph=graph)
        done += k
        print(f"[bench] {done}/{args.steps} steps queued", file=sys.stderr, flush=True)
    run.sync()
    t1 = time.perf_counter()
    s1 = run.status()
    elapsed = t1 - t0
    psteps = s1.particle_steps - s0.particle_steps
    value = psteps / elapsed
    if s1.err:
        print(f"warning: device error bits {s1.err:#x}", file=sys.stderr)

    if args.save_state:
        run.save_state(args.save_state)

    # per-kernel timing window (HIP events on the library stream)
    run.set_timing(True)
    run.solver(args.timing_steps, graph=False)
    kt, nl = run.kernel_times()
    run.set_timing(False)
    dom = max(kt, key=lambda k: kt[k]["ms"])
    achieved = kt[dom]["bytes"] / (kt[dom]["ms"] * 1e-3) / 1e9
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", f"{dom}_{cfg['name']}_traffic.txt")
    if os.path.exists(tf):
        kv = dict(l.split(" = ", 1) for l in open(tf).read().splitlines() if " = " in l)
        if kv.get("kernel") == dom:
            traffic = float(kv["traffic_bytes"])
            traffic_src = os.path.relpath(tf, ROOT)
    roofline = dict(bound="hbm", kernel=dom, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=traffic, traffic_source=traffic_src,
                    per_kernel={k: dict(ms=v["ms"], alg_bytes=v["bytes"],
                                        gbs=v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0)
                                for k, v in kt.items()})

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(run, info, args.cpu_seconds, threads)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / max(args.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": cfg["name"],
            "cells": info.ncell,
            "particles_initial": n_init,
            "particles_final": s1.n_live,
            "steps_total": args.warmup + args.steps,
            "sample_from_step": cfg["sample_from"],
            "launch": "eager" if args.eager else "hipgraph",
            "launch_settings": run.get_launch(),
            "autotuned": tuned is not None,
            "collisions": s1.collisions - s0.collisions,
            "seed": int(p.seed),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    run.close()


def bench_sweep(args):
    """The batched sweep: 64 uneven instances (cylinders and wedges, Mach
    3-12, Kn 0.005-0.1, T_w/T_inf 1-5, 64x32 ... 256x128 cells, about 110M
    particles) advanced together; launch settings autotuned in the warm-up;
    value = particle-steps of all instances per second of the timed steps."""
    from cudadsmc_amd import autotune, cases
    from cudadsmc_amd.batch import Batch

    steps = args.steps if args.steps != 19990 else 1000
    sw = cases.sweep(64)
    b = Batch([c["params"] for c in sw], [c["body"] for c in sw])
    b.initialize()
    n_init = b.status(-1).n_live
    print(f"[bench] sweep64: {len(sw)} instances, {n_init} particles, {int(b.cbase[-1])} cells",
          file=sys.stderr, flush=True)
    tuned, _ = autotune.tune(autotune.BatchTarget(b, lambda s: s.initialize(), warm_steps=5),
                             autotune.device_cus(), steps=6, graph_steps_measure=32,
                             log=lambda m: print(m, file=sys.stderr, flush=True))
    b.initialize()
    b.solver(args.warmup, graph=not args.eager)
    b.sync()
    s0 = b.status(-1)
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(250, steps - done)
        b.solver(k, graph=not args.eager)
        done += k
        print(f"[bench] {done}/{steps} steps queued", file=sys.stderr, flush=True)
    b.sync()
    elapsed = time.perf_counter() - t0
    s1 = b.status(-1)
    if s1.err:
        print(f"warning: device error bits {s1.err:#x}", file=sys.stderr)
    b.set_timing(True)
    b.solver(args.timing_steps // 2 or 1, graph=False)
    kt, _ = b.kernel_times()
    b.set_timing(False)
    dom = max(kt, key=lambda k: kt[k]["ms"])
    achieved = kt[dom]["bytes"] / (kt[dom]["ms"] * 1e-3) / 1e9
    psteps = s1.particle_steps - s0.particle_steps
    out = {
        "metric": METRIC,
        "value": psteps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "sweep64_batched",
            "instances": len(sw),
            "cells": int(b.cbase[-1]),
            "particles_initial": n_init,
            "particles_final": s1.n_live,
            "launch": "eager" if args.eager else "hipgraph",
            "launch_settings": tuned,
            "autotuned": True,
            "seed": 20251203,
        },
        "roofline": dict(bound="hbm", kernel=dom, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=achieved / HBM_PEAK_GBS, traffic=None,
                         per_kernel={k: dict(ms=v["ms"], alg_bytes=v["bytes"],
                                             gbs=v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0)
                                     for k, v in kt.items()}),
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    b.close()


if __name__ == "__main__":
    main()
