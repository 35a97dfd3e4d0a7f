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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the hot path on a single MI355X "
              "(multi-GPU sharding is a design note, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

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
            "collisions": s1.collisions - s0.collisions,
            "seed": int(p.seed),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    run.close()


if __name__ == "__main__":
    main()
