#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the DSMC hot path on one MI355X, and the
learned warm start's time to steady state against plain DSMC.

Default workload (sweep64): the batched sweep of 64 independent DSMC cases
-- cylinders and wedges (half-angle 10-30 deg), Mach 3-12, Kn 0.005-0.1,
T_w/T_inf 1-5, grids 64x32 to 256x128, about 110M particles in all (0.2M to
6M each, uneven) -- advanced together, one launch per phase per step.  The
cases start from freestream and run `--develop` untimed steps (1,500 by
default: the shock layers and wakes have formed, and every case samples
from then on), then W untimed warm-up steps, then the K timed steps.

A step is one pass of the hot path over every particle of every case: move
with walls and inflow and the per-cell count, scan, scatter, then per tile
of cells the stable cell order, the reorder into cell order (outflow
compaction), NTC/VHS collisions and FP64 sampling.  value = particle-steps
of the K timed steps / their wall time (inputs resident in HBM, barrier and
device sync on both sides).  The timed steps are launched eagerly with HIP
events between the four launches of each step on the library's stream; the
events give every kernel's mean duration over exactly those steps, and the
device counters read before and after give their algorithmic bytes, so
roofline.achieved comes from the timed region itself.  The same K steps are
then run again as hipGraph replays (graph_replay).

roofline.traffic is the HBM traffic per launch from rocprofv3 counters
(profiles/k_cell_<workload>_traffic.txt), used only when that profile was
taken on this workload at this state (same development steps and particle
count within 3%); otherwise null.  cpu_baseline times the CPU oracle on a
10-case subset of the same developed state on this host, on the threads
allotted to the process and on one thread.  time_to_steady runs the learned
warm start against plain DSMC on 32 held-out cases (cudadsmc_amd.surrogate).

Workload cylinder_mach10_400x200: the single Mach-10 cylinder (400x200
cells, about 2M particles), 20,000 steps with sampling from 10,000.
Workload cylinder_n2_mach10_400x200: the same case in nitrogen (zeta = 2,
Larsen-Borgnakke exchange in every collision, Mach 10 at gamma = 7/5).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s; time-to-steady warm-start vs plain at <=2% field error"


def log(m):
    print(m, file=sys.stderr, flush=True)


def host_threads():
    """Threads the process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (the share of the host a GPU job is
    given)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def roofline_of(kt, traffic=None):
    per = {k: dict(ms=v["ms"], alg_bytes=v["bytes"],
                   gbs=v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0) for k, v in kt.items()}
    dom = max(kt, key=lambda k: kt[k]["ms"])
    achieved = per[dom]["gbs"]
    out = dict(bound="hbm", kernel=dom, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
               frac=achieved / HBM_PEAK_GBS, traffic=None, per_kernel=per)
    if traffic is not None:
        out.update(traffic)
    return out


def matching_traffic(kernel, workload, develop, particles, sampling):
    """The committed counter profile of `kernel` for this workload when its
    state matches the timed one (same development steps, sampling alike,
    particle count within 3%); else a note why not."""
    f = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.txt")
    if not os.path.exists(f):
        return dict(traffic=None, traffic_note=f"no profile {os.path.relpath(f, ROOT)}")
    kv = dict(l.split(" = ", 1) for l in open(f).read().splitlines() if " = " in l and not l.startswith("#"))
    try:
        same = (kv.get("kernel") == kernel and kv.get("workload") == workload and int(kv["develop"]) == develop
                and kv.get("sampling") == ("on" if sampling else "off")
                and abs(float(kv["particles"]) / particles - 1.0) < 0.03)
    except (KeyError, ValueError):
        same = False
    if not same:
        return dict(traffic=None, traffic_note=f"{os.path.relpath(f, ROOT)} was taken at another state")
    return dict(traffic=float(kv["traffic_bytes"]), traffic_alg_bytes=float(kv["alg_bytes"]),
                traffic_ratio=float(kv["traffic_bytes"]) / float(kv["alg_bytes"]),
                traffic_source=os.path.relpath(f, ROOT))


# ------------------------------------------------------------------ CPU baseline
def _oracle_states(cases_, downloads):
    from oracle import oracle as O
    states = []
    for c, (parts, sm, mr, st) in zip(cases_, downloads):
        oc, info, _ = O.case_from_params(c["params"], c["body"])
        s = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0, step=st.step, next_pid=st.next_pid)
        s.sigcmax[:] = sm
        s.mrem[:] = mr
        states.append((oc, s))
    return states


def _time_oracle(states, seconds, threads, min_steps=1):
    psteps, steps, t0 = 0, 0, time.perf_counter()
    while True:
        for _, s in states:
            psteps += s.st.n
            s.step(sample=True, nthreads=threads)
        steps += 1
        dt = time.perf_counter() - t0
        if steps >= min_steps and dt >= seconds:
            return psteps / dt, steps, psteps, dt


def cpu_baseline_batch(b, sw, seconds, pick=10):
    """The oracle (the same algorithm and Philox streams, C with OpenMP) on
    `pick` cases of the sweep spread over its particle counts, from the
    GPU's state: on the allotted threads, then on one thread."""
    order = sorted(range(b.n), key=lambda i: b.status(i).n_live)
    sub = [order[round(k * (len(order) - 1) / (pick - 1))] for k in range(pick)]
    cases_ = [sw[i] for i in sub]
    downloads = []
    for i in sub:
        sm, mr, _ = b.cell_state(i)
        downloads.append((b.download(i), sm, mr, b.status(i)))
    n = host_threads()
    log(f"[bench] cpu baseline: oracle on {pick} cases, {n} threads")
    v, steps, ps, dt = _time_oracle(_oracle_states(cases_, downloads), seconds, n)
    log("[bench] cpu baseline: one thread")
    v1, steps1, ps1, dt1 = _time_oracle(_oracle_states(cases_, downloads), seconds * 0.5, 1)
    first = b.status(sub[0]).step
    return dict(value=v, unit="particle-steps/s", cores=n, kind="port",
                sample=(f"{pick} of the 64 sweep cases (instances {sub}, {sum(d[0]['x'].size for d in downloads)} "
                        f"particles) from the GPU state at step {first}, sampling on: {steps} oracle steps of "
                        f"each, {ps} particle-steps in {dt:.1f} s"),
                single_core=dict(value=v1, cores=1,
                                 sample=f"the same cases: {steps1} steps, {ps1} particle-steps in {dt1:.1f} s"),
                cpu_model=cpu_model(), host_cpus=os.cpu_count())


def cpu_baseline_run(run, cfg, seconds):
    sm, mr, _ = run.cell_state()
    downloads = [(run.download(), sm, mr, run.status())]
    n = host_threads()
    v, steps, ps, dt = _time_oracle(_oracle_states([cfg], downloads), seconds, n)
    v1, steps1, ps1, dt1 = _time_oracle(_oracle_states([cfg], downloads), seconds * 0.5, 1)
    return dict(value=v, unit="particle-steps/s", cores=n, kind="port",
                sample=(f"{cfg['name']} from the GPU state at step {downloads[0][3].step}: {steps} oracle steps, "
                        f"{ps} particle-steps in {dt:.1f} s"),
                single_core=dict(value=v1, cores=1, sample=f"{steps1} steps, {ps1} particle-steps in {dt1:.1f} s"),
                cpu_model=cpu_model(), host_cpus=os.cpu_count())


# ------------------------------------------------------------------ time to steady
def time_to_steady(args):
    """The learned warm start against plain DSMC: a surrogate trained in
    this run on DSMC runs of the sweep distribution, evaluated on held-out
    cases (cudadsmc_amd.surrogate.pipeline)."""
    from cudadsmc_amd import surrogate as S
    out = os.path.join(ROOT, "gpurun_out", "bench_warmstart")
    t0 = time.perf_counter()
    r = S.pipeline(out, runs=args.steady_runs, n_eval=args.steady_eval, fit_steps=args.steady_fit_steps,
                   average_steps=args.steady_average, settle_steps=args.steady_average // 2, log=log)
    e = lambda k: [max(d.values()) for d in r[k]]
    return dict(
        cases=args.steady_eval, training_runs=args.steady_runs, training_samples=r["train_samples"],
        data_s=r["data_s"], fit_s=r["fit_s"], pipeline_s=time.perf_counter() - t0,
        speedup_alone_median=r["speedup_alone_median"], speedup_alone_total=r["speedup_alone_total"],
        speedup_sweep_retiring=r["speedup_wall"], speedup_sweep_keep_all=r["speedup_wall_keep_all"],
        median_steps_plain=r["median_steps_plain"], median_steps_warm=r["median_steps_warm"],
        wall_alone_plain_s=sum(r["wall_alone_plain_s"]), wall_alone_warm_s=sum(r["wall_alone_warm_s"]),
        within_2pct_warm=r["within_2pct"], within_2pct_plain=r["within_2pct_plain"],
        misfires_warm=r["misfires_warm"], misfires_plain=r["misfires_plain"],
        max_field_error_warm=r["max_field_error"], median_field_error_warm=float(sorted(e("field_error"))[
            len(r["field_error"]) // 2]),
        rel_l2_warm={k: float(max(d[k] for d in r["field_error"])) for k in ("n", "u", "T")},
        fallback_rate=r["fallback_rate"], device_err=r["device_err"],
        # cases after which the one-off data generation and fit are paid back
        # by the time each warm-started case saves when run alone
        break_even_cases=(r["data_s"] + r["fit_s"]) / max(
            (sum(r["wall_alone_plain_s"]) - sum(r["wall_alone_warm_s"])) / len(r["wall_alone_plain_s"]), 1e-9))


# ------------------------------------------------------------------ workloads
def timed(sim, steps, chunk, events=True):
    """K eager steps with per-kernel HIP events (the library's timing mode;
    events=False: algorithmic bytes only); returns wall seconds and the
    kernel summary."""
    sim.set_timing(True, events)
    sim.sync()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        sim.solver(k, graph=False)
        if not events:
            sim.sync()     # counter passes: progress lines count finished steps
        done += k
        log(f"[bench] {done}/{steps} timed steps")
    sim.sync()
    dt = time.perf_counter() - t0
    kt, n = sim.kernel_times()
    sim.set_timing(False)
    return dt, kt


def graph_phase(sim, steps, chunk, status):
    """The same number of steps as hipGraph replays: particle-steps/s."""
    sim.sync()
    s0 = status()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        sim.solver(k, graph=True)
        done += k
    sim.sync()
    dt = time.perf_counter() - t0
    s1 = status()
    return dict(value=(s1.particle_steps - s0.particle_steps) / dt, ms_per_step=dt / steps * 1e3, steps=steps)


def bench_sweep(args):
    from cudadsmc_amd import autotune, cases
    from cudadsmc_amd.batch import Batch

    steps = args.steps if args.steps is not None else 200
    develop = args.develop if args.develop is not None else 1500
    sw = cases.sweep(64, sample_from=develop)
    b = Batch([c["params"] for c in sw], [c["body"] for c in sw])
    b.initialize()
    n_init = b.status(-1).n_live
    log(f"[bench] sweep64: {len(sw)} cases, {n_init} particles, {int(b.cbase[-1])} cells")
    tuned = None
    if args.tune:
        tuned, _ = autotune.tune(autotune.BatchTarget(b, lambda s: s.initialize(), warm_steps=5),
                                 autotune.device_cus(), steps=6, graph_steps_measure=32, log=log)
        b.initialize()
    dev_graph = not args.profile_run
    done = 0
    # a profiled run syncs every chunk, so its progress lines count finished
    # steps (the profiler slows the dispatches far behind the host)
    chunk = 50 if args.profile_run else 250
    while done < develop:
        k = min(chunk, develop - done)
        b.solver(k, graph=dev_graph)
        if args.profile_run:
            b.sync()
        done += k
        log(f"[bench] {done}/{develop} development steps")
    b.solver(args.warmup, graph=dev_graph)
    s0 = b.status(-1)
    dt, kt = timed(b, steps, 5 if args.no_events else 50, events=not args.no_events)
    s1 = b.status(-1)
    if s1.err:
        log(f"warning: device error bits {s1.err:#x}")
    value = (s1.particle_steps - s0.particle_steps) / dt
    sampling = all(b.status(i).nsamples > 0 for i in range(b.n))
    roof = roofline_of(kt, matching_traffic(max(kt, key=lambda k: kt[k]["ms"]), "sweep64", develop, s1.n_live,
                                            sampling))
    graph = None if args.profile_run else graph_phase(b, steps, 50, lambda: b.status(-1))
    if graph:
        log(f"[bench] graph replay: {graph['value']:.4g} particle-steps/s")
    cpu = None if (args.no_cpu or args.profile_run) else cpu_baseline_batch(b, sw, args.cpu_seconds)
    first_timed = develop + args.warmup
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": dt / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": "sweep64", "cases": len(sw), "cells": int(b.cbase[-1]),
            "particles_initial": n_init, "particles_timed_start": s0.n_live, "particles_final": s1.n_live,
            "develop_steps": develop, "timed_steps": [first_timed, first_timed + steps],
            "sampling_in_timed_steps": sampling, "collisions": s1.collisions - s0.collisions,
            "launch": "eager" if args.no_events else "eager with per-kernel HIP events",
            "launch_settings": b.get_launch(),
            "autotuned": tuned is not None, "seed": 20251203,
            "development_launch": "hipgraph" if dev_graph else "eager",
        },
        "graph_replay": graph,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    b.close()
    if not (args.no_steady or args.profile_run):
        out["time_to_steady"] = time_to_steady(args)
    print(json.dumps(out))


def bench_cylinder(args):
    from cudadsmc_amd import cases
    from cudadsmc_amd.run import Run

    if args.workload == "cylinder_n2_mach10_400x200":
        # the reference's gas: N2 with Larsen-Borgnakke exchange in every collision
        cfg = cases.cylinder_mach10(gas=cases.NITROGEN)
        cfg["name"] = args.workload
    else:
        cfg = cases.cylinder_mach10()
    p = cfg["params"]
    steps = args.steps if args.steps is not None else 19990
    develop = args.develop if args.develop is not None else 0
    run = Run(p, cfg["body"])
    run.initialize()
    n_init = run.status().n_live
    run.solver(develop + args.warmup, graph=True)
    if args.save_state:
        run.save_state(args.save_state)
    s0 = run.status()
    dt, kt = timed(run, steps, 2000)
    s1 = run.status()
    value = (s1.particle_steps - s0.particle_steps) / dt
    sampling = s1.nsamples > s0.nsamples
    roof = roofline_of(kt, matching_traffic(max(kt, key=lambda k: kt[k]["ms"]), cfg["name"], develop, s1.n_live,
                                            sampling))
    graph = graph_phase(run, min(steps, 2000), 1000, run.status)
    cpu = None if args.no_cpu else cpu_baseline_run(run, cfg, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": dt / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": cfg["name"], "cells": run.info.ncell, "particles_initial": n_init,
            "particles_final": s1.n_live, "develop_steps": develop,
            "timed_steps": [develop + args.warmup, develop + args.warmup + steps],
            "sample_from_step": cfg["sample_from"], "launch": "eager with per-kernel HIP events",
            "launch_settings": run.get_launch(), "collisions": s1.collisions - s0.collisions, "seed": int(p.seed),
            "gas": "N2 (zeta = 2, rot_prob %g)" % p.rot_prob if p.zeta else "Ar",
        },
        "graph_replay": graph,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    run.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (sweep64: 200; cylinder: 19,990)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--develop", type=int, default=None,
                    help="untimed development steps before the warm-up (sweep64: 1,500; cylinder: 0)")
    ap.add_argument("--workload", default="sweep64",
                    choices=["sweep64", "cylinder_mach10_400x200", "cylinder_n2_mach10_400x200"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-steady", action="store_true", help="skip the warm-start time-to-steady evaluation")
    ap.add_argument("--steady-runs", type=int, default=1024, help="DSMC runs that train the surrogate")
    ap.add_argument("--steady-eval", type=int, default=32, help="held-out cases")
    ap.add_argument("--steady-fit-steps", type=int, default=8000)
    ap.add_argument("--steady-average", type=int, default=12000,
                    help="steps the fields are averaged over for the field errors")
    ap.add_argument("--tune", action="store_true", help="run the launch-settings autotuner first (untimed)")
    ap.add_argument("--profile-run", action="store_true",
                    help="for rocprofv3: development and warm-up launched eagerly, no graph phase, no CPU "
                         "baseline, no time to steady -- the timed steps are then the last dispatches")
    ap.add_argument("--no-events", action="store_true",
                    help="with --profile-run: no HIP events around the timed kernels, algorithmic bytes only "
                         "(for counter passes, where the profiler times the kernels)")
    ap.add_argument("--save-state", default=None, help="cylinder: checkpoint of the state the timed steps start from")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path runs on a single MI355X "
              "(sharding across GPUs is designed, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.no_events and not args.profile_run:
        print("--no-events is for counter passes and needs --profile-run", file=sys.stderr)
        sys.exit(2)
    if args.workload == "sweep64":
        bench_sweep(args)
    else:
        bench_cylinder(args)


if __name__ == "__main__":
    main()
