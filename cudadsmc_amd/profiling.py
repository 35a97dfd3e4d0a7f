"""rocprofv3 helpers: summaries of kernel traces and counter runs, the plain
copy kernel that calibrates the HBM byte counters, and the per-launch HBM
traffic figure derived from them.

  python -m cudadsmc_amd.profiling copy [MiB]                  # calibration copy
  python -m cudadsmc_amd.profiling summary DIR [LAST]          # kernel-trace summary
  python -m cudadsmc_amd.profiling counters DIR [LAST]         # per-kernel counter means
  python -m cudadsmc_amd.profiling traffic KERNEL FETCH_DIR WRITE_DIR \
        COPY_FETCH_DIR COPY_WRITE_DIR COPY_MIB LAST OUT          # traffic per launch
  python -m cudadsmc_amd.profiling window STATE [STEPS]        # resumed timing window
  python -m cudadsmc_amd.profiling bench-traffic KERNEL BENCH_JSON FETCH_DIR WRITE_DIR \
        COPY_FETCH_DIR COPY_WRITE_DIR COPY_MIB OUT               # traffic of a profiled bench run

rocprofv3 is run with --output-format csv; DIR is its -d directory.  LAST
restricts means to the last LAST dispatches of each kernel (the bench's
timing window is its last 50 steps).
"""
import csv
import glob
import os
import sys


def copy_kernel(mib=512, reps=5):
    """Device-to-device copy of a known size: the reference reading for
    FETCH_SIZE / WRITE_SIZE (each replay reads and writes `mib` MiB)."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}")


def kernel_short_name(full):
    """'void dsmc::(anonymous namespace)::k_cell<...>(args)' -> 'k_cell'."""
    s = full.strip()
    if s.endswith(")"):
        depth = 0
        for i in range(len(s) - 1, -1, -1):
            depth += {")": 1, "(": -1}.get(s[i], 0)
            if depth == 0:
                s = s[:i]
                break
    depth, cut = 0, len(s)
    for i in range(len(s) - 1, -1, -1):       # drop a trailing template argument list
        depth += {">": 1, "<": -1}.get(s[i], 0)
        if depth == 0:
            cut = i if s[i] == "<" else cut
            break
    s = s[:cut] if s.endswith(">") else s
    return s.split("::")[-1].split(" ")[-1]


def _find(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no {pattern} under {d}")
    return hits[0]


def _per_kernel(rows, value):
    per = {}
    for r in sorted(rows, key=lambda r: int(r.get("Start_Timestamp", r.get("Dispatch_Id", 0)) or 0)):
        per.setdefault(kernel_short_name(r["Kernel_Name"]), []).append(value(r))
    return per


def trace_summary(d, last=None):
    """Per-kernel call count, mean duration (us) over all calls and over the
    last `last` calls."""
    rows = list(csv.DictReader(open(_find(d, "*kernel_trace.csv"))))
    per = _per_kernel(rows, lambda r: (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = {}
    for k, v in per.items():
        tail = v[-last:] if last else v
        out[k] = dict(calls=len(v), mean_us=sum(v) / len(v), last_mean_us=sum(tail) / len(tail))
    return out


def counter_summary(d, last=None):
    """{(kernel, counter): (mean over all dispatches, mean over the last
    `last`, dispatch count)} from counter_collection.csv."""
    rows = list(csv.DictReader(open(_find(d, "*counter_collection.csv"))))
    by = {}
    for r in rows:
        by.setdefault(r["Counter_Name"], []).append(r)
    out = {}
    for cname, rs in by.items():
        rs.sort(key=lambda r: int(r["Dispatch_Id"]))
        per = {}
        for r in rs:
            per.setdefault(kernel_short_name(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
        for k, v in per.items():
            tail = v[-last:] if last else v
            out[(k, cname)] = (sum(v) / len(v), sum(tail) / len(tail), len(v))
    return out


def traffic(kernel, fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_mib, last):
    """HBM bytes per launch of `kernel`: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch, mean over its last `last` dispatches), each scaled by
    known bytes / counter reading of the calibration copy."""
    known_kib = copy_mib * 1024.0

    def copy_reading(d, counter):
        cs = counter_summary(d)
        best = max(((m, k) for (k, c), (m, _, _) in cs.items() if c == counter), default=None)
        if best is None:
            raise ValueError(f"no {counter} reading under {d}")
        return best[0]
    f_scale = known_kib / copy_reading(copy_fetch_dir, "FETCH_SIZE")
    w_scale = known_kib / copy_reading(copy_write_dir, "WRITE_SIZE")
    f = counter_summary(fetch_dir, last)[(kernel, "FETCH_SIZE")][1]
    w = counter_summary(write_dir, last)[(kernel, "WRITE_SIZE")][1]
    return dict(kernel=kernel, fetch_kib=f, write_kib=w, fetch_scale=f_scale, write_scale=w_scale,
                traffic_bytes=(f * f_scale + w * w_scale) * 1024.0, window=last)


def bench_traffic(kernel, bench_json, fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_mib, out):
    """The traffic file bench.py pairs with its roofline: HBM bytes per
    launch of `kernel` over the timed steps of a `bench.py --profile-run`
    (its last `steps` dispatches), with the state that run timed --
    workload, development steps, sampling, particles at the end -- and the
    algorithmic bytes it reported for the same launches."""
    import json
    b = None
    for line in open(bench_json):
        line = line.strip()
        if line.startswith("{"):
            b = json.loads(line)
    if b is None:
        raise ValueError(f"no bench line in {bench_json}")
    cfg = b["config"]
    steps = int(b["steps"])
    t = traffic(kernel, fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_mib, steps)
    alg = b["roofline"]["per_kernel"][kernel]["alg_bytes"]
    lines = [
        f"# HBM traffic per launch of {kernel}, workload {cfg['workload']}, over the {steps} timed steps "
        f"{cfg['timed_steps']} of a bench.py --profile-run (the last {steps} dispatches of {kernel}).",
        f"# FETCH_SIZE and WRITE_SIZE (KiB per dispatch) each scaled by known bytes / reading of a {copy_mib:g} MiB "
        f"device copy.",
        f"# traffic / algorithmic = {t['traffic_bytes'] / alg:.3f}",
        f"kernel = {kernel}", f"workload = {cfg['workload']}", f"develop = {cfg['develop_steps']}",
        f"sampling = {'on' if cfg.get('sampling_in_timed_steps', True) else 'off'}",
        f"particles = {cfg['particles_final']}", f"alg_bytes = {alg}",
        f"fetch_kib = {t['fetch_kib']}", f"write_kib = {t['write_kib']}",
        f"fetch_scale = {t['fetch_scale']}", f"write_scale = {t['write_scale']}",
        f"traffic_bytes = {t['traffic_bytes']}", f"window = {steps}",
    ]
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def window(state, steps=50):
    """Resumes the Mach-10 cylinder bench case from a checkpoint written by
    `bench.py --save-state` and runs its eager timing window: the same
    kernel launches, on the same state, as the bench's last `steps` steps."""
    import json
    from . import cases
    from .run import Run
    cfg = cases.cylinder_mach10()
    r = Run(cfg["params"], cfg["body"])
    r.load_state(state)
    r.set_timing(True)
    r.solver(steps, graph=False)
    kt, n = r.kernel_times()
    print(json.dumps(dict(state=state, steps=steps, kernels=kt)))
    r.close()


def main(argv):
    cmd = argv[1]
    if cmd == "copy":
        copy_kernel(int(argv[2]) if len(argv) > 2 else 512)
    elif cmd == "summary":
        last = int(argv[3]) if len(argv) > 3 else None
        s = trace_summary(argv[2], last)
        for k, v in sorted(s.items(), key=lambda kv: -kv[1]["mean_us"] * kv[1]["calls"]):
            print(f"{k:32s} calls {v['calls']:7d}  mean {v['mean_us']:9.2f} us"
                  f"  last-{last or 'all'} mean {v['last_mean_us']:9.2f} us")
    elif cmd == "counters":
        last = int(argv[3]) if len(argv) > 3 else None
        for (k, c), (m, mt, n) in sorted(counter_summary(argv[2], last).items()):
            print(f"{k:32s} {c:12s} mean {m:14.1f}  last-{last or 'all'} mean {mt:14.1f}  ({n} dispatches)")
    elif cmd == "traffic":
        kernel, fd, wd, cfd, cwd, mib, last, out = argv[2:10]
        t = traffic(kernel, fd, wd, cfd, cwd, float(mib), int(last))
        lines = [f"{k} = {v}" for k, v in t.items()]
        open(out, "w").write("\n".join(lines) + "\n")
        print("\n".join(lines))
    elif cmd == "bench-traffic":
        kernel, bj, fd, wd, cfd, cwd, mib, out = argv[2:10]
        bench_traffic(kernel, bj, fd, wd, cfd, cwd, float(mib), out)
    elif cmd == "window":
        window(argv[2], int(argv[3]) if len(argv) > 3 else 50)
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv)
