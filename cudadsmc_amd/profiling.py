"""rocprofv3 helpers: summaries of kernel traces and counter runs, and the
plain copy kernel used to calibrate the HBM byte counters.

  python -m cudadsmc_amd.profiling copy [MiB]      # copy kernel for calibration
  python -m cudadsmc_amd.profiling summary DIR     # kernel-trace summary of a run
  python -m cudadsmc_amd.profiling counters DIR    # per-kernel counter means
"""
import csv
import glob
import os
import sys


def copy_kernel(mib=512, reps=5):
    """Device-to-device copy of a known size (torch's copy kernel): the
    reference reading for FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}")


def _find(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no {pattern} under {d}")
    return hits[0]


def trace_summary(d, last=None):
    """Per-kernel call count, mean duration (us) over all calls and over the
    last `last` calls of each kernel."""
    rows = list(csv.DictReader(open(_find(d, "*kernel_trace.csv"))))
    per = {}
    for r in rows:
        name = r["Kernel_Name"].split("(")[0]
        per.setdefault(name, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = {}
    for k, v in per.items():
        tail = v[-last:] if last else v
        out[k] = dict(calls=len(v), mean_us=sum(v) / len(v), last_mean_us=sum(tail) / len(tail))
    return out


def counter_summary(d):
    """Per-kernel mean of every collected counter (rocprofv3 counter_collection.csv)."""
    rows = list(csv.DictReader(open(_find(d, "*counter_collection.csv"))))
    acc = {}
    for r in rows:
        name = r["Kernel_Name"].split("(")[0]
        key = (name, r["Counter_Name"])
        acc.setdefault(key, []).append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


if __name__ == "__main__":
    cmd = sys.argv[1]
    if cmd == "copy":
        copy_kernel(int(sys.argv[2]) if len(sys.argv) > 2 else 512)
    elif cmd == "summary":
        last = int(sys.argv[3]) if len(sys.argv) > 3 else None
        for k, v in sorted(trace_summary(sys.argv[2], last).items(), key=lambda kv: -kv[1]["mean_us"] * kv[1]["calls"]):
            print(f"{k:60s} calls {v['calls']:7d} mean {v['mean_us']:9.2f} us  last-window mean {v['last_mean_us']:9.2f} us")
    elif cmd == "counters":
        for (k, c), (m, n) in sorted(counter_summary(sys.argv[2]).items()):
            print(f"{k:60s} {c:16s} mean {m:16.1f} over {n} dispatches")
