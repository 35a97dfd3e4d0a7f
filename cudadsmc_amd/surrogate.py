"""Learned warm start for DSMC: a coordinate MLP predicts per-cell
(log n, ux, uy, log T) at a new flow parameter point, particles start from
those local Maxwellians (dsmc_init_fields) instead of freestream, a short
DSMC residual check guards the prediction (freestream fallback), and the
steady-state detector measures how much sooner the run settles.

Features of a cell (7): x/L1, y/L2, Mach, log10 Kn, T_w/T_inf, geometry
parameter (-1 cylinder, half-angle/30 deg wedge), signed distance to the
body surface / L2 (positive in the fluid).  Targets (4): log(n/n_inf),
ux/U, uy/U, log(T/T_inf).  Both are standardised with the training set's
mean and standard deviation.

Command line (writes JSON and progress lines under OUT):
  python -m cudadsmc_amd.surrogate pipeline OUT [--runs 1024] [--eval 32]
"""
import json
import math
import os
import sys
import time

import numpy as np

from . import abi, cases
from .steady import SteadyDetector, half_fields

WIDTHS = [7, 256, 256, 256, 256, 4]
GRID = (128, 64)
# held-out evaluation grids: the Mach-10 cylinder bench resolution and below
EVAL_GRIDS = {"coarse": cases.SWEEP_GRIDS[:5], "fine": [(256, 128), (320, 160), (400, 200)]}


# ----------------------------------------------------------------- features
def draw_points(n, seed, grids=None, ppc=(30.0, 30.0)):
    """Flow parameter points of the sweep distribution (geometry, wedge
    half-angle, Mach, Kn, T_w/T_inf) with a grid per point."""
    rng = np.random.default_rng(seed)
    pts = []
    for _ in range(n):
        geom = "cylinder" if rng.random() < 0.5 else "wedge"
        pts.append(dict(geometry=geom, half_angle_deg=float(rng.uniform(10.0, 30.0)),
                        mach=float(rng.uniform(3.0, 12.0)),
                        kn=float(math.exp(rng.uniform(math.log(0.005), math.log(0.1)))),
                        tw_ratio=float(rng.uniform(1.0, 5.0)),
                        grid=GRID if grids is None else grids[int(rng.integers(len(grids)))],
                        ppc=float(math.exp(rng.uniform(math.log(ppc[0]), math.log(ppc[1])))),
                        seed=int(rng.integers(1, 2 ** 62))))
    return pts


def case_of(pt, sample_from=0, steps=0):
    nx, ny = pt["grid"]
    return cases.flow_case(pt["geometry"], pt["mach"], pt["kn"], pt["tw_ratio"], nx, ny, ppc=pt["ppc"],
                           half_angle_deg=pt["half_angle_deg"], sample_from=sample_from, steps=steps,
                           seed=pt["seed"], n_points=128)


def _signed_distance(px, py, poly):
    """Distance of points to the polygon boundary, negative inside."""
    d = np.full(px.shape, np.inf)
    inside = np.zeros(px.shape, bool)
    n = len(poly)
    for k in range(n):
        ax, ay = poly[k]
        bx, by = poly[(k + 1) % n]
        ex, ey = bx - ax, by - ay
        t = np.clip(((px - ax) * ex + (py - ay) * ey) / (ex * ex + ey * ey), 0.0, 1.0)
        d = np.minimum(d, np.hypot(px - ax - t * ex, py - ay - t * ey))
        cross = (ay > py) != (by > py)
        xint = ax + (py - ay) * ex / np.where(ey != 0, ey, 1.0)
        inside ^= cross & (px < xint)
    return np.where(inside, -d, d)


def cell_features(pt, cfg, wall=None):
    """Features of every cell of the case, in cell order c = j + i*ny.
    wall: the cells' signed distances to the body (m) when already known
    (Batch.wall_distance computes them on the device); else computed here."""
    p = cfg["params"]
    nx, ny = p.N1, p.N2
    i, j = np.divmod(np.arange(nx * ny), ny)
    x = (i + 0.5) * (p.L1 / nx)
    y = (j + 0.5) * (p.L2 / ny)
    geom = -1.0 if pt["geometry"] == "cylinder" else pt["half_angle_deg"] / 30.0
    wd = (_signed_distance(x, y, cfg["body"]) if wall is None else np.asarray(wall, np.float64)) / p.L2
    C = nx * ny
    return np.stack([x / p.L1, y / p.L2, np.full(C, pt["mach"]), np.full(C, math.log10(pt["kn"])),
                     np.full(C, pt["tw_ratio"]), np.full(C, geom), wd], 1).astype(np.float32)


def field_targets(cfg, f):
    p = cfg["params"]
    U = cfg["U"]
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.stack([np.log(f["n"].ravel() / p.n_inf), f["ux"].ravel() / U, f["uy"].ravel() / U,
                         np.log(f["T"].ravel() / p.T)], 1)


def fields_from_targets(cfg, y):
    p = cfg["params"]
    U = cfg["U"]
    y = np.asarray(y, np.float64)
    n = p.n_inf * np.exp(np.clip(y[:, 0], math.log(0.05), math.log(20.0)))
    T = p.T * np.exp(np.clip(y[:, 3], math.log(0.5), math.log(60.0)))
    return n, U * y[:, 1], U * y[:, 2], T


# ----------------------------------------------------------------- data
def generate(points, steady_steps=4000, sample_steps=2000, batch=64, log=print):
    """Runs the points as batched DSMC (freestream start, sampling after
    steady_steps) and returns features X, targets Y of the well-sampled
    fluid cells, and the run count."""
    from .batch import Batch
    from .run import describe
    Xs, Ys = [], []
    for b0 in range(0, len(points), batch):
        chunk = points[b0:b0 + batch]
        cfgs = [case_of(pt, sample_from=steady_steps) for pt in chunk]
        sim = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs])
        sim.initialize()
        total = steady_steps + sample_steps
        for k in range(0, total, 1000):
            sim.solver(min(1000, total - k))
            sim.sync()
            if log:
                log(f"[surrogate] data runs {b0}-{b0 + len(chunk) - 1}: step {min(k + 1000, total)}/{total}")
        err = sim.status(-1).err
        # any device flag (a particle stopped at the wall-event limit, an
        # inflow or capacity clamp, a failed fill) spoils the data
        if err:
            raise RuntimeError(f"device error bits {err:#x} in training runs {b0}-{b0 + len(chunk) - 1}")
        for i, (pt, c) in enumerate(zip(chunk, cfgs)):
            f = sim.fields(i)
            acc, ns = sim.samples(i)
            _, tab = describe(c["params"], c["body"])
            keep = (tab["frac"] >= 0.5) & (acc[0] >= 20)
            X = cell_features(pt, c)[keep]
            Y = field_targets(c, f)[keep]
            ok = np.all(np.isfinite(Y), 1)
            Xs.append(X[ok])
            Ys.append(Y[ok].astype(np.float32))
        sim.close()
    return np.concatenate(Xs), np.concatenate(Ys)


# ----------------------------------------------------------------- model
class Surrogate:
    def __init__(self, max_batch=65536, seed=0):
        from .mlp import MLP
        self.net = MLP(WIDTHS, max_batch=max_batch, seed=seed)
        self.xm = np.zeros(7, np.float32)
        self.xs = np.ones(7, np.float32)
        self.ym = np.zeros(4, np.float32)
        self.ys = np.ones(4, np.float32)

    def fit(self, X, Y, steps=3000, batch=65536, lr=2e-3, seed=1, log=print, chunk=500):
        self.xm, self.xs = X.mean(0), X.std(0) + 1e-6
        self.ym, self.ys = Y.mean(0), Y.std(0) + 1e-6
        self.net.set_data((X - self.xm) / self.xs, (Y - self.ym) / self.ys)
        batch = min(batch, self.net.max_batch, max(8, (X.shape[0] // 8) * 8))
        losses = []
        for k in range(0, steps, chunk):
            n = min(chunk, steps - k)
            # cosine decay of the learning rate over the whole fit
            lr_k = lr * 0.5 * (1 + math.cos(math.pi * k / max(steps, 1)))
            losses.append(self.net.train(n, batch=batch, lr=lr_k, seed=seed + k))
            if log:
                log(f"[surrogate] fit step {k + n}/{steps}: mean loss {float(np.mean(losses[-1])):.5f}")
        return np.concatenate(losses)

    def predict(self, X):
        return self.net.forward((X - self.xm) / self.xs) * self.ys + self.ym

    def fields_for(self, pt, cfg):
        return fields_from_targets(cfg, self.predict(cell_features(pt, cfg)))

    def fields_batch(self, sim, pts, cfgs):
        """Predicted (n, ux, uy, T) of every instance of a Batch: wall
        distances from the device, one forward pass over all cells."""
        wall = sim.wall_distance()
        cb = sim.cbase
        X = np.concatenate([cell_features(pt, c, wall[cb[i]:cb[i + 1]]) for i, (pt, c) in enumerate(zip(pts, cfgs))])
        Y = self.predict(X)
        return [fields_from_targets(c, Y[cb[i]:cb[i + 1]]) for i, c in enumerate(cfgs)]

    def save(self, path):
        np.savez(path, params=self.net.params(), xm=self.xm, xs=self.xs, ym=self.ym, ys=self.ys)

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.net.set_params(z["params"])
            self.xm, self.xs, self.ym, self.ys = z["xm"], z["xs"], z["ym"], z["ys"]


# ----------------------------------------------------------------- runs
def freestream_fields(cfg, C):
    p = cfg["params"]
    return np.full(C, p.n_inf), np.full(C, p.V_jet), np.full(C, p.V_y), np.full(C, p.T)


def residual(sim, i, pred, frac):
    """How far a short DSMC window started from a prediction moves away
    from it: the larger of the relative L2 gaps of density and temperature
    over mostly-fluid cells, each with the window's own sampling noise
    (Poisson counts for n; 2/(3N) relative variance for T) removed.  A
    prediction near the steady flow scores about 0.1-0.2; freestream or a
    reversed flow around a body scores above 1."""
    acc, _ = sim.samples(i)
    f = sim.fields(i)
    N = acc[0]
    ok = (frac.ravel() >= 0.5) & (N > 0)
    if not np.any(ok):
        return float("inf")
    out = 0.0
    for w, p, rel_var in ((f["n"].ravel(), pred[0], 1.0), (f["T"].ravel(), pred[3], 2.0 / 3.0)):
        gap = np.mean((w[ok] - p[ok]) ** 2) - np.mean(w[ok] ** 2 * rel_var / N[ok])
        out = max(out, math.sqrt(max(gap, 0.0)) / math.sqrt(np.mean(p[ok] ** 2)))
    return float(out)


def start(sim, pts, cfgs, tables, surrogate=None, check_steps=60, tol=0.5, force_bad=None, log=print,
          detectors=None):
    """Initialises a batch: freestream, or the surrogate's fields with a
    residual check per instance (instances that fail restart from
    freestream).  force_bad: indices whose prediction is replaced by a
    deliberately wrong one, the predicted flow reversed (tests of the
    fallback).  detectors: one SteadyDetector per instance; the check is then
    sampled as two halves and, when no instance falls back (a fallback
    refills the whole batch), fed to them as their first window.  Returns
    the fallback flags and the check's residuals."""
    C = [t["vol"].size for t in tables]
    if surrogate is None:
        sim.initialize()
        return np.zeros(len(pts), bool), np.zeros(len(pts))
    t0 = time.perf_counter()
    pred = [list(f) for f in surrogate.fields_batch(sim, pts, cfgs)]
    for i in (force_bad or []):
        pred[i][1], pred[i][2] = -pred[i][1], -pred[i][2]
    t1 = time.perf_counter()
    sim.initialize_fields(*[[p[k] for p in pred] for k in range(4)])
    sim.reset_samples()
    t2 = time.perf_counter()
    if detectors is None:
        sim.solver(check_steps, graph=False)
    else:
        if check_steps % 2:
            raise ValueError("a check fed to the detectors needs an even number of steps")
        half = check_steps // 2
        sim.solver(half)
        mid, _ = sim.moments()
        sim.solver(half)
        end, _ = sim.moments()
    res = np.array([residual(sim, i, pred[i], tables[i]["frac"]) for i in range(len(pts))])
    fb = res > tol
    if log:
        log(f"[surrogate] residual check: {int(fb.sum())} of {len(pts)} fall back (tol {tol}); prediction "
            f"{t1 - t0:.3f} s, fill {t2 - t1:.3f} s, check {time.perf_counter() - t2:.3f} s")
    if fb.any():
        f = [freestream_fields(c, C[i]) if fb[i] else pred[i] for i, c in enumerate(cfgs)]
        sim.initialize_fields(*[[x[k] for x in f] for k in range(4)])
    elif detectors is not None:
        cb = sim.cbase
        for i, d in enumerate(detectors):
            feed(d, sim, i, tables[i], mid[:, cb[i]:cb[i + 1]], end[:, cb[i]:cb[i + 1]] - mid[:, cb[i]:cb[i + 1]],
                 half)
    return fb, res


def feed(det, sim, i, table, first, second, half):
    """One window of instance i to its detector: density, velocity and
    temperature of the two halves (moment-sum differences first, second),
    over the mostly-fluid cells sampled in both halves; velocity and
    temperature weighted by the smaller half's sample count."""
    p = sim.params[i]
    vol, Fn = table["vol"], sim.info[i].Fn
    fa, na = half_fields(first, half, Fn, vol, p.mass, p.boltz)
    fb, nb = half_fields(second, half, Fn, vol, p.mass, p.boltz)
    ok = (table["frac"] >= 0.5) & (na > 0) & (nb > 0)
    g = np.minimum(na, nb)
    return det.update(fa, fb, ok, weights=[None, g, g])


# steady-state detector settings of every run here (cudadsmc_amd/steady.py)
STEADY = dict(threshold=1.5, patience=3, lag=4)


def run_to_steady(sim, tables, window=200, max_steps=30000, threshold=1.5, patience=3, lag=4, log=print, tag="",
                  retire=False, detectors=None):
    """Runs whole windows (each sampled as two halves) until every instance
    is steady (or max_steps); returns per-instance steps to steady (-1 if
    never) and the wall time to the last one.  The instances must be
    sampling every step; the half-window counts are differences of the
    running count sums, read for all instances in one transfer per half.
    retire: instances confirmed steady leave the batch (Batch.compact), so
    the rest run without them; the batch keeps the last ones to settle.
    Results are indexed by the instances the batch held at the call.
    detectors: the instances' detectors when they already hold windows
    (start's residual check); their windows count toward patience."""
    n = sim.n
    det = detectors if detectors is not None else [SteadyDetector(threshold, patience, lag) for _ in range(n)]
    ids = list(range(n))            # batch index -> index at the call
    steps_at = np.full(n, -1)
    half = window // 2
    t0 = time.perf_counter()
    steps = 0
    prev, ns0 = sim.moments()
    while steps < max_steps and (steps_at < 0).any():
        cb = sim.cbase
        sim.solver(half)
        mid, _ = sim.moments()
        sim.solver(half)
        end, ns = sim.moments()
        if np.any(ns - ns0 != 2 * half):
            raise RuntimeError("run_to_steady needs every instance sampling every step")
        steps += 2 * half
        for j, i in enumerate(ids):
            if steps_at[i] >= 0:
                continue
            first = mid[:, cb[j]:cb[j + 1]] - prev[:, cb[j]:cb[j + 1]]
            second = end[:, cb[j]:cb[j + 1]] - mid[:, cb[j]:cb[j + 1]]
            if feed(det[i], sim, j, tables[i], first, second, half):
                steps_at[i] = steps
        keep = [j for j, i in enumerate(ids) if steps_at[i] < 0]
        if retire and 0 < len(keep) < len(ids):
            sim.compact(keep)
            ids = [ids[j] for j in keep]
            end, ns = sim.moments()
        prev, ns0 = end, ns
        if log and steps % (10 * window) == 0:
            log(f"[surrogate] {tag} step {steps}: {int((steps_at >= 0).sum())}/{n} steady, {len(ids)} in the batch")
    sim.sync()
    return steps_at, time.perf_counter() - t0


def averaged_fields(sim, steps):
    sim.reset_samples()
    sim.solver(steps)
    sim.sync()
    return [sim.fields(i) for i in range(sim.n)]


def rel_l2(a, b, mask):
    return float(np.linalg.norm((a - b)[mask]) / max(np.linalg.norm(b[mask]), 1e-300))


def field_errors(fw, fp, frac):
    m = frac.ravel() >= 0.5
    e_n = rel_l2(fw["n"].ravel(), fp["n"].ravel(), m)
    e_T = rel_l2(fw["T"].ravel(), fp["T"].ravel(), m)
    uw = np.concatenate([fw["ux"].ravel()[m], fw["uy"].ravel()[m]])
    up = np.concatenate([fp["ux"].ravel()[m], fp["uy"].ravel()[m]])
    e_u = float(np.linalg.norm(uw - up) / np.linalg.norm(up))
    return dict(n=e_n, u=e_u, T=e_T)


def evaluate(surrogate, pts, window=100, average_steps=12000, settle_steps=6000, check_steps=None, log=print):
    """Warm start vs plain on held-out points (one batch each): steps and
    wall-clock to steady; field error of the warm run (averaged over
    average_steps right after it is steady) against the converged plain run
    (continued settle_steps past its steady point, then averaged over
    average_steps); the same error of the plain run at its own steady
    point; fallback rate.

    Each arm runs twice from the same seeds.  The timed pass retires
    instances from the batch as they are confirmed steady (what a sweep
    does), so its wall-clock is the time until all are steady.  The replay
    keeps every instance in the batch -- it must find the same steady steps,
    since a retired instance's neighbours evolve bit-identically -- and its
    state at the last steady step is what the fields are averaged from.
    The warm start's residual check (check_steps, one window by default) is
    the detector's first window; its steps count toward the warm run's."""
    from .batch import Batch
    from .run import describe
    check_steps = window if check_steps is None else check_steps
    cfgs = [case_of(pt) for pt in pts]
    tables = [describe(c["params"], c["body"])[1] for c in cfgs]
    out = {}
    for mode in ("plain", "warm"):
        sur = surrogate if mode == "warm" else None
        sim = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs])
        det = [SteadyDetector(**STEADY) for _ in pts]
        t0 = time.perf_counter()
        fb, res = start(sim, pts, cfgs, tables, sur, check_steps=check_steps, log=log, detectors=det)
        t_init = time.perf_counter() - t0
        steps_at, wall = run_to_steady(sim, tables, window=window, log=log, tag=mode, retire=True, detectors=det)
        chi_history = [[[round(c, 3) for c in f] for f in d.field_history] for d in det]
        sim.close()
        sim = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs])
        det = [SteadyDetector(**STEADY) for _ in pts]
        t0 = time.perf_counter()
        start(sim, pts, cfgs, tables, sur, check_steps=check_steps, log=None, detectors=det)
        t_init_keep = time.perf_counter() - t0
        steps_keep, wall_keep = run_to_steady(sim, tables, window=window, log=None, tag=mode, detectors=det)
        if not np.array_equal(steps_keep, steps_at):
            raise RuntimeError(f"{mode}: the replay that keeps every instance found other steady steps")
        if mode == "warm":
            steps_at = np.where(steps_at >= 0, steps_at + check_steps, steps_at)   # the check's steps count
        fields = averaged_fields(sim, average_steps)
        ref = None
        if mode == "plain":
            sim.solver(settle_steps)
            ref = averaged_fields(sim, average_steps)
        err = sim.status(-1).err
        out[mode] = dict(steps_to_steady=steps_at.tolist(), wall_to_steady_s=wall + t_init, init_s=t_init,
                         chi_history=chi_history,
                         wall_keep_all_s=wall_keep + t_init_keep,
                         fallback=fb.tolist(), residual=res.tolist(), fields=fields, ref=ref, err=int(err))
        sim.close()
        if log:
            log(f"[surrogate] {mode}: all steady after {int(steps_at.max())} steps, {wall + t_init:.2f} s retiring "
                f"steady instances, {wall_keep + t_init_keep:.2f} s keeping all")
    alone = run_alone(surrogate, pts, cfgs, tables, window=window, check_steps=check_steps)
    same = all(np.array_equal(alone[m]["steps"], out[m]["steps_to_steady"]) for m in ("plain", "warm"))
    ratio = np.array(alone["plain"]["wall"]) / np.array(alone["warm"]["wall"])
    if log:
        log(f"[surrogate] each case alone: wall-clock to steady, median ratio {np.median(ratio):.2f}x, "
            f"total {sum(alone['plain']['wall']):.2f} s -> {sum(alone['warm']['wall']):.2f} s; steps "
            f"{'the same as' if same else 'DIFFERENT from'} the batch's")
    errs = [field_errors(fw, fr, t["frac"].reshape(fw["n"].shape))
            for fw, fr, t in zip(out["warm"]["fields"], out["plain"]["ref"], tables)]
    errs_plain = [field_errors(fp, fr, t["frac"].reshape(fp["n"].shape))
                  for fp, fr, t in zip(out["plain"]["fields"], out["plain"]["ref"], tables)]
    sp = np.array(out["plain"]["steps_to_steady"], float)
    sw = np.array(out["warm"]["steps_to_steady"], float)
    return dict(points=pts,
                steps_plain=sp.tolist(), steps_warm=sw.tolist(),
                median_steps_plain=float(np.median(sp)), median_steps_warm=float(np.median(sw)),
                wall_plain_s=out["plain"]["wall_to_steady_s"], wall_warm_s=out["warm"]["wall_to_steady_s"],
                speedup_wall=out["plain"]["wall_to_steady_s"] / out["warm"]["wall_to_steady_s"],
                init_plain_s=out["plain"]["init_s"], init_warm_s=out["warm"]["init_s"],
                wall_keep_all_plain_s=out["plain"]["wall_keep_all_s"],
                wall_keep_all_warm_s=out["warm"]["wall_keep_all_s"],
                speedup_wall_keep_all=out["plain"]["wall_keep_all_s"] / out["warm"]["wall_keep_all_s"],
                speedup_median_steps=float(np.median(sp) / max(np.median(sw), 1.0)),
                fallback_rate=float(np.mean(out["warm"]["fallback"])),
                residual=out["warm"]["residual"],
                field_error=errs,
                max_field_error=float(max(max(e.values()) for e in errs)),
                plain_field_error=errs_plain,
                max_plain_field_error=float(max(max(e.values()) for e in errs_plain)),
                within_2pct=int(sum(max(e.values()) <= 0.02 for e in errs)),
                within_2pct_plain=int(sum(max(e.values()) <= 0.02 for e in errs_plain)),
                # detector misfires: arms declared steady with a field error above 2 %
                misfires_warm=[i for i, e in enumerate(errs) if max(e.values()) > 0.02],
                misfires_plain=[i for i, e in enumerate(errs_plain) if max(e.values()) > 0.02],
                window=window, average_steps=average_steps, settle_steps=settle_steps, detector=dict(STEADY),
                chi_history_plain=out["plain"]["chi_history"], chi_history_warm=out["warm"]["chi_history"],
                wall_alone_plain_s=alone["plain"]["wall"], wall_alone_warm_s=alone["warm"]["wall"],
                speedup_alone_median=float(np.median(ratio)), alone_steps_match_batch=bool(same),
                speedup_alone_total=float(sum(alone["plain"]["wall"]) / sum(alone["warm"]["wall"])),
                device_err=[out["plain"]["err"], out["warm"]["err"]])


def run_alone(surrogate, pts, cfgs, tables, window=100, check_steps=100):
    """Each case in a batch of its own, plain then warm: wall-clock from the
    start of initialisation (prediction, fill and residual check for the
    warm start) until the detector confirms it steady.  The steps must be
    the ones the same case reached in a shared batch (an instance evolves
    bit-identically alone or among others)."""
    from .batch import Batch
    res = {m: dict(steps=[], wall=[]) for m in ("plain", "warm")}
    for pt, c, t in zip(pts, cfgs, tables):
        for mode in ("plain", "warm"):
            sim = Batch([c["params"]], [c["body"]])
            det = [SteadyDetector(**STEADY)]
            t0 = time.perf_counter()
            start(sim, [pt], [c], [t], surrogate if mode == "warm" else None, check_steps=check_steps, log=None,
                  detectors=det)
            steps_at, _ = run_to_steady(sim, [t], window=window, log=None, detectors=det)
            wall = time.perf_counter() - t0
            sim.close()
            k = int(steps_at[0])
            res[mode]["steps"].append(k + check_steps if (mode == "warm" and k >= 0) else k)
            res[mode]["wall"].append(wall)
    return res


def pipeline(outdir, runs=1024, n_eval=32, fit_steps=8000, steady_steps=4000, sample_steps=2000,
             average_steps=12000, settle_steps=6000, grids="fine", log=print):
    os.makedirs(outdir, exist_ok=True)
    t0 = time.perf_counter()
    train_pts = draw_points(runs, seed=11)
    X, Y = generate(train_pts, steady_steps=steady_steps, sample_steps=sample_steps, log=log)
    t_data = time.perf_counter() - t0
    log(f"[surrogate] {X.shape[0]} training samples from {runs} runs in {t_data:.1f} s")
    sur = Surrogate(seed=3)
    t1 = time.perf_counter()
    losses = sur.fit(X, Y, steps=fit_steps, log=log)
    t_fit = time.perf_counter() - t1
    ms, flop = sur.net.train_stats()
    sur.save(os.path.join(outdir, "surrogate.npz"))
    eval_pts = draw_points(n_eval, seed=97, grids=EVAL_GRIDS[grids])
    res = evaluate(sur, eval_pts, average_steps=average_steps, settle_steps=settle_steps, log=log)
    res["grids"] = grids
    res.update(train_samples=int(X.shape[0]), train_runs=runs, data_s=t_data, fit_s=t_fit,
               fit_steps=fit_steps, fit_loss_first=float(losses[0]), fit_loss_last=float(losses[-1]),
               train_ms_per_step=ms, train_tflops=flop / ms / 1e9 if ms > 0 else 0.0)
    with open(os.path.join(outdir, "warmstart.json"), "w") as f:
        json.dump(res, f, indent=1)
    log(f"[surrogate] speedup (wall) {res['speedup_wall']:.2f}x (keeping all instances "
        f"{res['speedup_wall_keep_all']:.2f}x), median steps {res['median_steps_plain']:.0f} -> "
        f"{res['median_steps_warm']:.0f}, max field error {res['max_field_error']:.4f} "
        f"(plain {res['max_plain_field_error']:.4f}), within 2%: warm {res['within_2pct']}/{n_eval}, "
        f"plain {res['within_2pct_plain']}/{n_eval}, fallback {res['fallback_rate']:.2f}")
    return res


def evaluate_saved(model, outdir, n_eval=32, window=100, grids="fine", average_steps=12000, log=print):
    """The held-out evaluation of a saved surrogate (no data generation)."""
    os.makedirs(outdir, exist_ok=True)
    sur = Surrogate(seed=3)
    sur.load(model)
    eval_pts = draw_points(n_eval, seed=97, grids=EVAL_GRIDS[grids])
    res = evaluate(sur, eval_pts, window=window, average_steps=average_steps, settle_steps=average_steps // 2,
                   log=log)
    res["grids"] = grids
    with open(os.path.join(outdir, "warmstart.json"), "w") as f:
        json.dump(res, f, indent=1)
    log(f"[surrogate] window {window}: speedup (wall) {res['speedup_wall']:.2f}x (keeping all instances "
        f"{res['speedup_wall_keep_all']:.2f}x), median steps "
        f"{res['median_steps_plain']:.0f} -> {res['median_steps_warm']:.0f}, max field error "
        f"{res['max_field_error']:.4f} (plain {res['max_plain_field_error']:.4f}), within 2%: warm "
        f"{res['within_2pct']}/{n_eval}, plain {res['within_2pct_plain']}/{n_eval}, fallback {res['fallback_rate']:.2f}")
    return res


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["pipeline", "evaluate"])
    ap.add_argument("out")
    ap.add_argument("--runs", type=int, default=1024)
    ap.add_argument("--eval", type=int, default=32)
    ap.add_argument("--fit-steps", type=int, default=8000)
    ap.add_argument("--model", default=None, help="evaluate: a surrogate saved by pipeline")
    ap.add_argument("--window", type=int, default=100)
    ap.add_argument("--grids", default="fine", choices=sorted(EVAL_GRIDS))
    ap.add_argument("--threshold", type=float, default=STEADY["threshold"], help="steady-state detector threshold")
    ap.add_argument("--patience", type=int, default=STEADY["patience"])
    ap.add_argument("--lag", type=int, default=STEADY["lag"])
    ap.add_argument("--average", type=int, default=12000, help="evaluate: steps the fields are averaged over")
    a = ap.parse_args()
    STEADY.update(threshold=a.threshold, patience=a.patience, lag=a.lag)
    say = lambda m: print(m, flush=True)
    if a.cmd == "pipeline":
        pipeline(a.out, runs=a.runs, n_eval=a.eval, fit_steps=a.fit_steps, grids=a.grids, log=say)
    else:
        evaluate_saved(a.model, a.out, n_eval=a.eval, window=a.window, grids=a.grids, average_steps=a.average,
                       log=say)
