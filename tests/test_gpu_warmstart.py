"""GPU: the warm-start entry (dsmc_init_fields / Batch.initialize_fields)
places local Maxwellians with the requested moments; the residual check
falls back to freestream for a deliberately wrong prediction and keeps a
sound one; the steady-state runner reports every instance."""
import math

import numpy as np
import pytest

from cudadsmc_amd import cases, surrogate as S
from cudadsmc_amd.batch import Batch
from cudadsmc_amd.run import Run, describe
from cudadsmc_amd.steady import SteadyDetector

pytestmark = pytest.mark.gpu


def test_init_fields_moments(gpu_lib):
    """Left half of the box at 2 n0, 600 K, drift +300 m/s; right half at
    n0, 300 K, at rest: counts, mean velocity and temperature of each half
    match the request within 3 sigma."""
    cfg = cases.box_equilibrium(ppc=20, n_cells=32, seed=5)
    p = cfg["params"]
    r = Run(p, None)
    C = p.N1 * p.N2
    i = np.arange(C) // p.N2
    left = i < p.N1 // 2
    n = np.where(left, 2.0 * p.n_inf, p.n_inf)
    T = np.where(left, 600.0, 300.0)
    ux = np.where(left, 300.0, 0.0)
    r.initialize_fields(n, ux, np.zeros(C), T)
    q = r.download()
    L = p.L1
    k, m = cases.BOLTZ, cases.ARGON["mass"]
    for sel, nn, TT, uu in ((q["x"] < L / 2, 2.0, 600.0, 300.0), (q["x"] >= L / 2, 1.0, 300.0, 0.0)):
        N = int(sel.sum())
        expect = 20 * nn * C / 2
        assert abs(N - expect) < 3 * math.sqrt(expect) + 2, (N, expect)
        sig = math.sqrt(k * TT / m)
        vx = q["vx"][sel].astype(np.float64)
        assert abs(vx.mean() - uu) < 3 * sig / math.sqrt(N)
        var = np.var(q["vz"][sel].astype(np.float64))
        assert abs(var - sig * sig) < 3 * sig * sig * math.sqrt(2.0 / N)
    r.close()


class _RunPredictor:
    """Stands in for a trained surrogate: 'predicts' the fields a plain run
    of the same case reached after settling (a sound prediction)."""

    def __init__(self, fields):
        self.fields = fields

    def fields_for(self, pt, cfg):
        f = self.fields[pt["seed"]]
        p = cfg["params"]
        sampled = f["n"].ravel() > 0
        return (np.where(sampled, f["n"].ravel(), p.n_inf), np.where(sampled, f["ux"].ravel(), p.V_jet),
                np.where(sampled, f["uy"].ravel(), 0.0), np.where(sampled & (f["T"].ravel() > 0), f["T"].ravel(), p.T))

    def fields_batch(self, sim, pts, cfgs):
        return [self.fields_for(pt, c) for pt, c in zip(pts, cfgs)]


def test_residual_check_falls_back_on_a_bad_prediction(gpu_lib):
    """A sound prediction passes the check; the same prediction with the
    flow reversed (forced from the host) falls back to freestream."""
    pts = S.draw_points(2, seed=5, grids=[(64, 32)], ppc=(20.0, 20.0))
    cfgs = [S.case_of(pt) for pt in pts]
    tables = [describe(c["params"], c["body"])[1] for c in cfgs]
    plain = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs])
    plain.initialize()
    plain.solver(1500)
    f = S.averaged_fields(plain, 1000)
    plain.close()
    pred = _RunPredictor({pt["seed"]: f[i] for i, pt in enumerate(pts)})
    sim = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs])
    det = [SteadyDetector(**S.STEADY) for _ in pts]
    fb, res = S.start(sim, pts, cfgs, tables, pred, check_steps=60, force_bad=[1], log=None, detectors=det)
    print(f"[warm] residuals {res}", flush=True)
    assert list(fb) == [False, True]
    assert res[0] < 0.35 and res[1] > 1.0
    assert [d.windows for d in det] == [0, 0]       # the batch was refilled: the check is no window
    assert sim.status(-1).err == 0
    sim.close()
    # without a fallback the check is every detector's first window
    sim = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs])
    det = [SteadyDetector(**S.STEADY) for _ in pts]
    fb, res = S.start(sim, pts, cfgs, tables, pred, check_steps=60, log=None, detectors=det)
    assert not fb.any()
    assert [d.windows for d in det] == [1, 1]
    sim.close()


def test_run_to_steady_reports_every_instance(gpu_lib):
    """Every instance is reported steady; retiring steady instances from the
    batch (Batch.compact) finds the same steps, since the instances left
    evolve bit-identically."""
    pts = S.draw_points(4, seed=8, grids=[(64, 32), (48, 24)], ppc=(10.0, 10.0))
    cfgs = [S.case_of(pt) for pt in pts]
    tables = [describe(c["params"], c["body"])[1] for c in cfgs]
    got = {}
    for retire in (False, True):
        sim = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs])
        sim.initialize()
        got[retire], wall = S.run_to_steady(sim, tables, window=100, max_steps=20000, log=None, retire=retire)
        print(f"[warm] retire={retire}: steps to steady {got[retire]}, {sim.n} left in the batch", flush=True)
        if retire and len(set(got[retire].tolist())) > 1:
            assert sim.n < len(pts)
        sim.close()
    assert np.all(got[False] > 0)
    assert np.array_equal(got[False], got[True])


def test_wall_distance_matches_the_host_formula(gpu_lib):
    """dsmc_batch_wall_distance against the host's vectorised formula on a
    cylinder, a wedge and a body-free box of different grids."""
    pts = S.draw_points(6, seed=21, grids=[(64, 32), (96, 48), (40, 20)])
    pts = [next(p for p in pts if p["geometry"] == g) for g in ("cylinder", "wedge")]
    cfgs = [S.case_of(pt) for pt in pts] + [cases.box_equilibrium(ppc=5, n_cells=8, seed=2)]
    sim = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs[:2]] + [None])
    d = sim.wall_distance()
    cb = sim.cbase
    for i, c in enumerate(cfgs[:2]):
        p = c["params"]
        ii, jj = np.divmod(np.arange(p.N1 * p.N2), p.N2)
        ref = S._signed_distance((ii + 0.5) * (p.L1 / p.N1), (jj + 0.5) * (p.L2 / p.N2), c["body"])
        got = d[cb[i]:cb[i + 1]]
        print(f"[warm] wall distance {pts[i]['geometry']}: max |diff| {np.abs(got - ref).max():.3e}", flush=True)
        assert np.array_equal(np.sign(got), np.sign(ref))
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15 * p.L2)
        assert (got < 0).any() and (got > 0).any()
    assert np.all(np.isinf(d[cb[2]:cb[3]])) and np.all(d[cb[2]:cb[3]] > 0)
    sim.close()


def test_batch_counts_match_per_instance_samples(gpu_lib):
    """dsmc_batch_download_counts is the count row of every instance's
    sample sums, bit for bit, with each instance's sampled-step count; the
    half-window differences the steady runner takes are whole counts."""
    pts = S.draw_points(3, seed=8, grids=[(64, 32), (48, 24)], ppc=(10.0, 10.0))
    cfgs = [S.case_of(pt) for pt in pts]
    sim = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs])
    sim.initialize()
    sim.solver(37)
    c0, n0 = sim.counts()
    sim.solver(25)
    c1, n1 = sim.counts()
    assert n0.tolist() == [37] * 3 and n1.tolist() == [62] * 3
    for i in range(3):
        acc, ns = sim.samples(i)
        assert ns == 62
        assert np.array_equal(c1[sim.cbase[i]:sim.cbase[i + 1]], acc[0])
    diff = c1 - c0
    assert np.all(diff >= 0) and np.array_equal(diff, np.round(diff)) and diff.sum() > 0
    sim.close()


def test_trained_network_beats_freestream_start(gpu_lib):
    """What the learned part gains: a network trained on the GPU by the
    project's own kernels, on fields that batched DSMC runs produced,
    warm-starts 8 held-out cases.  Against the freestream start: the median
    over cases of the step ratio to a confirmed steady state is at least
    1.5, and the warm fields (averaged after steady) differ from the plain
    run's no more than a plain run with other seeds does."""
    grid, window = (96, 48), 100
    train = S.draw_points(32, seed=11, grids=[grid], ppc=(20.0, 20.0))
    X, Y = S.generate(train, steady_steps=2500, sample_steps=1500, batch=32, log=None)
    sur = S.Surrogate(max_batch=16384, seed=3)
    losses = sur.fit(X, Y, steps=1500, batch=16384, lr=2e-3, log=None)
    print(f"[warm] {X.shape[0]} training samples, loss {losses[0]:.3f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.05 * losses[0]
    pts = S.draw_points(8, seed=97, grids=[grid], ppc=(20.0, 20.0))
    other = [dict(p, seed=p["seed"] + 1) for p in pts]
    runs = {}
    for mode, points, net in (("plain", pts, None), ("warm", pts, sur), ("plain, other seeds", other, None)):
        cfgs = [S.case_of(p) for p in points]
        tables = [describe(c["params"], c["body"])[1] for c in cfgs]
        sim = Batch([c["params"] for c in cfgs], [c["body"] for c in cfgs])
        det = [SteadyDetector(**S.STEADY) for _ in points]
        fb, _ = S.start(sim, points, cfgs, tables, net, check_steps=window, log=None, detectors=det)
        steps, _ = S.run_to_steady(sim, tables, window=window, log=None, detectors=det)
        assert not fb.any() and np.all(steps > 0)
        if net is not None:
            steps = steps + window
        runs[mode] = (steps, S.averaged_fields(sim, 3000), tables)
        assert sim.status(-1).err == 0
        sim.close()
    ratio = runs["plain"][0] / runs["warm"][0]
    tables = runs["plain"][2]
    err = lambda a, b: [max(S.field_errors(u, v, t["frac"].reshape(u["n"].shape)).values())
                        for u, v, t in zip(runs[a][1], runs[b][1], tables)]
    e_warm, e_seed = err("warm", "plain"), err("plain, other seeds", "plain")
    print(f"[warm] steps plain {runs['plain'][0].tolist()} warm {runs['warm'][0].tolist()}; "
          f"median ratio {np.median(ratio):.2f}", flush=True)
    print(f"[warm] field gap warm-plain {np.round(e_warm, 4).tolist()}, plain-plain {np.round(e_seed, 4).tolist()}",
          flush=True)
    assert np.median(ratio) >= 1.5
    assert np.median(e_warm) <= 1.5 * np.median(e_seed) + 0.005
