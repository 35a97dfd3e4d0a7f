"""CPU: pins of the oracle (the checker the GPU kernels are compared with)
against closed forms, brute force, library routines and kinetic-theory
moments -- each chosen so that a plausible slip in the restatement (a
transposed index, a dropped weight, a wrong sign, an off-by-one pair count)
fails.  The reference itself cannot run here (CUDA, Eigen absent) and holds
no golden vectors, so these checks are what pins the oracle."""
import math

import numpy as np
import pytest
from scipy import integrate

from cudadsmc_amd import abi, cases
from oracle import oracle as O

import dsmc_testlib as T

K = cases.BOLTZ
M = cases.ARGON["mass"]


def _z(x, sigma):
    return abs(x) / sigma


def _box_case(n_cells=16, L=1.0, face=abi.BC_SPECULAR, face_T=300.0, tau=1e-6, V_jet=0.0, ppc=10.0, seed=3):
    p = cases.make_params(L1=L, L2=L, N1=n_cells, N2=n_cells, n_inf=1e20, T=300.0, V_jet=V_jet,
                          wall_bc=abi.BC_SPECULAR, face_bc=[face] * 4, face_T=[face_T] * 4,
                          tau=tau, ppc=ppc, seed=seed)
    return O.case_from_params(p, None)[:2] + (p,)


def _parts(n, rng, x, y, vx, vy, vz):
    f = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float32), (n,)).copy())
    return dict(x=f(x), y=f(y), vx=f(vx), vy=f(vy), vz=f(vz), pid=np.arange(n, dtype=np.uint32))


# ---------------------------------------------------------------- binning
def test_bin_matches_brute_force():
    """Cell of every particle contains its position; cell = j + i*ny
    (CartesianMesh.cpp:171-181 with nz = 1); counts = histogram; slot = rank
    among the cell's particles in index order (serial atomicAdd order)."""
    p, body = T.small_cylinder(40, 20, 4, 32, seed=5)
    oc, info, _ = O.case_from_params(p, body)
    rng = np.random.default_rng(1)
    n = 20000
    x = (rng.random(n) * info.L1).astype(np.float32)
    y = (rng.random(n) * info.L2).astype(np.float32)
    dead = (rng.random(n) < 0.1).astype(np.uint8)
    cell, count, slot = oc.bin(x, y, dead)
    live = dead == 0
    assert np.all(cell[~live] == -1)
    i, j = cell[live] // p.N2, cell[live] % p.N2
    dx, dy = np.float64(info.dx), np.float64(info.dy)
    assert np.all((x[live] >= i * dx * (1 - 1e-6)) & (x[live] <= (i + 1) * dx * (1 + 1e-6)))
    assert np.all((y[live] >= j * dy * (1 - 1e-6)) & (y[live] <= (j + 1) * dy * (1 + 1e-6)))
    np.testing.assert_array_equal(count, np.bincount(cell[live], minlength=info.ncell))
    order = np.argsort(cell[live], kind="stable")
    rank = np.empty(order.size, np.int64)
    cs = np.concatenate([[0], np.cumsum(np.bincount(cell[live], minlength=info.ncell))])
    rank[order] = np.arange(order.size) - cs[cell[live][order]]
    np.testing.assert_array_equal(slot[live], rank)


def test_bin_edges():
    """x = L1 and y = L2 (inclusive far faces) fall in the last cell."""
    p, body = T.small_cylinder(40, 20, 4, 32, seed=5)
    oc, info, _ = O.case_from_params(p, body)
    x = np.array([0.0, info.L1, info.L1, 0.0], np.float32)
    y = np.array([0.0, info.L2, 0.0, info.L2], np.float32)
    cell, count, _ = oc.bin(x, y, np.zeros(4, np.uint8))
    assert cell.tolist() == [0, info.ncell - 1, (p.N1 - 1) * p.N2, p.N2 - 1]


# ---------------------------------------------------------------- scan / scatter
def test_scan_closed_forms():
    oc, info, _ = _box_case()
    np.testing.assert_array_equal(oc.scan(np.ones(100, np.int32)), np.arange(101))
    np.testing.assert_array_equal(oc.scan(np.zeros(7, np.int32)), np.zeros(8))
    c = np.random.default_rng(2).integers(0, 50, 10007).astype(np.int32)
    np.testing.assert_array_equal(oc.scan(c), np.concatenate([[0], np.cumsum(c)]))


def test_scatter_is_the_stable_cell_sort():
    p, body = T.small_cylinder(40, 20, 4, 32, seed=5)
    oc, info, _ = O.case_from_params(p, body)
    rng = np.random.default_rng(3)
    n = 30000
    x = (rng.random(n) * info.L1).astype(np.float32)
    y = (rng.random(n) * info.L2).astype(np.float32)
    dead = (rng.random(n) < 0.05).astype(np.uint8)
    cell, count, slot = oc.bin(x, y, dead)
    start = oc.scan(count)
    perm = oc.scatter(cell, slot, start, int(count.sum()))
    live = np.flatnonzero(dead == 0)
    np.testing.assert_array_equal(perm, live[np.argsort(cell[live], kind="stable")])


# ---------------------------------------------------------------- move
def test_free_flight_is_ballistic():
    oc, info, _ = _box_case(tau=1e-6)
    rng = np.random.default_rng(4)
    n = 5000
    s = _parts(n, rng, 0.25 + 0.5 * rng.random(n), 0.25 + 0.5 * rng.random(n),
               rng.normal(0, 300, n), rng.normal(0, 300, n), rng.normal(0, 300, n))
    s0 = {k: v.copy() for k, v in s.items()}
    dead, err = oc.move(s, 7)
    assert err == 0 and not dead.any()
    dt = np.float64(info.dt)
    np.testing.assert_allclose(s["x"], s0["x"] + s0["vx"].astype(np.float64) * dt, rtol=0, atol=2e-7)
    np.testing.assert_allclose(s["y"], s0["y"] + s0["vy"].astype(np.float64) * dt, rtol=0, atol=2e-7)
    for k in ("vx", "vy", "vz"):
        np.testing.assert_array_equal(s[k], s0[k])


def test_specular_faces_fold_and_conserve_energy():
    """Specular walls: position folds about the wall (closed form), the
    normal component flips, speed is unchanged bit for bit; multi-bounce
    flights (|v| dt = several box lengths) stay inside."""
    oc, info, _ = _box_case(n_cells=8, L=1e-3, tau=1e-6)
    rng = np.random.default_rng(5)
    n = 4000
    x0 = rng.random(n) * 1e-4
    s = _parts(n, rng, x0, 0.5e-3, -(200 + rng.random(n) * 300), 0.0, 0.0)
    s0 = {k: v.copy() for k, v in s.items()}
    dead, err = oc.move(s, 1)
    assert err == 0 and not dead.any()
    fold = -(s0["x"].astype(np.float64) + s0["vx"].astype(np.float64) * 1e-6)
    np.testing.assert_allclose(s["x"], fold, rtol=0, atol=1e-9)
    np.testing.assert_array_equal(s["vx"], -s0["vx"])
    s = _parts(n, rng, rng.random(n) * 1e-3, rng.random(n) * 1e-3,
               rng.normal(0, 3000, n), rng.normal(0, 3000, n), rng.normal(0, 300, n))
    e0 = s["vx"] ** 2 + s["vy"] ** 2
    for step in range(5):
        dead, err = oc.move(s, step)
        assert err == 0 and not dead.any()
    np.testing.assert_array_equal(s["vx"] ** 2 + s["vy"] ** 2, e0)
    assert np.all((s["x"] >= 0) & (s["x"] <= 1e-3) & (s["y"] >= 0) & (s["y"] <= 1e-3))


def test_diffuse_face_emission_moments():
    """Diffuse wall at T_w: outgoing normal speed has the flux-weighted
    (Rayleigh) law, <v_n> = sqrt(pi k T_w / 2m); tangential and z components
    are N(0, k T_w / m) and independent of v_n (defect D7 in the reference
    reuses one magnitude for all three)."""
    Tw = 500.0
    oc, info, _ = _box_case(n_cells=16, L=1.0, face=abi.BC_DIFFUSE, face_T=Tw, tau=1e-6)
    rng = np.random.default_rng(6)
    n = 200000
    s = _parts(n, rng, 1e-5, 0.5, -100.0, 0.0, 0.0)
    dead, err = oc.move(s, 3)
    assert err == 0 and not dead.any()
    vn, vt, vz = s["vx"].astype(np.float64), s["vy"].astype(np.float64), s["vz"].astype(np.float64)
    sig2 = K * Tw / M
    assert np.all(vn > 0)
    mean_vn = math.sqrt(math.pi * K * Tw / (2 * M))
    sd_vn = math.sqrt(2 * sig2 - mean_vn ** 2)
    assert _z(vn.mean() - mean_vn, sd_vn / math.sqrt(n)) < 4
    assert _z(np.mean(vn ** 2) - 2 * sig2, 2 * sig2 / math.sqrt(n)) < 4     # <v_n^2> = 2kT/m
    for v in (vt, vz):
        assert _z(v.mean(), math.sqrt(sig2 / n)) < 4
        assert _z(np.mean(v * v) - sig2, sig2 * math.sqrt(2 / n)) < 4
    assert abs(np.corrcoef(vn, np.abs(vt))[0, 1]) < 0.02
    # the remaining flight after the hit: x = v_n (dt - 1e-7)
    np.testing.assert_allclose(s["x"], vn * (1e-6 - 1e-7), rtol=2e-4)


# ---------------------------------------------------------------- inflow
def test_inflow_count_sequence_is_exact():
    """Per-step counts floor((s+1)e) - floor(s e): over S steps exactly
    floor(S e) particles enter each stream face."""
    p, body = T.small_cylinder(40, 20, 10, 32, seed=9)
    oc, info, _ = O.case_from_params(p, body)
    tot = np.zeros(4, np.int64)
    S = 500
    for s in range(S):
        J = oc.inject_counts(s)
        assert np.all(np.abs(J - np.asarray(info.inj_e)) < 1.0 + 1e-9)
        tot += J
    for f in range(4):
        assert tot[f] == math.floor(S * info.inj_e[f])


@pytest.mark.parametrize("V_jet", [300.0, 2634.0])
def test_inflow_velocity_moments(V_jet):
    """Particles entering face x = 0: normal velocity with the flux-weighted
    drifting Maxwellian v f(v), v > 0 (Bird 1994 eq. 4.22 integrand; mean by
    quadrature); tangential and z components N(0, kT/m); entry times uniform
    in the step (x / v_x uniform on (0, dt))."""
    oc, info, p = _box_case(n_cells=32, L=1.0, face=abi.BC_STREAM, tau=2e-6, V_jet=V_jet, ppc=2000.0)
    sig = math.sqrt(K * p.T / M)
    vx, vy, vz, frac = [], [], [], []
    nxt, s = 0, 0
    while sum(len(a) for a in vx) < 150000:
        J = oc.inject_counts(s)
        out, dead, err = oc.inject(s, nxt, int(J.sum()))
        assert err == 0
        k = slice(0, int(J[0]))                       # face 0 particles come first
        keep = dead[k] == 0
        vx.append(out["vx"][k][keep]); vy.append(out["vy"][k][keep]); vz.append(out["vz"][k][keep])
        frac.append(out["x"][k][keep] / (out["vx"][k][keep] * np.float32(info.dt)))
        nxt += int(J.sum())
        s += 1
    vx, vy, vz, frac = (np.concatenate(a).astype(np.float64) for a in (vx, vy, vz, frac))
    n = vx.size
    w = lambda v: v * math.exp(-(v - V_jet) ** 2 / (2 * sig * sig))
    hi = max(V_jet, 0) + 10 * sig
    m1 = integrate.quad(lambda v: v * w(v), 0, hi)[0] / integrate.quad(w, 0, hi)[0]
    m2 = integrate.quad(lambda v: v * v * w(v), 0, hi)[0] / integrate.quad(w, 0, hi)[0]
    sd = math.sqrt(m2 - m1 * m1)
    assert np.all(vx > 0)
    assert _z(vx.mean() - m1, sd / math.sqrt(n)) < 4
    assert _z(vx.std() - sd, sd / math.sqrt(2 * n)) < 4
    for v in (vy, vz):
        assert _z(v.mean(), sig / math.sqrt(n)) < 4
        assert _z(np.mean(v * v) - sig * sig, sig * sig * math.sqrt(2 / n)) < 4
    assert _z(frac.mean() - 0.5, math.sqrt(1 / 12 / n)) < 4


# ---------------------------------------------------------------- collisions
def _box_state(ppc=20, n_cells=64, seed=20251203, aniso=None):
    cfg = cases.box_equilibrium(ppc=ppc, n_cells=n_cells, seed=seed)
    oc, info, tables = O.case_from_params(cfg["params"], None)
    parts, count, err = T.oracle_fill(oc, info, tables)
    assert err == 0
    if aniso is not None:
        sx, sz = aniso
        rng = np.random.default_rng(seed)
        n = parts["x"].size
        parts["vx"] = rng.normal(0, sx, n).astype(np.float32)
        parts["vy"] = rng.normal(0, sx, n).astype(np.float32)
        parts["vz"] = rng.normal(0, sz, n).astype(np.float32)
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    return cfg, oc, info, st


def test_collisions_conserve_momentum_and_energy():
    cfg, oc, info, st = _box_state(ppc=20, n_cells=32)
    p0 = st.particles()
    mom0 = [np.sum(p0[k].astype(np.float64)) for k in ("vx", "vy", "vz")]
    e0 = sum(np.sum(p0[k].astype(np.float64) ** 2) for k in ("vx", "vy", "vz"))
    for _ in range(20):
        st.step(nthreads=4)
    assert st.st.total_coll > 10000
    p1 = st.particles()
    scale = math.sqrt(e0 / p0["vx"].size)
    for k, m0 in zip(("vx", "vy", "vz"), mom0):
        # specular walls flip vx / vy, so only vz momentum is a wall invariant
        if k == "vz":
            assert abs(np.sum(p1[k].astype(np.float64)) - m0) < 1e-5 * scale * math.sqrt(st.st.total_coll)
    e1 = sum(np.sum(p1[k].astype(np.float64) ** 2) for k in ("vx", "vy", "vz"))
    assert abs(e1 / e0 - 1) < 1e-5


def test_equilibrium_collision_rate_matches_vhs_theory():
    """Closed box at equilibrium (64x64 cells, 20 per cell, dt = 0.2/nu):
    collisions per step = N nu dt / 2 = 8,192 (Bird eq. 4.64).  A pair count
    of np^2 instead of np(np-1) (+5%) or the reference's Gamma (+45%) fails."""
    cfg, oc, info, st = _box_state()
    n = st.st.n
    steps = 60
    c0 = st.st.total_coll
    for _ in range(steps):
        st.step(nthreads=8)
    rate = (st.st.total_coll - c0) / steps
    theory = 0.5 * n * cfg["nu"] * info.dt
    assert theory == pytest.approx(8192, rel=1e-3)
    assert _z(rate - theory, math.sqrt(theory / steps)) < 4


def test_anisotropic_state_relaxes_to_isotropy():
    """Energy in x, y only (vz = 0) relaxes to equipartition over ~20
    collisions per particle: the scattering is isotropic in 3D."""
    sig = math.sqrt(K * 300.0 / M)
    cfg, oc, info, st = _box_state(ppc=20, n_cells=32, aniso=(sig * math.sqrt(1.5), 0.0))
    for _ in range(120):
        st.step(nthreads=8)
    p = st.particles()
    ex, ey, ez = (np.mean(p[k].astype(np.float64) ** 2) for k in ("vx", "vy", "vz"))
    tot = ex + ey + ez
    assert abs(ez / tot - 1 / 3) < 0.01
    assert abs(ex / tot - 1 / 3) < 0.01


def test_scattering_law_is_isotropic_vhs():
    """One forced collision in each of 65,536 two-particle cells (remainder
    1, (sigma c)max below any pair's, so the single candidate is accepted):
    the post-collision relative velocity keeps |g| and the centre-of-mass
    velocity (Particle.cpp:84-85,99-100) and points uniformly over the
    sphere (cos chi = 2U - 1, phi = 2 pi U; Particle.cpp:87-95).  The law is
    one of the ordered pair (this, other): g' = v_this - v_other.  The pair's
    order is itself drawn at random, so g is oriented here by the first
    particle of the pair, whose index i1 = floor(2 w0 / 2^32) comes from
    the candidate's Philox draw (counter (0, step, collide substream, 0),
    key (cell, seed); the generator is pinned by its known answers).
    Every component of g'/|g'| has mean 0 and mean square 1/3, the
    components are uncorrelated, and the deflection from the incoming g
    has <cos chi> = 0, <cos^2 chi> = 1/3."""
    oc, info, _ = _box_case(n_cells=128)
    C = info.ncell
    start = (2 * np.arange(C + 1)).astype(np.int32)
    perm = np.arange(2 * C, dtype=np.int32)
    rng = np.random.default_rng(17)
    cells = np.arange(C)
    g0s, g1s = [], []
    for step in range(4):
        v = [rng.normal(0, 300, 2 * C).astype(np.float32) for _ in range(3)]
        w = [a.copy() for a in v]
        sm = np.full(C, 1e-30, np.float32)
        mr = np.ones(C, np.float32)
        nc = oc.collide(step, start, perm, *w, sm, mr)
        assert np.all(nc == 1)
        ctr = np.zeros((C, 4), np.uint32)
        ctr[:, 1] = step
        ctr[:, 2] = O.SUB_COLLIDE
        i1 = (O.philox(ctr, cells.astype(np.uint32), info.seed)[:, 0] >> 31).astype(np.int64)
        p1, p2 = 2 * cells + i1, 2 * cells + 1 - i1
        v64 = [a.astype(np.float64) for a in v]
        w64 = [a.astype(np.float64) for a in w]
        for a, b in zip(v64, w64):
            np.testing.assert_allclose(a[p1] + a[p2], b[p1] + b[p2], rtol=0, atol=1e-3)
        g0s.append(np.stack([a[p1] - a[p2] for a in v64], 1))
        g1s.append(np.stack([b[p1] - b[p2] for b in w64], 1))
    g0, g1 = np.concatenate(g0s), np.concatenate(g1s)
    n = g0.shape[0]
    m0, m1 = np.linalg.norm(g0, axis=1), np.linalg.norm(g1, axis=1)
    np.testing.assert_allclose(m1, m0, rtol=3e-6)
    d = g1 / m1[:, None]
    for k in range(3):
        assert _z(d[:, k].mean(), math.sqrt(1 / 3 / n)) < 4, k
        assert _z(np.mean(d[:, k] ** 2) - 1 / 3, math.sqrt(4 / 45 / n)) < 4, k
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert _z(np.mean(d[:, a] * d[:, b]), math.sqrt(1 / 15 / n)) < 4, (a, b)
    cosx = np.sum(g0 * g1, 1) / (m0 * m1)
    assert _z(cosx.mean(), math.sqrt(1 / 3 / n)) < 4
    assert _z(np.mean(cosx ** 2) - 1 / 3, math.sqrt(4 / 45 / n)) < 4


# ---------------------------------------------------------------- sampling / fill
def test_sample_sums_match_numpy():
    oc, info, _ = _box_case(n_cells=16)
    rng = np.random.default_rng(8)
    n = 50000
    cell = rng.integers(0, info.ncell, n)
    order = np.argsort(cell, kind="stable").astype(np.int32)
    count = np.bincount(cell, minlength=info.ncell).astype(np.int32)
    start = oc.scan(count)
    v = [rng.normal(0, 400, n).astype(np.float32) for _ in range(3)]
    ncoll = rng.integers(0, 9, info.ncell).astype(np.int32)
    acc = np.zeros(6 * info.ncell)
    oc.sample(start, order, *v, ncoll, acc)
    oc.sample(start, order, *v, ncoll, acc)        # accumulates
    acc = acc.reshape(6, -1)
    v64 = [a.astype(np.float64) for a in v]
    want = [np.bincount(cell, minlength=info.ncell).astype(np.float64)]
    want += [np.bincount(cell, weights=a, minlength=info.ncell) for a in v64]
    want += [np.bincount(cell, weights=sum(a * a for a in v64), minlength=info.ncell), ncoll.astype(np.float64)]
    for k in range(6):
        np.testing.assert_allclose(acc[k], 2 * want[k], rtol=1e-12, atol=1e-9)


def _outside(px, py, poly):
    inside = np.zeros(px.shape, bool)
    for k in range(len(poly)):
        ax, ay = poly[k]
        bx, by = poly[(k + 1) % len(poly)]
        cross = (ay > py) != (by > py)
        xint = ax + (py - ay) * (bx - ax) / np.where(by != ay, by - ay, 1.0)
        inside ^= cross & (px < xint)
    return ~inside


def test_fill_counts_positions_and_moments():
    """Freestream fill: count per cell = floor(e + U) (mean e), every
    particle inside its own cell and outside the body, velocities drifting
    Maxwellians with the requested mean and variance."""
    p, body = T.small_cylinder(80, 40, 30, 64, seed=12)
    oc, info, tables = O.case_from_params(p, body)
    parts, count, err = T.oracle_fill(oc, info, tables)
    assert err == 0
    e = info.ppc_eff * tables["frac"]
    assert np.all((count == np.floor(e)) | (count == np.floor(e) + 1))
    assert _z(count.sum() - e.sum(), math.sqrt(np.sum((e % 1) * (1 - e % 1))) + 1e-9) < 4
    cell_of_particle = np.repeat(np.arange(info.ncell), count)
    cell, _, _ = oc.bin(parts["x"], parts["y"], np.zeros(parts["x"].size, np.uint8))
    # x0 + w*u with u < 1 can round up onto the far edge, which bins into the
    # neighbour: allowed only for points on the shared edge
    bad = np.flatnonzero(cell != cell_of_particle)
    assert bad.size <= 4
    for q in bad:
        i, j = divmod(int(cell_of_particle[q]), p.N2)
        on_edge = (np.float32(parts["x"][q]) == np.float32((i + 1) * info.dx) or
                   np.float32(parts["y"][q]) == np.float32((j + 1) * info.dy))
        assert on_edge, q
    assert np.all(_outside(parts["x"].astype(np.float64), parts["y"].astype(np.float64), body))
    n = parts["x"].size
    sig = float(info.inj_sig)
    assert _z(parts["vx"].mean() - info.inj_ux, sig / math.sqrt(n)) < 4
    assert _z(parts["vy"].mean(), sig / math.sqrt(n)) < 4
    for k in ("vy", "vz"):
        assert _z(np.var(parts[k].astype(np.float64)) - sig * sig, sig * sig * math.sqrt(2 / n)) < 4
    # the three components are independent normals (Random.cpp:30-46 draws
    # each on its own; a Box-Muller pair gives two independent normals)
    vs = [parts[k].astype(np.float64) for k in ("vx", "vy", "vz")]
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert abs(np.corrcoef(vs[a], vs[b])[0, 1]) < 4 / math.sqrt(n), (a, b)
        assert abs(np.corrcoef(vs[a] ** 2, vs[b] ** 2)[0, 1]) < 4 / math.sqrt(n), (a, b)


def test_no_particle_enters_the_body():
    """100 full steps of the Mach-10 cylinder (coarse grid): every particle
    stays outside the polygon (reference defect D8 lets particles pass
    through cells other than the pre-move cell)."""
    p, body = T.small_cylinder(80, 40, 10, 64, seed=14)
    oc, info, tables = O.case_from_params(p, body)
    parts, count, err = T.oracle_fill(oc, info, tables)
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    for _ in range(100):
        st.step(nthreads=8)
    assert st.st.err == 0
    q = st.particles()
    assert np.all(_outside(q["x"].astype(np.float64), q["y"].astype(np.float64), body))
    assert np.all((q["x"] >= 0) & (q["x"] <= info.L1) & (q["y"] >= 0) & (q["y"] <= info.L2))


@pytest.mark.parametrize("index", [3, 54])
def test_fill_of_thin_fluid_slivers(index):
    """Wedge instances of the batched sweep whose back-face cells keep 1-3 %
    of their fluid bounding box as an L-shaped sliver: the sub-box masked
    sampler places every particle in the fluid (plain rejection in the
    bounding box ran out of tries here)."""
    c = cases.sweep(64)[index]
    oc, info, tables = O.case_from_params(c["params"], c["body"])
    cut = tables["cut"] == 1
    e = np.where(cut, info.ppc_eff * tables["frac"], 0.0)
    ux = np.full(info.ncell, info.inj_ux, np.float32)
    uy = np.full(info.ncell, info.inj_uy, np.float32)
    sg = np.full(info.ncell, info.inj_sig, np.float32)
    parts, count, err = oc.init_fill(e, tables["cut"], ux, uy, sg, int(e.sum() * 2 + 1000))
    assert err == 0
    assert np.all(_outside(parts["x"].astype(np.float64), parts["y"].astype(np.float64), c["body"]))



def test_acute_vertex_does_not_trap_a_particle():
    """A particle that hits a 17 deg wedge 0.6 um from the apex, where the
    body is thinner than the wall band: the re-emitted particle leaves with
    one wall event instead of bouncing between the two faces' bands until
    the 64-event limit (found in a batched sweep run)."""
    from cudadsmc_amd import surrogate as S
    pt = {"geometry": "wedge", "half_angle_deg": 17.054940866729332, "mach": 5.867278090943784,
          "kn": 0.02070981818448305, "tw_ratio": 3.3452777555517583, "grid": (128, 64),
          "ppc": 30.000000000000004, "seed": 3175139750848533173}
    cfg = S.case_of(pt, sample_from=4000)
    oc, info, _ = O.case_from_params(cfg["params"], cfg["body"])
    s = dict(x=np.array([0.14987538754940033], np.float32), y=np.array([0.15022362768650055], np.float32),
             vx=np.array([790.2767944335938], np.float32), vy=np.array([-1410.759033203125], np.float32),
             vz=np.array([80.26766967773438], np.float32), pid=np.array([422860], np.uint32))
    dead, err = oc.move(s, 667)
    assert err == 0 and not dead[0]
    assert np.all(_outside(s["x"].astype(np.float64), s["y"].astype(np.float64), cfg["body"]))
