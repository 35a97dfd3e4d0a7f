"""CPU: pins of the oracle's adaptive collision sub-cells and variable time
step -- the restatement of Cell::genAMRmesh (Cell.cpp:168-279), Cell::VTS
(Cell.cpp:117-136), Element::genAMRmesh (Element.cpp:158-182) and
Cell::collision over the leaves (Cell.cpp:335-352).  The reference defines
these on the host and never calls them, so they hold no outputs to compare
with: the grid rule is pinned by hand-worked cases and by its defining
inequalities, the mean free path and time step by their closed forms, the
leaf binning by brute force, and the collisions by conservation within every
leaf, a stable leaf order, the carried remainder and the equilibrium
collision rate (which sub-cells leave unchanged: with N particles shared by
L leaves of volume V/L, E[sum_l N_l (N_l - 1)] L / V = N (N - 1) / V)."""
import math

import numpy as np
import pytest

from cudadsmc_amd import cases
from oracle import oracle as O

import dsmc_testlib as T

K = cases.BOLTZ
M = cases.ARGON["mass"]


def pack(n2x, n2y, n3x=1, n3y=1):
    return n2x | (n2y << 8) | (n3x << 16) | (n3y << 24)


def unpack(s):
    return s & 0xff, (s >> 8) & 0xff, (s >> 16) & 0xff, (s >> 24) & 0xff


# ---------------------------------------------------------------- grid rule
@pytest.mark.parametrize("dx,dy,lam_max,lam_min,had,want", [
    # side lam/2 >= the cell: one sub-cell (int(L/side) = 0 -> 1)
    (1.0, 1.0, 2.0, 2.0, 0, (1, 1, 1, 1)),
    (1.0, 1.0, 5.0, 5.0, 1, (1, 1, 1, 1)),
    # L/side an integer: exactly that many (the reference's exact branch)
    (1.0, 1.0, 1.0, 1.0, 0, (2, 2, 1, 1)),
    (1.0, 1.0, 0.5, 0.5, 0, (4, 4, 1, 1)),
    # not an integer: ceil (1 / 0.35 = 2.86 -> 3)
    (1.0, 1.0, 0.7, 0.7, 0, (3, 3, 1, 1)),
    # a non-square cell: level 2 per axis (2 / 0.35 = 5.7 -> 6); without an
    # old grid level 3 comes from twice the level-2 x side (2/3 -> side 1/3)
    (1.0, 2.0, 0.7, 0.7, 0, (3, 6, 1, 1)),
    # with an old grid, from the smallest old mean free path: level-2 side
    # 1/3, level-3 side 0.175/2 = 0.0875 -> 3.81 -> 4 in both axes
    (1.0, 1.0, 0.7, 0.175, 1, (3, 3, 4, 4)),
    # capped at 16 per level and axis
    (1.0, 1.0, 0.01, 0.01, 0, (16, 16, 1, 1)),
    (1.0, 1.0, 1.0, 0.001, 1, (2, 2, 16, 16)),
])
def test_grid_rule_hand_worked(dx, dy, lam_max, lam_min, had, want):
    assert unpack(O.amr_grid(dx, dy, lam_max, lam_min, had)) == want


def test_grid_rule_inequalities():
    """Below the cap, n level-2 sub-cells of side L/n are no larger than
    lam/2, and n - 1 would be larger; with an old grid and lam_min = lam_max
    level 3 is one cell unless level 2 was capped, since the level-2 side is
    then already at most lam/2."""
    rng = np.random.default_rng(5)
    for _ in range(2000):
        dx, dy = rng.uniform(0.5, 2.0, 2)
        lam = math.exp(rng.uniform(math.log(0.08), math.log(6.0)))
        had = int(rng.integers(0, 2))
        n2x, n2y, n3x, n3y = unpack(O.amr_grid(dx, dy, lam, lam, had))
        for L, n in ((dx, n2x), (dy, n2y)):
            assert 1 <= n <= O.AMR_MAX
            if n < O.AMR_MAX:
                assert L / n <= 0.5 * lam * (1 + 1e-12)
            if n > 1:
                assert L / (n - 1) > 0.5 * lam
        l2x, l2y = dx / n2x, dy / n2y
        div = lambda L, side: min(O.AMR_MAX, max(1, math.ceil(L / side - 1e-12)))
        if had:
            # level-3 side lam/2: one cell unless level 2 was capped
            assert (n3x, n3y) == (div(l2x, 0.5 * lam), div(l2y, 0.5 * lam))
            if n2x < O.AMR_MAX and n2y < O.AMR_MAX:
                assert (n3x, n3y) == (1, 1)
        else:
            # no old grid: level 3 from twice the level-2 x side, in both axes
            # (the child's L1 at Cell.cpp:212): one cell along x, and along y
            # as many as the y side holds x sides
            assert n3x == 1
            assert n3y == div(l2y, l2x)


# ---------------------------------------------------- mean free path, VTS
def _gas(p, info):
    return dict(Fn=info.Fn, mass=p.mass, diam=p.diam, T_inf=p.T, omega=p.Vtl, boltz=p.boltz)


def _uniform_sums(C, N, n_per_sample, u, T, rng=None):
    """Sampling sums of a gas of fixed moments: N samples of n_per_sample
    particles a cell, velocity u, temperature T."""
    S0 = np.full(C, float(N * n_per_sample))
    acc = np.zeros(6 * C)
    acc[:C] = S0
    for k in range(3):
        acc[(k + 1) * C:(k + 2) * C] = S0 * u[k]
    acc[4 * C:5 * C] = S0 * (3 * K * T / M + sum(x * x for x in u))
    return acc


def test_mean_free_path_and_vts_closed_forms():
    cfg = cases.box_equilibrium(ppc=20, n_cells=16)
    p = cfg["params"]
    oc, info, tables = O.case_from_params(p, None)
    C = oc.ncell
    vol = np.asarray(tables["vol"], np.float64)
    u, Tg, N = (1200.0, -300.0, 50.0), 640.0, 10
    n_per = 37.0
    acc = _uniform_sums(C, N, n_per, u, Tg)
    sub, lam, dt = oc.amr_update(_gas(p, info), acc, N, vol, np.zeros(C, np.uint32))
    n = n_per * info.Fn / vol
    lam_th = 1.0 / (math.sqrt(2.0) * math.pi * p.diam ** 2 * n) * (Tg / p.T) ** (p.Vtl - 0.5)
    np.testing.assert_allclose(lam, lam_th, rtol=2e-6)
    umag = math.sqrt(sum(x * x for x in u))
    dt_th = 0.5 * lam_th / (umag + math.sqrt(2 * K * Tg / M))
    np.testing.assert_allclose(dt, dt_th, rtol=2e-6)
    # the output's MFP dataset (Cell.cpp:125) is the same quantity
    from cudadsmc_amd import output
    f = dict(n=n, ux=np.full(C, u[0]), uy=np.full(C, u[1]), uz=np.full(C, u[2]), T=np.full(C, Tg),
             p=n * K * Tg)
    np.testing.assert_allclose(output.datasets(f, p)["MFP"].ravel(), lam_th, rtol=1e-12)
    for c in range(C):
        assert sub[c] == O.amr_grid(info.dx, info.dy, float(lam[c]), float(lam[c]), 0)
    # a second update with the grid present: level 3 from lam, still one
    sub2, _, _ = oc.amr_update(_gas(p, info), acc, N, vol, sub)
    assert all(unpack(int(s))[2:] == (1, 1) for s in sub2)
    assert all(unpack(int(s))[:2] == unpack(int(t))[:2] for s, t in zip(sub, sub2))


def test_cells_without_gas_get_no_grid():
    cfg = cases.box_equilibrium(ppc=20, n_cells=16)
    p = cfg["params"]
    oc, info, tables = O.case_from_params(p, None)
    C = oc.ncell
    vol = np.asarray(tables["vol"], np.float64)
    acc = _uniform_sums(C, 10, 5.0, (0.0, 0.0, 0.0), 300.0)
    acc[:C:3] = 0.0                                 # empty cells
    acc[4 * C + 1:5 * C:7] = 0.0                    # T <= 0
    old = np.full(C, pack(2, 2), np.uint32)
    sub, lam, dt = oc.amr_update(_gas(p, info), acc, 10, vol, old)
    assert np.all(sub[::3] == 0) and np.all(lam[::3] == 0) and np.all(dt[::3] == 0)
    assert np.all(sub[1::7] == 0)
    _, _, _ = oc.amr_update(_gas(p, info), acc, 0, vol, old)
    assert np.all(oc.amr_update(_gas(p, info), acc, 0, vol, old)[0] == 0)


# ------------------------------------------------------------ leaf binning
def test_leaf_binning_brute_force():
    cfg = cases.box_equilibrium(ppc=20, n_cells=16)
    oc, info, _ = O.case_from_params(cfg["params"], None)
    rng = np.random.default_rng(11)
    dx, dy, ny = float(info.dx), float(info.dy), cfg["params"].N2
    checked = 0
    for _ in range(4000):
        c = int(rng.integers(0, oc.ncell))
        i, j = c // ny, c % ny
        nsx, nsy = int(rng.integers(1, 17)), int(rng.integers(1, 17))
        fx, fy = rng.uniform(0, 1, 2)
        x, y = np.float32((i + fx) * dx), np.float32((j + fy) * dy)
        gx, gy = (float(x) / dx - i) * nsx, (float(y) / dy - j) * nsy
        if min(abs(gx - round(gx)), abs(gy - round(gy))) < 1e-3:
            continue                                # FP32 rounding decides at a boundary
        want = min(max(int(math.floor(gy)), 0), nsy - 1) + min(max(int(math.floor(gx)), 0), nsx - 1) * nsy
        assert oc.amr_leaf(c, x, y, nsx, nsy) == want
        checked += 1
    assert checked > 3000


# -------------------------------------------------------------- collisions
def _binned_box(grid, ppc=24, n_cells=8, seed=4, drift=None):
    """A box state binned into cells, sub-cells set on the oracle case."""
    cfg = cases.box_equilibrium(ppc=ppc, n_cells=n_cells, seed=seed)
    p = cfg["params"]
    oc, info, tables = O.case_from_params(p, None)
    parts, _, err = T.oracle_fill(oc, info, tables)
    assert err == 0
    oc.set_subcells(np.asarray(grid, np.uint32) if np.ndim(grid) else np.full(oc.ncell, grid, np.uint32))
    cell, count, slot = oc.bin(parts["x"], parts["y"])
    start = oc.scan(count)
    perm = oc.scatter(cell, slot, start, int(start[-1]))
    return cfg, oc, info, parts, cell, start, perm


def _leaf_of(oc, info, parts, cell, grid_of_cell):
    out = np.empty(parts["x"].size, np.int64)
    for k in range(out.size):
        _, nsx, nsy = O.amr_leaves(int(grid_of_cell(cell[k])))
        out[k] = oc.amr_leaf(int(cell[k]), parts["x"][k], parts["y"][k], nsx, nsy)
    return out


def test_collisions_stay_in_their_leaf():
    """2 x 2 leaves whose particles drift apart (+-U in x by leaf): every
    leaf keeps its momentum and energy; one pair across leaves would move
    2U of momentum between them."""
    grid = pack(2, 2)
    cfg, oc, info, parts, cell, start, perm = _binned_box(grid, ppc=40)
    leaf = _leaf_of(oc, info, parts, cell, lambda c: grid)
    U = 2000.0
    vx = (parts["vx"] + np.where(leaf % 2 == 0, U, -U)).astype(np.float32)
    vy, vz = parts["vy"].copy(), parts["vz"].copy()
    key = cell.astype(np.int64) * 4 + leaf
    m0 = {k: np.bincount(key, weights=v.astype(np.float64), minlength=4 * oc.ncell)
          for k, v in (("x", vx), ("y", vy), ("z", vz))}
    e0 = np.bincount(key, weights=(vx.astype(np.float64) ** 2 + vy ** 2 + vz ** 2), minlength=4 * oc.ncell)
    sm = np.full(oc.ncell, info.sigcmax0 * 4, np.float32)
    mr = np.zeros(oc.ncell, np.float32)
    total = 0
    for step in range(40):
        nc = oc.collide(step, start, perm, vx, vy, vz, sm, mr, x=parts["x"], y=parts["y"])
        total += int(nc.sum())
    assert total > 3000
    for k, v in (("x", vx), ("y", vy), ("z", vz)):
        m1 = np.bincount(key, weights=v.astype(np.float64), minlength=4 * oc.ncell)
        assert np.max(np.abs(m1 - m0[k])) < 0.5, k
    e1 = np.bincount(key, weights=(vx.astype(np.float64) ** 2 + vy ** 2 + vz ** 2), minlength=4 * oc.ncell)
    np.testing.assert_allclose(e1, e0, rtol=2e-5)


def test_leaf_order_is_stable():
    rng = np.random.default_rng(3)
    grids = np.array([pack(int(a), int(b), int(c), int(d)) for a, b, c, d in
                      rng.integers(1, 4, size=(64, 4))], np.uint32)
    cfg, oc, info, parts, cell, start, perm = _binned_box(grids, ppc=30)
    before = perm.copy()
    vx, vy, vz = parts["vx"].copy(), parts["vy"].copy(), parts["vz"].copy()
    sm = np.full(oc.ncell, info.sigcmax0, np.float32)
    mr = np.zeros(oc.ncell, np.float32)
    oc.collide(0, start, perm, vx, vy, vz, sm, mr, x=parts["x"], y=parts["y"])
    leaf = _leaf_of(oc, info, parts, cell, lambda c: grids[c])
    for c in range(oc.ncell):
        seg = before[start[c]:start[c + 1]]
        nl = O.amr_leaves(int(grids[c]))[0]
        want = seg[np.argsort(leaf[seg], kind="stable")] if 1 < nl <= O.AMR_LEAVES else seg
        np.testing.assert_array_equal(perm[start[c]:start[c + 1]], want)


def test_candidate_remainder_runs_through_the_leaves():
    """The remainder after a cell's leaves is that of the sum of the leaves'
    expected candidate numbers, each with the leaf volume: r' = frac(r +
    sum_l N_l (N_l - 1) Fn dt/2 (sigma c)max / (V / L))."""
    grid = pack(3, 2)
    cfg, oc, info, parts, cell, start, perm = _binned_box(grid, ppc=36)
    leaf = _leaf_of(oc, info, parts, cell, lambda c: grid)
    vx, vy, vz = parts["vx"].copy(), parts["vy"].copy(), parts["vz"].copy()
    sm = np.full(oc.ncell, info.sigcmax0, np.float32)
    r0 = np.random.default_rng(1).uniform(0, 1, oc.ncell).astype(np.float32)
    mr = r0.copy()
    oc.collide(0, start, perm, vx, vy, vz, sm, mr, x=parts["x"], y=parts["y"])
    vleaf = np.asarray(oc.t["vol_ntc"], np.float64) / 6
    for c in range(oc.ncell):
        nl = np.bincount(leaf[cell == c], minlength=6).astype(np.float64)
        e = np.sum(np.where(nl >= 2, nl * (nl - 1), 0.0)) * float(info.ntc_coef) * float(info.sigcmax0) / vleaf[c]
        want = (float(r0[c]) + e) % 1.0
        d = abs(float(mr[c]) - want)
        assert min(d, 1 - d) < 2e-3, c


def test_equilibrium_collision_rate_with_subcells():
    """The closed box at equilibrium (64x64 cells, 20 per cell) with 2 x 2
    leaves in every cell collides at the VHS rate, N nu dt / 2 = 8,192 per
    step, as without sub-cells.  Leaf volumes equal to the cell's would give
    a quarter of it."""
    cfg = cases.box_equilibrium(ppc=20, n_cells=64)
    oc, info, tables = O.case_from_params(cfg["params"], None)
    parts, _, err = T.oracle_fill(oc, info, tables)
    oc.set_subcells(np.full(oc.ncell, pack(2, 2), np.uint32))
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    n = st.st.n
    for _ in range(5):
        st.step(nthreads=8)
    steps = 50
    c0 = st.st.total_coll
    for _ in range(steps):
        st.step(nthreads=8)
    rate = (st.st.total_coll - c0) / steps
    theory = 0.5 * n * cfg["nu"] * info.dt
    assert abs(rate - theory) / math.sqrt(theory / steps) < 4
