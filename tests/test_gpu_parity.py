"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on
identical inputs.  Integer results (cell ids, counts, scan offsets, the
stable permutation) must be bit-exact; because both sides follow the same
FP32 arithmetic contract, particle states, (sigma c_r)max and remainders are
required bit-exact too.  Sampling sums (FP64, different summation order)
must agree to 1e-12 relative."""
import numpy as np
import pytest

from cudadsmc_amd import abi, cases, run
from cudadsmc_amd.run import Run
from oracle import oracle as O

import dsmc_testlib as T

pytestmark = pytest.mark.gpu

def _wedge():
    w = cases.flow_case("wedge", mach=6.0, kn=0.02, tw_ratio=3.0, nx=64, ny=32, ppc=9.0, half_angle_deg=25.0,
                        seed=21)
    return w["params"], w["body"]


def _box_diffuse():
    """Closed box whose four faces are diffuse walls at 200, 300, 400 and
    500 K (the face-wall re-emission path of the move)."""
    p, body = T.small_box(16, 8, seed=9)
    for f, Tf in enumerate((200.0, 300.0, 400.0, 500.0)):
        p.face_bc[f] = abi.BC_DIFFUSE
        p.face_T[f] = Tf
    return p, body


CASES = {
    "box16": lambda: T.small_box(16, 8, seed=7),
    "cyl80": lambda: T.small_cylinder(80, 40, 10, 64, seed=11),
    "cyl80_specular": lambda: T.small_cylinder(80, 40, 10, 64, seed=13, wall=abi.BC_SPECULAR),
    "wedge64": _wedge,
    "box16_diffuse": _box_diffuse,
}


def _make(name):
    p, body = CASES[name]()
    print(f"[case] {name} N1={p.N1} N2={p.N2} ppc={p.ppc} seed={p.seed}", flush=True)
    oc, info, tables = O.case_from_params(p, body)
    return p, body, oc, info, tables


@pytest.mark.parametrize("name", list(CASES))
def test_fill_matches_oracle(gpu_lib, name):
    p, body, oc, info, tables = _make(name)
    r = Run(p, body)
    r.initialize()
    g = r.download()
    o, count, err = T.oracle_fill(oc, info, tables)
    assert err == 0 and r.status().err == 0
    T.assert_same_particles(g, o, what="fill ")
    r.close()


@pytest.mark.parametrize("name", list(CASES))
def test_phase_parity(gpu_lib, name):
    """move / reassign / collide each against the oracle, bit-exact."""
    p, body, oc, info, tables = _make(name)
    r = Run(p, body)
    r.initialize()
    r.set_keep_perm(True)
    for step in range(3):
        cur = r.download()
        sm, mr, _ = r.cell_state()
        st = r.status()
        ref = T.oracle_phases(oc, cur, step, st.next_pid, sm, mr, int(info.capacity))
        r.particlemove()
        moved = r.download(moved=True)
        T.assert_same_particles(moved, ref["moved"], what=f"step {step} move ")
        cell, count = r.bins()
        np.testing.assert_array_equal(cell, ref["cell"])
        np.testing.assert_array_equal(count, ref["count"])
        r.ressignParticle()
        np.testing.assert_array_equal(r.start(), ref["start"])
        r.collision()
        np.testing.assert_array_equal(r.perm(), ref["perm"])
        T.assert_same_particles(r.download(), ref["new"], what=f"step {step} collide ")
        sm2, mr2, nc2 = r.cell_state()
        np.testing.assert_array_equal(sm2, ref["sigcmax"])
        np.testing.assert_array_equal(mr2, ref["mrem"])
        np.testing.assert_array_equal(nc2, ref["ncoll"])
        s2 = r.status()
        assert s2.step == step + 1 and s2.err == 0
        assert s2.last_collisions == int(ref["ncoll"].sum())
    r.close()


@pytest.mark.parametrize("name,steps", [("box16", 30), ("cyl80", 30), ("wedge64", 60), ("box16_diffuse", 40)])
def test_trajectory_bitexact(gpu_lib, name, steps):
    """Many full steps on both sides from the same fill: identical states."""
    p, body, oc, info, tables = _make(name)
    p.sample_start = 5
    r = Run(p, body)
    r.initialize()
    r.solver(steps, graph=False)
    g = r.download()
    acc_g, ns_g = r.samples()
    parts, _, _ = T.oracle_fill(oc, info, tables)
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    for s in range(steps):
        st.step(sample=s >= 5, nthreads=4)
    o = st.particles()
    T.assert_same_particles(g, o, what=f"{name} after {steps} steps ")
    s = r.status()
    assert s.collisions == st.st.total_coll
    assert s.err == 0 and st.st.err == 0
    assert ns_g == steps - 5
    acc_o = st.acc.reshape(6, -1)
    np.testing.assert_allclose(acc_g, acc_o, rtol=1e-12, atol=1e-9)
    r.close()


def test_graph_equals_eager(gpu_lib):
    p, body = T.small_cylinder(80, 40, 10, 64, seed=21)
    p.sample_start = 0
    a, b = Run(p, body), Run(p, body)
    a.initialize()
    b.initialize()
    a.solver(37, graph=False)
    b.solver(37, graph=True)
    T.assert_same_particles(a.download(), b.download(), what="graph vs eager ")
    np.testing.assert_array_equal(a.samples()[0], b.samples()[0])
    a.close()
    b.close()


def _graph_vs_oracle(r, st, steps, sample_from, what):
    """Run `steps` through hipGraph replays (and the single steps that finish
    a count not divisible by the replay length) and as many oracle steps;
    compare particles, cell state, samples and counters exactly."""
    r.solver(steps, graph=True)
    for s in range(steps):
        st.step(sample=sample_from is not None and s >= sample_from, nthreads=4)
    T.assert_same_particles(r.download(), st.particles(), what=what)
    s = r.status()
    assert s.step == st.st.step and s.n_live == st.st.n and s.next_pid == st.st.next_pid, what
    assert s.collisions == st.st.total_coll, what
    assert s.err == 0 and st.st.err == 0, what
    sm_g, mr_g, nc_g = r.cell_state()
    np.testing.assert_array_equal(sm_g, st.sigcmax)
    np.testing.assert_array_equal(mr_g, st.mrem)
    np.testing.assert_array_equal(nc_g, st.ncoll)
    if sample_from is not None:
        acc_g, ns_g = r.samples()
        assert ns_g == steps - sample_from
        np.testing.assert_allclose(acc_g, st.acc.reshape(6, -1), rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("name,steps", [("cyl80", 27), ("wedge64", 35), ("box16_diffuse", 19)])
def test_graph_trajectory(gpu_lib, name, steps):
    """Graph replays on cases with inflow, outflow and specular or diffuse
    body walls: particles, cell state, samples and counters bit-exact
    against the oracle."""
    p, body, oc, info, tables = _make(name)
    p.sample_start = 3
    r = Run(p, body)
    r.initialize()
    parts, _, _ = T.oracle_fill(oc, info, tables)
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    print(f"[parity] graph {name}, {steps} steps", flush=True)
    _graph_vs_oracle(r, st, steps, 3, f"graph {name} ")
    r.close()


@pytest.mark.parametrize("tile", [2048, 1024, 256])
def test_graph_dense_cells(gpu_lib, tile):
    """Graph replays over dense cells: cells of 255 to 900 particles (each
    beyond a 256-particle tile), a 2,500-particle cell that stages in HBM,
    tiles cut into LDS runs (2,048-particle tiles) and sparse cells, over 19
    steps (two 8-step replays and 3 single steps): particles and collision
    state bit-exact against the oracle."""
    p, body = T.small_box(8, 1, seed=4)
    p.n_inf = 1.35e16
    oc, info, tables = O.case_from_params(p, body)
    rng = np.random.default_rng(13)
    dx = np.float32(info.dx)
    xs, ys = [], []
    for j, cnt in ((0, 40), (1, 300), (2, 260), (3, 20), (4, 900), (5, 2500), (6, 256), (7, 3)):
        xs.append((rng.random(cnt, dtype=np.float32) * dx * 0.99).astype(np.float32) + 2 * dx)
        ys.append((rng.random(cnt, dtype=np.float32) * dx * 0.99).astype(np.float32) + j * dx)
    n = sum(len(v) for v in xs)
    parts = dict(x=np.concatenate(xs), y=np.concatenate(ys),
                 vx=rng.normal(0, 1e-3, n).astype(np.float32), vy=rng.normal(0, 1e-3, n).astype(np.float32),
                 vz=rng.normal(0, 300, n).astype(np.float32), pid=np.arange(n, dtype=np.uint32))
    r = Run(p, body)
    r.set_launch(cell_tile=tile, graph_steps=8)
    r.upload(*[parts[k] for k in ("x", "y", "vx", "vy", "vz")], parts["pid"])
    sm, _, _ = r.cell_state()
    st = O.OracleState(oc, parts, int(info.capacity), sm)
    print(f"[parity] graph dense cells, tile {tile}", flush=True)
    _graph_vs_oracle(r, st, 19, None, f"graph dense tile {tile} ")
    counts = np.diff(st.start)
    assert (counts >= 256).sum() >= 3
    r.close()


@pytest.mark.parametrize("at", [20, 21])
def test_checkpoint_resume_bitexact(gpu_lib, tmp_path, at):
    """Save after 20 (even) or 21 (odd: the state then lives in the second
    particle buffer) steps, resume in a fresh handle: the next 17 steps
    (graph-launched, sampling on) match the uninterrupted run exactly."""
    p, body = T.small_cylinder(80, 40, 10, 64, seed=23)
    p.sample_start = 10
    a = Run(p, body)
    a.initialize()
    a.solver(at, graph=False)
    f = tmp_path / "state.npz"
    a.save_state(str(f))
    a.solver(17, graph=True)
    b = Run(p, body)
    b.load_state(str(f))
    b.solver(17, graph=True)
    T.assert_same_particles(a.download(), b.download(), what="resumed ")
    for u, v in zip(a.cell_state()[:2], b.cell_state()[:2]):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_array_equal(a.samples()[0], b.samples()[0])
    sa, sb = a.status(), b.status()
    for k in ("step", "n_live", "next_pid", "particle_steps", "collisions", "nsamples", "err"):
        assert getattr(sa, k) == getattr(sb, k), k
    a.close()
    b.close()


def test_heavy_cell_and_empty(gpu_lib):
    """All particles in a single cell (beyond the LDS tile capacity) and an
    empty state: the global-memory path and the n = 0 path."""
    p, body = T.small_box(8, 1, seed=3)
    p.n_inf = 1.35e16  # ~1000 NTC candidates for 6000 particles in one cell
    oc, info, tables = O.case_from_params(p, body)
    rng = np.random.default_rng(5)
    n = 6000
    dx = np.float32(info.dx)
    parts = dict(x=(rng.random(n, dtype=np.float32) * dx * 0.99).astype(np.float32) + 3 * dx,
                 y=(rng.random(n, dtype=np.float32) * dx * 0.99).astype(np.float32) + 5 * dx,
                 vx=rng.normal(0, 1e-3, n).astype(np.float32), vy=rng.normal(0, 1e-3, n).astype(np.float32),
                 vz=rng.normal(0, 300, n).astype(np.float32), pid=np.arange(n, dtype=np.uint32))
    r = Run(p, body)
    r.upload(*[parts[k] for k in ("x", "y", "vx", "vy", "vz")], parts["pid"])
    r.set_keep_perm(True)
    sm, mr, _ = r.cell_state()
    ref = T.oracle_phases(oc, parts, 0, n, sm, mr, int(info.capacity))
    r.particlemove()
    r.ressignParticle()
    r.collision()
    np.testing.assert_array_equal(r.perm(), ref["perm"])
    T.assert_same_particles(r.download(), ref["new"], what="heavy cell ")
    np.testing.assert_array_equal(r.cell_state()[0], ref["sigcmax"])
    assert ref["ncoll"].max() > 100
    # empty
    r.upload(*[np.zeros(0, np.float32)] * 5)
    r.solver(3, graph=False)
    s = r.status()
    assert s.n_live == 0 and s.step == 3 and s.err == 0
    r.close()


def test_tile_cut_into_lds_runs(gpu_lib):
    """A tile of several cells that together exceed the LDS capacity is cut
    into runs of whole cells that fit (cells of 700, 700 | 700 particles),
    next to a single 2,500-particle cell that stages in HBM and a small
    tile: order, particles and collision state bit-exact against the oracle."""
    p, body = T.small_box(8, 1, seed=3)
    p.n_inf = 1.35e16
    oc, info, tables = O.case_from_params(p, body)
    rng = np.random.default_rng(9)
    dx = np.float32(info.dx)
    xs, ys = [], []
    for j, cnt in ((2, 700), (3, 700), (4, 700), (5, 2500), (6, 300)):
        xs.append((rng.random(cnt, dtype=np.float32) * dx * 0.99).astype(np.float32) + 3 * dx)
        ys.append((rng.random(cnt, dtype=np.float32) * dx * 0.99).astype(np.float32) + j * dx)
    n = sum(len(v) for v in xs)
    # in-plane speeds tiny so that every particle stays in its cell
    parts = dict(x=np.concatenate(xs), y=np.concatenate(ys),
                 vx=rng.normal(0, 1e-3, n).astype(np.float32), vy=rng.normal(0, 1e-3, n).astype(np.float32),
                 vz=rng.normal(0, 300, n).astype(np.float32), pid=np.arange(n, dtype=np.uint32))
    r = Run(p, body)
    r.set_launch(cell_tile=2048)
    r.upload(*[parts[k] for k in ("x", "y", "vx", "vy", "vz")], parts["pid"])
    r.set_keep_perm(True)
    sm, mr, _ = r.cell_state()
    ref = T.oracle_phases(oc, parts, 0, n, sm, mr, int(info.capacity))
    r.particlemove()
    r.ressignParticle()
    r.collision()
    np.testing.assert_array_equal(r.perm(), ref["perm"])
    T.assert_same_particles(r.download(), ref["new"], what="cut tile ")
    sm_g, mr_g, nc_g = r.cell_state()
    np.testing.assert_array_equal(sm_g, ref["sigcmax"])
    np.testing.assert_array_equal(nc_g, ref["ncoll"])
    counts = np.diff(ref["start"])
    assert sorted(counts[counts > 0].tolist()) == [300, 700, 700, 700, 2500]
    assert ref["ncoll"].max() > 10
    r.close()


def test_dense_cells_over_steps(gpu_lib):
    """Cells of 255 to 900 particles next to sparse ones, over four steps
    (the dense cells stay dense): order, particles and collision state
    bit-exact against the oracle at every step, for the default tile and
    for a tile of 256 particles, where the dense cells each exceed a tile."""
    for tile in (1024, 256):
        p, body = T.small_box(8, 1, seed=4)
        p.n_inf = 1.35e16
        oc, info, tables = O.case_from_params(p, body)
        rng = np.random.default_rng(11)
        dx = np.float32(info.dx)
        xs, ys = [], []
        for j, cnt in ((0, 40), (1, 300), (2, 260), (3, 20), (4, 900), (5, 255), (6, 256), (7, 3)):
            xs.append((rng.random(cnt, dtype=np.float32) * dx * 0.99).astype(np.float32) + 2 * dx)
            ys.append((rng.random(cnt, dtype=np.float32) * dx * 0.99).astype(np.float32) + j * dx)
        n = sum(len(v) for v in xs)
        parts = dict(x=np.concatenate(xs), y=np.concatenate(ys),
                     vx=rng.normal(0, 1e-3, n).astype(np.float32), vy=rng.normal(0, 1e-3, n).astype(np.float32),
                     vz=rng.normal(0, 300, n).astype(np.float32), pid=np.arange(n, dtype=np.uint32))
        r = Run(p, body)
        r.set_launch(cell_tile=tile)
        r.upload(*[parts[k] for k in ("x", "y", "vx", "vy", "vz")], parts["pid"])
        r.set_keep_perm(True)
        for step in range(4):
            print(f"[parity] dense cells, tile {tile}, step {step}", flush=True)
            cur = r.download()
            sm, mr, _ = r.cell_state()
            st = r.status()
            ref = T.oracle_phases(oc, cur, step, st.next_pid, sm, mr, int(info.capacity))
            r.particlemove()
            r.ressignParticle()
            r.collision()
            np.testing.assert_array_equal(r.perm(), ref["perm"])
            T.assert_same_particles(r.download(), ref["new"], what=f"tile {tile} step {step} ")
            sm_g, mr_g, nc_g = r.cell_state()
            np.testing.assert_array_equal(sm_g, ref["sigcmax"])
            np.testing.assert_array_equal(nc_g, ref["ncoll"])
            counts = np.diff(ref["start"])
            assert (counts >= 256).sum() >= 3          # 4 at the start; collisions spread a few
        assert r.status().err == 0
        r.close()


def test_box_full_rate_and_conservation(gpu_lib):
    """Closed box at equilibrium, full size (64x64, 81,920 particles, 1000
    steps): collision rate against VHS theory, energy and particle
    conservation, T = 300 K."""
    cfg = cases.box_equilibrium()
    p = cfg["params"]
    r = Run(p)
    r.initialize()
    g0 = r.download()
    e0 = float((g0["vx"].astype(np.float64) ** 2 + g0["vy"] ** 2 + g0["vz"] ** 2).sum())
    r.solver(cfg["steps"], graph=True)
    s = r.status()
    assert s.err == 0 and s.n_live == 81920 and s.step == 1000
    rate = s.collisions / cfg["steps"]
    theory = 0.5 * 81920 * cfg["nu"] * r.info.dt
    # NTC collision counts per step are sub-Poisson; use the Poisson sigma of
    # the 1000-step mean as a conservative 3-sigma band.
    sigma = np.sqrt(theory) / np.sqrt(cfg["steps"])
    assert abs(rate - theory) < 3 * sigma + 0.002 * theory, (rate, theory)
    g = r.download()
    e1 = float((g["vx"].astype(np.float64) ** 2 + g["vy"] ** 2 + g["vz"] ** 2).sum())
    assert abs(e1 - e0) / e0 < 1e-4
    assert np.unique(g["pid"]).size == 81920
    f = r.fields()
    T_mean = float((f["T"] * f["n"]).sum() / f["n"].sum())
    assert abs(T_mean - 300.0) < 1.5, T_mean
    r.close()


def test_cylinder_full_invariants(gpu_lib):
    """Mach-10 cylinder at full size for 300 steps: no particle inside the
    body, cell-sorted state, unique ids, inflow count as specified, no
    device error."""
    cfg = cases.cylinder_mach10()
    p = cfg["params"]
    r = Run(p, cfg["body"])
    r.initialize()
    n0 = r.status().n_live
    r.solver(300, graph=True)
    s = r.status()
    assert s.err == 0
    g = r.download()
    assert g["x"].shape[0] == s.n_live
    assert np.unique(g["pid"]).size == s.n_live
    inj = sum(r.info.inj_e[f] for f in range(4)) * 300
    assert abs((s.next_pid - n0) - inj) <= 4
    cx, cy, R = 0.3, 0.15, 0.1
    rr = np.hypot(g["x"].astype(np.float64) - cx, g["y"].astype(np.float64) - cy)
    # polygon of 256 points is inscribed: its apothem is R cos(pi/256)
    assert (rr < R * np.cos(np.pi / 256) * (1 - 1e-5)).sum() == 0
    i = np.minimum((g["x"] / np.float32(r.info.dx)).astype(np.int64), p.N1 - 1)
    j = np.minimum((g["y"] / np.float32(r.info.dy)).astype(np.int64), p.N2 - 1)
    cell = j + i * p.N2
    assert (np.diff(cell) >= 0).all()
    r.close()


@pytest.mark.parametrize("name", ["cyl80", "cyl80_specular", "wedge64"])
def test_host_built_tables_equal_library_tables(gpu_lib, name):
    """A case created from host-built tables (dsmc_create_from_tables: the
    wall segments, fluid fractions and cut flags a host mesh computes, as
    GPUCells::UploadFromHost takes them) equals the case the library builds
    from the body outline: same tables, and 40 graph-replayed steps bit for
    bit -- also when the fractions are left to the library."""
    p, body, oc, info, tables = _make(name)
    p.sample_start = 10
    seg = run.outline_segments(body)
    a = Run(p, body)
    b = Run(p, tables=dict(seg=seg, frac=tables["frac"], cut=tables["cut"]))
    c = Run(p, tables=dict(seg=seg))
    ta = a.tables()
    for r in (b, c):
        for k, v in r.tables().items():
            np.testing.assert_array_equal(v, ta[k], err_msg=k)
    for r in (a, b, c):
        r.initialize()
        r.solver(40, graph=True)
    for r, what in ((b, "host tables "), (c, "host segments ")):
        T.assert_same_particles(r.download(), a.download(), what=what)
        for u, v in zip(r.cell_state(), a.cell_state()):
            np.testing.assert_array_equal(u, v)
        np.testing.assert_array_equal(r.samples()[0], a.samples()[0])
    for r in (a, b, c):
        r.close()


def test_standalone_sample_equals_fused_sampling(gpu_lib):
    """dsmc_sample after every step (GPUCells::Sample, called by the
    reference's solver loop) adds exactly what a step's own sampling adds;
    sampling every 100th step as the reference does gives the sums of those
    steps; before any step it is refused (the state is not in cell order)."""
    p, body, oc, info, tables = _make("cyl80")
    p.sample_start = 0
    a = Run(p, body)
    p2 = type(p).from_buffer_copy(p)
    p2.sample_start = -1
    b = Run(p2, body)
    a.initialize()
    b.initialize()
    with pytest.raises(abi.DSMCError, match="cell order"):
        b.sample()
    for _ in range(25):
        a.solver(1, graph=False)
        b.solver(1, graph=False)
        b.sample()
    acc_a, ns_a = a.samples()
    acc_b, ns_b = b.samples()
    assert ns_a == ns_b == 25
    np.testing.assert_array_equal(acc_b, acc_a)
    T.assert_same_particles(b.download(), a.download(), what="sampled run ")
    # every 100th step, as Run::solver does (Run.cpp:294-296)
    b.reset_samples()
    for it in range(300):
        b.solver(1, graph=False)
        if it % 100 == 0:
            b.sample()
    assert b.samples()[1] == 3
    a.close()
    b.close()
