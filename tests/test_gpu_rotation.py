"""GPU parity of the rotational-energy path (zeta = 2, the reference's N2 with
its Larsen-Borgnakke scheme, src/Param.h:14, src/cell/GPUCell.cu:256-275,
316-328) against the CPU oracle on identical inputs: rotational energies,
velocities, positions, ids and collision state bit-exact; sampling sums
(FP64, other summation order) to 1e-12 relative.  The closed-box relaxation
at full size is checked against the Larsen-Borgnakke rate."""
import ctypes as C

import numpy as np
import pytest

from cudadsmc_amd import abi, cases
from cudadsmc_amd.batch import Batch
from cudadsmc_amd.run import Run
from oracle import oracle as O

import dsmc_testlib as T

pytestmark = pytest.mark.gpu

N2 = cases.NITROGEN


def _box_n2(n_cells=16, ppc=8, seed=7, diffuse=False, rot_prob=1.0, n=1e20, T0=300.0):
    nu = cases.vhs_collision_frequency(n, T0, gas=N2)
    p = cases.make_params(L1=1e-3, L2=1e-3, N1=n_cells, N2=n_cells, n_inf=n, T=T0, V_jet=0.0,
                          wall_bc=abi.BC_SPECULAR, face_bc=[abi.BC_SPECULAR] * 4, tau=0.2 / nu, ppc=float(ppc),
                          seed=seed, sample_start=0, **cases.gas_fields(N2, rot_prob))
    if diffuse:
        for f, Tf in enumerate((200.0, 300.0, 400.0, 500.0)):
            p.face_bc[f] = abi.BC_DIFFUSE
            p.face_T[f] = Tf
    return p, None


CASES = {
    "cyl80_n2": lambda: T.small_cylinder(80, 40, 10, 64, seed=31, gas=N2),
    "cyl80_n2_zrot5": lambda: T.small_cylinder(80, 40, 10, 64, seed=33, gas=N2, rot_prob=0.2),
    "cyl80_n2_specular": lambda: T.small_cylinder(80, 40, 10, 64, seed=35, gas=N2, wall=abi.BC_SPECULAR),
    "box16_n2_diffuse": lambda: _box_n2(16, 8, seed=9, diffuse=True),
}


def _make(name):
    p, body = CASES[name]()
    print(f"[rot] {name} N1={p.N1} N2={p.N2} ppc={p.ppc} seed={p.seed} rot_prob={p.rot_prob}", flush=True)
    oc, info, tables = O.case_from_params(p, body)
    assert info.zeta == 2 and info.sample_blocks == 7
    return p, body, oc, info, tables


F = T.FIELDS + ("er",)


@pytest.mark.parametrize("name", list(CASES))
def test_fill_matches_oracle(gpu_lib, name):
    p, body, oc, info, tables = _make(name)
    r = Run(p, body)
    r.initialize()
    g = r.download()
    o, count, err = T.oracle_fill(oc, info, tables)
    assert err == 0 and r.status().err == 0
    T.assert_same_particles(g, o, fields=F, what="fill ")
    assert np.all(g["er"] >= 0)
    r.close()


@pytest.mark.parametrize("name", list(CASES))
def test_phase_parity(gpu_lib, name):
    """move (walls re-draw E_rot, inflow draws it) / reassign / collide
    (Larsen-Borgnakke exchange) against the oracle, bit-exact."""
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
        T.assert_same_particles(r.download(moved=True), ref["moved"], fields=F, what=f"step {step} move ")
        cell, count = r.bins()
        np.testing.assert_array_equal(cell, ref["cell"])
        np.testing.assert_array_equal(count, ref["count"])
        r.ressignParticle()
        np.testing.assert_array_equal(r.start(), ref["start"])
        r.collision()
        np.testing.assert_array_equal(r.perm(), ref["perm"])
        T.assert_same_particles(r.download(), ref["new"], fields=F, what=f"step {step} collide ")
        sm2, mr2, nc2 = r.cell_state()
        np.testing.assert_array_equal(sm2, ref["sigcmax"])
        np.testing.assert_array_equal(mr2, ref["mrem"])
        np.testing.assert_array_equal(nc2, ref["ncoll"])
        assert r.status().err == 0
    r.close()


@pytest.mark.parametrize("name,steps", [("cyl80_n2", 30), ("cyl80_n2_zrot5", 30), ("box16_n2_diffuse", 40)])
def test_trajectory_bitexact(gpu_lib, name, steps):
    p, body, oc, info, tables = _make(name)
    p.sample_start = 5
    r = Run(p, body)
    r.initialize()
    r.solver(steps, graph=False)
    parts, _, _ = T.oracle_fill(oc, info, tables)
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    for s in range(steps):
        st.step(sample=s >= 5, nthreads=4)
    T.assert_same_particles(r.download(), st.particles(), fields=F, what=f"{name} after {steps} steps ")
    s = r.status()
    assert s.collisions == st.st.total_coll and s.err == 0 and st.st.err == 0
    acc_g, ns_g = r.samples()
    assert acc_g.shape == (7, info.ncell) and ns_g == steps - 5
    np.testing.assert_allclose(acc_g, st.acc.reshape(7, -1), rtol=1e-12, atol=1e-9)
    r.close()


def test_graph_replays_bitexact(gpu_lib):
    """hipGraph replays (8-step graphs plus single steps) on the Z_rot = 5
    cylinder against the oracle."""
    p, body, oc, info, tables = _make("cyl80_n2_zrot5")
    p.sample_start = 3
    r = Run(p, body)
    r.set_launch(graph_steps=8)
    r.initialize()
    parts, _, _ = T.oracle_fill(oc, info, tables)
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    r.solver(27, graph=True)
    for s in range(27):
        st.step(sample=s >= 3, nthreads=4)
    T.assert_same_particles(r.download(), st.particles(), fields=F, what="graph n2 ")
    np.testing.assert_array_equal(r.cell_state()[0], st.sigcmax)
    np.testing.assert_allclose(r.samples()[0], st.acc.reshape(7, -1), rtol=1e-12, atol=1e-9)
    r.close()


def _dense_parts(info, counts, rng):
    dx = np.float32(info.dx)
    xs, ys = [], []
    for j, cnt in counts:
        xs.append((rng.random(cnt, dtype=np.float32) * dx * 0.99).astype(np.float32) + 2 * dx)
        ys.append((rng.random(cnt, dtype=np.float32) * dx * 0.99).astype(np.float32) + j * dx)
    n = sum(c for _, c in counts)
    return dict(x=np.concatenate(xs), y=np.concatenate(ys),
                vx=rng.normal(0, 1e-3, n).astype(np.float32), vy=rng.normal(0, 1e-3, n).astype(np.float32),
                vz=rng.normal(0, 300, n).astype(np.float32), pid=np.arange(n, dtype=np.uint32),
                er=rng.exponential(5e4, n).astype(np.float32))


@pytest.mark.parametrize("tile", [1024, 256])
def test_dense_cells(gpu_lib, tile):
    """Cells staged in LDS, cut tiles, a 2,500-particle cell staged in HBM
    with LDS conflict marks and a 9,000-particle one with shuffle conflict
    tests, over three steps with uploaded rotational energies: bit-exact."""
    p, _ = _box_n2(8, 1, seed=4, n=1.35e16)
    oc, info, tables = O.case_from_params(p, None)
    parts = _dense_parts(info, ((0, 40), (1, 300), (2, 700), (3, 9000), (4, 900), (5, 2500), (6, 256), (7, 3)),
                         np.random.default_rng(13))
    r = Run(p, None)
    r.set_launch(cell_tile=tile)
    r.upload(*[parts[k] for k in ("x", "y", "vx", "vy", "vz")], parts["pid"])
    assert np.all(r.download()["er"] == 0)
    r.upload_rotational(parts["er"])
    np.testing.assert_array_equal(r.download()["er"], parts["er"])
    r.set_keep_perm(True)
    for step in range(3):
        print(f"[rot] dense cells, tile {tile}, step {step}", flush=True)
        cur = r.download()
        sm, mr, _ = r.cell_state()
        st = r.status()
        ref = T.oracle_phases(oc, cur, step, st.next_pid, sm, mr, int(info.capacity))
        r.particlemove()
        r.ressignParticle()
        r.collision()
        np.testing.assert_array_equal(r.perm(), ref["perm"])
        T.assert_same_particles(r.download(), ref["new"], fields=F, what=f"dense step {step} ")
        np.testing.assert_array_equal(r.cell_state()[0], ref["sigcmax"])
        if step == 0:
            assert ref["ncoll"].max() > 100
    r.close()


@pytest.mark.parametrize("at", [20, 21])
def test_checkpoint_resume_bitexact(gpu_lib, tmp_path, at):
    p, body = T.small_cylinder(80, 40, 10, 64, seed=37, gas=N2, rot_prob=0.5)
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
    T.assert_same_particles(a.download(), b.download(), fields=F, what="resumed ")
    np.testing.assert_array_equal(a.samples()[0], b.samples()[0])
    a.close()
    b.close()


def test_standalone_sample_equals_fused_sampling(gpu_lib):
    """dsmc_sample on the state after a step adds exactly what the step's
    own sampling adds, the rotational block included."""
    p, body = T.small_cylinder(80, 40, 10, 64, seed=39, gas=N2)
    p.sample_start = -1
    a = Run(p, body)
    a.initialize()
    a.solver(6, graph=False)
    a.sample()
    p2, _ = T.small_cylinder(80, 40, 10, 64, seed=39, gas=N2)
    p2.sample_start = 5
    b = Run(p2, body)
    b.initialize()
    b.solver(6, graph=False)
    acc_a, ns_a = a.samples()
    acc_b, ns_b = b.samples()
    assert ns_a == ns_b == 1 and acc_a.shape[0] == 7
    np.testing.assert_array_equal(acc_a, acc_b)
    a.close()
    b.close()


def test_rotational_temperature_and_relaxation_at_size(gpu_lib):
    """A 128x128 closed box (about 820,000 particles) filled at 400 K with
    E_rot scaled to T_rot = 100 K and Z_rot = 5: the total energy
    3/2 T_tr + T_rot is conserved, T_tr - T_rot decays at the
    Larsen-Borgnakke rate per collision, (5/3) p a/(a + 2), within 3 %, and
    the library's T_rot field equals 2 m <e> / (zeta k) from the sums."""
    p, _ = _box_n2(128, 50, seed=17, rot_prob=0.2, T0=400.0)
    p.sample_start = -1
    r = Run(p, None)
    r.initialize()
    d = r.download()
    r.upload(d["x"], d["y"], d["vx"], d["vy"], d["vz"], d["pid"])
    r.upload_rotational((d["er"] * np.float32(0.25)).astype(np.float32))

    def temps():
        s = r.download()
        v = np.stack([s[k].astype(np.float64) for k in ("vx", "vy", "vz")])
        Ttr = p.mass / (3 * p.boltz) * np.mean(np.sum((v - v.mean(1, keepdims=True)) ** 2, 0))
        return Ttr, p.mass / p.boltz * np.mean(s["er"].astype(np.float64)), s["x"].shape[0]
    Ttr0, Trot0, N = temps()
    E0 = (3 * Ttr0 + 2 * Trot0) / 5
    c, dT = [0.0], [Ttr0 - Trot0]
    for k in range(30):
        r.solver(4, graph=False)
        Ttr, Trot, _ = temps()
        c.append(2.0 * r.status().collisions / N)
        dT.append(Ttr - Trot)
    Ttr, Trot, _ = temps()
    assert abs((3 * Ttr + 2 * Trot) / 5 / E0 - 1) < 1e-4
    c, dT = np.array(c), np.array(dT)
    use = dT > 0.1 * dT[0]
    a = 2.5 - p.Vtl
    k_th = (5 / 3) * 0.2 * a / (a + 2)
    slope = np.polyfit(c[use], np.log(dT[use]), 1)[0]
    print(f"[rot] relaxation slope / theory = {-slope / k_th:.4f} over {use.sum()} points", flush=True)
    assert abs(-slope / k_th - 1) < 0.03
    # the T_rot field
    r.solver(1, graph=False)
    r.sample()
    acc, ns = r.samples()
    f = r.fields()
    with np.errstate(invalid="ignore", divide="ignore"):
        want = np.where(acc[0] > 0, 2 * p.mass * (acc[6] / acc[0]) / (2 * p.boltz), 0.0)
    np.testing.assert_allclose(f["Trot"].ravel(), want, rtol=1e-12)
    np.testing.assert_allclose(f["Tov"], (3 * f["T"] + 2 * f["Trot"]) / 5, rtol=1e-12)
    assert abs(np.mean(f["Trot"]) / Trot - 1) < 0.02
    r.close()


def test_monatomic_and_batch_refuse_rotational_calls(gpu_lib):
    p, body = T.small_cylinder(40, 20, 5, 32, seed=3)
    r = Run(p, body)
    r.initialize()
    lib = r.lib
    n = C.c_int64()
    buf = np.zeros(4, np.float32)
    assert lib.dsmc_download_rotational(r.h, 0, 4, buf.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n)) == \
        abi.DSMC_ERR_ARG
    assert b"monatomic" in lib.dsmc_last_error()
    assert "er" not in r.download() and r.samples()[0].shape[0] == 6
    r.close()
    pn, bn = T.small_cylinder(40, 20, 5, 32, seed=3, gas=N2)
    with pytest.raises(abi.DSMCError, match="monatomic"):
        Batch([pn], [bn])
    # negative or non-finite energies are refused
    r = Run(pn, bn)
    r.initialize()
    s = r.status()
    bad = np.full(s.n_live, -1.0, np.float32)
    with pytest.raises(abi.DSMCError, match="finite and >= 0"):
        r.upload_rotational(bad)
    r.close()
