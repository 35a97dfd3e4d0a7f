"""GPU parity of the adaptive collision sub-cells and the variable time step
(Cell::genAMRmesh, Cell::VTS, Cell.cpp:117-279): the kernels through the C
ABI against the CPU oracle on identical inputs.  Particle states (in leaf
order within the cells with sub-cells), (sigma c_r)max, remainders,
collision counts and the grids, mean free paths and VTS steps must be
bit-exact; sampling sums (FP64, another summation order) agree to 1e-12."""
import numpy as np
import pytest

from cudadsmc_amd import abi, cases
from cudadsmc_amd.run import Run
from oracle import oracle as O

import dsmc_testlib as T

pytestmark = pytest.mark.gpu


def pack(n2x, n2y, n3x=1, n3y=1):
    return n2x | (n2y << 8) | (n3x << 16) | (n3y << 24)


def _grids(C, seed, big=False):
    """A mix: none, one leaf, small grids, two levels, large, beyond 256 leaves."""
    rng = np.random.default_rng(seed)
    choice = [0, pack(1, 1), pack(2, 2), pack(3, 1), pack(1, 4), pack(2, 3, 2, 1), pack(4, 4), pack(2, 2, 2, 2)]
    if big:
        choice += [pack(16, 16), pack(16, 16, 2, 1)]
    return np.array([choice[k] for k in rng.integers(0, len(choice), C)], np.uint32)


CASES = {
    "box16": lambda: T.small_box(16, 24, seed=7),
    "cyl80": lambda: T.small_cylinder(80, 40, 12, 64, seed=11),
    "cyl80_n2": lambda: T.small_cylinder(80, 40, 12, 64, seed=5, gas=cases.NITROGEN),
}


def _run_both(name, steps, grid_seed, sample_from=4, graph=False, big=False):
    p, body = CASES[name]()
    p.sample_start = sample_from
    print(f"[case] {name} steps={steps} grids={grid_seed} graph={graph}", flush=True)
    oc, info, tables = O.case_from_params(p, body)
    grids = _grids(info.ncell, grid_seed, big)
    r = Run(p, body)
    r.initialize()
    r.set_subcells(grids)
    parts, _, _ = T.oracle_fill(oc, info, tables)
    oc.set_subcells(grids)
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    r.solver(steps, graph=graph)
    for s in range(steps):
        st.step(sample=s >= sample_from, nthreads=4)
    return r, st, oc, info


@pytest.mark.parametrize("name", list(CASES))
def test_subcell_trajectory_bitexact(gpu_lib, name):
    r, st, oc, info = _run_both(name, 24, grid_seed=3, big=(name == "box16"))
    g = r.download()
    o = st.particles()
    T.assert_same_particles(g, o, fields=T.fields_of(oc), what=f"{name} sub-cells ")
    s = r.status()
    assert s.collisions == st.st.total_coll and s.collisions > 0
    assert s.err == 0 and st.st.err == 0
    sm, mr, nc = r.cell_state()
    np.testing.assert_array_equal(sm, st.sigcmax)
    np.testing.assert_array_equal(mr, st.mrem)
    np.testing.assert_array_equal(nc, st.ncoll)
    acc_g, ns_g = r.samples()
    assert ns_g == st.st.nsamples
    np.testing.assert_allclose(acc_g, st.acc.reshape(acc_g.shape), rtol=1e-12, atol=1e-9)
    r.close()


def test_subcell_graph_replays_bitexact(gpu_lib):
    r, st, oc, info = _run_both("cyl80", 19, grid_seed=8, graph=True)
    T.assert_same_particles(r.download(), st.particles(), what="graph replays with sub-cells ")
    assert r.status().collisions == st.st.total_coll
    r.close()


def test_amr_update_equals_oracle_and_continues_bitexact(gpu_lib):
    """The grids from the sampling sums, then more steps in them."""
    p, body = T.small_cylinder(80, 40, 12, 64, seed=17)
    p.sample_start = 0
    oc, info, tables = O.case_from_params(p, body)
    r = Run(p, body)
    r.initialize()
    parts, _, _ = T.oracle_fill(oc, info, tables)
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    r.solver(30, graph=False)
    for _ in range(30):
        st.step(sample=True, nthreads=4)
    T.assert_same_particles(r.download(), st.particles(), what="before the update ")
    gas = dict(Fn=info.Fn, mass=p.mass, diam=p.diam, T_inf=p.T, omega=p.Vtl, boltz=p.boltz)
    acc_g, ns = r.samples()
    for k in range(2):                      # the first update (no grids), then with grids
        r.amr_update()
        a = r.amr()
        old = np.zeros(info.ncell, np.uint32) if k == 0 else sub_o
        sub_o, lam_o, dt_o = oc.amr_update(gas, acc_g.ravel(), ns, tables["vol"], old)
        np.testing.assert_array_equal(a["sub"], sub_o)
        np.testing.assert_array_equal(a["lam"], lam_o)
        np.testing.assert_array_equal(a["dt"], dt_o)
    assert (a["Numchild"] > 1).sum() > 100          # Kn = 0.01 on 1.5 mm cells: 2 x 2 in the freestream
    oc.set_subcells(sub_o)
    r.solver(12, graph=False)
    for _ in range(12):
        st.step(sample=True, nthreads=4)
    T.assert_same_particles(r.download(), st.particles(), what="after the update ")
    assert r.status().collisions == st.st.total_coll
    r.set_subcells(None)
    r.close()


def test_subcell_equilibrium_rate_at_size(gpu_lib):
    """The 64x64 closed box at equilibrium (81,920 particles) with 2 x 2
    leaves everywhere collides at the VHS rate N nu dt / 2 = 8,192 per step."""
    cfg = cases.box_equilibrium()
    p = cfg["params"]
    r = Run(p)
    r.initialize()
    r.set_subcells(np.full(r.ncell, pack(2, 2), np.uint32))
    r.solver(10, graph=True)
    c0 = r.status().collisions
    steps = 200
    r.solver(steps, graph=True)
    rate = (r.status().collisions - c0) / steps
    n = r.status().n_live
    theory = 0.5 * n * cfg["nu"] * r.info.dt
    assert abs(rate - theory) / np.sqrt(theory / steps) < 4
    r.close()


def test_set_subcells_rejects_an_empty_level(gpu_lib):
    p, body = T.small_box(16, 4)
    r = Run(p, body)
    bad = np.zeros(r.ncell, np.uint32)
    bad[3] = 2 | (2 << 8)                     # level 3 with no cells
    with pytest.raises(RuntimeError):
        r.set_subcells(bad)
    with pytest.raises(RuntimeError):
        r.amr_update()                        # nothing sampled yet
    r.close()
