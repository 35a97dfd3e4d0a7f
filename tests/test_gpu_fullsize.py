"""GPU parity at the benchmark sizes themselves, and batched instances
against the oracle directly.

  - closed box, 64x64 cells, 81,920 particles: 50 full steps (hipGraph
    replays) bit-exact against the oracle, then one more step phase by phase
    with the cell ids, counts, scan offsets and the stable permutation
    compared (GPUCell.cu:169-198 binning, :104-120 scan, GPUParticle.cu:365-373
    scatter);
  - Mach-10 cylinder, 400x200 cells, about 2M particles: 100 full steps the
    same way, then the phase-by-phase step;
  - a batch of four uneven instances: every instance against an oracle run
    of its own case (not against a single-instance GPU run);
  - the 64-instance sweep at full size (about 110M particles): 20 graph
    steps without device errors, totals that add up, no particle inside a
    body, and three instances bit-identical to running alone."""
import os

import numpy as np
import pytest

from cudadsmc_amd import abi, cases
from cudadsmc_amd.batch import Batch
from cudadsmc_amd.run import Run
from oracle import oracle as O

import dsmc_testlib as T

pytestmark = pytest.mark.gpu

NT = max(1, min(16, os.cpu_count() or 1))


def _oracle_run(p, body, steps, sample_from=None):
    oc, info, tables = O.case_from_params(p, body)
    parts, _, err = T.oracle_fill(oc, info, tables)
    assert err == 0
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    for s in range(steps):
        st.step(sample=sample_from is not None and s >= sample_from, nthreads=NT)
    return oc, info, st


def _compare_run(r, st, what, samples=False):
    T.assert_same_particles(r.download(), st.particles(), what=what)
    s = r.status()
    assert (s.step, s.n_live, s.next_pid) == (st.st.step, st.st.n, st.st.next_pid), what
    assert s.collisions == st.st.total_coll, what
    assert s.err == 0 and st.st.err == 0, what
    sm, mr, nc = r.cell_state()
    np.testing.assert_array_equal(sm, st.sigcmax)
    np.testing.assert_array_equal(mr, st.mrem)
    np.testing.assert_array_equal(nc, st.ncoll)
    if samples:
        acc, ns = r.samples()
        assert ns == st.st.nsamples
        np.testing.assert_allclose(acc, st.acc.reshape(6, -1), rtol=1e-12, atol=1e-9)


def _phase_step(r, oc, info, what):
    """One more step phase by phase: move, bins and counts, scan offsets,
    stable permutation, state and collision state, all bit-exact."""
    cur = r.download()
    sm, mr, _ = r.cell_state()
    st = r.status()
    ref = T.oracle_phases(oc, cur, st.step, st.next_pid, sm, mr, int(info.capacity))
    r.set_keep_perm(True)
    r.particlemove()
    T.assert_same_particles(r.download(moved=True), ref["moved"], what=what + "move ")
    cell, count = r.bins()
    np.testing.assert_array_equal(cell, ref["cell"], err_msg=what + "cell ids")
    np.testing.assert_array_equal(count, ref["count"], err_msg=what + "counts")
    r.ressignParticle()
    np.testing.assert_array_equal(r.start(), ref["start"], err_msg=what + "scan")
    r.collision()
    np.testing.assert_array_equal(r.perm(), ref["perm"], err_msg=what + "permutation")
    T.assert_same_particles(r.download(), ref["new"], what=what + "collide ")
    sm2, mr2, nc2 = r.cell_state()
    np.testing.assert_array_equal(sm2, ref["sigcmax"])
    np.testing.assert_array_equal(mr2, ref["mrem"])
    np.testing.assert_array_equal(nc2, ref["ncoll"])
    return ref


def test_box_64x64_bitexact(gpu_lib):
    cfg = cases.box_equilibrium()
    p = cfg["params"]
    print("[fullsize] box 64x64, 81,920 particles, 50 steps", flush=True)
    r = Run(p)
    r.initialize()
    r.solver(50, graph=True)
    oc, info, st = _oracle_run(p, None, 50, sample_from=0)
    _compare_run(r, st, "box 64x64 after 50 steps ", samples=True)
    ref = _phase_step(r, oc, info, "box 64x64 step 50 ")
    assert int(ref["count"].sum()) == 81920 and ref["ncoll"].sum() > 7000
    r.close()


def test_grid_beyond_lds_tile_offsets(gpu_lib):
    """A 2,064 x 2,064 box (4.26M cells, 2,081 scan tiles of 2,048 cells:
    more tile offsets than k_place scans in LDS, so the one-block tile scan
    publishes them): a few steps, then one step phase by phase -- counts,
    scan offsets, stable permutation and state bit-exact against the
    oracle."""
    p, body = T.small_box(2064, 2, seed=3)
    print("[fullsize] box 2064x2064, 8.5M particles, scan tiles beyond the LDS", flush=True)
    r = Run(p, body)
    r.initialize()
    r.solver(3, graph=False)
    oc, info, _ = O.case_from_params(p, body)
    assert info.ncell > 2048 * 2048
    ref = _phase_step(r, oc, info, "box 2064x2064 ")
    assert int(ref["count"].sum()) == r.status().n_live
    r.close()


def test_cylinder_400x200_bitexact(gpu_lib):
    cfg = cases.cylinder_mach10()
    p = cfg["params"]
    p.sample_start = 50
    print("[fullsize] cylinder 400x200, about 2M particles, 100 steps", flush=True)
    r = Run(p, cfg["body"])
    r.initialize()
    r.solver(100, graph=True)
    oc, info, st = _oracle_run(p, cfg["body"], 100, sample_from=50)
    _compare_run(r, st, "cylinder 400x200 after 100 steps ", samples=True)
    ref = _phase_step(r, oc, info, "cylinder 400x200 step 100 ")
    assert int(ref["count"].sum()) > 1_900_000 and ref["dead"].sum() > 0 and ref["Jt"] > 1000
    r.close()


def _mixed():
    box, _ = T.small_box(16, 8, seed=31)
    box.sample_start = 0
    cyl, cbody = T.small_cylinder(80, 40, 10, 64, seed=32)
    cyl.sample_start = 10
    spec, sbody = T.small_cylinder(48, 24, 12, 48, seed=33, wall=abi.BC_SPECULAR)
    spec.sample_start = 5
    w = cases.flow_case("wedge", mach=6.0, kn=0.02, tw_ratio=3.0, nx=64, ny=32, ppc=9.0, half_angle_deg=25.0,
                        sample_from=12, seed=34, n_points=3)
    return [(box, None), (cyl, cbody), (spec, sbody), (w["params"], w["body"])]


@pytest.mark.parametrize("graph", [False, True])
def test_batch_instances_equal_oracle(gpu_lib, graph):
    """Each instance of a batched run (box, diffuse and specular cylinders,
    wedge; uneven sizes and sampling starts) against the oracle stepping
    that case alone: particles, cell state, samples, counters bit-exact."""
    specs = _mixed()
    steps = 29
    b = Batch([p for p, _ in specs], [bd for _, bd in specs])
    b.initialize()
    print(f"[fullsize] batch of {len(specs)} vs oracle, graph={graph}", flush=True)
    b.solver(steps, graph=graph)
    for i, (p, body) in enumerate(specs):
        oc, info, st = _oracle_run(p, body, steps, sample_from=p.sample_start)
        what = f"instance {i} "
        T.assert_same_particles(b.download(i), st.particles(), what=what)
        s = b.status(i)
        assert (s.step, s.n_live, s.next_pid, s.collisions) == (st.st.step, st.st.n, st.st.next_pid,
                                                                st.st.total_coll), what
        assert s.err == 0 and st.st.err == 0
        sm, mr, nc = b.cell_state(i)
        np.testing.assert_array_equal(sm, st.sigcmax)
        np.testing.assert_array_equal(mr, st.mrem)
        np.testing.assert_array_equal(nc, st.ncoll)
        acc, ns = b.samples(i)
        assert ns == st.st.nsamples
        np.testing.assert_allclose(acc, st.acc.reshape(6, -1), rtol=1e-12, atol=1e-9)
    b.close()


def _inside(px, py, poly):
    inside = np.zeros(px.shape, bool)
    for k in range(len(poly)):
        ax, ay = poly[k]
        bx, by = poly[(k + 1) % len(poly)]
        cross = (ay > py) != (by > py)
        xint = ax + (py - ay) * (bx - ax) / np.where(by != ay, by - ay, 1.0)
        inside ^= cross & (px < xint)
    return inside


def test_sweep64_full_size(gpu_lib):
    """The 64-instance sweep at full size (cylinders and wedges, Mach 3-12,
    Kn 0.005-0.1, grids 64x32 to 256x128, about 110M particles): 20 graph
    steps clean, totals equal the sums over instances, no particle inside a
    body, and the smallest, a middle and the largest instance equal a run
    of that case alone, bit for bit."""
    sw = cases.sweep(64, sample_from=10)
    b = Batch([c["params"] for c in sw], [c["body"] for c in sw])
    b.initialize()
    n0 = b.status(-1).n_live
    print(f"[fullsize] sweep64: {n0} particles", flush=True)
    assert 8e7 < n0 < 1.6e8
    b.solver(20, graph=True)
    tot = b.status(-1)
    assert tot.err == 0, hex(tot.err)
    per = [b.status(i) for i in range(64)]
    assert tot.n_live == sum(s.n_live for s in per)
    assert tot.collisions == sum(s.collisions for s in per)
    assert tot.particle_steps == sum(s.particle_steps for s in per)
    assert all(s.step == 20 and s.nsamples == 10 for s in per)
    order = np.argsort([s.n_live for s in per])
    for i in (int(order[0]), int(order[32]), int(order[-1])):
        c = sw[i]
        print(f"[fullsize] sweep64 instance {i} ({per[i].n_live} particles) alone", flush=True)
        r = Run(c["params"], c["body"])
        r.initialize()
        r.solver(20, graph=True)
        g = b.download(i)
        T.assert_same_particles(g, r.download(), what=f"sweep instance {i} ")
        for u, v in zip(b.cell_state(i), r.cell_state()):
            np.testing.assert_array_equal(u, v)
        np.testing.assert_array_equal(b.samples(i)[0], r.samples()[0])
        assert not _inside(g["x"].astype(np.float64), g["y"].astype(np.float64), c["body"]).any()
        r.close()
    b.close()
