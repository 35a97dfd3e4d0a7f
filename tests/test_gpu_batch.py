"""GPU: batched instances (one launch per phase over all of them) against
each instance run alone through its own handle -- particle states, cell
state, sample sums and counters must be bit-identical (every random stream
is keyed on the instance's own ids and seed)."""
import numpy as np
import pytest

from cudadsmc_amd import abi, cases
from cudadsmc_amd.batch import Batch
from cudadsmc_amd.run import Run

import dsmc_testlib as T

pytestmark = pytest.mark.gpu


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


def _compare(b, i, r, what):
    T.assert_same_particles(b.download(i), r.download(), what=what)
    bs, rs = b.cell_state(i), r.cell_state()
    for u, v in zip(bs, rs):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_array_equal(b.samples(i)[0], r.samples()[0])
    sb, sr = b.status(i), r.status()
    for k in ("step", "n_live", "next_pid", "particle_steps", "collisions", "nsamples", "err", "last_collisions"):
        assert getattr(sb, k) == getattr(sr, k), (what, k, getattr(sb, k), getattr(sr, k))


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_instances_alone(gpu_lib, graph):
    specs = _mixed()
    print(f"[batch] {len(specs)} instances graph={graph}", flush=True)
    b = Batch([p for p, _ in specs], [bd for _, bd in specs])
    b.initialize()
    steps = 27
    b.solver(steps, graph=graph)
    tot = b.status(-1)
    assert tot.err == 0
    psteps = 0
    for i, (p, body) in enumerate(specs):
        r = Run(p, body)
        r.initialize()
        r.solver(steps, graph=False)
        _compare(b, i, r, f"instance {i} ")
        psteps += r.status().particle_steps
        r.close()
    assert tot.particle_steps == psteps
    assert tot.n_live == sum(b.status(i).n_live for i in range(len(specs)))
    b.close()


def test_batch_warm_init_fields_equals_alone(gpu_lib):
    """init_fields over concatenated per-cell fields = each instance's own init_fields."""
    specs = _mixed()[1:3]
    b = Batch([p for p, _ in specs], [bd for _, bd in specs])
    rng = np.random.default_rng(3)
    fields = []
    for i, (p, _) in enumerate(specs):
        c = b.ncells[i]
        fields.append(dict(n=p.n_inf * rng.uniform(0.5, 2.0, c), ux=rng.normal(p.V_jet, 50.0, c),
                           uy=rng.normal(0.0, 50.0, c), T=p.T * rng.uniform(0.8, 3.0, c)))
    b.initialize_fields(*[[f[k] for f in fields] for k in ("n", "ux", "uy", "T")])
    b.solver(9, graph=True)
    for i, (p, body) in enumerate(specs):
        r = Run(p, body)
        r.initialize_fields(fields[i]["n"], fields[i]["ux"], fields[i]["uy"], fields[i]["T"])
        r.solver(9, graph=False)
        _compare(b, i, r, f"warm instance {i} ")
        r.close()
    b.close()


def test_sweep_batch_runs_clean(gpu_lib):
    """A reduced sweep (16 uneven instances, cylinders and wedges) steps
    without device errors; particle totals add up over instances."""
    sw = cases.sweep(n=16, seed=7, particles=(2e4, 2e5), sample_from=20)
    print(f"[batch] sweep of {len(sw)}", flush=True)
    b = Batch([c["params"] for c in sw], [c["body"] for c in sw])
    b.initialize()
    b.solver(40, graph=True)
    tot = b.status(-1)
    assert tot.err == 0, hex(tot.err)
    per = [b.status(i) for i in range(len(sw))]
    assert tot.n_live == sum(s.n_live for s in per)
    assert tot.collisions == sum(s.collisions for s in per)
    assert all(s.nsamples == 20 for s in per)
    b.close()


def _compare_batches(a, i, b, j, what):
    T.assert_same_particles(a.download(i), b.download(j), what=what)
    for u, v in zip(a.cell_state(i), b.cell_state(j)):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_array_equal(a.samples(i)[0], b.samples(j)[0])
    sa, sb = a.status(i), b.status(j)
    for k in ("step", "n_live", "next_pid", "particle_steps", "collisions", "nsamples", "err", "last_collisions"):
        assert getattr(sa, k) == getattr(sb, k), (what, k, getattr(sa, k), getattr(sb, k))


def test_compact_continues_bit_identically(gpu_lib):
    """Retiring instances mid-run (an odd step count, so the state sits in
    the second particle buffer): the kept ones go on exactly as in a batch
    that kept everything -- particles, cell state, sample sums, counters --
    through graph replays, and a second compaction down to one instance."""
    specs = _mixed()
    full = Batch([p for p, _ in specs], [bd for _, bd in specs])
    part = Batch([p for p, _ in specs], [bd for _, bd in specs])
    for b in (full, part):
        b.initialize()
        b.solver(13, graph=False)
    print("[batch] compact to [0, 2, 3] at step 13", flush=True)
    part.compact([0, 2, 3])
    assert part.n == 3 and part.origin == [0, 2, 3]
    assert part.ncells == [full.ncells[k] for k in (0, 2, 3)]
    for b in (full, part):
        b.solver(21)
    for j, i in enumerate(part.origin):
        _compare_batches(full, i, part, j, f"kept {i} ")
    tot = part.status(-1)
    assert tot.n_live == sum(part.status(j).n_live for j in range(3))
    assert tot.particle_steps == sum(part.status(j).particle_steps for j in range(3))
    print("[batch] compact to [1] (instance 2) at step 34", flush=True)
    part.compact([1])
    assert part.origin == [2]
    for b in (full, part):
        b.solver(9, graph=False)
    _compare_batches(full, 2, part, 0, "kept 2 alone ")
    assert full.status(-1).err == 0 and part.status(-1).err == 0
    full.close()
    part.close()


def test_compact_rejects_bad_lists(gpu_lib):
    specs = _mixed()[:2]
    b = Batch([p for p, _ in specs], [bd for _, bd in specs])
    b.initialize()
    for keep in ([], [1, 0], [0, 0], [2], [-1]):
        with pytest.raises(RuntimeError):
            b.compact(keep)
    assert b.n == 2 and b.origin == [0, 1]
    b.compact([0, 1])
    b.solver(3)
    assert b.status(-1).err == 0
    b.close()
