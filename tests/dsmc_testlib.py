"""Shared helpers of the tests: small cases and the oracle's phase-by-phase step."""
import numpy as np

from cudadsmc_amd import abi, cases
from oracle import oracle as O

FIELDS = ("x", "y", "vx", "vy", "vz", "pid")


def small_box(n_cells=16, ppc=8, seed=7):
    """The closed-box equilibrium case on a smaller grid (same physics)."""
    cfg = cases.box_equilibrium(ppc=ppc, n_cells=n_cells, seed=seed)
    return cfg["params"], None


def small_cylinder(nx=80, ny=40, ppc=10, n_points=64, seed=11, wall=abi.BC_DIFFUSE, gas=cases.ARGON, rot_prob=1.0):
    """The Mach-10 cylinder case on a coarse grid (same physics, fewer cells)."""
    cfg = cases.cylinder_mach10(ppc=ppc, nx=nx, ny=ny, n_points=n_points, seed=seed, gas=gas, rot_prob=rot_prob)
    p = cfg["params"]
    p.wall_bc = wall
    return p, cfg["body"]


def oracle_fill(oc, info, tables):
    C_ = info.ncell
    e = info.ppc_eff * tables["frac"]
    ux = np.full(C_, info.inj_ux, np.float32)
    uy = np.full(C_, info.inj_uy, np.float32)
    sg = np.full(C_, info.inj_sig, np.float32)
    parts, count, err = oc.init_fill(e, tables["cut"], ux, uy, sg, int(info.capacity))
    return parts, count, err


def fields_of(oc):
    """Particle fields of a case: with rotational energy (zeta = 2), er too."""
    return FIELDS + ("er",) if oc.zeta else FIELDS


def oracle_phases(oc, parts, step, next_pid, sigcmax, mrem, cap):
    """One oracle step split like the GPU phases.  parts is not modified."""
    FIELDS = fields_of(oc)
    s = {k: parts[k].copy() for k in FIELDS}
    n = s["x"].shape[0]
    dead, err = oc.move(s, step)
    J = oc.inject_counts(step)
    Jt = int(J.sum())
    if n + Jt > cap:
        Jt = cap - n
    inj, idead, ierr = oc.inject(step, next_pid, Jt)
    allp = {k: np.concatenate([s[k], inj[k]]) for k in FIELDS}
    alld = np.concatenate([dead, idead]).astype(np.uint8)
    cell, count, slot = oc.bin(allp["x"], allp["y"], alld)
    start = oc.scan(count)
    perm = oc.scatter(cell, slot, start, int(start[-1]))
    sm = sigcmax.copy()
    mr = mrem.copy()
    vx, vy, vz = allp["vx"].copy(), allp["vy"].copy(), allp["vz"].copy()
    er = allp["er"].copy() if oc.zeta else None
    ncoll = oc.collide(step, start, perm, vx, vy, vz, sm, mr, er=er, x=allp["x"], y=allp["y"])
    acc = np.zeros(oc.info.sample_blocks * oc.ncell)
    oc.sample(start, perm, vx, vy, vz, ncoll, acc, er=er)
    new = {k: allp[k][perm] for k in ("x", "y", "pid")}
    new["vx"], new["vy"], new["vz"] = vx[perm], vy[perm], vz[perm]
    if oc.zeta:
        new["er"] = er[perm]
    cell_alive = cell.copy()
    return dict(moved=allp, dead=alld, err=err | ierr, cell=cell_alive, count=count, start=start,
                perm=perm, new=new, sigcmax=sm, mrem=mr, ncoll=ncoll, acc=acc, Jt=Jt)


def assert_same_particles(a, b, fields=FIELDS, what=""):
    for k in fields:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, f"{what}{k}: shape {x.shape} vs {y.shape}"
        if x.dtype.kind == "f":
            bad = ~((x == y) | (np.isnan(x) & np.isnan(y)))
        else:
            bad = x != y
        nb = int(bad.sum())
        assert nb == 0, (f"{what}{k}: {nb} of {x.size} differ; first at {np.argmax(bad)}: "
                         f"{x[np.argmax(bad)]!r} vs {y[np.argmax(bad)]!r}")
