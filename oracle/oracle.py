"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

An OracleCase is built from the product's host-side case description
(dsmc_case_describe: derived constants and cut-cell tables), so the oracle
and the kernels run on identical inputs; those host values are pinned
separately against closed forms in tests/test_case_setup.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# A build of a deliberately altered copy of the oracle (the planted-mistake
# check, tests/test_oracle_mutations.py) is loaded in place of the real one
# when this names it.
LIB_ENV = "DSMC_ORACLE_LIB"

BC_STREAM, BC_SPECULAR, BC_DIFFUSE = 0, 1, 2
SUB_MOVE, SUB_INJECT, SUB_COLLIDE, SUB_INIT = 1, 2, 3, 4


class or_case(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("ncell", C.c_int),
        ("L1", C.c_float), ("L2", C.c_float), ("dx", C.c_float), ("dy", C.c_float),
        ("dt", C.c_float), ("ntc_coef", C.c_float), ("K1", C.c_float), ("e_exp", C.c_float),
        ("body_bc", C.c_int), ("body_vmp", C.c_float), ("body_sig", C.c_float),
        ("face_bc", C.c_int * 4), ("face_vmp", C.c_float * 4), ("face_sig", C.c_float * 4),
        ("inj_e", C.c_double * 4), ("inj_a", C.c_float * 4),
        ("inj_vmp", C.c_float), ("inj_sig", C.c_float), ("inj_ux", C.c_float), ("inj_uy", C.c_float),
        ("margin", C.c_float),
        ("nseg", C.c_int),
        ("seg", C.POINTER(C.c_float)),
        ("cseg_start", C.POINTER(C.c_int)),
        ("cseg_idx", C.POINTER(C.c_int)),
        ("vol_ntc", C.POINTER(C.c_float)),
        ("fbox", C.POINTER(C.c_float)),
        ("fmask", C.POINTER(C.c_uint64)),
        ("seed", C.c_uint32),
        ("zeta", C.c_int), ("rot_p", C.c_float), ("lb_ia", C.c_float), ("lb_ib", C.c_float),
        ("body_ekt", C.c_float), ("face_ekt", C.c_float * 4), ("inj_ekt", C.c_float),
        ("sub", C.POINTER(C.c_uint32)),
    ]


class or_amr_gas(C.Structure):
    _fields_ = [("Fn", C.c_double), ("mass", C.c_double), ("diam", C.c_double), ("T_inf", C.c_double),
                ("omega", C.c_double), ("boltz", C.c_double)]


AMR_MAX, AMR_CAP, AMR_LEAVES = 16, 2048, 256

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U = C.POINTER(C.c_uint32)
_B = C.POINTER(C.c_uint8)


class or_state(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("cap", C.c_int),
        ("step", C.c_uint32), ("next_pid", C.c_uint32), ("err", C.c_uint32),
        ("x", _F), ("y", _F), ("vx", _F), ("vy", _F), ("vz", _F), ("pid", _U),
        ("bx", _F), ("by", _F), ("bvx", _F), ("bvy", _F), ("bvz", _F), ("bpid", _U),
        ("dead", _B), ("cell", _I), ("slot", _I), ("perm", _I),
        ("count", _I), ("start", _I), ("ncoll_cell", _I),
        ("sigcmax", _F), ("mrem", _F),
        ("acc", _D), ("nsamples", C.c_longlong),
        ("total_coll", C.c_longlong),
        ("er", _F), ("ber", _F),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get(LIB_ENV)
    if not path:
        path = LIB_PATH
        if not os.path.exists(path):
            build()
    lib = C.CDLL(path)
    P = C.POINTER(or_case)
    sig = {
        "or_philox": [_U, C.c_uint32, C.c_uint32, _U],
        "or_philox_n": [C.c_int, _U, _U, _U, _U],
        "or_logf_n": [C.c_int, _F, _F],
        "or_expf_n": [C.c_int, _F, _F],
        "or_sincos2pi_n": [C.c_int, _F, _F, _F],
        "or_move": [P, C.c_int, _F, _F, _F, _F, _F, _F, _U, C.c_uint32, _B, _U],
        "or_inject_counts": [P, C.c_uint32, _I],
        "or_inject": [P, C.c_uint32, C.c_uint32, C.c_int, _F, _F, _F, _F, _F, _F, _U, _B, _U],
        "or_bin": [P, C.c_int, _F, _F, _B, _I, _I, _I],
        "or_scan": [_I, C.c_int, _I],
        "or_scatter": [C.c_int, _I, _I, _I, _I],
        "or_collide": [P, C.c_uint32, _I, _I, _F, _F, _F, _F, _F, _F, _F, _F, _I, C.c_int],
        "or_amr_leaves": [C.c_uint32, _I, _I],
        "or_amr_grid": [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int],
        "or_amr_update": [P, C.POINTER(or_amr_gas), _D, C.c_longlong, _D, _U, _F, _D],
        "or_amr_leaf": [P, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int],
        "or_sample": [C.c_int, _I, _I, _F, _F, _F, _F, _I, _D],
        "or_reorder": [C.c_int, _I, _F, _F, _F, _F, _F, _F, _U, _F, _F, _F, _F, _F, _F, _U],
        "or_init_fill": [P, _D, _B, _F, _F, _F, C.c_int, _F, _F, _F, _F, _F, _F, _U, _I, _U],
        "or_step": [P, C.POINTER(or_state), C.c_int, C.c_int],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int if name in ("or_init_fill", "or_step", "or_amr_leaves", "or_amr_leaf") else None
    lib.or_amr_grid.restype = C.c_uint32
    _lib = lib
    return lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


# ---------------------------------------------------------------- primitives
def philox(ctr, k0, k1):
    """Philox4x32-10 of arrays: ctr (n,4) uint32, k0, k1 (n,) -> (n,4)."""
    ctr = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32).reshape(-1, 4))
    n = ctr.shape[0]
    k0 = np.ascontiguousarray(np.broadcast_to(np.asarray(k0, dtype=np.uint32), (n,)))
    k1 = np.ascontiguousarray(np.broadcast_to(np.asarray(k1, dtype=np.uint32), (n,)))
    out = np.empty((n, 4), np.uint32)
    load().or_philox_n(n, _p(ctr, C.c_uint32), _p(k0, C.c_uint32), _p(k1, C.c_uint32), _p(out, C.c_uint32))
    return out


def logf(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    load().or_logf_n(x.size, _p(x, C.c_float), _p(out, C.c_float))
    return out


def expf(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    load().or_expf_n(x.size, _p(x, C.c_float), _p(out, C.c_float))
    return out


def sincos2pi(u):
    u = np.ascontiguousarray(u, dtype=np.float32)
    s = np.empty_like(u)
    c = np.empty_like(u)
    load().or_sincos2pi_n(u.size, _p(u, C.c_float), _p(s, C.c_float), _p(c, C.c_float))
    return s, c


# ---------------------------------------------------------------- the case
class OracleCase:
    """or_case built from a product case description (info + tables)."""

    def __init__(self, info, tables, ncell_xy):
        self.info = info
        self.t = {k: np.ascontiguousarray(v) for k, v in tables.items()}
        c = or_case()
        c.nx, c.ny = ncell_xy
        c.ncell = info.ncell
        for k in ("L1", "L2", "dx", "dy", "ntc_coef", "K1", "e_exp", "body_vmp", "body_sig",
                  "inj_vmp", "inj_sig", "inj_ux", "inj_uy", "margin", "body_bc", "seed"):
            setattr(c, k, getattr(info, k))
        c.dt = np.float32(info.dt)
        for k in ("zeta", "rot_p", "lb_ia", "lb_ib", "body_ekt", "inj_ekt"):
            setattr(c, k, getattr(info, k))
        for f in range(4):
            c.face_ekt[f] = info.face_ekt[f]
            c.face_bc[f] = info.face_bc[f]
            c.face_vmp[f] = info.face_vmp[f]
            c.face_sig[f] = info.face_sig[f]
            c.inj_e[f] = info.inj_e[f]
            c.inj_a[f] = info.inj_a[f]
        c.nseg = info.nseg
        c.seg = _p(self.t["seg"], C.c_float)
        c.cseg_start = _p(self.t["cseg_start"], C.c_int)
        c.cseg_idx = _p(self.t["cseg_idx"], C.c_int)
        c.vol_ntc = _p(self.t["vol_ntc"], C.c_float)
        c.fbox = _p(self.t["fbox"], C.c_float)
        c.fmask = _p(self.t["fmask"], C.c_uint64)
        c.sub = None
        self._sub = None
        self.c = c
        self.ptr = C.byref(c)

    @property
    def ncell(self):
        return self.info.ncell

    @property
    def zeta(self):
        return self.info.zeta

    # --- phases ---
    def move(self, s, step):
        """s: dict of x,y,vx,vy,vz (float32), pid and, zeta = 2, er; moved in
        place; returns dead, err."""
        n = s["x"].shape[0]
        dead = np.zeros(n, np.uint8)
        err = C.c_uint32(0)
        er = s["er"] if self.zeta else None
        load().or_move(self.ptr, n, *[_p(s[k], C.c_float) for k in ("x", "y", "vx", "vy", "vz")],
                       _p(er, C.c_float), _p(s["pid"], C.c_uint32), step, _p(dead, C.c_uint8), C.byref(err))
        return dead, err.value

    def inject_counts(self, step):
        J = np.zeros(4, np.int32)
        load().or_inject_counts(self.ptr, step, _p(J, C.c_int))
        return J

    def inject(self, step, next_pid, Jt):
        names = ("x", "y", "vx", "vy", "vz", "er") if self.zeta else ("x", "y", "vx", "vy", "vz")
        out = {k: np.zeros(Jt, np.float32) for k in names}
        out["pid"] = np.zeros(Jt, np.uint32)
        dead = np.zeros(Jt, np.uint8)
        err = C.c_uint32(0)
        load().or_inject(self.ptr, step, next_pid, Jt, *[_p(out[k], C.c_float) for k in ("x", "y", "vx", "vy", "vz")],
                         _p(out.get("er"), C.c_float), _p(out["pid"], C.c_uint32), _p(dead, C.c_uint8),
                         C.byref(err))
        return out, dead, err.value

    def bin(self, x, y, dead=None):
        n = x.shape[0]
        cell = np.zeros(n, np.int32)
        slot = np.zeros(n, np.int32)
        count = np.zeros(self.ncell, np.int32)
        load().or_bin(self.ptr, n, _p(x, C.c_float), _p(y, C.c_float), _p(dead, C.c_uint8),
                      _p(cell, C.c_int), _p(count, C.c_int), _p(slot, C.c_int))
        return cell, count, slot

    def scan(self, count):
        count = np.ascontiguousarray(count, dtype=np.int32)
        start = np.zeros(count.shape[0] + 1, np.int32)
        load().or_scan(_p(count, C.c_int), count.shape[0], _p(start, C.c_int))
        return start

    def scatter(self, cell, slot, start, n_out):
        perm = np.full(n_out, -1, np.int32)
        load().or_scatter(cell.shape[0], _p(cell, C.c_int), _p(slot, C.c_int), _p(start, C.c_int), _p(perm, C.c_int))
        return perm

    def collide(self, step, start, perm, vx, vy, vz, sigcmax, mrem, nthreads=1, er=None, x=None, y=None):
        """er: rotational energies (zeta = 2), updated in place; required then.
        x, y: positions, required once the case has sub-cells (set_subcells),
        whose cells' perm segments are then reordered by leaf in place."""
        if self.zeta and er is None:
            raise ValueError("a diatomic case collides with its rotational energies")
        if self._sub is not None and (x is None or y is None):
            raise ValueError("collisions in sub-cells need the positions")
        ncoll = np.zeros(self.ncell, np.int32)
        load().or_collide(self.ptr, step, _p(start, C.c_int), _p(perm, C.c_int), _p(x, C.c_float),
                          _p(y, C.c_float), _p(vx, C.c_float),
                          _p(vy, C.c_float), _p(vz, C.c_float), _p(er if self.zeta else None, C.c_float),
                          _p(sigcmax, C.c_float), _p(mrem, C.c_float), _p(ncoll, C.c_int), nthreads)
        return ncoll

    # --- adaptive collision sub-cells (Cell::genAMRmesh, Cell::VTS) ---
    def set_subcells(self, sub):
        """Packed per-cell grids (n2x | n2y << 8 | n3x << 16 | n3y << 24;
        0 = none), or None for none at all."""
        self._sub = None if sub is None else np.ascontiguousarray(sub, dtype=np.uint32)
        self.c.sub = _p(self._sub, C.c_uint32)

    def amr_update(self, gas, acc, nsamples, vol, sub):
        """gas: dict Fn, mass, diam, T_inf, omega, boltz.  Returns the new
        packed grids (sub is not modified), mean free paths and VTS steps."""
        g = or_amr_gas(*(float(gas[k]) for k in ("Fn", "mass", "diam", "T_inf", "omega", "boltz")))
        acc = np.ascontiguousarray(acc, dtype=np.float64)
        vol = np.ascontiguousarray(vol, dtype=np.float64)
        out = np.array(sub, dtype=np.uint32, copy=True)
        lam = np.zeros(self.ncell, np.float32)
        dt = np.zeros(self.ncell, np.float64)
        load().or_amr_update(self.ptr, C.byref(g), _p(acc, C.c_double), int(nsamples), _p(vol, C.c_double),
                             _p(out, C.c_uint32), _p(lam, C.c_float), _p(dt, C.c_double))
        return out, lam, dt

    def amr_leaf(self, c, x, y, nsx, nsy):
        return load().or_amr_leaf(self.ptr, int(c), float(x), float(y), int(nsx), int(nsy))


    def sample(self, start, perm, vx, vy, vz, ncoll, acc, er=None):
        load().or_sample(self.ncell, _p(start, C.c_int), _p(perm, C.c_int), _p(vx, C.c_float), _p(vy, C.c_float),
                         _p(vz, C.c_float), _p(er if self.zeta else None, C.c_float), _p(ncoll, C.c_int),
                         _p(acc, C.c_double))

    def init_fill(self, expected, cut, ux, uy, sig, cap):
        names = ("x", "y", "vx", "vy", "vz", "er") if self.zeta else ("x", "y", "vx", "vy", "vz")
        arrs = {k: np.zeros(cap, np.float32) for k in names}
        arrs["pid"] = np.zeros(cap, np.uint32)
        count = np.zeros(self.ncell, np.int32)
        err = C.c_uint32(0)
        e = np.ascontiguousarray(expected, dtype=np.float64)
        cut = np.ascontiguousarray(cut, dtype=np.uint8)
        ux, uy, sig = (np.ascontiguousarray(v, dtype=np.float32) for v in (ux, uy, sig))
        n = load().or_init_fill(self.ptr, _p(e, C.c_double), _p(cut, C.c_uint8), _p(ux, C.c_float),
                                _p(uy, C.c_float), _p(sig, C.c_float), cap,
                                *[_p(arrs[k], C.c_float) for k in ("x", "y", "vx", "vy", "vz")],
                                _p(arrs.get("er"), C.c_float), _p(arrs["pid"], C.c_uint32), _p(count, C.c_int),
                                C.byref(err))
        if n < 0:
            raise RuntimeError("oracle fill exceeds capacity")
        return {k: v[:n].copy() for k, v in arrs.items()}, count, err.value


class OracleState:
    """Full-step state for or_step (buffers owned here)."""

    def __init__(self, case, parts, cap, sigcmax0, step=0, next_pid=None):
        self.case = case
        C_ = case.ncell
        n = parts["x"].shape[0]
        self.bufs = {}
        st = or_state()
        st.n, st.cap, st.step = n, cap, step
        st.next_pid = int(parts["pid"].max()) + 1 if next_pid is None and n else (next_pid or 0)
        self.keys = ("x", "y", "vx", "vy", "vz", "er") if case.zeta else ("x", "y", "vx", "vy", "vz")
        for k in self.keys:
            a = np.zeros(cap, np.float32)
            a[:n] = parts[k]
            b = np.zeros(cap, np.float32)
            self.bufs[k], self.bufs["b" + k] = a, b
            setattr(st, k, _p(a, C.c_float))
            setattr(st, "b" + k, _p(b, C.c_float))
        a = np.zeros(cap, np.uint32)
        a[:n] = parts["pid"]
        b = np.zeros(cap, np.uint32)
        self.bufs["pid"], self.bufs["bpid"] = a, b
        st.pid, st.bpid = _p(a, C.c_uint32), _p(b, C.c_uint32)
        self.dead = np.zeros(cap, np.uint8)
        self.cell = np.zeros(cap, np.int32)
        self.slot = np.zeros(cap, np.int32)
        self.perm = np.zeros(cap, np.int32)
        self.count = np.zeros(C_, np.int32)
        self.start = np.zeros(C_ + 1, np.int32)
        self.ncoll = np.zeros(C_, np.int32)
        self.sigcmax = np.full(C_, np.float32(sigcmax0), np.float32)
        self.mrem = np.zeros(C_, np.float32)
        self.acc = np.zeros(case.info.sample_blocks * C_, np.float64)
        st.dead = _p(self.dead, C.c_uint8)
        st.cell, st.slot, st.perm = _p(self.cell, C.c_int), _p(self.slot, C.c_int), _p(self.perm, C.c_int)
        st.count, st.start, st.ncoll_cell = _p(self.count, C.c_int), _p(self.start, C.c_int), _p(self.ncoll, C.c_int)
        st.sigcmax, st.mrem = _p(self.sigcmax, C.c_float), _p(self.mrem, C.c_float)
        st.acc = _p(self.acc, C.c_double)
        self.st = st

    def step(self, sample=False, nthreads=1):
        return load().or_step(self.case.ptr, C.byref(self.st), 1 if sample else 0, nthreads)

    def _cur(self, k):
        addr = C.cast(getattr(self.st, k), C.c_void_p).value
        for name in (k, "b" + k):
            if self.bufs[name].ctypes.data == addr:
                return self.bufs[name]
        raise RuntimeError("buffer lookup failed")

    def particles(self):
        n = self.st.n
        return {k: self._cur(k)[:n].copy() for k in self.keys + ("pid",)}


def case_from_params(params, body=None):
    """OracleCase from the product's host-side case description."""
    import cudadsmc_amd.run as run
    info, tables = run.describe(params, body)
    return OracleCase(info, tables, (params.N1, params.N2)), info, tables


def amr_grid(dx, dy, lam_max, lam_min, had):
    """Packed two-level grid of one cell (or_amr_grid)."""
    return int(load().or_amr_grid(float(dx), float(dy), float(lam_max), float(lam_min), 1 if had else 0))


def amr_leaves(sub):
    """(leaves, nsx, nsy) of a packed grid."""
    a, b = C.c_int(), C.c_int()
    n = load().or_amr_leaves(int(sub), C.byref(a), C.byref(b))
    return n, a.value, b.value
