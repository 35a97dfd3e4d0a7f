"""Run: the driver of one DSMC instance over the C ABI.

Mirrors the reference driver's interface (src/Run.h:30-44 --
initialize, solver, particlemove, ressignParticle, collision, finalize and
the Transfer* methods) with the state on the GPU; every call goes to
libcudadsmc.so.
"""
import ctypes as C

import numpy as np

from . import abi
from .abi import check, ptr


def _body_struct(body):
    if body is None:
        return None, None
    xy = np.ascontiguousarray(np.asarray(body, dtype=np.float64).reshape(-1, 2))
    return abi.dsmc_body(xy.shape[0], xy.ctypes.data_as(C.POINTER(C.c_double))), xy


def describe(params, body=None):
    """Host-only case description (no device needed): derived constants and
    cut-cell tables, as dsmc_case_describe returns them."""
    lib = abi.load()
    b, _xy = _body_struct(body)
    bptr = C.byref(b) if b is not None else None
    info = abi.dsmc_case_info()
    check(lib.dsmc_case_describe(C.byref(params), bptr, C.byref(info), None, None, None, None, None, None, None, None,
                                 None))
    nc = info.ncell
    t = dict(seg=np.empty(6 * info.nseg, np.float32), cseg_start=np.empty(nc + 1, np.int32),
             cseg_idx=np.empty(max(info.cseg_len, 1), np.int32), vol=np.empty(nc, np.float64),
             vol_ntc=np.empty(nc, np.float32), cut=np.empty(nc, np.uint8), frac=np.empty(nc, np.float64),
             fbox=np.empty(4 * nc, np.float32), fmask=np.empty(nc, np.uint64))
    check(lib.dsmc_case_describe(C.byref(params), bptr, C.byref(info), ptr(t["seg"], C.c_float),
                                 ptr(t["cseg_start"], C.c_int), ptr(t["cseg_idx"], C.c_int),
                                 ptr(t["vol"], C.c_double), ptr(t["vol_ntc"], C.c_float),
                                 ptr(t["cut"], C.c_uint8), ptr(t["frac"], C.c_double),
                                 ptr(t["fbox"], C.c_float), ptr(t["fmask"], C.c_uint64)))
    return info, t


def outline_segments(body):
    """Wall segments of a counter-clockwise body outline, 6 per edge (ax ay bx
    by nx ny, n = (ey, -ex)/|e| into the fluid), in float64: the segment table
    a host mesh layer hands to dsmc_create_from_tables."""
    xy = np.asarray(body, dtype=np.float64).reshape(-1, 2)
    a, b = xy, np.roll(xy, -1, axis=0)
    e = b - a
    L = np.hypot(e[:, 0], e[:, 1])
    return np.ascontiguousarray(np.stack([a[:, 0], a[:, 1], b[:, 0], b[:, 1], e[:, 1] / L, -e[:, 0] / L], 1).ravel())


def subcell_counts(sub):
    """Numchild (level-2 cells) and Numgrandchild (level-3 cells in all of
    them) of packed grids, as the reference's Output counts them
    (Output.cpp:45-63); 0 for a cell without a grid."""
    s = np.asarray(sub, dtype=np.uint32)
    n2 = (s & 0xff) * ((s >> 8) & 0xff)
    n3 = ((s >> 16) & 0xff) * ((s >> 24) & 0xff)
    return dict(Numchild=n2.astype(np.int32), Numgrandchild=(n2 * n3).astype(np.int32))


class Run:
    def __init__(self, params, body=None, tables=None):
        """body: the body outline (x, y rows); or tables: host-built cut-cell
        tables, a dict with seg (6 per wall segment, float64) and optional
        frac (fluid fraction per cell) and cut (flags) -- GPUCells::
        UploadFromHost's role (dsmc_create_from_tables)."""
        self.lib = abi.load()
        self.params = params
        self.h = C.c_void_p()
        self._body_xy = None
        if tables is not None:
            seg = np.ascontiguousarray(tables["seg"], dtype=np.float64)
            self._tab = dict(seg=seg)
            t = abi.dsmc_cut_tables(seg.size // 6, ptr(seg, C.c_double), None, None)
            if tables.get("frac") is not None:
                self._tab["frac"] = np.ascontiguousarray(tables["frac"], dtype=np.float64)
                t.frac = ptr(self._tab["frac"], C.c_double)
            if tables.get("cut") is not None:
                self._tab["cut"] = np.ascontiguousarray(tables["cut"], dtype=np.uint8)
                t.cut = ptr(self._tab["cut"], C.c_uint8)
            if seg.size:
                self._body_xy = seg.reshape(-1, 6)[:, :2].copy()
            check(self.lib.dsmc_create_from_tables(C.byref(params), C.byref(t), C.byref(self.h)))
        else:
            bptr = None
            if body is not None:
                xy = np.ascontiguousarray(np.asarray(body, dtype=np.float64).reshape(-1, 2))
                self._body_xy = xy
                b = abi.dsmc_body(xy.shape[0], xy.ctypes.data_as(C.POINTER(C.c_double)))
                self._body = b
                bptr = C.byref(b)
            check(self.lib.dsmc_create(C.byref(params), bptr, C.byref(self.h)))
        self.info = abi.dsmc_case_info()
        check(self.lib.dsmc_case_info_get(self.h, C.byref(self.info)))
        self.ncell = self.info.ncell

    # -- lifecycle (Run::initialize / finalize, Run.cpp:29-106,306-311) --
    def initialize(self):
        """Freestream fill of every fluid cell (Run::assignParticle)."""
        check(self.lib.dsmc_init_freestream(self.h))

    def initialize_fields(self, n, ux, uy, T):
        """Fill from per-cell number density, velocity and temperature."""
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (n, ux, uy, T)]
        check(self.lib.dsmc_init_fields(self.h, *[ptr(v, C.c_double) for v in a]))

    def finalize(self):
        if self.h:
            self.lib.dsmc_destroy(self.h)
            self.h = C.c_void_p()

    close = finalize

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass

    # -- the three phases of Run::solver (Run.cpp:159-162,245-256) --
    def particlemove(self):
        check(self.lib.dsmc_move(self.h))

    def ressignParticle(self):
        check(self.lib.dsmc_reassign(self.h))

    def collision(self):
        check(self.lib.dsmc_collide(self.h))

    def solver(self, nsteps, graph=True):
        check(self.lib.dsmc_run(self.h, int(nsteps), 1 if graph else 0))

    def sync(self):
        check(self.lib.dsmc_sync(self.h))

    # -- state transfer (Run::Transfer*, Run.cpp:164-238) --
    def status(self):
        s = abi.dsmc_status()
        check(self.lib.dsmc_status_get(self.h, C.byref(s)))
        return s

    def upload(self, x, y, vx, vy, vz, pid=None):
        a = [abi.f32(v) for v in (x, y, vx, vy, vz)]
        n = a[0].shape[0]
        p = None if pid is None else np.ascontiguousarray(pid, dtype=np.uint32)
        check(self.lib.dsmc_upload_particles(self.h, n, *[ptr(v, C.c_float) for v in a],
                                             ptr(p, C.c_uint32)))

    def download(self, moved=False):
        """Current state (n_live particles), or with moved=True the n_move
        particles processed by the last particlemove(); a diatomic gas
        (zeta = 2) adds er, the rotational energy per unit mass (J/kg)."""
        s = self.status()
        n = s.n_move if moved else s.n_live
        out = [np.empty(max(n, 1), np.float32) for _ in range(5)]
        pid = np.empty(max(n, 1), np.uint32)
        nn = C.c_int64()
        fn = self.lib.dsmc_download_moved if moved else self.lib.dsmc_download_particles
        check(fn(self.h, n, *[ptr(v, C.c_float) for v in out], ptr(pid, C.c_uint32), C.byref(nn)))
        d = dict(x=out[0][:n], y=out[1][:n], vx=out[2][:n], vy=out[3][:n], vz=out[4][:n], pid=pid[:n])
        if self.info.zeta:
            er = np.empty(max(n, 1), np.float32)
            check(self.lib.dsmc_download_rotational(self.h, 1 if moved else 0, n, ptr(er, C.c_float), C.byref(nn)))
            d["er"] = er[:n]
        return d

    def upload_rotational(self, er):
        """Rotational energies per unit mass of the current state (zeta = 2)."""
        e = abi.f32(er)
        check(self.lib.dsmc_upload_rotational(self.h, e.shape[0], ptr(e, C.c_float)))

    def tables(self):
        i = self.info
        seg = np.empty(6 * i.nseg, np.float32)
        cs = np.empty(i.ncell + 1, np.int32)
        ci = np.empty(max(i.cseg_len, 1), np.int32)
        vol = np.empty(i.ncell, np.float64)
        vn = np.empty(i.ncell, np.float32)
        cut = np.empty(i.ncell, np.uint8)
        fb = np.empty(4 * i.ncell, np.float32)
        fm = np.empty(i.ncell, np.uint64)
        check(self.lib.dsmc_case_tables(self.h, ptr(seg, C.c_float), ptr(cs, C.c_int), ptr(ci, C.c_int),
                                        ptr(vol, C.c_double), ptr(vn, C.c_float), ptr(cut, C.c_uint8),
                                        ptr(fb, C.c_float), ptr(fm, C.c_uint64)))
        return dict(seg=seg, cseg_start=cs, cseg_idx=ci, vol=vol, vol_ntc=vn, cut=cut, fbox=fb, fmask=fm)

    def bins(self):
        s = self.status()
        cell = np.empty(max(s.n_move, 1), np.int32)
        count = np.empty(self.ncell, np.int32)
        check(self.lib.dsmc_download_bins(self.h, s.n_move, ptr(cell, C.c_int), ptr(count, C.c_int)))
        return cell[:s.n_move], count

    def start(self):
        st = np.empty(self.ncell + 1, np.int32)
        check(self.lib.dsmc_download_start(self.h, ptr(st, C.c_int)))
        return st

    def perm(self):
        s = self.status()
        p = np.empty(max(s.n_live, 1), np.int32)
        check(self.lib.dsmc_download_perm(self.h, s.n_live, ptr(p, C.c_int)))
        return p[:s.n_live]

    def set_keep_perm(self, on=True):
        check(self.lib.dsmc_set_keep_perm(self.h, 1 if on else 0))

    # -- adaptive collision sub-cells and the variable time step
    #    (Cell::genAMRmesh, Cell::VTS, Cell.cpp:117-279) --
    def set_subcells(self, sub):
        """Packed per-cell grids (n2x | n2y << 8 | n3x << 16 | n3y << 24; 0:
        none), or None to switch sub-cells off."""
        if sub is None:
            check(self.lib.dsmc_set_subcells(self.h, None))
            return
        a = np.ascontiguousarray(sub, dtype=np.uint32)
        if a.shape != (self.ncell,):
            raise ValueError("one grid per cell")
        check(self.lib.dsmc_set_subcells(self.h, ptr(a, C.c_uint32)))

    def amr_update(self):
        """Grids, mean free paths and VTS time steps from the sampling sums."""
        check(self.lib.dsmc_amr_update(self.h))

    def amr(self):
        """dict sub (packed grids), lam (mean free path, m), dt (VTS, s),
        Numchild and Numgrandchild (the output datasets' counts)."""
        sub = np.empty(self.ncell, np.uint32)
        lam = np.empty(self.ncell, np.float32)
        dt = np.empty(self.ncell, np.float64)
        check(self.lib.dsmc_download_amr(self.h, ptr(sub, C.c_uint32), ptr(lam, C.c_float),
                                         ptr(dt, C.c_double)))
        return dict(sub=sub, lam=lam, dt=dt, **subcell_counts(sub))

    def cell_state(self):
        sm = np.empty(self.ncell, np.float32)
        mr = np.empty(self.ncell, np.float32)
        nc = np.empty(self.ncell, np.int32)
        check(self.lib.dsmc_cell_state(self.h, ptr(sm, C.c_float), ptr(mr, C.c_float), ptr(nc, C.c_int)))
        return sm, mr, nc

    def set_cell_state(self, sigcmax=None, mrem=None):
        a = None if sigcmax is None else abi.f32(sigcmax)
        b = None if mrem is None else abi.f32(mrem)
        check(self.lib.dsmc_set_cell_state(self.h, ptr(a, C.c_float), ptr(b, C.c_float)))

    # -- checkpoint / resume (no counterpart in the reference) --
    def save_state(self, path):
        """Writes everything a bit-identical continuation needs to an .npz:
        particles in stored order with ids, cell state, counters, samples."""
        self.sync()
        s = self.status()
        p = self.download()
        sm, mr, _ = self.cell_state()
        acc, ns = self.samples()
        np.savez(path, **p, sigcmax=sm, mrem=mr, acc=acc, nsamples=np.uint64(ns),
                 counters=np.array([s.step, s.next_pid, s.particle_steps, s.collisions,
                                    s.last_collisions], np.uint64),
                 ncell=np.int64(self.ncell))

    def load_state(self, path):
        """Restores a state written by save_state into this (same-case) run."""
        with np.load(path, allow_pickle=False) as z:
            if int(z["ncell"]) != self.ncell:
                raise ValueError(f"checkpoint has {int(z['ncell'])} cells, this case {self.ncell}")
            self.upload(z["x"], z["y"], z["vx"], z["vy"], z["vz"], z["pid"])
            if self.info.zeta:
                self.upload_rotational(z["er"])
            self.set_cell_state(z["sigcmax"], z["mrem"])
            c = [int(v) for v in z["counters"]]
            s = abi.dsmc_status()
            s.step, s.next_pid, s.particle_steps, s.collisions, s.last_collisions = c
            s.nsamples = int(z["nsamples"])
            check(self.lib.dsmc_restore_counters(self.h, C.byref(s)))
            check(self.lib.dsmc_upload_samples(self.h, ptr(np.ascontiguousarray(z["acc"], np.float64).ravel(),
                                                           C.c_double), int(z["nsamples"])))

    # -- sampling (GPUCells::Sample / DownloadToHost) --
    def sample(self):
        """One sample of the current state (GPUCells::Sample, which the
        reference's Run::solver calls every 100 steps)."""
        check(self.lib.dsmc_sample(self.h))

    def reset_samples(self):
        check(self.lib.dsmc_reset_samples(self.h))

    def samples(self):
        """Sample sums, (sample_blocks, ncell): n, vx, vy, vz, v^2, collisions
        and, for zeta = 2, the rotational energy per unit mass."""
        nb = self.info.sample_blocks
        acc = np.empty(nb * self.ncell, np.float64)
        ns = C.c_uint64()
        check(self.lib.dsmc_download_samples(self.h, ptr(acc, C.c_double), C.byref(ns)))
        return acc.reshape(nb, self.ncell), ns.value

    def fields(self):
        """n, ux, uy, uz, T (translational), p; with zeta = 2 also Trot and
        Tov = (3 T + zeta Trot) / (3 + zeta)."""
        out = [np.empty(self.ncell, np.float64) for _ in range(6)]
        check(self.lib.dsmc_download_fields(self.h, *[ptr(v, C.c_double) for v in out]))
        shape = (self.params.N1, self.params.N2)
        names = ("n", "ux", "uy", "uz", "T", "p")
        f = {k: v.reshape(shape) for k, v in zip(names, out)}
        z = self.info.zeta
        if z:
            tr = np.empty(self.ncell, np.float64)
            check(self.lib.dsmc_download_rot_temperature(self.h, ptr(tr, C.c_double)))
            f["Trot"] = tr.reshape(shape)
            f["Tov"] = (3.0 * f["T"] + z * f["Trot"]) / (3.0 + z)
        return f

    # -- launch settings (results never depend on them) --
    def get_launch(self):
        L = abi.dsmc_launch()
        check(self.lib.dsmc_get_launch(self.h, C.byref(L)))
        return L.as_dict()

    def set_launch(self, **kw):
        L = abi.dsmc_launch()
        check(self.lib.dsmc_get_launch(self.h, C.byref(L)))
        for k, v in kw.items():
            if not hasattr(L, k):
                raise KeyError(k)
            setattr(L, k, int(v))
        check(self.lib.dsmc_set_launch(self.h, C.byref(L)))

    # -- per-kernel timing --
    def set_timing(self, on=True, events=True):
        """Per-kernel durations (HIP events) and algorithmic bytes of the
        eager steps that follow; events=False keeps only the bytes."""
        check(self.lib.dsmc_set_timing(self.h, 0 if not on else (1 if events else 2)))

    def kernel_times(self):
        ms = np.zeros(4)
        by = np.zeros(4)
        n = C.c_int64()
        check(self.lib.dsmc_kernel_times(self.h, ptr(ms, C.c_double), ptr(by, C.c_double), C.byref(n)))
        names = ("k_move", "k_scan_runs", "k_place+k_rank", "k_cell")
        return {k: dict(ms=float(m), bytes=float(b)) for k, m, b in zip(names, ms, by)}, n.value
