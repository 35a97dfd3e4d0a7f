"""Batched DSMC instances through the C ABI (dsmc_batch_*): many independent
cases advanced together, one launch per phase per step.

The reference runs one case per process (Run::solver, Run.cpp:259-304);
Batch keeps Run's method names so a sweep reads like a loop of runs.
"""
import ctypes as C

import numpy as np

from . import abi
from .abi import check, ptr


class Batch:
    def __init__(self, params_list, bodies=None):
        self.lib = abi.load()
        self.params = list(params_list)
        n = len(self.params)
        if n < 1:
            raise ValueError("a batch needs at least one case")
        P = (abi.dsmc_params * n)(*self.params)
        self._xy = []
        B = None
        if bodies is not None:
            if len(bodies) != n:
                raise ValueError("one body (or None) per case")
            B = (abi.dsmc_body * n)()
            for i, body in enumerate(bodies):
                if body is None:
                    B[i] = abi.dsmc_body(0, None)
                    continue
                xy = np.ascontiguousarray(np.asarray(body, dtype=np.float64).reshape(-1, 2))
                self._xy.append(xy)
                B[i] = abi.dsmc_body(xy.shape[0], xy.ctypes.data_as(C.POINTER(C.c_double)))
        self.h = C.c_void_p()
        check(self.lib.dsmc_batch_create(P, B, n, C.byref(self.h)))
        self.origin = list(range(n))     # index each instance had when the batch was made
        self._describe()

    def _describe(self):
        self.n = len(self.params)
        self.info = []
        for i in range(self.n):
            inf = abi.dsmc_case_info()
            check(self.lib.dsmc_batch_case_info(self.h, i, C.byref(inf)))
            self.info.append(inf)
        self.ncells = [inf.ncell for inf in self.info]
        self.cbase = np.concatenate([[0], np.cumsum(self.ncells)]).astype(np.int64)

    def compact(self, keep):
        """Keeps the instances listed in keep (current indices, ascending)
        and frees the rest; each kept one continues bit-identically.
        self.origin maps the new indices to the ones the batch was made with."""
        keep = [int(k) for k in keep]
        arr = np.ascontiguousarray(keep, dtype=np.int32)
        check(self.lib.dsmc_batch_compact(self.h, ptr(arr, C.c_int), len(keep)))
        self.params = [self.params[k] for k in keep]
        self.origin = [self.origin[k] for k in keep]
        self._describe()

    # -- lifecycle (Run::initialize / finalize) --
    def initialize(self):
        check(self.lib.dsmc_batch_init_freestream(self.h))

    def initialize_fields(self, n, ux, uy, T):
        """Per-cell n, ux, uy, T of every instance: lists of arrays (one per
        instance) or flat arrays over all cells in instance order."""
        def flat(v):
            if isinstance(v, (list, tuple)):
                v = np.concatenate([np.asarray(a, np.float64).ravel() for a in v])
            v = np.ascontiguousarray(v, dtype=np.float64).ravel()
            if v.size != self.cbase[-1]:
                raise ValueError(f"expected {self.cbase[-1]} cell values, got {v.size}")
            return v
        a = [flat(v) for v in (n, ux, uy, T)]
        check(self.lib.dsmc_batch_init_fields(self.h, *[ptr(v, C.c_double) for v in a]))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.dsmc_batch_destroy(self.h)
            self.h = C.c_void_p()

    finalize = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping (Run::solver) --
    def solver(self, nsteps, graph=True):
        check(self.lib.dsmc_batch_run(self.h, int(nsteps), 1 if graph else 0))

    def sync(self):
        check(self.lib.dsmc_batch_sync(self.h))

    def status(self, i=-1):
        s = abi.dsmc_status()
        check(self.lib.dsmc_batch_status(self.h, int(i), C.byref(s)))
        return s

    # -- per-instance readout --
    def download(self, i):
        n = self.status(i).n_live
        out = [np.empty(max(n, 1), np.float32) for _ in range(5)]
        pid = np.empty(max(n, 1), np.uint32)
        nn = C.c_int64()
        check(self.lib.dsmc_batch_download_particles(self.h, int(i), n, *[ptr(v, C.c_float) for v in out],
                                                     ptr(pid, C.c_uint32), C.byref(nn)))
        return dict(x=out[0][:n], y=out[1][:n], vx=out[2][:n], vy=out[3][:n], vz=out[4][:n], pid=pid[:n])

    def cell_state(self, i):
        c = self.ncells[i]
        sm, mr, nc = np.empty(c, np.float32), np.empty(c, np.float32), np.empty(c, np.int32)
        check(self.lib.dsmc_batch_cell_state(self.h, int(i), ptr(sm, C.c_float), ptr(mr, C.c_float),
                                             ptr(nc, C.c_int)))
        return sm, mr, nc

    def reset_samples(self):
        check(self.lib.dsmc_batch_reset_samples(self.h))

    def samples(self, i):
        acc = np.empty(6 * self.ncells[i], np.float64)
        ns = C.c_uint64()
        check(self.lib.dsmc_batch_download_samples(self.h, int(i), ptr(acc, C.c_double), C.byref(ns)))
        return acc.reshape(6, -1), ns.value

    def counts(self):
        """Particle-count sums of every cell of every instance (instance
        order, one transfer) and the per-instance sampled-step counts."""
        out = np.empty(int(self.cbase[-1]), np.float64)
        ns = np.empty(self.n, np.uint64)
        check(self.lib.dsmc_batch_download_counts(self.h, ptr(out, C.c_double), ptr(ns, C.c_uint64)))
        return out, ns

    def moments(self):
        """The five moment sums of every cell of every instance -- rows sum 1,
        sum vx, sum vy, sum vz, sum v^2; columns in instance order, one
        transfer -- and the per-instance sampled-step counts."""
        out = np.empty((5, int(self.cbase[-1])), np.float64)
        ns = np.empty(self.n, np.uint64)
        check(self.lib.dsmc_batch_download_moments(self.h, ptr(out, C.c_double), ptr(ns, C.c_uint64)))
        return out, ns

    def wall_distance(self):
        """Signed distance (m) of every cell centre to its instance's body
        outline, negative inside (instance order; +inf without a body)."""
        out = np.empty(int(self.cbase[-1]), np.float64)
        check(self.lib.dsmc_batch_wall_distance(self.h, ptr(out, C.c_double)))
        return out

    def fields(self, i):
        c = self.ncells[i]
        out = [np.empty(c, np.float64) for _ in range(6)]
        check(self.lib.dsmc_batch_download_fields(self.h, int(i), *[ptr(v, C.c_double) for v in out]))
        shape = (self.params[i].N1, self.params[i].N2)
        return {k: v.reshape(shape) for k, v in zip(("n", "ux", "uy", "uz", "T", "p"), out)}

    # -- launch settings (results never depend on them) --
    def get_launch(self):
        L = abi.dsmc_launch()
        check(self.lib.dsmc_batch_get_launch(self.h, C.byref(L)))
        return L.as_dict()

    def set_launch(self, **kw):
        L = abi.dsmc_launch()
        check(self.lib.dsmc_batch_get_launch(self.h, C.byref(L)))
        for k, v in kw.items():
            if not hasattr(L, k):
                raise KeyError(k)
            setattr(L, k, int(v))
        check(self.lib.dsmc_batch_set_launch(self.h, C.byref(L)))

    # -- per-kernel timing --
    def set_timing(self, on=True, events=True):
        """Per-kernel durations (HIP events) and algorithmic bytes of the
        eager steps that follow; events=False keeps only the bytes."""
        check(self.lib.dsmc_batch_set_timing(self.h, 0 if not on else (1 if events else 2)))

    def kernel_times(self):
        ms, by = np.zeros(4), np.zeros(4)
        n = C.c_int64()
        check(self.lib.dsmc_batch_kernel_times(self.h, ptr(ms, C.c_double), ptr(by, C.c_double), C.byref(n)))
        names = ("k_move", "k_scan_runs", "k_place+k_rank", "k_cell")
        return {k: dict(ms=float(m), bytes=float(b)) for k, m, b in zip(names, ms, by)}, n.value
