"""Autotuner of the launch settings (block counts, k_cell tile size, steps
per captured graph) of a Run or a Batch.

Every candidate starts from the same state and runs the same steps; the
per-kernel times come from HIP events on the library's stream, the graph
length from the wall time of graph replays.  The settings never change the
results, only the time -- the tuner checks that as it goes: the final
state of every candidate must equal the first candidate's, bit for bit,
or the tuner raises.

Search: coordinate descent, one setting at a time in the order tile size
(k_place + k_rank + k_cell), k_cell blocks, k_move blocks, the blocks of
k_place and k_rank,
graph length; each keeps the best value found before moving on.
"""
import hashlib
import io
import time

import numpy as np

TILES = (512, 768, 1024, 1536, 2048)
CELL_BLOCKS_PER_CU = (1, 2, 3)
MOVE_BLOCKS_PER_CU = (1, 2, 4, 8)
SCATTER_BLOCKS_PER_CU = (1, 2, 4, 8)
GRAPH_STEPS = (4, 8, 16, 32)


def device_cus():
    """Compute units of the current HIP device."""
    import ctypes as C
    from . import abi
    n = C.c_int(0)
    abi.check(abi.load().dsmc_device_cus(C.byref(n)))
    return n.value


class RunTarget:
    """A Run: snapshots through save_state / load_state (bit-exact resume)."""

    def __init__(self, run):
        self.sim = run

    def snapshot(self):
        buf = io.BytesIO()
        self.sim.save_state(buf)
        return buf.getvalue()

    def restore(self, snap):
        self.sim.load_state(io.BytesIO(snap))

    def digest(self):
        p = self.sim.download()
        h = hashlib.sha256()
        for k in ("x", "y", "vx", "vy", "vz", "pid"):
            h.update(np.ascontiguousarray(p[k]).tobytes())
        s = self.sim.status()
        h.update(np.array([s.n_live, s.collisions, s.step], np.uint64).tobytes())
        return h.hexdigest()


class BatchTarget:
    """A Batch: the snapshot is a deterministic re-start (init + warm steps);
    the digest covers every instance's counters and the particles of the
    first and last instance."""

    def __init__(self, batch, init, warm_steps):
        self.sim = batch
        self.init = init
        self.warm = warm_steps

    def snapshot(self):
        return None

    def restore(self, _):
        self.init(self.sim)
        if self.warm:
            self.sim.solver(self.warm, graph=False)

    def digest(self):
        h = hashlib.sha256()
        for i in range(self.sim.n):
            s = self.sim.status(i)
            h.update(np.array([s.n_live, s.collisions, s.next_pid, s.nsamples], np.uint64).tobytes())
        for i in {0, self.sim.n - 1}:
            p = self.sim.download(i)
            for k in ("x", "y", "vx", "vy", "vz", "pid"):
                h.update(np.ascontiguousarray(p[k]).tobytes())
        return h.hexdigest()


def _measure(target, snap, cfg, steps):
    sim = target.sim
    target.restore(snap)
    sim.set_launch(**cfg)
    sim.set_timing(True)
    sim.solver(steps, graph=False)
    kt, _ = sim.kernel_times()
    sim.set_timing(False)
    return {k: v["ms"] for k, v in kt.items()}, target.digest()


def _graph_time(target, snap, cfg, steps):
    sim = target.sim
    target.restore(snap)
    sim.set_launch(**cfg)
    sim.solver(0, graph=True)                       # capture outside the timed region
    sim.sync()
    t0 = time.perf_counter()
    sim.solver(steps, graph=True)
    sim.sync()
    return (time.perf_counter() - t0) / steps * 1e3, target.digest()


def tune(target, cus, steps=8, graph_steps_measure=64, log=print):
    """Returns (best settings, table of measurements).  The target's state
    afterwards is the snapshot state with the best settings applied."""
    sim = target.sim
    snap = target.snapshot()
    best = dict(sim.get_launch())
    table = []
    ref = None

    def check(dig, cfg):
        nonlocal ref
        if ref is None:
            ref = dig
        elif dig != ref:
            raise RuntimeError(f"launch settings changed the results: {cfg}")

    def sweep(key, values, cost):
        nonlocal best
        scores = []
        for v in values:
            cfg = dict(best, **{key: v})
            t, dig = _measure(target, snap, cfg, steps)
            check(dig, cfg)
            c = cost(t)
            scores.append((c, v))
            table.append(dict(cfg, **{f"ms_{k}": x for k, x in t.items()}))
            if log:
                log(f"[tune] {key}={v}: {c:.4f} ms")
        best[key] = min(scores)[1]

    sweep("cell_tile", TILES, lambda t: t["k_cell"] + t["k_place+k_rank"])
    sweep("cell_blocks", [cus * k for k in CELL_BLOCKS_PER_CU], lambda t: t["k_cell"])
    sweep("move_blocks", [cus * k for k in MOVE_BLOCKS_PER_CU], lambda t: t["k_move"])
    sweep("scatter_blocks", [cus * k for k in SCATTER_BLOCKS_PER_CU], lambda t: t["k_place+k_rank"])
    # graph length: wall time of replays; the number of timed steps is a
    # multiple of every candidate so that each replays whole graphs only
    gscores = []
    ref = None
    for g in GRAPH_STEPS:
        cfg = dict(best, graph_steps=g)
        ms, dig = _graph_time(target, snap, cfg, graph_steps_measure)
        check(dig, cfg)
        table.append(dict(cfg, ms_step_graph=ms))
        gscores.append((ms, g))
        if log:
            log(f"[tune] graph_steps={g}: {ms:.4f} ms/step")
    best["graph_steps"] = min(gscores)[1]
    target.restore(snap)
    sim.set_launch(**best)
    return best, table
