"""GPU: binning on caller device arrays (dsmc_bin_sort) against the oracle's
binning, scan and scatter (or_bin / or_scan / or_scatter, which restate
GPUCell.cu:169-198, :104-120 and GPUParticle.cu:365-373): cell ids, start
offsets, the stable permutation and local ids, bit-exact -- at the
benchmark sizes (the Mach-10 cylinder's 400x200 fill, about 2M particles;
the closed box's 64x64 fill), on shuffled random states, and on the edge
cases (no particles, the far faces x = L1 and y = L2, positions outside the
grid, cells of more than 64 and more than 4,096 particles, empty cells)."""
import ctypes as C

import numpy as np
import pytest

from cudadsmc_amd import abi, cases
from oracle import oracle as O

import dsmc_testlib as T
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def bin_sort(lib, grid, x, y, fn="dsmc_bin_sort"):
    n = x.size
    dx, dy = DeviceArray(x.astype(np.float32)), DeviceArray(y.astype(np.float32))
    cell = DeviceArray(shape=(n,), dtype=np.int32)
    perm = DeviceArray(shape=(n,), dtype=np.int32)
    lid = DeviceArray(shape=(n,), dtype=np.int32)
    start = DeviceArray(shape=(grid.N1 * grid.N2 + 1,), dtype=np.int32)
    abi.check(getattr(lib, fn)(C.byref(grid), dx.ptr, dy.ptr, n, cell.ptr, start.ptr, perm.ptr, lid.ptr, None))
    out = dict(cell=cell.get(), start=start.get(), perm=perm.get(), local_id=lid.get())
    for a in (dx, dy, cell, perm, lid, start):
        a.free()
    return out


def oracle_bin_sort(N1, N2, L1, L2, x, y):
    p = cases.make_params(L1=L1, L2=L2, N1=N1, N2=N2, n_inf=1e20, T=300.0, V_jet=0.0, tau=1e-6, ppc=1.0,
                          face_bc=[abi.BC_SPECULAR] * 4, wall_bc=abi.BC_SPECULAR)
    oc, info, _ = O.case_from_params(p, None)
    x, y = x.astype(np.float32), y.astype(np.float32)
    cell, count, slot = oc.bin(x, y, None)
    start = oc.scan(count)
    perm = oc.scatter(cell, slot, start, x.size)
    return dict(cell=cell, start=start, perm=perm, local_id=slot)


def _check(g, o, what):
    for k in ("cell", "start", "perm", "local_id"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=f"{what}: {k}")


def _grid(N1, N2, L1, L2):
    return abi.dsmc_grid(N1, N2, L1, L2)


def test_bin_sort_cylinder_400x200_fill(gpu_lib):
    cfg = cases.cylinder_mach10()
    p = cfg["params"]
    oc, info, tables = O.case_from_params(p, cfg["body"])
    parts, _, err = T.oracle_fill(oc, info, tables)
    assert err == 0
    rng = np.random.default_rng(1)
    order = rng.permutation(parts["x"].size)          # not cell-sorted: a hard order for the atomics
    x, y = parts["x"][order], parts["y"][order]
    print(f"[bin_sort] cylinder 400x200 fill, {x.size} particles", flush=True)
    g = bin_sort(gpu_lib, _grid(p.N1, p.N2, p.L1, p.L2), x, y)
    _check(g, oracle_bin_sort(p.N1, p.N2, p.L1, p.L2, x, y), "cylinder 400x200")
    assert g["start"][-1] == x.size and x.size > 1_900_000


def test_bin_sort_box_64x64_fill(gpu_lib):
    p = cases.box_equilibrium()["params"]
    oc, info, tables = O.case_from_params(p, None)
    parts, _, _ = T.oracle_fill(oc, info, tables)
    g = bin_sort(gpu_lib, _grid(p.N1, p.N2, p.L1, p.L2), parts["x"], parts["y"])
    _check(g, oracle_bin_sort(p.N1, p.N2, p.L1, p.L2, parts["x"], parts["y"]), "box 64x64")
    # 20 per cell, but for the few fill points that round onto a far edge
    assert g["start"][-1] == 81920 and np.abs(np.diff(g["start"]) - 20).sum() <= 8


def test_bin_sort_edge_cases(gpu_lib):
    """Far faces, outside points, dense cells (100, 3,000 and 9,000
    particles: the wave, LDS and HBM sort paths), empty cells, and n = 0."""
    N1, N2, L1, L2 = 37, 23, 0.37, 0.23
    rng = np.random.default_rng(2)
    xs = [rng.random(5000) * L1, np.full(50, L1), np.full(20, -0.01), np.full(20, L1 + 0.5),
          0.05 + 0.001 * rng.random(100), 0.15 + 0.001 * rng.random(3000), 0.25 + 0.001 * rng.random(9000)]
    ys = [rng.random(5000) * L2, rng.random(50) * L2, rng.random(20) * L2, rng.random(20) * L2,
          0.05 + 0.001 * rng.random(100), 0.15 + 0.001 * rng.random(3000), 0.2 + 0.001 * rng.random(9000)]
    xs.append(rng.random(30) * L1); ys.append(np.full(30, L2))
    x, y = np.concatenate(xs), np.concatenate(ys)
    order = rng.permutation(x.size)
    x, y = x[order].astype(np.float32), y[order].astype(np.float32)
    grid = _grid(N1, N2, L1, L2)
    g = bin_sort(gpu_lib, grid, x, y)
    _check(g, oracle_bin_sort(N1, N2, L1, L2, x, y), "edge cases")
    counts = np.diff(g["start"])
    assert counts.max() >= 9000 and ((counts > 64) & (counts <= 4096)).sum() >= 2 and (counts == 0).any()
    e = bin_sort(gpu_lib, grid, np.zeros(0, np.float32), np.zeros(0, np.float32))
    assert np.all(e["start"] == 0)


def test_bin_sort_rejects_bad_arguments(gpu_lib):
    x = DeviceArray(np.zeros(4, np.float32))
    for grid, n in ((_grid(0, 4, 1.0, 1.0), 4), (_grid(4, 4, -1.0, 1.0), 4), (_grid(4, 4, 1.0, 1.0), -1)):
        rc = gpu_lib.dsmc_bin_sort(C.byref(grid), x.ptr, x.ptr, n, x.ptr, x.ptr, x.ptr, None, None)
        assert rc == abi.DSMC_ERR_ARG
    x.free()
