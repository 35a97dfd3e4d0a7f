"""GPU: the allocation-free binning entry point (dsmc_bin_sort_noalloc)
against dsmc_bin_sort, the entry point it is a variant of, on two calls in
a row -- the Mach-10 cylinder's 400x200 fill (about 2M particles, shuffled)
and then a smaller 37x23 grid with dense and empty cells -- bit-exact, and
against the oracle.  The second call runs on scratch the first one left."""
import ctypes as C

import numpy as np
import pytest

from cudadsmc_amd import abi, cases
from oracle import oracle as O

import dsmc_testlib as T
from hipmem import DeviceArray
from test_gpu_bin_sort import _check, _grid, bin_sort, oracle_bin_sort

pytestmark = pytest.mark.gpu


def test_noalloc_equals_allocating_entry_point(gpu_lib):
    cfg = cases.cylinder_mach10()
    p = cfg["params"]
    oc, info, tables = O.case_from_params(p, cfg["body"])
    parts, _, _ = T.oracle_fill(oc, info, tables)
    order = np.random.default_rng(7).permutation(parts["x"].size)
    big = (_grid(p.N1, p.N2, p.L1, p.L2), parts["x"][order], parts["y"][order])
    rng = np.random.default_rng(8)
    N1, N2, L1, L2 = 37, 23, 0.37, 0.23
    x = np.concatenate([rng.random(3000) * L1, 0.15 + 0.001 * rng.random(5000), np.full(10, L1)])
    y = np.concatenate([rng.random(3000) * L2, 0.15 + 0.001 * rng.random(5000), rng.random(10) * L2])
    small = (_grid(N1, N2, L1, L2), x.astype(np.float32), y.astype(np.float32))
    for k, (grid, x, y) in enumerate((big, small)):
        print(f"[noalloc] call {k}: {grid.N1}x{grid.N2} cells, {x.size} particles", flush=True)
        a = bin_sort(gpu_lib, grid, x, y, "dsmc_bin_sort")
        b = bin_sort(gpu_lib, grid, x, y, "dsmc_bin_sort_noalloc")
        _check(b, a, f"call {k}: noalloc vs allocating")
        _check(b, oracle_bin_sort(grid.N1, grid.N2, grid.L1, grid.L2, x, y), f"call {k}: noalloc vs oracle")


def test_noalloc_rejects_a_grid_above_its_scratch(gpu_lib):
    grid = _grid(1025, 1024, 1.0, 1.0)          # 1,049,600 cells > DSMC_BIN_SORT_MAX_CELLS
    start = DeviceArray(shape=(grid.N1 * grid.N2 + 1,), dtype=np.int32)
    rc = gpu_lib.dsmc_bin_sort_noalloc(C.byref(grid), None, None, 0, None, start.ptr, None, None, None)
    start.free()
    assert rc == abi.DSMC_ERR_ARG
    assert b"at most 1048576 cells" in gpu_lib.dsmc_last_error()
