"""CPU: libcudadsmc.so loads without a GPU, exports every function that
include/dsmc_abi.h declares, and reports argument errors through status
codes and dsmc_last_error (no device calls)."""
import ctypes as C
import os
import re

import pytest

from cudadsmc_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            src = open(os.path.join(ROOT, "include", f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dsmc_\w+)\s*\(", src, flags=re.M))
    return names


def test_every_declared_symbol_is_exported():
    lib = abi.load()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    from cudadsmc_amd import mlp
    names = _declared() - {"dsmc_last_error", "dsmc_mlp_last_error"}
    covered = set(abi.SIGNATURES) | set(mlp._SIG)
    assert names <= covered, sorted(names - covered)


def test_null_arguments_are_errors():
    lib = abi.load()
    assert lib.dsmc_run(None, 1, 0) == abi.DSMC_ERR_ARG
    assert b"handle" in lib.dsmc_last_error() or b"argument" in lib.dsmc_last_error()
    assert lib.dsmc_destroy(None) == abi.DSMC_OK        # like free(NULL)
    assert lib.dsmc_case_describe(None, None, None, None, None, None, None, None, None, None, None,
                                  None) != abi.DSMC_OK
    with pytest.raises(abi.DSMCError):
        abi.check(lib.dsmc_sync(None))


def test_batch_null_arguments_are_errors():
    """The sweep-monitoring and retirement entry points reject a null batch
    or null outputs before touching any device state."""
    lib = abi.load()
    keep = (abi.C.c_int * 1)(0)
    out = (abi.C.c_double * 4)()
    assert lib.dsmc_batch_download_counts(None, out, None) == abi.DSMC_ERR_ARG
    assert lib.dsmc_batch_wall_distance(None, out) == abi.DSMC_ERR_ARG
    assert lib.dsmc_batch_compact(None, keep, 1) == abi.DSMC_ERR_ARG
    assert lib.dsmc_batch_destroy(None) == abi.DSMC_OK


@pytest.mark.parametrize("n", ["2", "8"])
def test_bench_refuses_more_than_one_gpu(n):
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    so, before touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", n], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
