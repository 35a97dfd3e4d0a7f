"""CPU: the oracle's phase functions (used by the GPU parity tests) compose
to exactly its full step, and the OpenMP step equals the serial one."""
import numpy as np
import pytest

from oracle import oracle as O

import dsmc_testlib as T


@pytest.mark.parametrize("case", ["box", "cyl"])
def test_phases_compose_to_step(case):
    p, body = T.small_box(16, 8, seed=7) if case == "box" else T.small_cylinder(80, 40, 10, 64, seed=11)
    oc, info, tables = O.case_from_params(p, body)
    parts, _, err = T.oracle_fill(oc, info, tables)
    assert err == 0
    st = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    cur = parts
    sm = np.full(info.ncell, np.float32(info.sigcmax0), np.float32)
    mr = np.zeros(info.ncell, np.float32)
    next_pid = parts["pid"].size
    for step in range(4):
        ref = T.oracle_phases(oc, cur, step, next_pid, sm, mr, int(info.capacity))
        st.step(sample=True)
        T.assert_same_particles(st.particles(), ref["new"], what=f"step {step} ")
        np.testing.assert_array_equal(st.sigcmax, ref["sigcmax"])
        np.testing.assert_array_equal(st.mrem, ref["mrem"])
        cur, sm, mr = ref["new"], ref["sigcmax"], ref["mrem"]
        next_pid += ref["Jt"]
    assert st.st.err == 0


def test_openmp_step_is_bit_identical():
    p, body = T.small_cylinder(80, 40, 10, 64, seed=17)
    oc, info, tables = O.case_from_params(p, body)
    parts, _, _ = T.oracle_fill(oc, info, tables)
    a = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    b = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0)
    for _ in range(10):
        a.step(sample=True, nthreads=1)
        b.step(sample=True, nthreads=4)
    T.assert_same_particles(a.particles(), b.particles())
    np.testing.assert_array_equal(a.acc, b.acc)
    assert a.st.total_coll == b.st.total_coll


def test_openmp_step_under_a_thread_cap():
    """A team smaller than the requested thread count (OMP_THREAD_LIMIT, as
    on a host that caps threads) still bins every particle: 8 threads asked
    for, 2 allowed, result equal to the serial step."""
    import os
    import subprocess
    import sys
    code = (
        "import sys; sys.path[:0] = ['tests', '.'];"
        "import numpy as np, dsmc_testlib as T; from oracle import oracle as O;"
        "p, body = T.small_cylinder(80, 40, 10, 64, seed=17);"
        "oc, info, tables = O.case_from_params(p, body);"
        "parts, _, _ = T.oracle_fill(oc, info, tables);"
        "a = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0);"
        "b = O.OracleState(oc, parts, int(info.capacity), info.sigcmax0);"
        "[(a.step(sample=True, nthreads=1), b.step(sample=True, nthreads=8)) for _ in range(5)];"
        "T.assert_same_particles(a.particles(), b.particles());"
        "assert a.st.total_coll == b.st.total_coll > 0; print('ok')"
    )
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                       env=dict(os.environ, OMP_THREAD_LIMIT="2"), timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
