"""CPU: the C++ host layer (cudadsmc_amd/cpp, the dsmc_run driver): its
field writers against cudadsmc_amd/output.py, and its command-line and
no-device error paths.  The run itself is tested on the GPU
(test_gpu_cpp_host.py)."""
import math
import os
import subprocess

import numpy as np
import pytest

from cudadsmc_amd import abi, build, output

EXE = build.HOST


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        build.build_host()
    return EXE


def _device_count():
    n = abi.C.c_int(0)
    lib = abi.load()
    return n.value if lib.dsmc_device_count(abi.C.byref(n)) == 0 else 0


def test_field_writers_match_the_python_writers(exe, tmp_path):
    """dsmc_run --write-fields: the .npz holds the Write2HDF5 dataset names
    with [N1][N2] shapes and dtypes, and its .vts is byte-identical to
    output.write_vts of the same arrays (NaN, -0, inf and negative ints
    included)."""
    r = subprocess.run([exe, "--write-fields", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    with np.load(tmp_path / "fields.npz", allow_pickle=False) as z:
        ds = {k: z[k] for k in z.files}
    assert sorted(ds) == sorted(output.NAMES)
    for k, a in ds.items():
        assert a.shape == (7, 5), k
        assert a.dtype == (np.int32 if k in ("Numchild", "Numgrandchild") else np.float64), k
    c = np.arange(35, dtype=np.float64).reshape(7, 5)
    np.testing.assert_array_equal(ds["V"], -3.25e-3 * c)
    np.testing.assert_allclose(ds["U"], 1000.0 * np.sin(c), rtol=1e-15, atol=1e-12)
    assert np.isnan(ds["T"].ravel()[4]) and np.isinf(ds["MFP"].ravel()[6])
    assert math.copysign(1.0, ds["W"].ravel()[3]) < 0
    np.testing.assert_array_equal(ds["Numchild"].ravel(), np.arange(35) - 17)

    class P:
        N1, N2, L1, L2 = 7, 5, 0.7, 0.25
    output.write_vts(str(tmp_path / "py.vts"), ds, P)
    assert (tmp_path / "fields.vts").read_text() == (tmp_path / "py.vts").read_text()


def test_bad_command_line_is_exit_2(exe):
    for args in (["--nx"], ["--nx", "abc"], ["--case", "sphere"], ["--frobnicate"]):
        r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.stderr)
        assert "dsmc_run:" in r.stderr


def test_no_device_is_a_clean_error(exe, tmp_path):
    if _device_count() > 0:
        pytest.skip("a HIP device is present")
    r = subprocess.run([exe, "--nx", "40", "--ny", "20", "--steps", "5", "--out", str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "no HIP device" in r.stderr


@pytest.mark.parametrize("case", [["--case", "cylinder", "--nx", "120", "--ny", "60"],
                                  ["--case", "wedge", "--half-angle", "28", "--nx", "96", "--ny", "48"]])
def test_host_mesh_tables_equal_the_library_tables(exe, tmp_path, case):
    """The C++ host mesh (CartesianMesh::cutcell, the reference's
    meshes/CartesianMesh.cpp:100-169 role) builds the same fluid fractions,
    cut flags and wall segments as the library's own case set-up, bit for
    bit -- what dsmc_create_from_tables then receives (no device needed);
    getIndex follows the j + i*N2 convention."""
    import json
    from cudadsmc_amd import cases
    from cudadsmc_amd.run import describe, outline_segments
    f = tmp_path / "mesh.bin"
    r = subprocess.run([exe, "--write-mesh", str(f)] + case + ["--out", str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    raw = f.read_bytes()
    N1, N2, nseg = np.frombuffer(raw[:12], np.int32)
    C = int(N1) * int(N2)
    frac = np.frombuffer(raw[12:12 + 8 * C], np.float64)
    cut = np.frombuffer(raw[12 + 8 * C:12 + 9 * C], np.uint8)
    seg = np.frombuffer(raw[12 + 9 * C:], np.float64)
    geom = case[1]
    half = float(case[3]) if geom == "wedge" else 20.0
    c = cases.flow_case(geom, nx=int(N1), ny=int(N2), half_angle_deg=half)
    info, t = describe(c["params"], c["body"])
    np.testing.assert_array_equal(frac, t["frac"])
    np.testing.assert_array_equal(cut, t["cut"])
    assert nseg == c["body"].shape[0]
    np.testing.assert_array_equal(seg, outline_segments(c["body"]))
    np.testing.assert_array_equal(seg.astype(np.float32), t["seg"])
