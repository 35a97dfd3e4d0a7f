"""GPU: the C++ driver (dsmc_run, cudadsmc_amd/cpp) runs a case through the
C ABI exactly as the Python Run does: the fields it writes at each output
step equal the Python run of the case it recorded (case.json), and the
reference-style per-phase loop (--phases) gives the same files as graph
replay."""
import json
import os
import subprocess

import numpy as np
import pytest

from cudadsmc_amd import build, cases, output
from cudadsmc_amd.run import Run

pytestmark = pytest.mark.gpu


def _run(args, out):
    if not os.path.exists(build.HOST):
        build.build_host()
    cmd = [build.HOST, "--nx", "80", "--ny", "40", "--ppc", "10", "--steps", "300", "--sample-from", "100",
           "--output-every", "150", "--out", str(out)] + args
    print(f"[cpp] {' '.join(cmd[1:])}", flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_cpp_driver_equals_python_run(gpu_lib, tmp_path):
    log = _run([], tmp_path / "graph")
    assert "Simulation Finished" in log
    case = json.load(open(tmp_path / "graph" / "case.json"))
    p = cases.make_params(**case["params"])
    body = np.array(case["body"], np.float64)
    r = Run(p, body)
    r.initialize()
    for it in (150, 300):
        r.solver(150)
        ds = output.datasets(r.fields(), p)
        got = _load(tmp_path / "graph" / f"result_{it}.npz")
        for k in ("U", "V", "W", "P", "T", "Rho", "Numchild", "Numgrandchild"):
            np.testing.assert_array_equal(got[k], ds[k], err_msg=f"step {it} {k}")
        np.testing.assert_allclose(got["MFP"], ds["MFP"], rtol=1e-14, err_msg=f"step {it} MFP")
        vts = output.read_vts(str(tmp_path / "graph" / f"result_{it}.vts"))
        np.testing.assert_allclose(vts["T"], ds["T"], rtol=1e-8)
    s = r.status()
    assert s.err == 0 and s.step == 300
    assert np.nanmax(ds["T"]) > 1000.0          # a shock layer formed
    r.close()


def test_cpp_phase_loop_equals_graph_replay(gpu_lib, tmp_path):
    _run([], tmp_path / "graph")
    _run(["--phases"], tmp_path / "phases")
    for it in (150, 300):
        a = (tmp_path / "graph" / f"result_{it}.npz").read_bytes()
        b = (tmp_path / "phases" / f"result_{it}.npz").read_bytes()
        assert a == b, it


def test_cpp_sample_every_equals_python_sample(gpu_lib, tmp_path):
    """dsmc_run --sample-every 100 (GPUCells::Sample after steps 0, 100, 200,
    as the reference's Run::solver does, Run.cpp:294-296) writes the fields
    of the Python Run sampling at the same steps with Run.sample()."""
    out = tmp_path / "every"
    _run(["--sample-every", "100", "--sample-from", "-1", "--output-every", "300"], out)
    case = json.load(open(out / "case.json"))
    p = cases.make_params(**case["params"])
    assert p.sample_start == -1
    r = Run(p, np.array(case["body"], np.float64))
    r.initialize()
    for it in range(300):
        r.solver(1, graph=False)
        if it % 100 == 0:
            r.sample()
    assert r.samples()[1] == 3
    ds = output.datasets(r.fields(), p)
    got = _load(out / "result_300.npz")
    for k in ("U", "V", "W", "P", "T", "Rho"):
        np.testing.assert_array_equal(got[k], ds[k], err_msg=k)
    r.close()


def test_cpp_nitrogen_equals_python_run(gpu_lib, tmp_path):
    """dsmc_run --gas nitrogen (zeta = 2, Z_rot = 5): the fields written,
    the rotational temperature among them, equal the Python Run of the case
    it recorded."""
    out = tmp_path / "n2"
    _run(["--gas", "nitrogen", "--zrot", "5", "--output-every", "300"], out)
    case = json.load(open(out / "case.json"))
    assert case["params"]["zeta"] == 2 and abs(case["params"]["rot_prob"] - 0.2) < 1e-15
    p = cases.make_params(**case["params"])
    r = Run(p, np.array(case["body"], np.float64))
    r.initialize()
    r.solver(300)
    ds = output.datasets(r.fields(), p)
    got = _load(out / "result_300.npz")
    for k in ("U", "V", "W", "P", "T", "Rho", "Trot"):
        np.testing.assert_array_equal(got[k], ds[k], err_msg=k)
    vts = output.read_vts(str(out / "result_300.vts"))
    np.testing.assert_allclose(vts["Trot"], ds["Trot"], rtol=1e-8)
    assert np.nanmax(ds["Trot"]) > 500.0          # rotational energy heated in the shock layer
    r.close()


def test_cpp_subcells_equal_python_run(gpu_lib, tmp_path):
    """--amr 50: the driver renews the sub-cell grids after every 50th step
    once sampled (dsmc_amr_update, Cell::genAMRmesh); its fields and
    Numchild/Numgrandchild equal a Python run doing the same."""
    _run(["--amr", "50"], tmp_path / "amr")
    case = json.load(open(tmp_path / "amr" / "case.json"))
    p = cases.make_params(**case["params"])
    body = np.array(case["body"], np.float64)
    r = Run(p, body)
    r.initialize()
    step = 0
    for it in (150, 300):
        while step < it:
            r.solver(50)
            step += 50
            if r.status().nsamples > 0:
                r.amr_update()
        a = r.amr()
        ds = output.datasets(r.fields(), p, a)
        got = _load(tmp_path / "amr" / f"result_{it}.npz")
        for k in ("U", "V", "W", "P", "T", "Rho", "Numchild", "Numgrandchild"):
            np.testing.assert_array_equal(got[k], ds[k], err_msg=f"step {it} {k}")
    assert (ds["Numchild"] > 1).sum() > 100
    r.close()

