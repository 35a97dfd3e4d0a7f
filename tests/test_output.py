"""CPU: field output -- the reference's dataset names, shape and index
order, the mean free path closed form, and a .vts round trip."""
import math

import numpy as np

from cudadsmc_amd import cases, output


def _fields(p):
    rng = np.random.default_rng(0)
    shape = (p.N1, p.N2)
    n = p.n_inf * rng.uniform(0.5, 2.0, shape)
    T = p.T * rng.uniform(0.8, 3.0, shape)
    return dict(n=n, ux=rng.normal(0, 100, shape), uy=rng.normal(0, 100, shape), uz=rng.normal(0, 10, shape),
                T=T, p=n * cases.BOLTZ * T)


def test_datasets_names_and_closed_forms():
    p = cases.box_equilibrium(n_cells=8)["params"]
    f = _fields(p)
    ds = output.datasets(f, p)
    assert tuple(ds) == output.NAMES
    assert all(ds[k].shape == (p.N1, p.N2) for k in ds)
    i, j = 3, 5
    assert ds["Rho"][i, j] == f["n"][i, j] * p.mass
    # Cell.cpp:125 with density rho: m / (sqrt(2) rho pi d^2) (T/T_inf)^(omega - 1/2)
    mfp = p.mass / (math.sqrt(2) * ds["Rho"][i, j] * math.pi * p.diam ** 2) * (f["T"][i, j] / p.T) ** (p.Vtl - 0.5)
    assert math.isclose(ds["MFP"][i, j], mfp, rel_tol=1e-12)
    # at T = T_inf the hard-sphere mean free path 1/(sqrt(2) pi d^2 n)
    f["T"][:] = p.T
    ds = output.datasets(f, p)
    assert math.isclose(ds["MFP"][0, 0], 1 / (math.sqrt(2) * math.pi * p.diam ** 2 * f["n"][0, 0]), rel_tol=1e-12)


def test_flat_cell_order_is_i_times_n2_plus_j():
    p = cases.box_equilibrium(n_cells=4)["params"]
    flat = np.arange(p.N1 * p.N2, dtype=np.float64)
    f = dict(n=np.full(flat.size, 1e20), ux=flat, uy=flat, uz=flat, T=np.full(flat.size, 300.0), p=flat)
    ds = output.datasets(f, p)
    assert ds["U"][2, 3] == 2 * p.N2 + 3


def test_vts_round_trip(tmp_path):
    p = cases.box_equilibrium(n_cells=6)["params"]
    ds = output.write_result(str(tmp_path), 100, _fields(p), p)
    back = output.read_vts(str(tmp_path / "result_100.vts"))
    for k in output.NAMES:
        np.testing.assert_allclose(back[k], ds[k], rtol=1e-8)
    z = np.load(str(tmp_path / "result_100.npz"))
    assert sorted(z.files) == sorted(output.NAMES)
