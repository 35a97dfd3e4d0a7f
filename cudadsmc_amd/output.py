"""Field output: the datasets of the reference's Output::Write2HDF5 and
Output::Write2VTK (Output.cpp:12-79, 82-154), written as a VTK XML
structured grid (.vts, cell data) and as an .npz with the HDF5 dataset
names (no HDF5 or VTK library needed).

Datasets, shape [N1][N2], index i*N2 + j (the cell order of the solver):
  U, V, W  mean velocity (m/s)
  P        pressure n k T (Pa)
  T        temperature (K)
  Rho      mass density n m (kg/m^3)
  Numchild, Numgrandchild  adaptive sub-cell counts: level-2 cells and level-3
           cells in all of them (int; 0: no sub-cells)
  MFP      mean free path m / (sqrt(2) rho pi d^2) (T/T_inf)^(omega - 1/2)
           (Cell.cpp:125)
  Trot     rotational temperature (K), a diatomic gas (zeta = 2) only; T is
           then the translational temperature
"""
import os

import numpy as np

NAMES = ("U", "V", "W", "P", "T", "Rho", "Numchild", "Numgrandchild", "MFP")


def datasets(fields, params, counts=None):
    """fields: dict n, ux, uy, uz, T, p (arrays [N1][N2] or flat) from
    Run.fields() / Batch.fields(i); params: the case's dsmc_params; counts:
    Numchild and Numgrandchild of the adaptive sub-cells (Run.amr()), else
    zeros."""
    shape = (params.N1, params.N2)
    g = {k: np.asarray(v, np.float64).reshape(shape) for k, v in fields.items()}
    rho = g["n"] * params.mass
    with np.errstate(divide="ignore", invalid="ignore"):
        mfp = np.where(rho > 0, params.mass / (np.sqrt(2.0) * rho * np.pi * params.diam ** 2)
                       * np.power(np.where(g["T"] > 0, g["T"], 1.0) / params.T, params.Vtl - 0.5), 0.0)
    nc = np.zeros(shape, np.int32) if counts is None else np.asarray(counts["Numchild"], np.int32).reshape(shape)
    ng = np.zeros(shape, np.int32) if counts is None else np.asarray(counts["Numgrandchild"], np.int32).reshape(shape)
    ds = dict(U=g["ux"], V=g["uy"], W=g["uz"], P=g["p"], T=g["T"], Rho=rho,
              Numchild=nc, Numgrandchild=ng, MFP=mfp)
    if "Trot" in g:
        ds["Trot"] = g["Trot"]
    return ds


def write_npz(path, ds):
    np.savez(path, **ds)


def write_vts(path, ds, params):
    """VTK XML StructuredGrid: (N1+1) x (N2+1) points on the cell corners,
    one cell value per dataset (VTK orders cells with x fastest)."""
    n1, n2 = params.N1, params.N2
    dx, dy = params.L1 / n1, params.L2 / n2
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="StructuredGrid" version="0.1" byte_order="LittleEndian">\n')
        f.write(f'<StructuredGrid WholeExtent="0 {n1} 0 {n2} 0 0">\n<Piece Extent="0 {n1} 0 {n2} 0 0">\n')
        f.write('<Points>\n<DataArray type="Float64" NumberOfComponents="3" format="ascii">\n')
        for j in range(n2 + 1):
            f.write(" ".join(f"{i * dx:.9g} {j * dy:.9g} 0" for i in range(n1 + 1)))
            f.write("\n")
        f.write("</DataArray>\n</Points>\n<CellData>\n")
        for name in NAMES + (("Trot",) if "Trot" in ds else ()):
            a = ds[name]
            typ = "Int32" if a.dtype.kind == "i" else "Float64"
            f.write(f'<DataArray type="{typ}" Name="{name}" format="ascii">\n')
            vals = a.T.ravel()                     # x fastest
            fmt = (lambda v: str(int(v))) if typ == "Int32" else (lambda v: f"{v:.9g}")
            for s in range(0, vals.size, 512):
                f.write(" ".join(fmt(v) for v in vals[s:s + 512]))
                f.write("\n")
            f.write("</DataArray>\n")
        f.write("</CellData>\n</Piece>\n</StructuredGrid>\n</VTKFile>\n")


def write_result(directory, iteration, fields, params, vtk=True, counts=None):
    """res/result_<iter>.npz (+ .vts), the reference's file naming."""
    os.makedirs(directory, exist_ok=True)
    ds = datasets(fields, params, counts)
    base = os.path.join(directory, f"result_{iteration}")
    write_npz(base + ".npz", ds)
    if vtk:
        write_vts(base + ".vts", ds, params)
    return ds


def read_vts(path):
    """Cell data of a .vts written by write_vts, as [N1][N2] arrays."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("StructuredGrid").find("Piece")
    e = [int(v) for v in piece.get("Extent").split()]
    n1, n2 = e[1], e[3]
    out = {}
    for da in piece.find("CellData"):
        vals = np.array(da.text.split(), np.float64 if da.get("type") == "Float64" else np.int64)
        out[da.get("Name")] = vals.reshape(n2, n1).T
    return out
