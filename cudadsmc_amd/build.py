"""Builds the native libraries in-tree.

libcudadsmc.so  -- the HIP kernels, the C ABI and the C++ host layer, for gfx950.
dsmc_run        -- the C++ driver (cudadsmc_amd/cpp: Run, Param, Geom, Output),
                   linked against libcudadsmc.so through the C ABI only.
oracle/liboracle.so is built by oracle/Makefile (test infrastructure).
"""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cudadsmc_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libcudadsmc.so")

SOURCES = [
    os.path.join(CSRC, "dsmc_kernels.hip"),
    os.path.join(CSRC, "dsmc_cell.hip"),
    os.path.join(CSRC, "dsmc_abi.hip"),
    os.path.join(CSRC, "dsmc_batch.hip"),
    os.path.join(CSRC, "mlp.hip"),
    os.path.join(CSRC, "dsmc_sort.hip"),
    os.path.join(CSRC, "host", "case_setup.cpp"),
]

# -ffp-contract=off: every product and sum is rounded on its own, which is
# the arithmetic contract the bit-exact parity tests rely on.
# -packed-fp32-ops off: no v_pk_*_f32 instructions.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build_library(verbose=False):
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + SOURCES
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(res.stderr)
    os.replace(tmp, LIB)
    return LIB


HOST = os.path.join(PKG, "dsmc_run")
HOST_SOURCES = [os.path.join(PKG, "cpp", f)
                for f in ("Param.cpp", "Geom.cpp", "CartesianMesh.cpp", "Output.cpp", "Run.cpp", "main.cpp")]


def build_host():
    """The C++ driver: plain g++, the C ABI header and libcudadsmc.so (found
    next to the executable at run time)."""
    tmp = HOST + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-o", tmp] + HOST_SOURCES + [
        "-L" + PKG, "-lcudadsmc", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, HOST)
    return HOST


def build_oracle():
    res = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")],
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)


if __name__ == "__main__":
    print(build_library())
    print(build_host())
    build_oracle()
