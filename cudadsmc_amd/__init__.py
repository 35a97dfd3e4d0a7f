"""cudadsmc_amd -- MI355X-native DSMC hot path (move, walls, binning, NTC/VHS
collisions, sampling, Philox RNG) as hand-written HIP kernels for gfx950
behind a C ABI (include/dsmc_abi.h, libcudadsmc.so)."""
from . import abi, cases  # noqa: F401
from .run import Run  # noqa: F401

__all__ = ["abi", "cases", "Run"]
