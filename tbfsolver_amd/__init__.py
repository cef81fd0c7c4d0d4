"""tbfsolver_amd -- MI355X pressure-projection engine for TBFsolver.

Drop-in for the reference's FAST_MODE (fastPoissonSolver) and MG_MODE
(pcg + multiGrid) pressure solvers, with every stage a HIP kernel for gfx950
in libtbfgpu.so (C ABI: include/tbfgpu.h).
"""
from ._lib import GridDesc, TbfgError, lib, LIB_PATH  # noqa: F401
from .solvers import FastPoissonSolver, PCG, channel_grid, default_params, read_parameter_file  # noqa: F401
from . import fields  # noqa: F401
from .poisson import PoissonEqn  # noqa: F401

__all__ = ["FastPoissonSolver", "PCG", "PoissonEqn", "channel_grid", "default_params", "read_parameter_file", "fields",
           "GridDesc", "TbfgError", "lib"]
