"""MI355X-native RHS + Runge-Kutta hot path of an octree-AMR finite-volume
compressible (multiphase) Euler / Navier-Stokes solver.

The native library (host forest, flattener and gfx950 HIP kernels behind a C
ABI) must be built first; importing the package does not build anything and
fails loudly when the library is missing.
"""
from ._capi import (FIELD_FACEFLUX, FIELD_RHS, FIELD_S, FIELD_SF, FIELD_SG, FIELD_SN, FIELD_STAGE_IN, FVG_MINMOD,
                    FVG_MIXED, FVG_PLAIN, FVG_VANALBADA, FvError, lib, lib_path)
from .core import BC_INLET, BC_OUTFLOW, FlatMesh, Forest, Solver, neq_of, params

__all__ = [
    "FIELD_FACEFLUX", "FIELD_RHS", "FIELD_S", "FIELD_SF", "FIELD_SG", "FIELD_SN", "FIELD_STAGE_IN", "FVG_MINMOD",
    "FVG_MIXED", "FVG_PLAIN", "FVG_VANALBADA", "FvError", "lib", "lib_path", "BC_INLET",
    "BC_OUTFLOW", "FlatMesh", "Forest", "Solver", "neq_of", "params",
]
