"""ogs5_amd -- MI355X-native implicit FEM step for OGS-5 LIQUID_FLOW and
HEAT_TRANSPORT: element assembly + Jacobi-preconditioned Krylov solve on
hand-written gfx950 HIP kernels, behind the C ABI in include/ogsx.h.

The Python layer is a thin host binding (ctypes) over libogsx.so; all
arithmetic of the path runs in the library's HIP kernels.
"""
from ._lib import (BICGSTAB, CG, DIRICHLET_ROW, DIRICHLET_SYMMETRIC, FP32_IR, FP64,
                   GROUNDWATER_FLOW, HEAT_TRANSPORT, HEX8, LINE2, LIQUID_FLOW, PRECOND_JACOBI, PRECOND_NONE,
                   IR_RELIABLE, IR_RELIABLE_SR, IR_RESTART, QUAD4, TET4, VEC_B, VEC_IC, VEC_UOLD, VEC_X, Group, OgsxError, Pattern, System,
                   NOALLOC_MAX_ROWS, device_count, get_persistent, lib, persistent_giveup_probe,
                   set_persistent, solve_crs_noalloc, solve_status, solver_config)

__all__ = ["Pattern", "System", "Group", "solve_status", "solver_config", "OgsxError", "device_count", "lib", "CG", "BICGSTAB",
           "FP64", "FP32_IR", "LIQUID_FLOW", "GROUNDWATER_FLOW", "HEAT_TRANSPORT", "LINE2", "QUAD4", "HEX8",
           "TET4", "DIRICHLET_ROW", "DIRICHLET_SYMMETRIC", "PRECOND_JACOBI", "PRECOND_NONE",
           "VEC_X", "VEC_B", "VEC_UOLD", "VEC_IC", "solve_crs_noalloc", "NOALLOC_MAX_ROWS",
           "IR_RESTART", "IR_RELIABLE", "IR_RELIABLE_SR", "set_persistent", "get_persistent",
           "persistent_giveup_probe"]
