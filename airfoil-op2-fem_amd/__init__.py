"""MI355X-native nodal DG Euler solver: Python host interface.

Mirrors the reference's operator interface for the DG Euler hot path: the
op_decl_const constants, the mesh sets and maps, and one call per OP2
par_loop of ``src/airfoil.cpp`` (init_grid, set_ic, neighbour_zero,
get_neighbour_q, get_bedge_q, euler_rhs, set_workingQ, update_Q, calc_dt),
each running the library's HIP kernels on device tensors laid out as the
reference's op_dats.  The SSPRK3 driver itself is native (``Solver``).

The package loads ``libdgeuler.so`` (built in this directory by ``make``)
with ctypes and fails loudly when it is missing: there is no CPU fallback.

The directory name is not an identifier; import it with
``importlib.import_module("airfoil-op2-fem_amd")``.
"""
from ._lib import DGError, lib, lib_path  # noqa: F401
from .dg import (  # noqa: F401
    BC_EXACT, BC_FARFIELD, BC_INFLOW, BC_OUTFLOW, BC_WALL, GEOM_CELL, GEOM_NODE, MODE_FUSED, MODE_REFERENCE,
    BatchSolver, Context, Mesh, Sensitivity, Solver, generate_operators, naca0012_ogrid, periodic_grid, rect_grid,
    mesh_check, read_vtk, sweep_cases)
