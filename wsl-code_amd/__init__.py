"""vemd: MI355X-native matrix-free mixed VEM Darcy solver (Python host binding).

A thin ctypes layer over libvemd's C ABI (include/vemd.h).  The classes mirror
the reference project's C++ interface for this path so that code written
against it reads the same:

    reader = StraightMeshReader(); reader.read_mesh("square_2x2.msh")
    problem = DarcyProblem(); problem.set_problem_index(1); problem.init_problem()
    solver = DarcySolver(reader, problem, 9, 9, 1)      # CPU/solver/DarcySolver.h:21-25
    solver.solve()                                       # :38
    x = solver.getSolution()                             # :107
    solver.computeL2ErrorPressure(9), solver.computeL2ErrorFlux(9)   # :122-123

The solve runs only on a HIP device; without one it raises VemdError
(VEMD_ERR_NO_DEVICE) -- there is no CPU fallback.
"""
from ._lib import (VemdError, lib, device_count, BC_DIRICHLET_P, BC_FLUX_MEANZERO, MAP_ARCS, MAP_SIN,
                   SOL_SINCOS, SOL_SINSIN, FP32, FP64, MIXED, DOMAIN_ANNULUS, DOMAIN_DISK, PC_AUTO, PC_JACOBI,
                   PC_MULTIGRID)
from .mesh import Mesh, StraightMeshReader, generate_mesh
from .solver import DarcyProblem, DarcySolver, PdeData, Solver, SolveOptions

__all__ = [
    "VemdError", "lib", "device_count", "Mesh", "StraightMeshReader", "generate_mesh", "DarcyProblem",
    "DarcySolver", "PdeData", "Solver", "SolveOptions", "BC_DIRICHLET_P", "BC_FLUX_MEANZERO", "MAP_ARCS", "MAP_SIN",
    "SOL_SINCOS", "SOL_SINSIN", "FP32", "FP64", "MIXED", "DOMAIN_DISK", "DOMAIN_ANNULUS",
]
