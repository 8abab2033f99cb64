"""Caller-defined problem data on the device (GPU): solves of problems whose
permeability tensor, solution, source and map come from callbacks against
the oracle's sparse direct solve of the same problem."""
import numpy as np
import pytest

from conftest import golden
from test_problem_data import aniso_pde

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bc", [0, 1])
def test_anisotropic_tensor_solve_matches_oracle(vemd, oracle, bc):
    """Reference mesh square_8x8 (256 cells), sinusoidal map, k = 1: the
    reference boundary condition (essential flux, mean-zero multiplier) and
    Dirichlet pressure; FP64 and mixed solves."""
    path = golden("square_8x8.msh")
    s = vemd.Solver(vemd.Mesh.read(path), k=1, eps=0.05, bc=bc, pde=aniso_pde(vemd), device=0)
    op = oracle.Problem(oracle.Mesh.read(path), k=1, eps=0.05, bc=bc, pde=aniso_pde(vemd))
    xo = op.solve_direct()
    x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12))
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo), info
    eo, ep = op.errors(), s.errors(x)
    for key in ("p_comp", "u_comp", "p_phys", "u_phys", "div_phys"):
        assert abs(ep[key] - eo[key]) <= 1e-9 * eo[key], key
    xm, _ = s.solve(vemd.SolveOptions(precision=vemd.MIXED, rtol=1e-10))
    assert np.linalg.norm(xm - xo) <= 1e-8 * np.linalg.norm(xo)


def test_tensor_with_arc_map_on_disk(vemd, oracle):
    """Generated 600-cell disk with exact boundary arcs (built-in per-cell map),
    the anisotropic tensor and its solution from callbacks, Dirichlet
    pressure, k = 1 and k = 0."""
    m = vemd.generate_mesh(600, lloyd_iters=20, seed=31, min_edge_frac=0.05)
    a = m.arrays()
    om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
    for k in (0, 1):
        s = vemd.Solver(m, k=k, map_kind=vemd.MAP_ARCS, bc=vemd.BC_DIRICHLET_P, pde=aniso_pde(vemd), device=0)
        op = oracle.Problem(om, k=k, map_kind=oracle.MAP_ARCS, bc=oracle.BC_DIRICHLET_P, pde=aniso_pde(vemd))
        xo = op.solve_direct()
        x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12))
        assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo), (k, info)
