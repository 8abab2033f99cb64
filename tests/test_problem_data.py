"""Caller-defined problem data (CPU): the callbacks of vemd_problem_funcs --
the reference's DarcyPdeData fields kappa, solution_p, solution_u, source_g,
dirichlet_p, neumann_un and its IsoparametricMapping (CPU/examples/
darcy_problem.h:39-63,162-180) -- against the built-in data, the oracle and
the mathematics:

* callbacks that restate a built-in problem (the reference's example 1: the
  sinusoidal map, K = I, p = sin(pi x) cos(pi y)) give the built-in element
  operators, right-hand side and errors;
* an anisotropic, spatially varying tensor K(x, y) with its manufactured
  solution: element operators, right-hand side and errors equal the oracle's,
  which builds the reference's computational tensor det J J^-1 K J^-T and
  inverts it (the product forms J^T K^-1 J / det J directly); the oracle's
  direct solution converges at the orders of the method (k = 1: 2, k = 0: 1)
  in both error functionals, and the source callback is the divergence of
  the flux callback (central differences);
* K = c I equals the built-in theta = log c (every element operator / c);
* neumann_un with the physical normal equals the pulled-back solution_u flux,
  and dirichlet_p equals the built-in Dirichlet data;
* an exception in a callback reaches the caller (Python and C++).
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden

PI = math.pi


def sincos_pde(vemd, eps=0.05, with_map=True):
    """The reference's example 1 as callbacks (CPU/examples/darcy_problem.cpp:106-149)."""
    def p(x, y):
        return math.sin(PI * x) * math.cos(PI * y)

    def u(x, y):
        return (-(PI * math.cos(PI * x) * math.cos(PI * y)), -(-PI * math.sin(PI * x) * math.sin(PI * y)))

    def g(x, y):
        return 2 * PI * PI * (math.sin(PI * x) * math.cos(PI * y))

    def fmap(xi, eta):
        return ((xi + eps * math.sin(2 * PI * eta), eta + eps * math.sin(2 * PI * xi)),
                ((1.0, eps * 2 * PI * math.cos(2 * PI * eta)), (eps * 2 * PI * math.cos(2 * PI * xi), 1.0)))

    return vemd.PdeData(kappa=lambda x, y: ((1.0, 0.0), (0.0, 1.0)), solution_p=p, solution_u=u, source_g=g,
                        map=fmap if with_map else None)


class Aniso:
    """K(x, y) = [[2 + x, a], [a, 1 + y^2]], a = 0.3 + 0.2 y (symmetric
    positive definite on the mapped unit square), p = sin(pi x) cos(pi y)
    (mean zero on the unit square, as the reference's multiplier condition
    needs), u = -K grad p, g = div u."""

    @staticmethod
    def kappa(x, y):
        a = 0.3 + 0.2 * y
        return ((2.0 + x, a), (a, 1.0 + y * y))

    @staticmethod
    def solution_p(x, y):
        return math.sin(PI * x) * math.cos(PI * y)

    @staticmethod
    def grad(x, y):
        return PI * math.cos(PI * x) * math.cos(PI * y), -PI * math.sin(PI * x) * math.sin(PI * y)

    @classmethod
    def solution_u(cls, x, y):
        k = cls.kappa(x, y)
        px, py = cls.grad(x, y)
        return -(k[0][0] * px + k[0][1] * py), -(k[1][0] * px + k[1][1] * py)

    @classmethod
    def source_g(cls, x, y):
        px, py = cls.grad(x, y)
        s, c = math.sin(PI * x) * math.cos(PI * y), math.cos(PI * x) * math.sin(PI * y)
        pxx, pyy, pxy = -PI * PI * s, -PI * PI * s, -PI * PI * c
        a = 0.3 + 0.2 * y
        return -((px + (2.0 + x) * pxx + a * pxy) + (0.2 * px + a * pxy + 2.0 * y * py + (1.0 + y * y) * pyy))


def aniso_pde(vemd):
    return vemd.PdeData(kappa=Aniso.kappa, solution_p=Aniso.solution_p, solution_u=Aniso.solution_u,
                        source_g=Aniso.source_g)


def _elements_equal(s1, s2, ncell, tol, scale2=None):
    for c in range(ncell):
        e1, e2 = s1.element(c), s2.element(c)
        f = 1.0 if scale2 is None else scale2[c]
        A2 = e2["A"] * f
        assert np.abs(e1["A"] - A2).max() <= tol * np.abs(A2).max(), c
        assert np.abs(e1["W"] - e2["W"]).max() <= tol * max(1.0, np.abs(e2["W"]).max()), c
        assert np.allclose(e1["L"], e2["L"], rtol=tol, atol=1e-300), c
        assert np.allclose(e1["f"], e2["f"], rtol=tol, atol=tol * np.abs(e2["f"]).max()), c


@pytest.mark.parametrize("k,bc", [(0, 0), (1, 0), (1, 1)])
def test_callbacks_reproduce_builtin_problem(vemd, oracle, k, bc):
    path = golden("square_2x2.msh")
    m = vemd.Mesh.read(path)
    s_cb = vemd.Solver(m, k=k, bc=bc, pde=sincos_pde(vemd))
    s_bi = vemd.Solver(m, k=k, bc=bc, eps=0.05)
    _elements_equal(s_cb, s_bi, 16, 1e-14)
    b1, b2 = s_cb.rhs(), s_bi.rhs()
    assert np.abs(b1 - b2).max() <= 1e-14 * np.abs(b2).max()
    op = oracle.Problem(oracle.Mesh.read(path), k=k, eps=0.05, bc=bc)
    op_cb = oracle.Problem(oracle.Mesh.read(path), k=k, eps=0.05, bc=bc, pde=sincos_pde(vemd))
    x = op.solve_direct()
    x_cb = op_cb.solve_direct()
    assert np.abs(x - x_cb).max() <= 1e-12 * np.abs(x).max()
    e1, e2 = s_cb.errors(x), s_bi.errors(x)
    for key in e1:
        assert abs(e1[key] - e2[key]) <= 1e-13 * abs(e2[key]) + 1e-15, key


def test_manufactured_source_is_the_flux_divergence():
    rng = np.random.default_rng(5)
    h = 1e-5
    for x, y in rng.uniform(-0.05, 1.05, size=(20, 2)):
        div = ((Aniso.solution_u(x + h, y)[0] - Aniso.solution_u(x - h, y)[0]) +
               (Aniso.solution_u(x, y + h)[1] - Aniso.solution_u(x, y - h)[1])) / (2 * h)
        assert abs(div - Aniso.source_g(x, y)) <= 1e-6 * (1 + abs(div))
        k = np.array(Aniso.kappa(x, y))
        assert np.all(np.linalg.eigvalsh(k) > 0.5)


@pytest.mark.parametrize("k,bc", [(0, 0), (1, 0), (1, 1)])
def test_anisotropic_tensor_matches_oracle(vemd, oracle, k, bc):
    path = golden("square_2x2.msh")
    s = vemd.Solver(vemd.Mesh.read(path), k=k, bc=bc, eps=0.05, pde=aniso_pde(vemd))
    op = oracle.Problem(oracle.Mesh.read(path), k=k, eps=0.05, bc=bc, pde=aniso_pde(vemd))
    for c in range(16):
        e = s.element(c)
        n = e["n"]
        d = op.element(c, 13)[:n]
        Ao = op.element(c, 7, (n, n)) * np.outer(d, d)
        assert np.abs(e["A"] - Ao).max() <= 1e-13 * np.abs(Ao).max(), c
        assert np.allclose(e["f"], op.element(c, 9), rtol=1e-12, atol=1e-14)
    A, rr = op.csr(True)
    fi, ub = op.free()
    b = s.rhs()
    assert np.abs(b[fi] - rr).max() <= 1e-13 * np.abs(rr).max()
    x = op.solve_direct()
    eo, ep = op.errors(), s.errors(x)
    for key in eo:
        assert abs(ep[key] - eo[key]) <= 1e-12 * abs(eo[key]) + (1e-15 if key == "mass" else 0.0), key


@pytest.mark.parametrize("k", [0, 1])
def test_anisotropic_tensor_converges(vemd, oracle, k):
    """Errors of the oracle's direct solution (reference boundary condition,
    sinusoidal map) on the reference's square meshes 16, 64, 256 cells."""
    errs = []
    for n in (2, 4, 8):
        path = golden("square_%dx%d.msh" % (n, n))
        s = vemd.Solver(vemd.Mesh.read(path), k=k, eps=0.05, pde=aniso_pde(vemd))
        op = oracle.Problem(oracle.Mesh.read(path), k=k, eps=0.05, pde=aniso_pde(vemd))
        e = s.errors(op.solve_direct())
        errs.append((e["p_comp"], e["u_comp"]))
    errs = np.array(errs)
    orders = np.log2(errs[:-1] / errs[1:])
    want = 1.8 if k == 1 else 0.9
    assert np.all(orders[-1] >= want), orders
    assert np.all(orders[-1] <= want + 0.6), orders


def test_constant_tensor_equals_theta(vemd):
    c = 3.7
    path = golden("square_2x2.msh")
    m = vemd.Mesh.read(path)
    s_cb = vemd.Solver(m, k=1, eps=0.05, pde=vemd.PdeData(kappa=lambda x, y: ((c, 0.0), (0.0, c))))
    s_th = vemd.Solver(m, k=1, eps=0.05, theta=np.full(16, math.log(c)))
    _elements_equal(s_cb, s_th, 16, 1e-14, scale2=np.full(16, 1.0 / c))


def test_neumann_un_equals_pulled_back_flux(vemd):
    path = golden("square_2x2.msh")
    m = vemd.Mesh.read(path)
    pde1 = aniso_pde(vemd)
    pde2 = aniso_pde(vemd)
    pde2.neumann_un = lambda x, y, nx, ny: Aniso.solution_u(x, y)[0] * nx + Aniso.solution_u(x, y)[1] * ny
    b1 = vemd.Solver(m, k=1, eps=0.05, pde=pde1).rhs()
    b2 = vemd.Solver(m, k=1, eps=0.05, pde=pde2).rhs()
    assert np.abs(b1 - b2).max() <= 1e-13 * np.abs(b1).max()


def test_dirichlet_p_callback_equals_builtin(vemd):
    path = golden("square_2x2.msh")
    m = vemd.Mesh.read(path)
    p = lambda x, y: math.sin(PI * x) * math.cos(PI * y)
    b1 = vemd.Solver(m, k=1, eps=0.05, bc=1, pde=vemd.PdeData(dirichlet_p=p)).rhs()
    b2 = vemd.Solver(m, k=1, eps=0.05, bc=1).rhs()
    assert np.abs(b1 - b2).max() <= 1e-14 * np.abs(b2).max()


def test_callback_exception_reaches_caller(vemd):
    def bad(x, y):
        raise ValueError("no permeability here")
    with pytest.raises(ValueError, match="no permeability"):
        vemd.Solver(vemd.Mesh.read(golden("square_1x1.msh")), k=1, pde=vemd.PdeData(kappa=bad))


def test_non_symmetric_tensor_rejected(vemd):
    with pytest.raises(vemd.VemdError, match="symmetric positive definite"):
        vemd.Solver(vemd.Mesh.read(golden("square_1x1.msh")), k=1,
                    pde=vemd.PdeData(kappa=lambda x, y: ((1.0, 0.5), (0.0, 1.0))))


def test_device_setup_rejects_callbacks(vemd):
    with pytest.raises(vemd.VemdError):
        vemd.Solver(vemd.Mesh.read(golden("square_1x1.msh")), k=1, pde=aniso_pde(vemd), setup_device=0)


def test_cpp_drop_in_with_pde_data(vemd, tmp_path):
    """The C++ wrapper (include/vemd.hpp): a DarcyProblem with DarcyPdeData
    std::function fields, as a caller of the reference's DarcyProblem sets
    them, builds the same element operators as the Python binding; an
    exception thrown in a field surfaces from the DarcySolver constructor."""
    exe = tmp_path / "pde_data"
    src = os.path.join(ROOT, "tests", "cpp", "pde_data.cpp")
    libdir = os.path.join(ROOT, "wsl-code_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", str(exe),
                    "-L", libdir, "-lvemd", "-Wl,-rpath," + libdir], check=True)
    out = subprocess.run([str(exe), golden("square_2x2.msh")], check=True, capture_output=True, text=True).stdout
    lines = out.strip().splitlines()
    assert lines[-1] == "caught: kappa out of range"
    vals = np.array([float(v) for v in lines[0].split()])
    s = vemd.Solver(vemd.Mesh.read(golden("square_2x2.msh")), k=1, eps=0.05, pde=aniso_pde(vemd))
    A = s.element(3)["A"]
    assert np.abs(vals - A.ravel()).max() <= 1e-12 * np.abs(A).max()
