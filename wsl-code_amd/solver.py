"""Solver objects: the low-level Solver over the C ABI and the reference-style
DarcyProblem / DarcySolver pair."""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import (BC_DIRICHLET_P, BC_FLUX_MEANZERO, FP32, FP64, KAPPA_FN, MAP_ARCS, MAP_FN, MAP_SIN, MIXED,
                   NORMAL_FN, SCALAR_FN, SOL_SINCOS, SOL_SINSIN, VECTOR_FN, ProblemDesc, ProblemFuncs, SolveInfo,
                   SolveOptsC, VemdError, c_dp, c_ip, c_lp, c_vp, check, lib)


def _d(a):
    return a.ctypes.data_as(c_dp)


def _vec(a, n, what):
    """Contiguous float64 copy of a 1-D array of exactly n values: the C ABI
    takes bare pointers and reads n values."""
    a = np.ascontiguousarray(a, np.float64)
    if a.shape != (n,):
        raise VemdError(1, "%s: expected %d values, got shape %s" % (what, n, a.shape))
    return a


@dataclass
class SolveOptions:
    precision: int = FP64
    rtol: float = 1e-8
    max_iters: int = 1000000
    inner_rtol: float = 1e-4
    check_every: int = 32
    use_graph: int = 1
    precond: int = 0  # PC_AUTO: multigrid where built (Dirichlet pressure), else Jacobi
    stream: int = 0   # caller's HIP stream handle for the device-resident solves (0: the null stream)

    def c(self):
        return SolveOptsC(self.precision, self.rtol, self.max_iters, self.inner_rtol, self.check_every,
                          self.use_graph, self.precond, self.stream or None)


@dataclass
class PdeData:
    """Problem data as Python callables, the fields of the reference's
    DarcyPdeData plus its IsoparametricMapping (CPU/examples/darcy_problem.h:
    39-63,162-180).  Any field left None keeps the built-in choice.

    kappa(x, y) -> 2x2 permeability tensor (symmetric positive definite)
    solution_p(x, y) -> exact pressure;  solution_u(x, y) -> (ux, uy) = -K grad p
    source_g(x, y) -> div u;  dirichlet_p(x, y) -> boundary pressure
    neumann_un(x, y, nx, ny) -> u . n on the boundary
    map(xi, eta) -> ((x, y), ((dx/dxi, dx/deta), (dy/dxi, dy/deta)))
    """
    kappa: object = None
    solution_p: object = None
    solution_u: object = None
    source_g: object = None
    dirichlet_p: object = None
    neumann_un: object = None
    map: object = None

    def c_funcs(self):
        """The vemd_problem_funcs struct; keep the returned object alive while
        the solver exists (it holds the ctypes callbacks).  An exception in a
        callable is kept (the C side then sees NaN) and re-raised by
        raise_pending() after the library call."""
        self._error = None

        def guard(f, nan):
            def g(*a):
                try:
                    return f(*a)
                except BaseException as e:  # reported after the C call returns
                    if self._error is None:
                        self._error = e
                    return nan
            return g

        def kappa(_, x, y, out):
            k = np.asarray(self.kappa(x, y), np.float64).reshape(2, 2)
            out[0], out[1], out[2], out[3] = k[0, 0], k[0, 1], k[1, 0], k[1, 1]

        def vec(_, x, y, out):
            u = self.solution_u(x, y)
            out[0], out[1] = u[0], u[1]

        def mp(_, xi, eta, xy, j):
            (x, y), ((a, b), (c, d)) = self.map(xi, eta)
            xy[0], xy[1] = x, y
            j[0], j[1], j[2], j[3] = a, b, c, d

        def fill_nan(out, n):
            for i in range(n):
                out[i] = float("nan")

        def vguard(f, n, pos):
            def g(*a):
                try:
                    f(*a)
                except BaseException as e:
                    if self._error is None:
                        self._error = e
                    for p in pos:
                        fill_nan(a[p], n)
            return g

        def scal(f):
            return SCALAR_FN(guard(lambda _, x, y: float(f(x, y)), float("nan"))) if f else SCALAR_FN()

        nrm = guard(lambda _, x, y, nx, ny: float(self.neumann_un(x, y, nx, ny)), float("nan"))
        return ProblemFuncs(None, KAPPA_FN(vguard(kappa, 4, [3])) if self.kappa else KAPPA_FN(),
                            scal(self.solution_p), VECTOR_FN(vguard(vec, 2, [3])) if self.solution_u else VECTOR_FN(),
                            scal(self.source_g), scal(self.dirichlet_p),
                            NORMAL_FN(nrm) if self.neumann_un else NORMAL_FN(),
                            MAP_FN(vguard(mp, 2, [3])) if self.map else MAP_FN())

    def raise_pending(self):
        e, self._error = getattr(self, "_error", None), None
        if e is not None:
            raise e


class Solver:
    """Element setup (host FP64) and device solve for one mesh and problem."""

    def __init__(self, mesh, k=1, nq2d=9, nq1d=9, map_kind=MAP_SIN, eps=0.05, solution=SOL_SINCOS,
                 bc=BC_FLUX_MEANZERO, theta=None, device=None, setup_device=None, pde=None):
        """setup_device: build the element operators on that GPU instead of
        the host cores (k <= 1, polygons of up to 10 edges).  pde: a PdeData
        of callables replacing the built-in problem data (host setup only)."""
        self.mesh = mesh
        self._theta = None if theta is None else _vec(theta, mesh.num_cells, "theta")
        self._funcs = None if pde is None else pde.c_funcs()  # callbacks stay referenced by the solver
        d = ProblemDesc(k, nq2d, nq1d, map_kind, eps, solution, bc,
                        None if self._theta is None else _d(self._theta),
                        None if self._funcs is None else ctypes.pointer(self._funcs))
        self.h = c_vp()
        self._pde = pde
        try:
            if setup_device is None:
                check(lib().vemd_solver_create(mesh.h, ctypes.byref(d), ctypes.byref(self.h)))
            else:
                check(lib().vemd_solver_create_device(mesh.h, ctypes.byref(d), int(setup_device),
                                                      ctypes.byref(self.h)))
        finally:
            if pde is not None:
                pde.raise_pending()
        s = np.zeros(6, np.int64)
        check(lib().vemd_solver_sizes(self.h, s.ctypes.data_as(c_lp)))
        self.nflux, self.npres, self.ntot, self.nvec, self.ngroups, self.max_ne = (int(v) for v in s)
        self.k, self.bc = k, bc
        self.uploaded = False
        if device is not None:
            self.upload(device)

    def upload(self, device=0):
        check(lib().vemd_solver_upload(self.h, int(device)))
        self.uploaded = True

    def tuning(self):
        """Apply variant picked by the autotuner per valence: {ne: (fp64, fp32)}."""
        out = np.zeros(16, np.int32)
        check(lib().vemd_solver_tuning(self.h, out.ctypes.data_as(c_ip)))
        return {ne: (int(out[2 * (ne - 3)]), int(out[2 * (ne - 3) + 1])) for ne in range(3, 11)
                if out[2 * (ne - 3)] >= 0}

    def edge_sides(self):
        """(nside, table): the number of cell-side slots of the split apply and
        the slots of side 0 and side 1 of every internal edge (2 per edge)."""
        n = ctypes.c_int64()
        check(lib().vemd_edge_sides(self.h, ctypes.byref(n), None))
        t = np.zeros(2 * self.mesh.num_edges, np.int32)
        check(lib().vemd_edge_sides(self.h, ctypes.byref(n), t.ctypes.data_as(c_ip)))
        return int(n.value), t

    def set_theta(self, theta):
        self._theta = _vec(theta, self.mesh.num_cells, "theta")
        check(lib().vemd_set_theta(self.h, _d(self._theta)))

    def rhs(self):
        b = np.zeros(self.ntot)
        check(lib().vemd_rhs(self.h, _d(b)))
        return b

    def apply(self, x, precision=FP64):
        x = _vec(x, self.ntot, "x")
        y = np.zeros(self.ntot)
        check(lib().vemd_apply(self.h, precision, _d(x), _d(y)))
        return y

    def solve(self, opts=None, rhs=None):
        opts = opts or SolveOptions()
        x = np.zeros(self.ntot)
        info = SolveInfo()
        rb = None
        if rhs is not None:
            rhs = _vec(rhs, self.ntot, "rhs")
            rb = _d(rhs)
        oc = opts.c()
        check(lib().vemd_solve(self.h, ctypes.byref(oc), rb, _d(x), ctypes.byref(info)))
        return x, dict(iterations=info.iterations, refinements=info.refinements, rel_residual=info.rel_residual,
                       seconds=info.seconds)

    def instance_vectors(self, x):
        """Splits a reference-order vector of a batch into the instances'
        own reference-order vectors (edge moments r * E + e, internal flux
        dofs, pressure moments and, with the mean-zero condition, the
        multiplier; edges and cells of instance i follow those of instances
        < i, and the multipliers close the vector in instance order)."""
        parts = self.mesh.parts
        nlam = len(parts) if self.ntot - self.nflux - self.npres == len(parts) else 0
        kk = self.k + 1
        q = (self.k + 1) * (self.k + 2) // 2
        ni = (q - 1) + (q - kk)
        E = self.mesh.num_edges
        C = self.mesh.num_cells
        nflux = kk * E + ni * C
        out, e0, c0 = [], 0, 0
        for m in parts:
            e, c = m.num_edges, m.num_cells
            v = [x[r * E + e0:r * E + e0 + e] for r in range(kk)]
            v.append(x[kk * E + ni * c0:kk * E + ni * (c0 + c)])
            v.append(x[nflux + q * c0:nflux + q * (c0 + c)])
            if nlam:
                v.append(x[self.ntot - nlam + len(out):self.ntot - nlam + len(out) + 1])
            out.append(np.concatenate(v))
            e0 += e
            c0 += c
        return out

    def batch_info(self):
        """Per-instance iterations and true relative residuals of the last solve."""
        n = len(self.mesh.instance_cells) - 1 if hasattr(self.mesh, "instance_cells") else 1
        arr = (SolveInfo * n)()
        check(lib().vemd_batch_info(self.h, n, arr))
        return [dict(iterations=a.iterations, rel_residual=a.rel_residual) for a in arr]

    def adjoint_theta(self, x, dldx, opts=None):
        opts = opts or SolveOptions()
        x = _vec(x, self.ntot, "x")
        g = _vec(dldx, self.ntot, "dldx")
        grad = np.zeros(self.mesh.num_cells)
        info = SolveInfo()
        oc = opts.c()
        check(lib().vemd_adjoint_theta(self.h, ctypes.byref(oc), _d(x), _d(g), _d(grad), ctypes.byref(info)))
        return grad, dict(iterations=info.iterations, rel_residual=info.rel_residual)

    def errors(self, x, nq=9):
        """L2 errors of a reference-order solution: pressure and flux in the
        computational and the physical measure, the physical divergence error
        and the mass int p_h det J."""
        x = _vec(x, self.ntot, "x")
        out = np.zeros(6)
        try:
            check(lib().vemd_errors_full(self.h, _d(x), nq, _d(out)))
        finally:
            if self._pde is not None:
                self._pde.raise_pending()
        return dict(p_comp=out[0], u_comp=out[1], p_phys=out[2], u_phys=out[3], div_phys=out[4], mass=out[5])

    def errors_device(self, x_dev, nq=9, stream=0):
        """The six error functionals of errors() for an internal-layout FP64
        solution in device memory (a torch tensor or a device pointer),
        evaluated on the GPU after the work queued on `stream`."""
        ptr = x_dev.data_ptr() if hasattr(x_dev, "data_ptr") else int(x_dev)
        out = np.zeros(6)
        check(lib().vemd_errors_device(self.h, c_vp(ptr), nq, _d(out), c_vp(stream or None)))
        return dict(p_comp=out[0], u_comp=out[1], p_phys=out[2], u_phys=out[3], div_phys=out[4], mass=out[5])

    def element(self, cell):
        """FP64 element operator of one cell: dict with A (n x n), W (q x n), L, f, sign, gdof."""
        q = (self.k + 1) * (self.k + 2) // 2
        n = ctypes.c_int()
        A = np.zeros(64 * 64)
        W = np.zeros(q * 64)
        L = np.zeros(q)
        f = np.zeros(q)
        sg = np.zeros(64)
        gd = np.zeros(64, np.int64)
        check(lib().vemd_element_matrices(self.h, int(cell), ctypes.byref(n), _d(A), _d(W), _d(L), _d(f), _d(sg),
                                          gd.ctypes.data_as(c_lp)))
        n = n.value
        return dict(n=n, A=A[: n * n].reshape(n, n), W=W[: q * n].reshape(q, n), L=L, f=f, sign=sg[:n],
                    gdof=gd[:n])

    def __del__(self):
        try:
            lib().vemd_solver_destroy(self.h)
        except Exception:
            pass


class DarcyProblem:
    """Mirror of Darcy::DarcyProblem (CPU/examples/darcy_problem.h:185-229):
    example 1 = sinusoidal curved map eps = 0.05, example 2 = eps = 0, both with
    p = sin(pi x) cos(pi y), K = I, essential normal flux + mean-zero pressure.
    Extensions: map_kind=MAP_ARCS (per-cell arc maps), solution=SOL_SINSIN,
    bc=BC_DIRICHLET_P and a per-cell log-permeability theta."""

    def __init__(self, map_kind=MAP_SIN, eps=0.05, solution=SOL_SINCOS, bc=BC_FLUX_MEANZERO, theta=None, pde=None):
        """pde: a PdeData of callables (the reference's DarcyPdeData fields and
        its map) replacing the built-in data."""
        self.problem_index = 0
        self.map_kind, self.eps, self.solution, self.bc, self.theta = map_kind, eps, solution, bc, theta
        self.pde = pde

    def set_problem_index(self, index):
        self.problem_index = index

    def init_problem(self):
        if self.problem_index == 1:
            self.map_kind, self.eps = MAP_SIN, 0.05
        elif self.problem_index == 2:
            self.map_kind, self.eps = MAP_SIN, 0.0
        elif self.problem_index != 0:
            return False
        return True


class DarcySolver:
    """Mirror of DarcySolver (CPU/solver/DarcySolver.h:17-186) on the device path."""

    def __init__(self, mesh_reader, problem, gauss_point_num=9, gauss_point_num_1d=9, k=1, device=0):
        mesh = getattr(mesh_reader, "mesh", mesh_reader)
        self.solver = Solver(mesh, k=k, nq2d=gauss_point_num, nq1d=gauss_point_num_1d, map_kind=problem.map_kind,
                             eps=problem.eps, solution=problem.solution, bc=problem.bc, theta=problem.theta,
                             pde=getattr(problem, "pde", None))
        self.device = device
        self.solution = None
        self.info = None
        self.opts = SolveOptions(precision=FP64, rtol=1e-10)

    def getTotalDofFlux(self):
        return self.solver.nflux

    def getTotalDofConc(self):
        return self.solver.npres

    def getTotalDof(self):
        return self.solver.nflux + self.solver.npres

    def solve(self):
        try:
            if not self.solver.uploaded:
                self.solver.upload(self.device)
            self.solution, self.info = self.solver.solve(self.opts)
        except VemdError as e:
            if e.status in ("ERR_NO_DEVICE",):
                raise
            return False
        return True

    def getSolution(self):
        return self.solution

    def computeL2ErrorPressure(self, gauss_point_num=9):
        return self.solver.errors(self.solution, gauss_point_num)["p_comp"]

    def computeL2ErrorFlux(self, gauss_point_num=9):
        return self.solver.errors(self.solution, gauss_point_num)["u_comp"]


__all__ = ["Solver", "SolveOptions", "DarcyProblem", "DarcySolver", "FP32", "FP64", "MIXED", "MAP_ARCS", "MAP_SIN",
           "SOL_SINCOS", "SOL_SINSIN", "BC_FLUX_MEANZERO", "BC_DIRICHLET_P", "c_ip"]
