"""ctypes loader for the CPU oracle (test infrastructure only).

The oracle restates the reference project's CPU algorithm (vem/CurvedVEM_Darcy)
in oracle/oracle.cpp.  Only tests, the smoke check and the CPU-baseline leg of
bench.py use it, always as the checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.orc_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _check(rc):
    if rc < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return rc


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Mesh:
    """Oracle mesh: restates vem::StraightMeshData conventions."""

    def __init__(self, handle):
        self.h = handle
        cnt = np.zeros(4, np.int32)
        lib().orc_mesh_counts(self.h, _i(cnt))
        self.nn, self.nc, self.ne, nnodes = (int(v) for v in cnt)
        self.cx = np.zeros(self.nc)
        self.cy = np.zeros(self.nc)
        self.area = np.zeros(self.nc)
        self.diam = np.zeros(self.nc)
        self.off = np.zeros(self.nc + 1, np.int32)
        self.nodes = np.zeros(nnodes, np.int32)
        self.cell_edges = np.zeros(nnodes, np.int32)
        self.edge_nodes = np.zeros(2 * self.ne, np.int32)
        self.edge_occ = np.zeros(self.ne, np.int32)
        lib().orc_mesh_get(self.h, _d(self.cx), _d(self.cy), _d(self.area), _d(self.diam), _i(self.off),
                           _i(self.nodes), _i(self.cell_edges), _i(self.edge_nodes), _i(self.edge_occ))

    @classmethod
    def read(cls, path, vtk=False):
        h = ctypes.c_void_p()
        _check(lib().orc_mesh_read(str(path).encode(), int(vtk), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, xy, off, nodes, arc_nodes=None, arc_r=None):
        xy = np.ascontiguousarray(xy, np.float64).reshape(-1)
        off = np.ascontiguousarray(off, np.int32)
        nodes = np.ascontiguousarray(nodes, np.int32)
        if arc_nodes is None:
            arc_nodes = np.zeros((0, 2), np.int32)
            arc_r = np.zeros(0)
        arc_nodes = np.ascontiguousarray(arc_nodes, np.int32).reshape(-1)
        arc_r = np.ascontiguousarray(arc_r, np.float64)
        h = ctypes.c_void_p()
        _check(lib().orc_mesh_from_arrays(len(xy) // 2, _d(xy), len(off) - 1, _i(off), _i(nodes), len(arc_r),
                                          _i(arc_nodes), _d(arc_r), ctypes.byref(h)))
        return cls(h)

    def triangulate(self, cell):
        buf = np.zeros(6 * 64)
        n = _check(lib().orc_triangulate(self.h, int(cell), _d(buf), 64))
        return buf[: 6 * n].reshape(n, 6)

    def __del__(self):
        try:
            lib().orc_mesh_free(self.h)
        except Exception:
            pass


MAP_SIN, MAP_ARCS = 0, 1
SOL_SINCOS, SOL_SINSIN = 0, 1
BC_FLUX_MEANZERO, BC_DIRICHLET_P = 0, 1


_KAPPA = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _dp)
_SCAL = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_double, ctypes.c_double)
_VEC = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _dp)
_NRM = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                        ctypes.c_double)
_MAP = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _dp, _dp)


class _OrcFuncs(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("kappa", _KAPPA), ("solution_p", _SCAL), ("solution_u", _VEC),
                ("source_g", _SCAL), ("dirichlet_p", _SCAL), ("neumann_un", _NRM), ("map", _MAP)]


def _orc_funcs(pde):
    """ctypes callbacks for an object with the optional attributes kappa,
    solution_p, solution_u, source_g, dirichlet_p, neumann_un, map (Python
    callables of physical (x, y), or (xi, eta) for map)."""
    g = lambda name: getattr(pde, name, None)

    def kappa(_, x, y, out):
        k = np.asarray(g("kappa")(x, y), np.float64).reshape(2, 2)
        out[0], out[1], out[2], out[3] = k[0, 0], k[0, 1], k[1, 0], k[1, 1]

    def vec(_, x, y, out):
        u = g("solution_u")(x, y)
        out[0], out[1] = u[0], u[1]

    def mp(_, xi, eta, xy, j):
        (x, y), ((a, b), (c, d)) = g("map")(xi, eta)
        xy[0], xy[1] = x, y
        j[0], j[1], j[2], j[3] = a, b, c, d

    def scal(name):
        f = g(name)
        return _SCAL(lambda _, x, y: float(f(x, y))) if f else _SCAL()

    return _OrcFuncs(None, _KAPPA(kappa) if g("kappa") else _KAPPA(), scal("solution_p"),
                     _VEC(vec) if g("solution_u") else _VEC(), scal("source_g"), scal("dirichlet_p"),
                     _NRM(lambda _, x, y, nx, ny: float(g("neumann_un")(x, y, nx, ny))) if g("neumann_un")
                     else _NRM(), _MAP(mp) if g("map") else _MAP())


class Problem:
    """Oracle discretisation: restates HdivMatrix + DarcySolver (CPU/solver/).
    pde: optional object of callables (DarcyPdeData fields and map) replacing
    the built-in data."""

    def __init__(self, mesh, k=1, nq2=9, nq1=9, map_kind=MAP_SIN, eps=0.05, sol=SOL_SINCOS,
                 bc=BC_FLUX_MEANZERO, theta=None, pde=None):
        self.mesh = mesh
        self.k = k
        h = ctypes.c_void_p()
        th = None
        if theta is not None:
            self._theta = np.ascontiguousarray(theta, np.float64)
            th = _d(self._theta)
        self._funcs = None if pde is None else _orc_funcs(pde)
        _check(lib().orc_problem_new2(mesh.h, k, nq2, nq1, map_kind, ctypes.c_double(eps), sol, bc, th,
                                      None if self._funcs is None else ctypes.byref(self._funcs),
                                      ctypes.byref(h)))
        self.h = h
        self.assembled = False

    def sizes(self):
        s = np.zeros(6, np.int64)
        lib().orc_sizes(self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        return dict(nf=int(s[0]), npres=int(s[1]), ntot=int(s[2]), nfree=int(s[3]), nnz=int(s[4]),
                    nnz_free=int(s[5]))

    def element(self, cell, what, shape=None):
        buf = np.zeros(64 * 64)
        n = _check(lib().orc_element(self.h, int(cell), int(what), _d(buf)))
        out = buf[:n].copy()
        return out.reshape(shape) if shape else out

    def assemble(self):
        _check(lib().orc_assemble(self.h))
        self.assembled = True

    def csr(self, reduced=True):
        import scipy.sparse as sp
        if not self.assembled:
            self.assemble()
        s = self.sizes()
        n = s["nfree"] if reduced else s["ntot"]
        nnz = s["nnz_free"] if reduced else s["nnz"]
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        rhs = np.zeros(n)
        lib().orc_csr(self.h, int(reduced), _i(rp), _i(ci), _d(v), _d(rhs))
        return sp.csr_matrix((v, ci, rp), shape=(n, n)), rhs

    def free(self):
        s = self.sizes()
        fi = np.zeros(s["nfree"], np.int32)
        ub = np.zeros(s["ntot"])
        lib().orc_free(self.h, _i(fi), _d(ub))
        return fi, ub

    def solve_direct(self):
        """Tight reference solution: sparse LU of the reduced system."""
        import scipy.sparse.linalg as sla
        A, b = self.csr(True)
        xf = sla.spsolve(A.tocsc(), b)
        fi, ub = self.free()
        x = ub.copy()
        x[fi] = xf
        self.set_solution(x)
        return x

    def solve_gmres_ilu(self, rtol=1e-10, restart=100, maxit=1000000):
        it = ctypes.c_int()
        rel = ctypes.c_double()
        _check(lib().orc_solve_gmres_ilu(self.h, ctypes.c_double(rtol), restart, maxit, ctypes.byref(it),
                                         ctypes.byref(rel)))
        return self.solution(), it.value, rel.value

    def set_solution(self, x):
        x = np.ascontiguousarray(x, np.float64)
        lib().orc_set_solution(self.h, _d(x))

    def solution(self):
        x = np.zeros(self.sizes()["ntot"])
        _check(lib().orc_get_solution(self.h, _d(x)))
        return x

    def errors(self, nq=9):
        out = np.zeros(6)
        _check(lib().orc_errors(self.h, nq, _d(out)))
        return dict(p_comp=out[0], u_comp=out[1], p_phys=out[2], u_phys=out[3], div_phys=out[4], mass=out[5])

    def cell_errors(self, nq=9):
        """Each cell's squared pressure and flux errors (computational measure), shape (nc, 2)."""
        out = np.zeros(6)
        sq = np.zeros(2 * self.mesh.nc)
        _check(lib().orc_errors_cells(self.h, nq, _d(out), _d(sq)))
        return sq.reshape(-1, 2)

    def problem_data(self, cell, xi, eta):
        """[p_hat, u_hat_x, u_hat_y, g_hat, K_hat xx, xy, yx, yy, det J] at a computational point."""
        out = np.zeros(9)
        _check(lib().orc_problem_data(self.h, int(cell), ctypes.c_double(xi), ctypes.c_double(eta), _d(out)))
        return out

    def map(self, cell, xi, eta):
        out = np.zeros(7)
        _check(lib().orc_map(self.h, int(cell), ctypes.c_double(xi), ctypes.c_double(eta), _d(out)))
        return out

    def __del__(self):
        try:
            lib().orc_problem_free(self.h)
        except Exception:
            pass
