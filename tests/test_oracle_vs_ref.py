"""Pins the oracle's restatement of the reference's PETSc-free modules
(quadrature, basis, mesh reader conventions, ear clipping, problem data)
against those modules themselves, compiled from the reference tree into
oracle/_ref/libref.so.  Skipped where that library was not built (it needs the
reference tree, which exists only in the build container)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, golden

REF = os.path.join(ROOT, "oracle", "_ref", "libref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (no reference tree)")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


@pytest.fixture(scope="module")
def ref():
    L = ctypes.CDLL(REF)
    L.ref_edge_mono.restype = ctypes.c_double
    for f in ("ref_edge_mono",):
        getattr(L, f).argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 6
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _ref_mesh(ref, path):
    cnt = np.zeros(4, np.int32)
    assert ref.ref_read(path.encode(), 0, _i(cnt)) == 0
    nn, nc, ne, nt = (int(v) for v in cnt)
    out = dict(cx=np.zeros(nc), cy=np.zeros(nc), area=np.zeros(nc), diam=np.zeros(nc),
               off=np.zeros(nc + 1, np.int32), nodes=np.zeros(nt, np.int32), cell_edges=np.zeros(nt, np.int32),
               edge_nodes=np.zeros(2 * ne, np.int32), edge_occ=np.zeros(ne, np.int32))
    ref.ref_mesh_get(_d(out["cx"]), _d(out["cy"]), _d(out["area"]), _d(out["diam"]), _i(out["off"]),
                     _i(out["nodes"]), _i(out["cell_edges"]), _i(out["edge_nodes"]), _i(out["edge_occ"]))
    return out


@pytest.mark.parametrize("n", [1, 2, 4, 8])
def test_mesh_conventions_and_triangulation(ref, oracle, n):
    path = golden("square_%dx%d.msh" % (n, n))
    r = _ref_mesh(ref, path)
    m = oracle.Mesh.read(path)
    for key in ("cx", "cy", "area", "diam"):
        assert np.array_equal(getattr(m, key), r[key]), key
    for key in ("off", "nodes", "cell_edges", "edge_nodes", "edge_occ"):
        assert np.array_equal(getattr(m, key), r[key]), key
    for cell in range(m.nc):
        buf = np.zeros(6 * 16)
        nt = ref.ref_triangulate(cell, _d(buf), 16)
        assert np.array_equal(m.triangulate(cell), buf[: 6 * nt].reshape(nt, 6))


@pytest.mark.parametrize("npts", [1, 3, 4, 7, 9, 12])
def test_triangle_rules(ref, oracle, npts):
    v = np.array([0.1, 0.2, 0.9, 0.35, 0.4, 1.1])
    out = np.zeros(3 * npts)
    assert ref.ref_triangle_points(npts, _d(v), _d(out)) == npts
    out = out.reshape(npts, 3)
    xi, eta, w = np.zeros(npts), np.zeros(npts), np.zeros(npts)
    assert oracle.lib().orc_triangle_rule(npts, _d(xi), _d(eta), _d(w)) == npts
    J = abs((v[2] - v[0]) * (v[5] - v[1]) - (v[4] - v[0]) * (v[3] - v[1]))
    mine = np.stack([v[0] + (v[2] - v[0]) * xi + (v[4] - v[0]) * eta,
                     v[1] + (v[3] - v[1]) * xi + (v[5] - v[1]) * eta, w * J], 1)
    # the oracle generates the collapsed-Gauss rules in its own point order
    key = lambda a: np.lexsort((a[:, 1], a[:, 0]))
    # The reference tabulates the second weight orbit of the 7-point rule as
    # 0.0661970763907399 (CPU/core/gauss_quadrature.cpp:285); Radon's closed
    # form (155+sqrt(15))/2400 = 0.06619707639425309, so its weights sum to
    # 0.49999999998946.  The oracle keeps the closed form; no Darcy case here
    # uses the 7-point rule.
    atol = 1e-11 if npts == 7 else 2e-15
    np.testing.assert_allclose(mine[key(mine)], out[key(out)], rtol=0, atol=atol)


@pytest.mark.parametrize("n", range(1, 11))
def test_gauss_legendre(ref, oracle, n):
    out = np.zeros(3 * n)
    assert ref.ref_line_points(n, ctypes.c_double(-1.0), ctypes.c_double(0.0), ctypes.c_double(1.0),
                               ctypes.c_double(0.0), _d(out)) == n
    x, w = np.zeros(n), np.zeros(n)
    oracle.lib().orc_gauss_legendre(n, _d(x), _d(w))
    out = out.reshape(n, 3)
    np.testing.assert_allclose(x, out[:, 0], rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, out[:, 2], rtol=0, atol=2e-15)


@pytest.mark.parametrize("k", [0, 1, 2, 3, 4])
def test_basis(ref, oracle, k):
    rng = np.random.default_rng(k)
    lib = oracle.lib()
    q = (k + 1) * (k + 2) // 2
    p = (k + 2) * (k + 3) // 2 - 1 + (k * (k + 1) // 2)
    for _ in range(20):
        xc, yc, h, x, y = rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.1, 2), rng.uniform(-1, 1), rng.uniform(-1, 1)
        a, b = np.zeros(2), np.zeros(2)
        for nk in range(q):
            args = [ctypes.c_double(v) for v in (xc, yc, h, x, y)]
            ref.ref_basis_eval(k, 0, nk, *args, _d(a))
            lib.orc_basis_eval(k, 0, nk, *args, _d(b))
            assert abs(a[0] - b[0]) <= 1e-14 * max(1, abs(a[0]))
        for nk in range(p):
            args = [ctypes.c_double(v) for v in (xc, yc, h, x, y)]
            ref.ref_basis_eval(k, 1, nk, *args, _d(a))
            lib.orc_basis_eval(k, 1, nk, *args, _d(b))
            np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-14)
        x1, y1, x2, y2 = rng.uniform(-1, 1, 4)
        for nk in range(q):
            ca, cb = np.zeros(k + 1), np.zeros(k + 1)
            args = [ctypes.c_double(v) for v in (xc, yc, h, x1, y1, x2, y2)]
            ref.ref_edge_coeffs(k, nk, *args, _d(ca))
            lib.orc_edge_coeffs(k, nk, *args, _d(cb))
            np.testing.assert_allclose(ca, cb, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("example,eps", [(1, 0.05), (2, 0.0)])
def test_problem_data(ref, oracle, example, eps):
    """Pulled-back p, u, g, the computational permeability tensor and det J of
    reference examples 1 and 2 (Darcy::DarcyProblem::computational_*), each
    against the reference's own accessor at the same computational point."""
    m = oracle.Mesh.read(golden("square_1x1.msh"))
    prob = oracle.Problem(m, k=1, eps=eps)
    rng = np.random.default_rng(example)
    out = np.zeros(9)
    names = ("p", "u_x", "u_y", "g", "K_xx", "K_xy", "K_yx", "K_yy", "det J")
    for _ in range(30):
        xi, eta = rng.uniform(0, 1, 2)
        assert ref.ref_problem_data(example, ctypes.c_double(xi), ctypes.c_double(eta), _d(out)) == 0
        mine = prob.problem_data(0, xi, eta)
        for i, name in enumerate(names):
            assert abs(mine[i] - out[i]) <= 1e-14 * max(1.0, abs(out[i])), (name, mine[i], out[i])
        g = prob.map(0, xi, eta)
        assert abs(g[6] - out[8]) <= 1e-15
    if eps == 0.0:
        # example 2 is the identity map: K_hat = K = I exactly
        assert np.array_equal(out[4:8], [1.0, 0.0, 0.0, 1.0])


@pytest.mark.parametrize("seed", [5, 6])
def test_vtk_reader(ref, oracle, vemd, tmp_path, seed):
    """A generated Voronoi disk (mostly 5- to 7-gons after Lloyd relaxation,
    quads from the edge collapse) written by vemd_mesh_write_vtk and read
    back by the reference's read_mesh_vtk (CPU/mesh/straight_mesh.cpp:92-214),
    the oracle's and the product's VTK readers: coordinates, centroids,
    diameters, areas, node lists, edge numbering and the ear-clipping
    triangulation of every cell agree bitwise."""
    m = vemd.generate_mesh(300, lloyd_iters=20, seed=seed, min_edge_frac=0.05)
    path = str(tmp_path / "disk.vtk")
    m.write_vtk(path)
    r = _ref_mesh_vtk(ref, path)
    o = oracle.Mesh.read(path, vtk=True)
    p = vemd.Mesh.read(path, vtk=True).arrays()
    valences = np.diff(r["off"])
    assert valences.min() <= 4 and valences.max() >= 7
    for key in ("cx", "cy", "area", "diam"):
        assert np.array_equal(getattr(o, key), r[key]), key
        assert np.array_equal(p[key], r[key]), key
    for key, pk in (("off", "cell_off"), ("nodes", "cell_nodes"), ("cell_edges", "cell_edges"),
                    ("edge_nodes", "edge_nodes"), ("edge_occ", "edge_occ")):
        assert np.array_equal(getattr(o, key), r[key]), key
        assert np.array_equal(p[pk], r[key]), key
    for cell in range(o.nc):
        buf = np.zeros(6 * 16)
        nt = ref.ref_triangulate(cell, _d(buf), 16)
        assert np.array_equal(o.triangulate(cell), buf[: 6 * nt].reshape(nt, 6)), cell


def _ref_mesh_vtk(ref, path):
    cnt = np.zeros(4, np.int32)
    assert ref.ref_read(path.encode(), 1, _i(cnt)) == 0
    nn, nc, ne, nt = (int(v) for v in cnt)
    out = dict(cx=np.zeros(nc), cy=np.zeros(nc), area=np.zeros(nc), diam=np.zeros(nc),
               off=np.zeros(nc + 1, np.int32), nodes=np.zeros(nt, np.int32), cell_edges=np.zeros(nt, np.int32),
               edge_nodes=np.zeros(2 * ne, np.int32), edge_occ=np.zeros(ne, np.int32))
    ref.ref_mesh_get(_d(out["cx"]), _d(out["cy"]), _d(out["area"]), _d(out["diam"]), _i(out["off"]),
                     _i(out["nodes"]), _i(out["cell_edges"]), _i(out["edge_nodes"]), _i(out["edge_occ"]))
    return out
