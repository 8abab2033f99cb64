"""C-ABI boundary (CPU): the library loads, exports every declared symbol, the
host FP64 setup equals the oracle, and the solve path fails loudly without a
HIP device."""
import os
import re

import numpy as np
import pytest

from conftest import rosette, PHYS, ROOT, check_phys, golden


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "vemd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vemd_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol(vemd):
    import ctypes
    L = ctypes.CDLL(os.path.join(ROOT, "wsl-code_amd", "libvemd.so"))
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(vemd._lib.SIGNATURES), set(names) ^ set(vemd._lib.SIGNATURES)


@pytest.mark.parametrize("n", [1, 2, 4])
def test_mesh_reader_matches_oracle(vemd, oracle, n):
    path = golden("square_%dx%d.msh" % (n, n))
    a = vemd.Mesh.read(path).arrays()
    o = oracle.Mesh.read(path)
    assert np.array_equal(a["cell_off"], o.off) and np.array_equal(a["cell_nodes"], o.nodes)
    assert np.array_equal(a["cell_edges"], o.cell_edges) and np.array_equal(a["edge_nodes"], o.edge_nodes)
    assert np.array_equal(a["edge_occ"], o.edge_occ)
    for key in ("cx", "cy", "area", "diam"):
        assert np.array_equal(a[key], getattr(o, key))


def test_vtk_roundtrip(vemd, tmp_path):
    m = vemd.Mesh.read(golden("square_2x2.msh"))
    p = tmp_path / "m.vtk"
    m.write_vtk(p)
    a, b = m.arrays(), vemd.Mesh.read(p, vtk=True).arrays()
    for key in ("cell_nodes", "edge_nodes", "cx", "diam"):
        assert np.array_equal(a[key], b[key])


CASES = [(1, 0.05, 0), (0, 0.05, 0), (1, 0.0, 0), (1, 0.05, 1), (2, 0.05, 0), (3, 0.05, 0)]


@pytest.mark.parametrize("k,eps,bc", CASES)
def test_setup_matches_oracle(vemd, oracle, k, eps, bc):
    """Element operators, direction signs, dof map, lifted rhs and the L2 error
    functionals of the product's FP64 setup against the oracle."""
    path = golden("square_2x2.msh")
    nq = 12 if k == 3 else 9
    s = vemd.Solver(vemd.Mesh.read(path), k=k, nq2d=nq, eps=eps, bc=bc)
    op = oracle.Problem(oracle.Mesh.read(path), k=k, nq2=nq, eps=eps, bc=bc)
    for c in range(16):
        e = s.element(c)
        n = e["n"]
        d = op.element(c, 13)[:n]
        Ao = op.element(c, 7, (n, n)) * np.outer(d, d)
        Wo = op.element(c, 2).reshape(-1, n) * d
        assert np.abs(e["A"] - Ao).max() <= 1e-13 * np.abs(Ao).max()
        assert np.abs(e["W"] - Wo).max() <= 1e-13
        assert np.allclose(e["L"], op.element(c, 8), rtol=1e-13, atol=1e-16)
        assert np.allclose(e["f"], op.element(c, 9), rtol=1e-12, atol=1e-14)
        assert np.array_equal(e["sign"], d) and np.array_equal(e["gdof"], op.element(c, 12)[:n].astype(int))
    A, rr = op.csr(True)
    fi, ub = op.free()
    b = s.rhs()
    assert np.abs(b[fi] - rr).max() <= 1e-13 * np.abs(rr).max()
    mask = np.ones(len(b), bool)
    mask[fi] = False
    assert np.allclose(b[mask], ub[mask], rtol=1e-14, atol=1e-15)
    x = op.solve_direct()
    eo, ep = op.errors(nq), s.errors(x, nq)
    for key in eo:  # the mass is zero to round-off under the mean-zero condition
        assert abs(ep[key] - eo[key]) <= 1e-13 * abs(eo[key]) + (1e-15 if key == "mass" else 0.0), key


@pytest.mark.parametrize("case", PHYS, ids=lambda c: "%s-ex%d" % (c["mesh"][:-4], c["example"]))
def test_physical_errors_match_reference_reports(vemd, oracle, case):
    """The product's error functionals (host FP64) on the oracle's direct
    solution: physical-domain pressure, flux and divergence errors and the
    mass against the reference GPU driver's report.json values."""
    path = golden(case["mesh"])
    eps = 0.05 if case["example"] == 1 else 0.0
    s = vemd.Solver(vemd.Mesh.read(path), k=case["k"], nq2d=case["nq2"], nq1d=case["nq1"], eps=eps)
    op = oracle.Problem(oracle.Mesh.read(path), k=case["k"], nq2=case["nq2"], nq1=case["nq1"], eps=eps)
    check_phys(s.errors(op.solve_direct(), case["nq2"]), case["phys"])


@pytest.mark.parametrize("k", [0, 1])
def test_errors_on_curved_cells_match_oracle(vemd, oracle, k):
    """All six error functionals (both measures, physical divergence, mass)
    of the product's host routine against the oracle's, on a generated annulus
    with exact boundary arcs, Dirichlet pressure and a cellwise permeability,
    for the oracle's direct solution and for a perturbed one (so that the
    errors are not small by construction)."""
    from wsl_code_amd.fields import grf_theta
    m = vemd.generate_mesh(400, domain=vemd.DOMAIN_ANNULUS, r_inner=0.5, lloyd_iters=20, seed=21,
                           min_edge_frac=0.05)
    a = m.arrays()
    th = grf_theta(a["cx"], a["cy"], sigma=1.0, ell=0.1, seed=22)
    s = vemd.Solver(m, k=k, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P, theta=th)
    om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
    op = oracle.Problem(om, k=k, map_kind=oracle.MAP_ARCS, sol=oracle.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P, theta=th)
    x = op.solve_direct()
    for y in (x, x + 1e-3 * np.random.default_rng(23).standard_normal(len(x))):
        op.set_solution(y)
        eo, ep = op.errors(9), s.errors(y, 9)
        for key in eo:
            assert abs(ep[key] - eo[key]) <= 1e-12 * abs(eo[key]) + 1e-15, (key, ep[key], eo[key])
    assert eo["div_phys"] > 0 and abs(eo["mass"]) > 0


@pytest.mark.parametrize("k", [0, 1])
def test_edge_side_table(vemd, k):
    """The side table of the one-launch apply (built with the layout, on the
    host): every edge has a real first side, exactly the boundary edges have
    the zero slot as second side, and the real slots are distinct and as many
    as the cells' edges."""
    m = vemd.generate_mesh(500, domain=vemd.DOMAIN_ANNULUS, r_inner=0.5, lloyd_iters=10, seed=31,
                           min_edge_frac=0.05)
    occ = m.arrays()["edge_occ"]
    s = vemd.Solver(m, k=k, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P)
    nside, t = s.edge_sides()
    zero = nside - 1
    assert t.shape == (2 * m.num_edges,) and t.min() >= 0 and t.max() <= zero
    assert (t[0::2] < zero).all()
    assert (t[1::2] == zero).sum() == (occ == 1).sum()
    real = t[t < zero]
    assert len(real) == occ.sum() == len(np.unique(real))


def test_errors_reported_not_raised(vemd):
    with pytest.raises(vemd.VemdError) as ei:
        vemd.Mesh.read("/nonexistent/mesh.msh")
    assert ei.value.status == "ERR_IO"
    m = vemd.Mesh.read(golden("square_1x1.msh"))
    with pytest.raises(vemd.VemdError) as ei:
        vemd.Solver(m, k=5)
    assert ei.value.status == "ERR_ARG"
    with pytest.raises(vemd.VemdError) as ei:
        vemd.Solver(m, k=3, nq2d=9)
    assert ei.value.status == "ERR_ARG"


@pytest.mark.parametrize("k", [0, 1])
def test_largest_polygon_setup_matches_oracle(vemd, oracle, k):
    """A 10-gon ringed by triangles (both ends of the valence range): element
    operators, lifted right-hand side and the direct solution's error
    functionals of the product's setup against the oracle."""
    xy, off, nodes = rosette()
    m = vemd.Mesh.from_arrays(xy, off, nodes)
    a = m.arrays()
    assert abs(a["area"].sum() - 5 * 0.45 ** 2 * np.sin(np.pi / 5)) <= 1e-14
    s = vemd.Solver(m, k=k, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P)
    om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
    op = oracle.Problem(om, k=k, map_kind=oracle.MAP_ARCS, sol=oracle.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P)
    for c in range(m.num_cells):
        e = s.element(c)
        n = e["n"]
        assert n == (k + 1) * (10 if c == 0 else 3) + ((k + 1) * (k + 2) // 2 - 1) + (k * (k + 1) // 2)
        d = op.element(c, 13)[:n]
        Ao = op.element(c, 7, (n, n)) * np.outer(d, d)
        assert np.abs(e["A"] - Ao).max() <= 1e-13 * np.abs(Ao).max()
        assert np.abs(e["W"] - op.element(c, 2).reshape(-1, n) * d).max() <= 1e-13
    A, rr = op.csr(True)
    fi, _ = op.free()
    assert np.abs(s.rhs()[fi] - rr).max() <= 1e-13 * np.abs(rr).max()
    x = op.solve_direct()
    eo, ep = op.errors(9), s.errors(x, 9)
    for key in ("p_comp", "u_comp"):
        assert abs(ep[key] - eo[key]) <= 1e-12 * abs(eo[key]), key


def test_malformed_inputs_rejected(vemd, tmp_path):
    """Inputs the element setup cannot integrate, and arrays whose lengths do
    not match what the C ABI reads, fail with a status instead of reading out
    of bounds or producing NaN operators."""
    xy = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    one = np.array([0, 4], np.int32)

    def status(f):
        with pytest.raises(vemd.VemdError) as ei:
            f()
        return ei.value.status

    assert status(lambda: vemd.Mesh.from_arrays(xy, one, np.array([0, 1, 2, 7], np.int32))) == "ERR_MESH"
    assert status(lambda: vemd.Mesh.from_arrays(xy, one, np.array([0, 1, 2, -1], np.int32))) == "ERR_MESH"
    assert status(lambda: vemd.Mesh.from_arrays(xy, np.array([0, 2], np.int32), np.array([0, 1], np.int32))) == "ERR_MESH"
    assert status(lambda: vemd.Mesh.from_arrays(xy, one, np.array([0, 1, 1, 3], np.int32))) == "ERR_MESH"
    assert status(lambda: vemd.Mesh.from_arrays(np.array([[0, 0], [1, 0], [2, 0], [3, 0]], float), one,
                                                np.arange(4, dtype=np.int32))) == "ERR_MESH"  # zero area
    bad = xy.copy()
    bad[2, 1] = np.nan
    assert status(lambda: vemd.Mesh.from_arrays(bad, one, np.arange(4, dtype=np.int32))) == "ERR_MESH"
    # offsets past the end of cell_nodes, and decreasing offsets
    assert status(lambda: vemd.Mesh.from_arrays(xy, np.array([0, 5], np.int32), np.arange(4, dtype=np.int32))) == "ERR_ARG"
    assert status(lambda: vemd.Mesh.from_arrays(xy, np.array([0, 4, 2], np.int32), np.arange(4, dtype=np.int32))) == "ERR_ARG"
    # a clockwise cell is reordered, as the reference's readers do
    cw = vemd.Mesh.from_arrays(xy, one, np.array([3, 2, 1, 0], np.int32))
    ccw = vemd.Mesh.from_arrays(xy, one, np.arange(4, dtype=np.int32))
    assert np.allclose(cw.arrays()["area"], ccw.arrays()["area"])
    e = vemd.Solver(ccw, k=1).element(0)
    assert np.allclose(vemd.Solver(cw, k=1).element(0)["A"], e["A"])
    # Gmsh 4.1 file with one quad: valid, then with a repeated node
    def msh(quad):
        return ("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 4 1 4\n2 1 0 4\n1\n2\n3\n4\n0 0 0\n1 0 0\n"
                "1 1 0\n0 1 0\n$EndNodes\n$Elements\n1 1 1 1\n2 1 3 1\n1 %s\n$EndElements\n" % quad)
    (tmp_path / "ok.msh").write_text(msh("1 2 3 4"))
    assert vemd.Mesh.read(tmp_path / "ok.msh").num_cells == 1
    (tmp_path / "rep.msh").write_text(msh("1 2 2 4"))
    assert status(lambda: vemd.Mesh.read(tmp_path / "rep.msh")) == "ERR_MESH"
    m = vemd.Mesh.read(golden("square_2x2.msh"))
    assert status(lambda: vemd.Solver(m, k=1, theta=np.zeros(3))) == "ERR_ARG"
    theta = np.zeros(m.num_cells)
    theta[1] = np.inf
    assert status(lambda: vemd.Solver(m, k=1, theta=theta, bc=vemd.BC_DIRICHLET_P)) == "ERR_ARG"
    s = vemd.Solver(m, k=1, bc=vemd.BC_DIRICHLET_P)
    assert status(lambda: s.set_theta(np.zeros(m.num_cells + 1))) == "ERR_ARG"
    assert status(lambda: s.set_theta(theta)) == "ERR_ARG"
    assert status(lambda: s.apply(np.zeros(s.ntot - 1))) == "ERR_ARG"
    assert status(lambda: s.errors(np.zeros(s.ntot + 1))) == "ERR_ARG"


def test_solve_without_device_fails_loudly(vemd):
    if vemd.device_count() > 0:
        pytest.skip("a HIP device is present")
    s = vemd.Solver(vemd.Mesh.read(golden("square_1x1.msh")))
    with pytest.raises(vemd.VemdError) as ei:
        s.upload(0)
    assert ei.value.status == "ERR_NO_DEVICE"
    with pytest.raises(vemd.VemdError) as ei:
        s.solve()
    assert ei.value.status == "ERR_STATE"


def test_device_setup_limits_checked_on_the_host(vemd):
    """The GPU element setup rejects k > 2 and polygons of more than 10 edges
    before touching a device, and without a device it fails loudly."""
    m = vemd.Mesh.read(golden("square_2x2.msh"))
    with pytest.raises(vemd.VemdError) as ei:
        vemd.Solver(m, k=3, nq2d=12, setup_device=0)
    assert ei.value.status == "ERR_ARG"
    # one regular 11-gon
    t = np.linspace(0.0, 2.0 * np.pi, 12)[:-1]
    xy = np.stack([np.cos(t), np.sin(t)], 1)
    big = vemd.Mesh.from_arrays(xy, np.array([0, 11], np.int32), np.arange(11, dtype=np.int32))
    with pytest.raises(vemd.VemdError) as ei:
        vemd.Solver(big, k=1, setup_device=0)
    assert ei.value.status == "ERR_ARG" and "10 edges" in str(ei.value)
    if vemd.device_count() == 0:
        with pytest.raises(vemd.VemdError) as ei:
            vemd.Solver(m, k=1, setup_device=0)
        assert ei.value.status == "ERR_DEVICE"


def test_vector_range_limit_fails_loudly(vemd, monkeypatch):
    """The apply addresses the solve's vectors with 32-bit byte offsets (a
    record past the range reads as zero), so a layout whose FP64 vector would
    not fit is rejected at setup with VEMD_ERR_MESH instead of solving with
    silently zeroed entries.  The real limit (2^31 bytes, 268M entries) needs
    a mesh of ~20M cells; the environment lowers it here to force the error."""
    m = vemd.Mesh.read(golden("square_2x2.msh"))
    nbytes = 8 * vemd.Solver(m, k=1).nvec
    monkeypatch.setenv("VEMD_MAX_VECTOR_BYTES", str(nbytes - 1))
    with pytest.raises(vemd.VemdError) as e:
        vemd.Solver(m, k=1)
    assert e.value.code == 3 and "32-bit offsets" in str(e.value)
    monkeypatch.setenv("VEMD_MAX_VECTOR_BYTES", str(nbytes))
    assert vemd.Solver(m, k=1).nvec * 8 == nbytes  # exactly at the limit
    monkeypatch.setenv("VEMD_MAX_VECTOR_BYTES", str(1 << 40))  # the variable cannot raise the limit
    assert vemd.Solver(m, k=1).nvec * 8 == nbytes


def test_meanzero_batch_has_one_multiplier_per_instance(vemd):
    """A batch under the reference's condition (essential normal flux and a
    mean-zero pressure multiplier) carries one multiplier per instance, the
    last unknowns in instance order: the sizes, the right-hand side (zero
    multiplier rows), the split into instance vectors and the reference <->
    internal maps agree with the instances solved alone."""
    import ctypes
    parts = [vemd.Mesh.read(golden("square_2x2.msh")), vemd.Mesh.read(golden("square_4x4.msh")),
             vemd.Mesh.read(golden("square_2x2.msh"))]
    u = vemd.Mesh.concat(parts)
    kw = dict(k=1, bc=vemd.BC_FLUX_MEANZERO)
    sb = vemd.Solver(u, **kw)
    singles = [vemd.Solver(m, **kw) for m in parts]
    assert sb.ntot == sb.nflux + sb.npres + len(parts)
    assert sb.ntot == sum(s.ntot for s in singles)
    for bi, s in zip(sb.instance_vectors(sb.rhs()), singles):
        assert np.array_equal(bi, s.rhs())
        assert bi[-1] == 0.0
    rng = np.random.default_rng(5)
    x = rng.standard_normal(sb.ntot)
    dp = ctypes.POINTER(ctypes.c_double)
    xi = np.zeros(sb.nvec)
    xr = np.zeros(sb.ntot)
    vemd._lib.check(vemd.lib().vemd_to_internal(sb.h, x.ctypes.data_as(dp), xi.ctypes.data_as(dp), 0))
    vemd._lib.check(vemd.lib().vemd_to_reference(sb.h, xi.ctypes.data_as(dp), xr.ctypes.data_as(dp), 0))
    assert np.array_equal(x, xr)
    # the multipliers sit in distinct internal slots
    lam = xi[np.isin(xi, x[-len(parts):])]
    assert len(lam) == len(parts)
