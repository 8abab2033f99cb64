"""Device parity (MI355X): the HIP apply and MINRES solve against the oracle."""
import json

import numpy as np
import pytest

from conftest import PHYS, ROOT, check_phys, golden, rosette

pytestmark = pytest.mark.gpu

CASES = [("square_2x2.msh", 1, 0.05, 0), ("square_4x4.msh", 0, 0.05, 0), ("square_4x4.msh", 1, 0.05, 1),
         ("square_8x8.msh", 1, 0.05, 0), ("square_8x8.msh", 1, 0.0, 1)]
IDS = ["%s-k%d-eps%g-bc%d" % c for c in CASES]


def _pair(vemd, oracle, mesh, k, eps, bc):
    print("case", mesh, "k", k, "eps", eps, "bc", bc, flush=True)
    assert vemd.device_count() > 0, "no HIP device"
    s = vemd.Solver(vemd.Mesh.read(golden(mesh)), k=k, eps=eps, bc=bc, device=0)
    op = oracle.Problem(oracle.Mesh.read(golden(mesh)), k=k, eps=eps, bc=bc)
    return s, op


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_apply_matches_oracle(vemd, oracle, case):
    s, op = _pair(vemd, oracle, *case)
    A, _ = op.csr(False)
    fi, _ = op.free()
    rng = np.random.default_rng(7)
    x = np.zeros(s.ntot)
    x[fi] = rng.standard_normal(len(fi))
    y_ref = A @ x
    y = s.apply(x, vemd.FP64)
    scale = np.abs(y_ref[fi]).max()
    assert np.abs(y[fi] - y_ref[fi]).max() <= 1e-13 * scale
    bnd = np.setdiff1d(np.arange(s.ntot), fi)
    assert np.all(y[bnd] == 0.0)
    y32 = s.apply(x, vemd.FP32)
    assert np.abs(y32[fi] - y_ref[fi]).max() <= 2e-5 * scale
    # one writer per output slot and fixed-order edge sums: bitwise repeatable
    assert np.array_equal(s.apply(x, vemd.FP64), y)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_fp64_solve_matches_oracle_direct(vemd, oracle, case):
    s, op = _pair(vemd, oracle, *case)
    xo = op.solve_direct()
    x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12))
    assert info["rel_residual"] <= 1e-12
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    assert rel <= 1e-10, (rel, info)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_mixed_solve(vemd, oracle, case):
    """FP32 operator and preconditioner inside FP64 refinement reach the same
    1e-8 true residual; the error functionals then agree with those of a tight
    FP64 solve to 1e-9 absolute or 1e-7 relative (a 1e-8 residual moves an O(1)
    error functional of a coarse mesh by a few 1e-9)."""
    s, op = _pair(vemd, oracle, *case)
    xt, _ = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12))
    x64, i64 = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-8))
    xm, im = s.solve(vemd.SolveOptions(precision=vemd.MIXED, rtol=1e-8, inner_rtol=1e-4))
    assert i64["rel_residual"] <= 1e-8
    assert im["rel_residual"] <= 1e-8 and im["refinements"] >= 1
    et, e64, em = s.errors(xt), s.errors(x64), s.errors(xm)
    for key in ("u_comp", "p_comp"):
        assert abs(em[key] - et[key]) <= max(1e-9, 1e-7 * et[key]), (key, em[key], et[key])
        assert abs(e64[key] - et[key]) <= max(1e-9, 1e-7 * et[key]), (key, e64[key], et[key])
    assert np.linalg.norm(xm - xt) <= 1e-5 * np.linalg.norm(xt)


# Generated Voronoi meshes with arc boundaries: every valence group of the store
# (squares have only quadrilaterals), log-permeability fields, both boundary
# conditions; the 1024-cell k=0 disk is the lowest-order Dirichlet workload.
GEN = [("disk", 1024, 0, 1, False), ("annulus", 3000, 1, 1, True), ("disk", 2000, 1, 0, True)]
GEN_IDS = ["%s%d-k%d-bc%d-theta%d" % g for g in GEN]


def _gen_pair(vemd, oracle, domain, cells, k, bc, theta):
    print("generated", domain, cells, "k", k, "bc", bc, "theta", theta, flush=True)
    assert vemd.device_count() > 0, "no HIP device"
    from wsl_code_amd.fields import grf_theta
    m = vemd.generate_mesh(cells, domain=vemd.DOMAIN_DISK if domain == "disk" else vemd.DOMAIN_ANNULUS,
                           r_inner=0.5, lloyd_iters=20, seed=11, min_edge_frac=0.05)
    a = m.arrays()
    th = grf_theta(a["cx"], a["cy"], sigma=1.0, ell=0.1, seed=12) if theta else None
    s = vemd.Solver(m, k=k, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=bc, theta=th, device=0)
    om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
    op = oracle.Problem(om, k=k, map_kind=oracle.MAP_ARCS, sol=oracle.SOL_SINSIN, bc=bc, theta=th)
    return s, op


@pytest.mark.parametrize("case", GEN, ids=GEN_IDS)
def test_generated_apply_and_solve(vemd, oracle, case):
    s, op = _gen_pair(vemd, oracle, *case)
    assert s.ngroups >= 3  # several valences in one launch
    A, _ = op.csr(False)
    fi, _ = op.free()
    rng = np.random.default_rng(8)
    x = np.zeros(s.ntot)
    x[fi] = rng.standard_normal(len(fi))
    y_ref = A @ x
    y = s.apply(x, vemd.FP64)
    scale = np.abs(y_ref[fi]).max()
    assert np.abs(y[fi] - y_ref[fi]).max() <= 1e-13 * scale
    assert np.abs(s.apply(x, vemd.FP32)[fi] - y_ref[fi]).max() <= 2e-5 * scale
    assert np.array_equal(s.apply(x, vemd.FP64), y)
    xo = op.solve_direct()
    xs, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12))
    assert np.linalg.norm(xs - xo) <= 1e-10 * np.linalg.norm(xo), info


@pytest.mark.parametrize("k", [0, 1])
def test_largest_polygon_apply_and_solve(vemd, oracle, k):
    """A 10-gon (the largest polygon of the device path, alone in its
    valence group: one live cell in a 64-cell tile) ringed by 20 triangles:
    the apply against the oracle's CSR, the FP64 solve against its direct
    solve, and the mixed solve to the same true residual."""
    print("rosette k", k, flush=True)
    assert vemd.device_count() > 0, "no HIP device"
    xy, off, nodes = rosette()
    m = vemd.Mesh.from_arrays(xy, off, nodes)
    a = m.arrays()
    s = vemd.Solver(m, k=k, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P, device=0)
    assert s.ngroups == 2 and s.max_ne == 10
    om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
    op = oracle.Problem(om, k=k, map_kind=oracle.MAP_ARCS, sol=oracle.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P)
    A, _ = op.csr(False)
    fi, _ = op.free()
    x = np.zeros(s.ntot)
    x[fi] = np.random.default_rng(9).standard_normal(len(fi))
    y_ref = A @ x
    y = s.apply(x, vemd.FP64)
    scale = np.abs(y_ref[fi]).max()
    assert np.abs(y[fi] - y_ref[fi]).max() <= 1e-13 * scale
    assert np.abs(s.apply(x, vemd.FP32)[fi] - y_ref[fi]).max() <= 2e-5 * scale
    assert np.array_equal(s.apply(x, vemd.FP64), y)
    xo = op.solve_direct()
    xs, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12))
    assert np.linalg.norm(xs - xo) <= 1e-10 * np.linalg.norm(xo), info
    xm, im = s.solve(vemd.SolveOptions(precision=vemd.MIXED, rtol=1e-8))
    assert im["rel_residual"] <= 1e-8
    assert np.linalg.norm(xm - xo) <= 1e-5 * np.linalg.norm(xo), im


@pytest.mark.parametrize("cells", [3000, 12000])
def test_multigrid_preconditioner(vemd, oracle, cells):
    """The smoothed-aggregation pressure cycle is a fixed SPD preconditioner:
    the FP64 solve matches the oracle's direct solve, MINRES needs a small
    fraction of the iterations of the diagonal preconditioner, and the mixed
    solve reaches the same true residual."""
    s, op = _gen_pair(vemd, oracle, "annulus", cells, 1, 1, True)
    xo = op.solve_direct()
    x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12, precond=vemd.PC_MULTIGRID))
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo), info
    _, ij = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-8, precond=vemd.PC_JACOBI))
    _, im = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-8, precond=vemd.PC_MULTIGRID))
    print("iterations jacobi", ij["iterations"], "multigrid", im["iterations"], flush=True)
    assert im["rel_residual"] <= 1e-8 and ij["rel_residual"] <= 1e-8
    assert 4 * im["iterations"] <= ij["iterations"]
    xm, imx = s.solve(vemd.SolveOptions(precision=vemd.MIXED, rtol=1e-8, precond=vemd.PC_MULTIGRID))
    assert imx["rel_residual"] <= 1e-8
    assert np.linalg.norm(xm - xo) <= 1e-5 * np.linalg.norm(xo)


def test_multigrid_scalar_level_paths(vemd, oracle, monkeypatch):
    """The scalar multigrid levels run either 8 lanes per row (levels below
    VEMD_MG_LANE_ROWS rows, the default for small levels) or one lane per row
    with 16 gathers in flight (larger levels): with every scalar level forced
    to each form, the FP64 and mixed solves match the oracle's direct solve,
    and the two forms need the same iterations (they differ by rounding)."""
    res = {}
    for lane_rows in ("1000000000", "0"):
        monkeypatch.setenv("VEMD_MG_LANE_ROWS", lane_rows)
        s, op = _gen_pair(vemd, oracle, "annulus", 6000, 1, 1, True)
        xo = op.solve_direct()
        x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12, precond=vemd.PC_MULTIGRID))
        assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo), (lane_rows, info)
        xm, im = s.solve(vemd.SolveOptions(precision=vemd.MIXED, rtol=1e-10, precond=vemd.PC_MULTIGRID))
        assert im["rel_residual"] <= 1e-10 and np.linalg.norm(xm - xo) <= 1e-7 * np.linalg.norm(xo), (lane_rows, im)
        res[lane_rows] = (info["iterations"], im["iterations"])
    print("iterations (fp64, mixed): 8 lanes per row", res["1000000000"], "one lane", res["0"], flush=True)
    assert abs(res["0"][0] - res["1000000000"][0]) <= 2


def test_mixed_cycle_with_fp32_finest_level(vemd, oracle, monkeypatch):
    """The FP32 cycle's finest level in half precision (default, scaled to a
    unit diagonal) and in FP32 (VEMD_MG_HALF=0) both precondition the mixed
    solve to the same solution; the half-precision level costs at most a few
    iterations."""
    s, op = _gen_pair(vemd, oracle, "annulus", 3000, 1, 1, True)
    o = vemd.SolveOptions(precision=vemd.MIXED, rtol=1e-10, precond=vemd.PC_MULTIGRID)
    xh, ih = s.solve(o)
    monkeypatch.setenv("VEMD_MG_HALF", "0")
    s32, _ = _gen_pair(vemd, oracle, "annulus", 3000, 1, 1, True)
    x32, i32 = s32.solve(o)
    print("iterations: half-precision finest level", ih["iterations"], "FP32", i32["iterations"], flush=True)
    assert ih["rel_residual"] <= 1e-10 and i32["rel_residual"] <= 1e-10
    assert np.linalg.norm(xh - x32) <= 1e-8 * np.linalg.norm(x32)
    assert ih["iterations"] <= 1.1 * i32["iterations"] + 5


@pytest.mark.parametrize("mesh", ["square_8x8.msh", "square_16x16.msh"])
def test_multigrid_reference_bc_golden(vemd, oracle, mesh):
    """The reference's boundary condition (essential normal flux and a
    mean-zero pressure multiplier; singular pressure Schur approximation) with
    the multigrid cycle on S + rho L L^T: on the reference's meshes the FP64
    solve matches the oracle's direct solve to 1e-10, and the mixed solve
    reaches 1e-8."""
    s, op = _pair(vemd, oracle, mesh, 1, 0.05, 0)
    xo = op.solve_direct()
    x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12, precond=vemd.PC_MULTIGRID))
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo), info
    xm, im = s.solve(vemd.SolveOptions(precision=vemd.MIXED, rtol=1e-8, precond=vemd.PC_MULTIGRID))
    assert im["rel_residual"] <= 1e-8
    assert np.linalg.norm(xm - xo) <= 1e-5 * np.linalg.norm(xo), im


def test_multigrid_reference_bc_iterations(vemd, oracle):
    """Reference boundary condition on generated disks of 1k, 4k and 16k
    cells (k = 1, permeability field): the multigrid-preconditioned FP64
    solve matches the oracle's direct solve to 1e-10 (1k, 4k), its iteration
    count does not grow with the mesh the way the diagonal preconditioner's
    does, and it needs a small fraction of those iterations."""
    its, jac = {}, {}
    for cells in (1000, 4000, 16000):
        s, op = _gen_pair(vemd, oracle, "disk", cells, 1, 0, True)
        assert s.ntot == s.nflux + s.npres + 1
        x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-10, precond=vemd.PC_MULTIGRID))
        assert info["rel_residual"] <= 1e-10
        if cells <= 4000:
            xo = op.solve_direct()
            assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo), (cells, info)
        its[cells] = info["iterations"]
        _, ij = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-10, precond=vemd.PC_JACOBI))
        jac[cells] = ij["iterations"]
    print("iterations multigrid", its, "jacobi", jac, flush=True)
    assert its[16000] <= 1.5 * its[1000]
    assert jac[16000] >= 2.5 * jac[1000]
    assert 4 * its[16000] <= jac[16000]


@pytest.mark.parametrize("case", GEN, ids=GEN_IDS)
def test_split_apply_sums_to_apply(vemd, oracle, case):
    """The solver's one-launch form: the two cell sides of every edge dof,
    gathered through the side table, sum to the complete apply bitwise (same
    single addition); every cell side of an edge appears in exactly one slot."""
    import torch
    s, _ = _gen_pair(vemd, oracle, *case)
    L = vemd.lib()
    import ctypes
    nside, t = s.edge_sides()
    nedge, kk = s.mesh.num_edges, s.k + 1
    # side 0 of every edge is a real slot, slots are distinct, the zero slot is last
    assert t.min() >= 0 and t.max() == nside - 1 and (t[0::2] < nside - 1).all()
    real = t[t < nside - 1]
    assert len(np.unique(real)) == len(real)
    tt = torch.from_numpy(t.astype(np.int64)).cuda()
    for prec, dt in ((vemd.FP64, torch.float64), (vemd.FP32, torch.float32)):
        g = torch.Generator(device="cuda").manual_seed(3)
        x = torch.randn(s.nvec, dtype=dt, device="cuda", generator=g)
        y, y0 = torch.empty_like(x), torch.full_like(x, 7.0)
        h = torch.zeros(nside * kk, dtype=dt, device="cuda")
        p = [ctypes.c_void_p(v.data_ptr()) for v in (x, y, y0, h)]
        assert L.vemd_apply_device(s.h, prec, p[0], p[1], None) == 0
        assert L.vemd_apply_device_split(s.h, prec, p[0], p[2], p[3], None) == 0
        torch.cuda.synchronize()
        hv = h.view(nside, kk)
        edge = (hv[tt[0::2]] + hv[tt[1::2]]).reshape(-1)
        ne = kk * nedge
        assert torch.equal(y[:ne], edge)
        assert torch.equal(y0[:ne], torch.full_like(y0[:ne], 7.0))  # edge block untouched
        # cell-owned and multiplier rows; the padding lanes' slot of k + 1
        # entries before the multiplier is no row of K
        tr = s.nvec - kk - (1 if s.bc == vemd.BC_FLUX_MEANZERO else 0)
        keep = torch.ones(s.nvec, dtype=torch.bool, device="cuda")
        keep[:ne] = False
        keep[tr:tr + kk] = False
        assert torch.equal(y[keep], y0[keep])


GOLD = [c for c in json.load(open(golden("golden_errors.json")))["cases"] if c["k"] <= 1]


@pytest.mark.parametrize("case", GOLD, ids=lambda c: "%s-ex%d" % (c["mesh"][:-4], c["example"]))
def test_reference_errors_on_device(vemd, case):
    print("golden", case["mesh"], case["example"], flush=True)
    s = vemd.Solver(vemd.Mesh.read(golden(case["mesh"])), k=case["k"], nq2d=case["nq2"], nq1d=case["nq1"],
                    eps=0.05 if case["example"] == 1 else 0.0, device=0)
    x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12))
    e = s.errors(x, case["nq2"])
    for key in ("p_comp", "u_comp"):
        v = case[key]
        if case["digits"] >= 16:  # printed from runs solved to ~1e-11
            tol = 5e-10 * v
        else:  # half a unit of the 7th digit plus the printing run's rtol 1e-10
            tol = 0.5 * 10.0 ** (np.floor(np.log10(v)) - 6) + 1e-7 * v
        assert abs(e[key] - v) <= tol, (key, e[key], v)
    if "phys" in case:
        check_phys(e, case["phys"])


def test_apply_autotuner(vemd, oracle, monkeypatch):
    """The autotuner picks a variant per valence present in the mesh, and the
    variants only reshape the operator stream's consumption: the apply with
    the tuned variants equals the apply with variant 0 everywhere, bitwise."""
    s, _ = _gen_pair(vemd, oracle, "annulus", 12000, 1, 1, True)
    t = s.tuning()
    print("apply variants (fp64, fp32) per valence", t, flush=True)
    a = s.mesh.arrays()
    present = set(int(v) for v in np.unique(np.diff(a["cell_off"])))
    assert set(t) == present
    assert all(0 <= v < 4 for pair in t.values() for v in pair)
    monkeypatch.setenv("VEMD_AUTOTUNE", "0")
    s0, _ = _gen_pair(vemd, oracle, "annulus", 12000, 1, 1, True)
    assert all(pair == (0, 0) for pair in s0.tuning().values())
    x = np.random.default_rng(4).standard_normal(s.ntot)
    for prec in (vemd.FP64, vemd.FP32):
        assert np.array_equal(s.apply(x, prec), s0.apply(x, prec))


def test_cpp_driver_reports_reference_errors():
    """The C++ drop-in (include/vemd.hpp: the reference's class names and call
    order) through its driver vemd_darcy, example 1, k = 1, on square_2x2 and
    square_4x4: the computational and physical errors, the divergence error
    and the mass it prints against the reference reports' 16-digit values."""
    import os
    import subprocess
    cases = {}
    for c in PHYS:
        if c["example"] == 1 and c["mesh"] in ("square_2x2.msh", "square_4x4.msh"):
            cases.setdefault(c["mesh"], c)
    meshes = sorted(cases)
    exe = os.path.join(ROOT, "wsl-code_amd", "vemd_darcy")
    print("running", exe, meshes, flush=True)
    out = subprocess.run([exe, "--example", "1", "--k", "1", "--report"] + [golden(m) for m in meshes],
                         capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    reports = [ln.split() for ln in out.stdout.splitlines() if ln.startswith("report ")]
    assert len(reports) == len(meshes), out.stdout
    for m, r in zip(meshes, reports):
        f = dict(zip(r[1::2], map(float, r[2::2])))
        c = cases[m]
        e = dict(p_comp=f["scalar_comp"], u_comp=f["flux_comp"], p_phys=f["scalar_phys"], u_phys=f["flux_phys"],
                 div_phys=f["div_phys"], mass=f["mass"])
        for key in ("p_comp", "u_comp"):
            assert abs(e[key] - c[key]) <= 5e-10 * c[key], (m, key, e[key], c[key])
        check_phys(e, c["phys"])
