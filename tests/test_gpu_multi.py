"""Multi-right-hand-side path: the MFMA element apply on 32 interleaved FP32
columns against the oracle's assembled operator and the single-column FP32
apply, and simultaneous solves of 32 right-hand sides (FP32 MFMA MINRES inside
FP64 refinement to 1e-10) against a sparse LU solve of the oracle's system
and one FP64 solve per right-hand side."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NR = 32


@pytest.fixture(scope="module", params=["annulus", "rosette"])
def setup(vemd, oracle, request):
    """A 3000-cell Voronoi annulus with a permeability field, and the 10-gon
    rosette (the largest cell block of the MFMA apply: 26 rows, 348 floats)."""
    import torch
    from conftest import rosette
    from wsl_code_amd.fields import grf_theta
    assert vemd.device_count() > 0, "no HIP device"
    print("multi-rhs", request.param, flush=True)
    if request.param == "annulus":
        m = vemd.generate_mesh(3000, domain=vemd.DOMAIN_ANNULUS, r_inner=0.5, lloyd_iters=20, seed=41,
                               min_edge_frac=0.05)
    else:
        m = vemd.Mesh.from_arrays(*rosette())
    a = m.arrays()
    th = grf_theta(a["cx"], a["cy"], sigma=1.0, ell=0.1, seed=42)
    s = vemd.Solver(m, k=1, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P, theta=th,
                    device=0)
    om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
    op = oracle.Problem(om, k=1, map_kind=oracle.MAP_ARCS, sol=oracle.SOL_SINSIN, bc=oracle.BC_DIRICHLET_P,
                        theta=th)
    A, _ = op.csr(False)  # Dirichlet pressure: nothing eliminated, reference-form rows
    return s, torch, A


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_mfma_apply_matches_oracle_operator(vemd, setup):
    """Every column of the MFMA apply, taken back to the reference order and
    row form, against the oracle's assembled operator times the same column:
    FP32 products with FP32 accumulation, so 5e-5 of the column's largest
    entry (the single-column FP32 apply meets 2e-5 against the same matrix)."""
    s, torch, A = setup
    L = vemd.lib()
    g = torch.Generator(device="cuda").manual_seed(19)
    X = torch.randn(s.nvec, NR, dtype=torch.float32, device="cuda", generator=g)
    Y, Y1 = torch.empty_like(X), torch.zeros_like(X)
    assert L.vemd_apply_multi_device(s.h, _ptr(X), _ptr(Y), _ptr(Y1), None) == 0
    torch.cuda.synchronize()
    ne = (s.k + 1) * s.mesh.num_edges
    Y[:ne] += Y1[:ne]
    Xh, Yh = X.double().cpu().numpy(), Y.double().cpu().numpy()
    worst = 0.0
    for j in range(NR):
        xr, yr = np.zeros(s.ntot), np.zeros(s.ntot)
        L.vemd_to_reference(s.h, _dp(np.ascontiguousarray(Xh[:, j])), _dp(xr), 0)
        L.vemd_to_reference(s.h, _dp(np.ascontiguousarray(Yh[:, j])), _dp(yr), 1)
        y_ref = A @ xr
        err = np.abs(yr - y_ref).max() / np.abs(y_ref).max()
        worst = max(worst, err)
        assert err <= 5e-5, (j, err)
    print("MFMA apply vs oracle operator: largest relative difference", worst, flush=True)


def test_mfma_apply_matches_single_column_apply(vemd, setup):
    s, torch, _ = setup
    L = vemd.lib()
    g = torch.Generator(device="cuda").manual_seed(9)
    X = torch.randn(s.nvec, NR, dtype=torch.float32, device="cuda", generator=g)
    Y, Y1 = torch.empty_like(X), torch.zeros_like(X)
    assert L.vemd_apply_multi_device(s.h, _ptr(X), _ptr(Y), _ptr(Y1), None) == 0
    torch.cuda.synchronize()
    ne = (s.k + 1) * s.mesh.num_edges
    Ym = Y.clone()
    Ym[:ne] += Y1[:ne]
    # slots of real dofs (the MFMA apply leaves padding positions and the trash
    # slot untouched)
    ones = np.ones(s.ntot)
    live = np.zeros(s.nvec)
    L.vemd_to_internal(s.h, ones.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                       live.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0)
    live = torch.from_numpy(live != 0).cuda()
    for j in range(NR):
        # both FP32 products against the FP64 apply of the same column: the
        # MFMA path is as accurate as the single-column FP32 path
        x = X[:, j].contiguous()
        y = torch.empty_like(x)
        assert L.vemd_apply_device(s.h, vemd.FP32, _ptr(x), _ptr(y), None) == 0
        x64 = x.double()
        y64 = torch.empty_like(x64)
        assert L.vemd_apply_device(s.h, vemd.FP64, _ptr(x64), _ptr(y64), None) == 0
        torch.cuda.synchronize()
        y64, y = y64[live], y[live]
        scale = y64.abs().max().item()
        err_single = (y.double() - y64).abs().max().item()
        err_mfma = (Ym[:, j][live].double() - y64).abs().max().item()
        assert err_single <= 1e-4 * scale, (j, err_single, scale)
        assert err_mfma <= 2.0 * err_single + 1e-6 * scale, (j, err_mfma, err_single, scale)


def test_multi_rhs_solve_matches_single_solves(vemd, setup):
    """Every column against a sparse LU solve of the oracle's system with the
    same right-hand side (1e-7: the columns stop at a 1e-10 true residual)
    and against the product's FP64 single solve."""
    import scipy.sparse.linalg as sla
    s, torch, A = setup
    lu = sla.splu(A.tocsc())
    L = vemd.lib()
    rng = np.random.default_rng(43)
    base = s.rhs()
    B_ref = []
    for j in range(NR):
        if j % 2 == 0:
            b = base * (1.0 + 0.1 * j)
        else:  # random pressure sources
            b = np.zeros(s.ntot)
            b[s.nflux:] = rng.standard_normal(s.ntot - s.nflux)
        B_ref.append(b)
    Bi = np.zeros((NR, s.nvec))
    for j in range(NR):
        L.vemd_to_internal(s.h, B_ref[j].ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           Bi[j].ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1)
    Bd = torch.from_numpy(Bi).cuda()
    Xd = torch.zeros_like(Bd)
    o = vemd.SolveOptions(precision=vemd.MIXED, rtol=1e-10, check_every=16).c()
    info = (vemd._lib.SolveInfo * NR)()
    rc = L.vemd_solve_multi_device(s.h, ctypes.byref(o), NR, _ptr(Bd), _ptr(Xd), info)
    assert rc == 0, L.vemd_last_error().decode()
    print("multi-rhs iterations", [i.iterations for i in info][:8], "residuals", max(i.rel_residual for i in info),
          "seconds", info[0].seconds, flush=True)
    assert all(i.rel_residual <= 1e-10 for i in info)
    Xi = Xd.cpu().numpy()
    for j in range(NR):
        xr = np.zeros(s.ntot)
        L.vemd_to_reference(s.h, np.ascontiguousarray(Xi[j]).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            xr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0)
        xo = lu.solve(B_ref[j])
        rel_o = np.linalg.norm(xr - xo) / np.linalg.norm(xo)
        assert rel_o <= 1e-7, (j, rel_o)
        xs, _ = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12), rhs=B_ref[j])
        rel = np.linalg.norm(xr - xs) / np.linalg.norm(xs)
        assert rel <= 1e-7, (j, rel)


def test_multi_rhs_needs_mixed(vemd, setup):
    s, torch, _ = setup
    L = vemd.lib()
    Bd = torch.zeros(1, s.nvec, dtype=torch.float64, device="cuda")
    o = vemd.SolveOptions(precision=vemd.FP64).c()
    info = (vemd._lib.SolveInfo * 1)()
    assert L.vemd_solve_multi_device(s.h, ctypes.byref(o), 1, _ptr(Bd), _ptr(Bd), info) == 1  # VEMD_ERR_ARG
