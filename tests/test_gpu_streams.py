"""Stream ordering of the device-resident entry points and recovery of the
multi-right-hand-side state after a failed set-up (C ABI).

* A solve given the caller's stream (vemd_solve_opts.stream) waits for work
  queued there before it reads b_dev, and work queued there after the call
  sees its solution: torch writes b on a side stream behind a long device
  spin, with no synchronisation, and the solution must be the one of the
  synchronised solve (bitwise: the solve is deterministic).
* Calls that share a solver's device scratch run one after the other: an
  apply queued on a side stream behind a spin, followed at once by a solve on
  the library's own stream (mean-zero problem: both use the multiplier-row
  partials), leaves both results as they are alone, and the solve returns only
  after the apply has run.
* A multi-right-hand-side set-up that fails after its first allocations (a
  lowered group cap) leaves no half-built state: the next call fails the same
  way, and with the cap lifted the state is built afresh and applies."""
import ctypes
import time

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

SPIN = 400_000_000  # device cycles of the side stream's spin (~0.2 s)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _spin(torch):
    torch.cuda._sleep(SPIN)


@pytest.fixture(scope="module")
def torch_mod(vemd):
    import torch
    assert vemd.device_count() > 0, "no HIP device"
    return torch


def _spin_seconds(torch):
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    t = time.perf_counter()
    with torch.cuda.stream(side):
        _spin(torch)
    side.synchronize()
    return time.perf_counter() - t


def test_device_solve_waits_for_the_callers_stream(vemd, torch_mod):
    torch = torch_mod
    L = vemd.lib()
    print("stream order: solve on b written behind a spin on the caller's stream", flush=True)
    m = vemd.generate_mesh(3000, domain=vemd.DOMAIN_ANNULUS, r_inner=0.5, lloyd_iters=20, seed=51,
                           min_edge_frac=0.05)
    s = vemd.Solver(m, k=1, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P, device=0)
    bi = np.zeros(s.nvec)
    L.vemd_to_internal(s.h, _dp(s.rhs()), _dp(bi), 1)
    b_ready = torch.from_numpy(bi).cuda()
    x_ref = torch.zeros_like(b_ready)
    torch.cuda.synchronize()
    o = vemd.SolveOptions(precision=vemd.FP64, rtol=1e-10).c()
    info = vemd._lib.SolveInfo()
    assert L.vemd_solve_device(s.h, ctypes.byref(o), _ptr(b_ready), _ptr(x_ref), ctypes.byref(info)) == 0
    assert info.iterations > 0 and x_ref.abs().max().item() > 0
    side = torch.cuda.Stream()
    b = torch.zeros_like(b_ready)
    x = torch.zeros_like(b_ready)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        _spin(torch)
        b.copy_(b_ready)
    o.stream = ctypes.c_void_p(side.cuda_stream)
    assert L.vemd_solve_device(s.h, ctypes.byref(o), _ptr(b), _ptr(x), ctypes.byref(info)) == 0, \
        L.vemd_last_error().decode()
    with torch.cuda.stream(side):
        y = x.clone()  # queued after the call on the caller's stream
    side.synchronize()
    assert torch.equal(y, x_ref)


def test_apply_then_solve_share_scratch_in_order(vemd, torch_mod):
    torch = torch_mod
    L = vemd.lib()
    print("stream order: apply on a side stream, then a solve at once (mean-zero problem)", flush=True)
    s = vemd.Solver(vemd.Mesh.read(golden("square_8x8.msh")), k=1, eps=0.05, device=0)
    assert s.ntot == s.nflux + s.npres + 1  # the multiplier row: its partials are shared scratch
    g = torch.Generator(device="cuda").manual_seed(3)
    xin = torch.randn(s.nvec, dtype=torch.float64, device="cuda", generator=g)
    y_ref = torch.empty_like(xin)
    assert L.vemd_apply_device(s.h, vemd.FP64, _ptr(xin), _ptr(y_ref), None) == 0
    torch.cuda.synchronize()
    x_sol_ref, _ = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-10))
    spin = _spin_seconds(torch)
    side = torch.cuda.Stream()
    y = torch.zeros_like(xin)
    torch.cuda.synchronize()
    t = time.perf_counter()
    with torch.cuda.stream(side):
        _spin(torch)
    assert L.vemd_apply_device(s.h, vemd.FP64, _ptr(xin), _ptr(y), ctypes.c_void_p(side.cuda_stream)) == 0
    x_sol, _ = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-10))
    waited = time.perf_counter() - t
    side.synchronize()
    print("spin %.3f s, apply + solve returned after %.3f s" % (spin, waited), flush=True)
    assert waited >= 0.8 * spin  # the solve ran after the queued apply
    assert torch.equal(y, y_ref)
    assert np.array_equal(x_sol, x_sol_ref)


def test_failed_multi_setup_leaves_no_state(vemd, torch_mod, monkeypatch):
    torch = torch_mod
    L = vemd.lib()
    print("multi-RHS set-up failing after its first allocations, then recovering", flush=True)
    m = vemd.generate_mesh(2000, domain=vemd.DOMAIN_DISK, lloyd_iters=20, seed=52, min_edge_frac=0.05)
    s = vemd.Solver(m, k=1, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P, device=0)
    assert s.ngroups > 1
    X = torch.randn(s.nvec, 32, dtype=torch.float32, device="cuda")
    Y, Y1 = torch.zeros_like(X), torch.zeros_like(X)
    monkeypatch.setenv("VEMD_MR_MAX_GROUPS", "1")
    for _ in range(2):
        rc = L.vemd_apply_multi_device(s.h, _ptr(X), _ptr(Y), _ptr(Y1), None)
        assert rc == 1 and "too many valence groups" in L.vemd_last_error().decode()
    monkeypatch.delenv("VEMD_MR_MAX_GROUPS")
    assert L.vemd_apply_multi_device(s.h, _ptr(X), _ptr(Y), _ptr(Y1), None) == 0
    torch.cuda.synchronize()
    # the rebuilt state applies like the single-column FP32 apply (column 0)
    x0 = X[:, 0].contiguous()
    y0 = torch.empty_like(x0)
    assert L.vemd_apply_device(s.h, vemd.FP32, _ptr(x0), _ptr(y0), None) == 0
    torch.cuda.synchronize()
    ne = 2 * m.num_edges
    ym = Y[:, 0].clone()
    ym[:ne] += Y1[:ne, 0]
    ones, live = np.ones(s.ntot), np.zeros(s.nvec)
    L.vemd_to_internal(s.h, _dp(ones), _dp(live), 0)
    live = torch.from_numpy(live != 0).cuda()
    scale = y0[live].abs().max().item()
    assert (ym[live] - y0[live]).abs().max().item() <= 1e-4 * scale
