"""Allocation-free apply (vemd_apply_device_static) through the C ABI: the
same results as vemd_apply_device, bit for bit (the same kernels in the same
order, with the side slots and multiplier-row partials in the library's static
device array), on two calls in a row of which the second has the smaller
input -- so the second call finds the first call's values in the shared array
and must clear its zero slot itself -- in FP64 and FP32, with and without the
mean-zero multiplier row; device memory in use does not change across the
calls; a mesh whose side slots do not fit the array is refused."""
import ctypes

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def solvers(vemd):
    import torch
    assert vemd.device_count() > 0, "no HIP device"
    print("allocation-free apply: 6000-cell annulus (mean-zero multiplier), then square_4x4 k=0", flush=True)
    big = vemd.Solver(vemd.generate_mesh(6000, domain=vemd.DOMAIN_ANNULUS, r_inner=0.5, lloyd_iters=20, seed=61,
                                         min_edge_frac=0.05),
                      k=1, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINCOS, bc=vemd.BC_FLUX_MEANZERO,
                      theta=np.random.default_rng(62).standard_normal(6000), device=0)
    small = vemd.Solver(vemd.Mesh.read(golden("square_4x4.msh")), k=0, eps=0.05, bc=vemd.BC_DIRICHLET_P, device=0)
    assert big.nvec > 4 * small.nvec and big.ntot == big.nflux + big.npres + 1
    return torch, big, small


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_static_apply_equals_apply_on_two_calls(vemd, solvers, prec):
    torch, big, small = solvers
    L = vemd.lib()
    p = vemd.FP64 if prec == "fp64" else vemd.FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    g = torch.Generator(device="cuda").manual_seed(63)
    xs = [torch.randn(s.nvec, dtype=dt, device="cuda", generator=g) for s in (big, small)]
    refs = []
    for s, x in zip((big, small), xs):
        y = torch.zeros_like(x)
        assert L.vemd_apply_device(s.h, p, _ptr(x), _ptr(y), None) == 0
        refs.append(y)
    ys = [torch.full_like(x, 7.0) for x in xs]
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for s, x, y in zip((big, small), xs, ys):  # the larger input first, then the smaller one
        rc = L.vemd_apply_device_static(s.h, p, _ptr(x), _ptr(y), None)
        assert rc == 0, L.vemd_last_error().decode()
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    for s, y, r in zip((big, small), ys, refs):
        print(prec, "cells", s.mesh.num_cells, "entries", s.nvec, "largest |y|", r.abs().max().item(), flush=True)
        assert torch.equal(y, r)
    # a third call on the larger input after the smaller one
    y = torch.zeros_like(xs[0])
    assert L.vemd_apply_device_static(big.h, p, _ptr(xs[0]), _ptr(y), None) == 0
    torch.cuda.synchronize()
    assert torch.equal(y, refs[0])


def test_static_apply_refuses_a_mesh_beyond_its_array(vemd, solvers, monkeypatch):
    torch, big, _ = solvers
    L = vemd.lib()
    x = torch.zeros(big.nvec, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    nside, _ = big.edge_sides()
    need = -(-(nside * 2 * 8) // 256) * 256 + 8 * -(-big.mesh.num_cells // 64)  # at least: side slots + tiles
    monkeypatch.setenv("VEMD_STATIC_SCRATCH_LIMIT", str(nside * 2 * 8 - 1))
    assert L.vemd_apply_device_static(big.h, vemd.FP64, _ptr(x), _ptr(y), None) == 1  # VEMD_ERR_ARG
    msg = L.vemd_last_error().decode()
    assert "static scratch" in msg and "use vemd_apply_device" in msg, msg
    monkeypatch.setenv("VEMD_STATIC_SCRATCH_LIMIT", str(64 << 20))
    assert L.vemd_apply_device_static(big.h, vemd.FP64, _ptr(x), _ptr(y), None) == 0
    torch.cuda.synchronize()
    assert need < (64 << 20)
