"""The multigrid values rebuilt on the device after a theta change.

vemd_set_theta on an uploaded solver recomputes the hierarchy's values on the
GPU (the Schur approximation for the new flux diagonal, inverse diagonals,
power-iteration spectral bounds, smoothed prolongators, restrictions, R A and
the Galerkin operators, the rank-one weight of the mean-zero problem; the
host inverts only the dense coarsest matrix) with the host build's operation
order and no fused multiply-adds.  Checks:

* VEMD_MG_VERIFY=1 rebuilds every value on the host as well and requires each
  one (operators, diagonals, power-iteration vectors, spectral bounds, P, R,
  R A, the rank-one weight, every entry of the coarsest inverse) to be
  bitwise equal: k = 0, 1 and 2, Dirichlet pressure on a curved annulus and
  the reference's mean-zero problem on a reference square, three successive
  theta fields each;
* solves after a device rebuild equal, bitwise, the solves of a twin solver
  whose hierarchy was rebuilt on the host (VEMD_MG_DEVICE_REBUILD=0), FP64
  and mixed precision;
* a solver uploaded again after device rebuilds (its host values are then
  rebuilt on the host first) solves to the oracle's direct solution."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _annulus(vemd, n=3000, seed=41):
    return vemd.generate_mesh(n, domain=vemd.DOMAIN_ANNULUS, r_inner=0.5, lloyd_iters=20, seed=seed,
                              min_edge_frac=0.05)


CASES = [("annulus", 1, 1), ("annulus", 0, 1), ("square", 1, 0), ("square", 0, 0), ("annulus", 2, 1), ("square", 2, 0)]


def _case(vemd, name, k, bc):
    if name == "annulus":
        m = _annulus(vemd)
        kw = dict(k=k, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=bc)
    else:
        m = vemd.Mesh.read(golden("square_16x16.msh"))
        kw = dict(k=k, bc=bc)
    return m, kw


@pytest.mark.parametrize("name,k,bc", CASES, ids=[f"{c[0]}-k{c[1]}-bc{c[2]}" for c in CASES])
def test_device_rebuild_is_bitwise_the_host_build(vemd, monkeypatch, name, k, bc):
    print("device multigrid rebuild vs host:", name, "k", k, "bc", bc, flush=True)
    m, kw = _case(vemd, name, k, bc)
    rng = np.random.default_rng(7 + k)
    ths = [rng.standard_normal(m.num_cells) for _ in range(3)]
    s = vemd.Solver(m, theta=ths[0], device=0, **kw)
    monkeypatch.setenv("VEMD_MG_VERIFY", "1")
    for th in ths[1:] + [0.5 * ths[0]]:
        s.set_theta(th)  # raises if any rebuilt value differs from the host's
    x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-10))
    assert info["rel_residual"] <= 1e-10


SOLVE_CASES = CASES[:3] + CASES[4:5]


@pytest.mark.parametrize("name,k,bc", SOLVE_CASES, ids=[f"{c[0]}-k{c[1]}-bc{c[2]}" for c in SOLVE_CASES])
def test_solves_after_device_and_host_rebuilds_are_identical(vemd, monkeypatch, name, k, bc):
    print("device vs host rebuild, solves:", name, "k", k, "bc", bc, flush=True)
    m, kw = _case(vemd, name, k, bc)
    rng = np.random.default_rng(11)
    t0, t1, t2 = (rng.standard_normal(m.num_cells) for _ in range(3))
    dev = vemd.Solver(m, theta=t0, device=0, **kw)
    host = vemd.Solver(m, theta=t0, device=0, **kw)
    for th in (t1, t2):
        dev.set_theta(th)
        monkeypatch.setenv("VEMD_MG_DEVICE_REBUILD", "0")
        host.set_theta(th)
        monkeypatch.delenv("VEMD_MG_DEVICE_REBUILD")
        for prec, rtol in ((vemd.FP64, 1e-11), (vemd.MIXED, 1e-9)):
            xd, idev = dev.solve(vemd.SolveOptions(precision=prec, rtol=rtol))
            xh, ihost = host.solve(vemd.SolveOptions(precision=prec, rtol=rtol))
            assert idev["iterations"] == ihost["iterations"], (idev, ihost)
            assert np.array_equal(xd, xh)


def test_upload_after_device_rebuilds(vemd, oracle):
    print("re-upload after device rebuilds", flush=True)
    m = _annulus(vemd, 2000, seed=5)
    a = m.arrays()
    rng = np.random.default_rng(3)
    t0, t1 = rng.standard_normal(m.num_cells), rng.standard_normal(m.num_cells)
    kw = dict(k=1, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=1)
    s = vemd.Solver(m, theta=t0, device=0, **kw)
    s.set_theta(t1)
    s.upload(0)  # the host hierarchy is rebuilt from the current theta before the upload
    x, info = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12))
    om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
    xo = oracle.Problem(om, k=1, map_kind=oracle.MAP_ARCS, sol=oracle.SOL_SINSIN, bc=1, theta=t1).solve_direct()
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
