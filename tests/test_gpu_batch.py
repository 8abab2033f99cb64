"""Batched solves: several independent problems on meshes of uneven size in
one operator store and one captured Krylov loop, each with its own
recurrences, stopping test and refinement.  Every instance must agree with the
oracle's direct solve of that instance's mesh (FP64 to 1e-10; mixed to 1e-5
with error functionals within 1e-9) and with a solve of that instance alone,
under Dirichlet pressure and under the reference's mean-zero condition."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES, SEEDS = [400, 1500, 3000, 6000], [31, 32, 33, 34]


@pytest.fixture(scope="module")
def batch(vemd):
    from wsl_code_amd.fields import grf_theta
    assert vemd.device_count() > 0, "no HIP device"
    print("batch of", SIZES, "and the 10-gon rosette", flush=True)
    meshes, thetas = [], []
    for n, sd in zip(SIZES, SEEDS):
        m = vemd.generate_mesh(n, domain=vemd.DOMAIN_DISK, lloyd_iters=20, seed=sd, min_edge_frac=0.05)
        a = m.arrays()
        sigma, ell = 0.5 + 1.5 * (sd % 4) / 3.0, 0.02 + 0.08 * (sd % 3) / 2.0
        meshes.append(m)
        thetas.append(grf_theta(a["cx"], a["cy"], sigma=sigma, ell=ell, seed=sd + 100))
    # a 10-gon ringed by triangles: a valence group of its own in the shared store
    from conftest import rosette
    m = vemd.Mesh.from_arrays(*rosette())
    a = m.arrays()
    meshes.append(m)
    thetas.append(grf_theta(a["cx"], a["cy"], sigma=1.0, ell=0.1, seed=135))
    u = vemd.Mesh.concat(meshes)
    kw = dict(k=1, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINSIN, bc=vemd.BC_DIRICHLET_P)
    sb = vemd.Solver(u, theta=np.concatenate(thetas), device=0, **kw)
    singles = [vemd.Solver(m, theta=t, device=0, **kw) for m, t in zip(meshes, thetas)]
    return sb, singles, thetas


@pytest.fixture(scope="module")
def direct(batch, oracle):
    """The oracle's direct solution of every instance on its own mesh."""
    _, singles, thetas = batch
    out = []
    for s, th in zip(singles, thetas):
        a = s.mesh.arrays()
        om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
        op = oracle.Problem(om, k=1, map_kind=oracle.MAP_ARCS, sol=oracle.SOL_SINSIN, bc=oracle.BC_DIRICHLET_P,
                            theta=th)
        out.append((op.solve_direct(), op.errors(9)))
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_matches_oracle_direct_solves(vemd, batch, direct, prec):
    sb, singles, _ = batch
    p = vemd.FP64 if prec == "fp64" else vemd.MIXED
    xb, info = sb.solve(vemd.SolveOptions(precision=p, rtol=1e-12 if prec == "fp64" else 1e-10))
    print("batch vs oracle", prec, info, flush=True)
    for xi, s, (xo, eo) in zip(sb.instance_vectors(xb), singles, direct):
        rel = np.linalg.norm(xi - xo) / np.linalg.norm(xo)
        e = s.errors(xi)
        print("instance", s.mesh.num_cells, "cells: |x - x_oracle| / |x_oracle| =", rel, flush=True)
        if prec == "fp64":
            assert rel <= 1e-10
        else:
            assert rel <= 1e-5
        for key in ("p_comp", "u_comp", "p_phys", "u_phys", "div_phys"):
            assert abs(e[key] - eo[key]) <= 1e-9 * max(1.0, eo[key]), (key, e[key], eo[key])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_matches_single_solves(vemd, batch, prec):
    sb, singles, _ = batch
    p = vemd.FP64 if prec == "fp64" else vemd.MIXED
    xb, info = sb.solve(vemd.SolveOptions(precision=p, rtol=1e-10))
    per = sb.batch_info()
    print("batch", info, per, flush=True)
    assert len(per) == len(singles)
    assert all(r["rel_residual"] <= 1e-10 for r in per)
    for xi, s, r in zip(sb.instance_vectors(xb), singles, per):
        xs, i1 = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12))
        rel = np.linalg.norm(xi - xs) / np.linalg.norm(xs)
        print("instance", s.mesh.num_cells, "cells: its", r["iterations"], "alone", i1["iterations"], "diff", rel,
              flush=True)
        assert rel <= 1e-7
        # the error functionals of each instance agree with its own solve
        eb, es = s.errors(xi), s.errors(xs)
        assert abs(eb["u_phys"] - es["u_phys"]) <= 1e-9 * max(1.0, es["u_phys"])


def test_batch_of_one_equals_plain_solve(vemd, batch):
    _, singles, _ = batch
    s = singles[1]
    x1, i1 = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-10))
    x2, i2 = s.solve(vemd.SolveOptions(precision=vemd.FP64, rtol=1e-10))
    assert i1["iterations"] == i2["iterations"] and np.array_equal(x1, x2)  # deterministic
    assert s.batch_info()[0]["iterations"] == i2["iterations"]


@pytest.mark.parametrize("pc", ["multigrid", "jacobi"])
def test_meanzero_batch_matches_oracle(vemd, oracle, pc):
    """The reference's own condition in a batch (essential normal flux and one
    mean-zero pressure multiplier per instance; the multigrid cycle then
    carries one rank-one term per instance): every instance of a batch of
    reference squares and curved disks and annuli, each with its own theta,
    equals the oracle's direct solve of that instance (FP64 to 1e-10, mixed to
    1e-7), multiplier included."""
    from conftest import golden
    print("mean-zero batch,", pc, flush=True)
    parts = [vemd.Mesh.read(golden("square_8x8.msh")),
             vemd.generate_mesh(1500, domain=vemd.DOMAIN_DISK, lloyd_iters=20, seed=71, min_edge_frac=0.05),
             vemd.Mesh.read(golden("square_16x16.msh")),
             vemd.generate_mesh(2500, domain=vemd.DOMAIN_ANNULUS, r_inner=0.5, lloyd_iters=20, seed=72,
                                min_edge_frac=0.05)]
    rng = np.random.default_rng(73)
    thetas = [rng.standard_normal(m.num_cells) for m in parts]
    kw = dict(k=1, map_kind=vemd.MAP_ARCS, solution=vemd.SOL_SINCOS, bc=vemd.BC_FLUX_MEANZERO)
    sb = vemd.Solver(vemd.Mesh.concat(parts), theta=np.concatenate(thetas), device=0, **kw)
    ref = []
    for m, th in zip(parts, thetas):
        a = m.arrays()
        om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
        ref.append(oracle.Problem(om, k=1, map_kind=oracle.MAP_ARCS, sol=oracle.SOL_SINCOS,
                                  bc=oracle.BC_FLUX_MEANZERO, theta=th).solve_direct())
    p = vemd.PC_MULTIGRID if pc == "multigrid" else vemd.PC_JACOBI
    for prec, rtol, tol in ((vemd.FP64, 1e-12, 1e-10), (vemd.MIXED, 1e-10, 1e-7)):
        xb, info = sb.solve(vemd.SolveOptions(precision=prec, rtol=rtol, precond=p))
        per = sb.batch_info()
        print("  precision", prec, info, [r["iterations"] for r in per], flush=True)
        assert all(r["rel_residual"] <= rtol for r in per), per
        for xi, xo in zip(sb.instance_vectors(xb), ref):
            assert len(xi) == len(xo)
            assert np.linalg.norm(xi - xo) <= tol * np.linalg.norm(xo)
            assert abs(xi[-1] - xo[-1]) <= tol * np.linalg.norm(xo)
