"""Pins the CPU oracle against the reference's published error functionals and
against identities the mathematics fixes (closed forms, projector algebra)."""
import json
import math

import numpy as np
import pytest

from conftest import PHYS, check_phys, golden

CASES = json.load(open(golden("golden_errors.json")))["cases"]


def _tol(case, value):
    if case["digits"] >= 16:
        # printed to 16 digits by runs solved to ~1e-11: agreement is limited by
        # that run's solve tolerance, not by the printed digits
        return 5e-10 * abs(value)
    # half a unit of the last printed digit, plus 1e-7 relative for the
    # printing run's own GMRES stop at rtol 1e-10 (its k=3 CPU and GPU runs
    # differ by up to 2e-6 relative: CPU/results_darcy_convergence_k3_cpu_vs_gpu.txt:28)
    return 0.5 * 10.0 ** (math.floor(math.log10(abs(value))) - case["digits"] + 1) + 1e-7 * abs(value)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-ex%d-k%d-%s" % (c["mesh"][:-4], c["example"], c["k"],
                                                                       c["source"].split("/")[-1][:18]))
def test_golden_errors(oracle, case):
    m = oracle.Mesh.read(golden(case["mesh"]))
    p = oracle.Problem(m, k=case["k"], nq2=case["nq2"], nq1=case["nq1"],
                       eps=0.05 if case["example"] == 1 else 0.0)
    assert p.sizes()["ntot"] == case["dofs"]
    p.solve_direct()
    e = p.errors(case["nq2"])
    assert abs(e["p_comp"] - case["p_comp"]) <= _tol(case, case["p_comp"]), (e, case)
    assert abs(e["u_comp"] - case["u_comp"]) <= _tol(case, case["u_comp"]), (e, case)


@pytest.mark.parametrize("case", PHYS, ids=lambda c: "%s-ex%d-%s" % (c["mesh"][:-4], c["example"],
                                                                   c["phys"]["source"].split("/")[-2][:18]))
def test_golden_physical_errors(oracle, case):
    m = oracle.Mesh.read(golden(case["mesh"]))
    p = oracle.Problem(m, k=case["k"], nq2=case["nq2"], nq1=case["nq1"],
                       eps=0.05 if case["example"] == 1 else 0.0)
    p.solve_direct()
    check_phys(p.errors(case["nq2"]), case["phys"])


def test_divergence_error_depends_on_flux_only(oracle):
    """div u_h = H^{-1} W u uses the flux dofs alone: moving the pressure and
    the multiplier changes the pressure errors and the mass but leaves the
    divergence error and the flux errors as they were."""
    m = oracle.Mesh.read(golden("square_2x2.msh"))
    p = oracle.Problem(m, k=1, nq2=9, nq1=9, eps=0.05)
    p.solve_direct()
    e0 = p.errors(9)
    x = p.solution().copy()
    nf = p.sizes()["nf"]
    rng = np.random.default_rng(4)
    x[nf:] += rng.standard_normal(len(x) - nf)  # pressure moments and multiplier
    p.set_solution(x)
    e1 = p.errors(9)
    for key in ("div_phys", "u_phys", "u_comp"):
        assert e1[key] == e0[key], key
    assert abs(e1["p_phys"] - e0["p_phys"]) > 1e-3 * e0["p_phys"]
    assert abs(e1["mass"]) > 1e-6


@pytest.mark.parametrize("k", [0, 1, 2, 3])
@pytest.mark.parametrize("mesh", ["square_1x1.msh", "square_2x2.msh"])
def test_projector_identities(oracle, k, mesh):
    """B D = G, P D = I, (D P)^2 = D P, A symmetric positive semidefinite,
    and A = P^T G_K P on polynomial dof vectors (consistency)."""
    m = oracle.Mesh.read(golden(mesh))
    p = oracle.Problem(m, k=k, nq2=12 if k == 3 else 9, nq1=9, eps=0.05)
    for cell in range(min(m.nc, 4)):
        info = p.element(cell, 10)
        n, pp, q = int(info[2]), int(info[3]), int(info[4])
        D = p.element(cell, 0, (n, pp))
        G = p.element(cell, 1, (pp, pp))
        B = p.element(cell, 3, (pp, n))
        P = p.element(cell, 4, (pp, n))
        Pd = p.element(cell, 5, (n, n))
        A = p.element(cell, 7, (n, n))
        GK = p.element(cell, 6, (pp, pp))
        scale = np.abs(G).max()
        assert np.abs(B @ D - G).max() <= 1e-11 * scale
        assert np.abs(P @ D - np.eye(pp)).max() <= 1e-10
        assert np.abs(Pd @ Pd - Pd).max() <= 1e-10
        assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
        assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > -1e-10 * np.abs(A).max()
        # consistency: on dof vectors of polynomials (v = D c) the stabilisation vanishes
        c = np.random.default_rng(cell).standard_normal(pp)
        v = D @ c
        assert abs(v @ A @ v - c @ GK @ c) <= 1e-10 * abs(c @ GK @ c)


def test_gk_equals_g_for_identity_map(oracle):
    m = oracle.Mesh.read(golden("square_2x2.msh"))
    p = oracle.Problem(m, k=1, eps=0.0)
    for cell in range(m.nc):
        G = p.element(cell, 1, (6, 6))
        GK = p.element(cell, 6, (6, 6))
        assert np.abs(G - GK).max() <= 1e-14 * np.abs(G).max()


def test_w_k0_is_divergence_theorem(oracle):
    """k=0: W is the row of signed edge lengths' moments -- the sum over the
    cell's edges of int_e phi.n equals int_E div phi, i.e. W = [1,...,1]."""
    m = oracle.Mesh.read(golden("square_2x2.msh"))
    p = oracle.Problem(m, k=0, eps=0.05)
    W = p.element(0, 2)
    assert np.allclose(W, 1.0, atol=0, rtol=1e-15)


def test_edge_sign_parity(oracle):
    """Reversed interior edges flip even-degree moments only: (-1)^(r+1)."""
    m = oracle.Mesh.read(golden("square_2x2.msh"))
    p = oracle.Problem(m, k=2, eps=0.05)
    for cell in range(m.nc):
        g = p.element(cell, 12).astype(int)
        d = p.element(cell, 13)
        ne = 4
        for r in range(3):
            for j in range(ne):
                e = g[r * ne + j] - r * m.ne
                if m.edge_occ[e] == 1:
                    assert d[r * ne + j] == 1.0
                elif d[0 * ne + j] == -1.0:
                    assert d[r * ne + j] == (-1.0 if r % 2 == 0 else 1.0)


def test_gmres_ilu_matches_direct(oracle):
    m = oracle.Mesh.read(golden("square_4x4.msh"))
    p = oracle.Problem(m, k=1, eps=0.05)
    xd = p.solve_direct()
    ed = p.errors()
    xg, it, rel = p.solve_gmres_ilu(rtol=1e-12)
    assert rel <= 1e-12 and 0 < it < 2000
    assert np.linalg.norm(xg - xd) <= 1e-8 * np.linalg.norm(xd)
    eg = p.errors()
    assert abs(eg["p_comp"] - ed["p_comp"]) <= 1e-8 * ed["p_comp"]


def test_dirichlet_pressure_converges(oracle):
    """EXTENSION (Dirichlet p, sin*sin solution): second-order convergence of
    pressure and flux for k=1 on the reference's curved map."""
    errs = []
    for n in (2, 4, 8):
        m = oracle.Mesh.read(golden("square_%dx%d.msh" % (n, n)))
        p = oracle.Problem(m, k=1, eps=0.05, sol=oracle.SOL_SINSIN, bc=oracle.BC_DIRICHLET_P)
        p.solve_direct()
        errs.append(p.errors())
    for key in ("p_comp", "u_comp", "p_phys", "u_phys"):
        o1 = math.log2(errs[0][key] / errs[1][key])
        o2 = math.log2(errs[1][key] / errs[2][key])
        assert 1.8 < o2 < 2.3, (key, o1, o2)
