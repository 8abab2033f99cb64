"""Element setup on the GPU (vemd_solver_create_device) against the oracle
and the host setup: every cell's signed A_E, W_E, L_E, f_E, the assembled
right-hand side (boundary data included) and the projector (through the flux
error), on straight, sinusoidally mapped and curved-arc meshes, k = 0, 1 and
2, both boundary conditions.  The device threads follow the host arithmetic step
by step, so the two agree to the rounding of fused multiply-adds and the math
library: relative 1e-11 per cell; the oracle computes the same quantities
in its own order, and the device values meet it to 1e-11 too."""
import numpy as np
import pytest

from conftest import golden, rosette

pytestmark = pytest.mark.gpu

CASES = [
    # (mesh, k, map_kind, eps, solution, bc)
    ("square_8x8.msh", 1, "sin", 0.05, "sincos", "flux"),
    ("square_4x4.msh", 0, "sin", 0.05, "sincos", "dirichlet"),
    ("disk1500", 0, "arcs", 0.0, "sinsin", "dirichlet"),
    ("annulus2500", 1, "arcs", 0.0, "sinsin", "dirichlet"),
    ("annulus2500", 1, "arcs", 0.0, "sincos", "flux"),
    ("rosette", 1, "sin", 0.05, "sincos", "flux"),  # a 10-gon: the largest polygon of the GPU setup
    ("rosette", 0, "arcs", 0.0, "sinsin", "dirichlet"),
    ("square_4x4.msh", 2, "sin", 0.05, "sincos", "flux"),
    ("disk1500", 2, "arcs", 0.0, "sinsin", "dirichlet"),
    ("rosette", 2, "arcs", 0.0, "sinsin", "dirichlet"),
]
IDS = ["%s-k%d-%s-%s" % (c[0], c[1], c[2], c[5]) for c in CASES]


def _mesh(vemd, name):
    if name == "rosette":
        return vemd.Mesh.from_arrays(*rosette())
    if name.endswith(".msh"):
        return vemd.Mesh.read(golden(name))
    if name.startswith("disk"):
        return vemd.generate_mesh(int(name[4:]), domain=vemd.DOMAIN_DISK, lloyd_iters=20, seed=13,
                                  min_edge_frac=0.05)
    return vemd.generate_mesh(int(name[7:]), domain=vemd.DOMAIN_ANNULUS, r_inner=0.5, lloyd_iters=20, seed=14,
                              min_edge_frac=0.05)


def _pair(vemd, case):
    name, k, mk, eps, sol, bc = case
    print("setup case", case, flush=True)
    assert vemd.device_count() > 0, "no HIP device"
    m = _mesh(vemd, name)
    rng = np.random.default_rng(3)
    kw = dict(k=k, map_kind=vemd.MAP_SIN if mk == "sin" else vemd.MAP_ARCS, eps=eps,
              solution=vemd.SOL_SINCOS if sol == "sincos" else vemd.SOL_SINSIN,
              bc=vemd.BC_FLUX_MEANZERO if bc == "flux" else vemd.BC_DIRICHLET_P,
              theta=rng.standard_normal(m.num_cells))
    return m, vemd.Solver(m, **kw), vemd.Solver(m, setup_device=0, **kw)


def _oracle_problem(oracle, m, case, theta):
    name, k, mk, eps, sol, bc = case
    if name.endswith(".msh"):
        om = oracle.Mesh.read(golden(name))
    else:
        a = m.arrays()
        om = oracle.Mesh.from_arrays(a["xy"], a["cell_off"], a["cell_nodes"], a["arc_nodes"], a["arc_radius"])
    return oracle.Problem(om, k=k, map_kind=oracle.MAP_SIN if mk == "sin" else oracle.MAP_ARCS, eps=eps,
                          sol=oracle.SOL_SINCOS if sol == "sincos" else oracle.SOL_SINSIN,
                          bc=oracle.BC_FLUX_MEANZERO if bc == "flux" else oracle.BC_DIRICHLET_P, theta=theta)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_device_setup_matches_oracle(vemd, oracle, case):
    m, _, sd = _pair(vemd, case)
    theta = np.random.default_rng(3).standard_normal(m.num_cells)  # the permeability _pair used
    op = _oracle_problem(oracle, m, case, theta)
    worst = 0.0
    for c in range(m.num_cells):
        ed = sd.element(c)
        n = ed["n"]
        d = op.element(c, 13)[:n]
        ref = dict(A=op.element(c, 7, (n, n)) * np.outer(d, d) * np.exp(theta[c]),  # oracle A carries exp(-theta)
                   W=op.element(c, 2).reshape(-1, n) * d, L=op.element(c, 8), f=op.element(c, 9))
        for key, v in ref.items():
            scale = max(np.abs(v).max(), 1e-300)
            worst = max(worst, np.abs(ed[key] - v).max() / scale)
        assert np.array_equal(ed["sign"], d) and np.array_equal(ed["gdof"], op.element(c, 12)[:n].astype(int))
    print("largest relative difference from the oracle's element data", worst, flush=True)
    assert worst <= 1e-11
    _, rr = op.csr(True)
    fi, ub = op.free()
    b = sd.rhs()
    assert np.abs(b[fi] - rr).max() <= 1e-11 * np.abs(rr).max()


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_device_setup_matches_host(vemd, case):
    m, sh, sd = _pair(vemd, case)
    assert (sh.ntot, sh.nvec) == (sd.ntot, sd.nvec)
    worst = 0.0
    for c in range(m.num_cells):
        eh, ed = sh.element(c), sd.element(c)
        for key in ("A", "W", "L", "f"):
            scale = max(np.abs(eh[key]).max(), 1e-300)
            worst = max(worst, np.abs(eh[key] - ed[key]).max() / scale)
        assert np.array_equal(eh["gdof"], ed["gdof"]) and np.array_equal(eh["sign"], ed["sign"])
    print("largest relative difference of the element data", worst, flush=True)
    assert worst <= 1e-11
    bh, bd = sh.rhs(), sd.rhs()
    assert np.abs(bh - bd).max() <= 1e-11 * np.abs(bh).max()
    # the projector enters the flux error of a fixed solution vector
    x = np.random.default_rng(5).standard_normal(sh.ntot)
    eh, ed = sh.errors(x), sd.errors(x)
    for key in ("p_comp", "u_comp", "p_phys", "u_phys"):
        assert abs(eh[key] - ed[key]) <= 1e-11 * abs(eh[key]), key


def test_device_setup_solves_like_host(vemd):
    case = ("annulus2500", 1, "arcs", 0.0, "sinsin", "dirichlet")
    m, sh, sd = _pair(vemd, case)
    sh.upload(0)
    sd.upload(0)
    o = vemd.SolveOptions(precision=vemd.FP64, rtol=1e-12)
    xh, ih = sh.solve(o)
    xd, idv = sd.solve(o)
    assert np.linalg.norm(xh - xd) <= 1e-9 * np.linalg.norm(xh)


def test_device_setup_rejects_k3(vemd):
    m = vemd.Mesh.read(golden("square_2x2.msh"))
    with pytest.raises(RuntimeError, match="k = 0, 1 and 2"):
        vemd.Solver(m, k=3, nq2d=12, setup_device=0)
