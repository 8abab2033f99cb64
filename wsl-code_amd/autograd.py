"""torch.autograd binding of the adjoint permeability gradient.

`PressureSampler` evaluates the discrete pressure at interior points (the
observation operator C of an inverse problem); `DarcyPressureAtPoints` is a
torch.autograd.Function mapping the per-cell log-permeability theta to C p_h:
the forward pass is a device solve, the backward pass one adjoint solve with
the same operator (it is symmetric in the solver's internal form) followed by
the per-cell contraction g_E = lambda_E^T (exp(-theta_E) A_E) x_E on the GPU.
"""
import numpy as np

from .solver import SolveOptions
from ._lib import FP64


def _inside(px, py, poly):
    """Point-in-polygon test for a counter-clockwise polygon."""
    x, y = poly[:, 0], poly[:, 1]
    x2, y2 = np.roll(x, -1), np.roll(y, -1)
    cross = (x2 - x) * (py - y) - (y2 - y) * (px - x)
    return bool(np.all(cross >= -1e-14))


class _Locator:
    """Cell containing a point, by nearest centroids and a polygon test;
    -1 outside the mesh, -2 inside a cell with a boundary arc."""

    def __init__(self, arrays):
        from scipy.spatial import cKDTree
        self.a = arrays
        self.arc_nodes = set(map(int, arrays["arc_nodes"].ravel()))
        self.tree = cKDTree(np.stack([arrays["cx"], arrays["cy"]], 1))

    def __call__(self, px, py):
        a = self.a
        off, nodes = a["cell_off"], a["cell_nodes"]
        for c in self.tree.query([px, py], k=min(12, len(a["cx"])))[1]:
            vs = nodes[off[c]:off[c + 1]]
            if _inside(px, py, a["xy"][vs]):
                return -2 if any(int(v) in self.arc_nodes for v in vs) else int(c)
        return -1


class PressureSampler:
    """C: reference-order solution vector -> p_h at the given points.

    Points must lie in cells without boundary arcs, where the cell map is the
    identity; there p_h(x) = sum_r p_r m_r(x) with the scaled monomials of the
    cell (1, (x - xc)/h, (y - yc)/h for k = 1), the pressure basis of
    CPU/lib/polynomial_basis.cpp:92-101."""

    def __init__(self, solver, points):
        import scipy.sparse as sp
        a = solver.mesh.arrays()
        loc = _Locator(a)
        self.points = np.asarray(points, np.float64).reshape(-1, 2)
        q = (solver.k + 1) * (solver.k + 2) // 2
        rows, cols, vals = [], [], []
        self.cells = np.zeros(len(self.points), np.int64)
        for i, (px, py) in enumerate(self.points):
            cell = loc(px, py)
            if cell < 0:
                raise ValueError("point %d (%g, %g) is %s" % (i, px, py, "outside the mesh" if cell == -1
                                                             else "in a boundary cell (curved map)"))
            self.cells[i] = cell
            h = a["diam"][cell]
            mx, my = (px - a["cx"][cell]) / h, (py - a["cy"][cell]) / h
            mono = [1.0, mx, my, mx * mx, mx * my, my * my][:q]
            base = solver.nflux + cell * q
            for r in range(q):
                rows.append(i)
                cols.append(base + r)
                vals.append(mono[r])
        self.C = sp.csr_matrix((vals, (rows, cols)), shape=(len(self.points), solver.ntot))

    def __call__(self, x):
        return self.C @ x


def interior_points(solver, npts, seed, r_min=0.0, r_max=1.0):
    """Seeded points, uniform in the annulus r_min < r < r_max, kept when they
    fall in a cell without boundary arcs."""
    loc = _Locator(solver.mesh.arrays())
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < npts:
        r = np.sqrt(rng.uniform(r_min ** 2, r_max ** 2))
        t = rng.uniform(0.0, 2.0 * np.pi)
        p = (r * np.cos(t), r * np.sin(t))
        if loc(*p) >= 0:
            pts.append(p)
    return np.array(pts)


def _torch():
    import torch
    return torch


def make_function(solver, sampler, opts=None):
    """Returns f(theta) -> C p_h(theta) as a differentiable torch function."""
    torch = _torch()
    opts = opts or SolveOptions(precision=FP64, rtol=1e-12)

    class DarcyPressureAtPoints(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            solver.set_theta(theta.detach().cpu().double().numpy())
            x, info = solver.solve(opts)
            ctx.x = x
            ctx.info = info
            return torch.from_numpy(sampler(x)).to(theta.dtype)

        @staticmethod
        def backward(ctx, gy):
            dldx = sampler.C.T @ gy.detach().cpu().double().numpy()
            g, _ = solver.adjoint_theta(ctx.x, dldx, opts)
            return torch.from_numpy(g).to(gy.dtype)

    return DarcyPressureAtPoints.apply
