// FP64 element setup: projector, stabilisation, divergence and data terms of the
// mixed VEM on one polygon, for every cell in parallel.
//
// Mathematics follows the reference's CPU path:
//   quadrature      CPU/core/gauss_quadrature.cpp:30-50,252-363
//   basis           CPU/lib/polynomial_basis.cpp:92-101,163-226,245-325
//   ear clipping    CPU/mesh/polygon_triangulator.cpp:50-145 (tolerances made
//                   relative to the cell diameter so that cells of any size
//                   triangulate; identical decisions on the reference meshes)
//   D,G,W,H,B,P     CPU/solver/HdivMatrix.cpp:94-652
//   G_K, alpha, A   CPU/solver/DarcySolver.cpp:1133-1269,1476-1509
//   rhs, multiplier CPU/solver/DarcySolver.cpp:1271-1452
//   boundary        CPU/solver/DarcySolver.cpp:502-663
// but evaluates every basis function once per quadrature point and builds all
// matrices from those tables instead of re-triangulating per matrix entry.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "common.h"

namespace vemd {

static int dimP(int k) { return k < 0 ? 0 : (k + 1) * (k + 2) / 2; }

Shape::Shape(int kk) : k(kk) {
    q = dimP(k);
    grad = dimP(k + 1) - 1;
    curl = k == 0 ? 0 : dimP(k - 1);
    p = grad + curl;
    ni = (q - 1) + (q - (k + 1));
}

// ----------------------------------------------------------------- quadrature
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    if (n < 1 || n > 10) throw Error(VEMD_ERR_ARG, "nq1d must be between 1 and 10");
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        long double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 1, p1 = 1, p2 = 0;
        for (int it = 0; it < 100; ++it) {
            p1 = 1;
            p2 = 0;
            for (int j = 1; j <= n; ++j) {
                long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1) * z * p2 - (j - 1.0L) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            long double dz = p1 / pp;
            z -= dz;
            if (std::fabs((double)dz) < 1e-19) break;
        }
        p1 = 1;
        p2 = 0;
        for (int j = 1; j <= n; ++j) {
            long double p3 = p2;
            p2 = p1;
            p1 = ((2.0L * j - 1) * z * p2 - (j - 1.0L) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1);
        double zi = (2 * i + 1 == n) ? 0.0 : (double)z;
        x[n - 1 - i] = zi;
        x[i] = -zi;
        w[i] = w[n - 1 - i] = (double)(2.0L / ((1 - z * z) * pp * pp));
    }
}

struct Rule {
    std::vector<double> a, b, w;  // reference triangle (0,0),(1,0),(0,1)
};

static Rule tri_rule(int n) {
    Rule r;
    auto collapsed = [&](int m) {
        std::vector<double> x, w;
        gauss_legendre(m, x, w);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                double px = 0.5 * (1 + x[i]);
                r.a.push_back(px);
                r.b.push_back((1 - px) * (1 + x[j]) * 0.5);
                r.w.push_back(w[i] * w[j] * (1 - px) * 0.25);
            }
    };
    switch (n) {
        case 1: r.a = {1.0 / 3}; r.b = {1.0 / 3}; r.w = {0.5}; break;
        case 3: r.a = {0.5, 0.5, 0}; r.b = {0, 0.5, 0.5}; r.w = {1.0 / 6, 1.0 / 6, 1.0 / 6}; break;
        case 4: collapsed(2); break;
        case 9: collapsed(3); break;
        case 7: {
            double s = std::sqrt(15.0), a = (9 + 2 * s) / 21, b = (6 - s) / 21, c = (9 - 2 * s) / 21,
                   d = (6 + s) / 21, wa = (155 - s) / 2400, wc = (155 + s) / 2400;
            r.a = {1.0 / 3, a, b, b, c, d, d};
            r.b = {1.0 / 3, b, a, b, d, c, d};
            r.w = {9.0 / 80, wa, wa, wa, wc, wc, wc};
            break;
        }
        case 12: {
            // degree-6 rule of the reference (CPU/core/gauss_quadrature.cpp:302-316)
            const double o[3][4] = {{0.5611400349004341, 0.2194299825497830, 0, 0.1713331241529811},
                                    {0.0397240717755698, 0.4801379641122151, 0, 0.0807310895930309},
                                    {0.0193717243612408, 0.1416190159239682, 0.8390092597147911, 0.0406345597936607}};
            for (int t = 0; t < 2; ++t) {
                double a = o[t][0], b = o[t][1];
                double pa[3] = {a, b, b}, pb[3] = {b, b, a};
                for (int i = 0; i < 3; ++i) {
                    r.a.push_back(pa[i]);
                    r.b.push_back(pb[i]);
                    r.w.push_back(0.5 * o[t][3]);
                }
            }
            double a = o[2][0], b = o[2][1], c = o[2][2];
            double pa[6] = {a, b, c, b, a, c}, pb[6] = {b, c, a, a, c, b};
            for (int i = 0; i < 6; ++i) {
                r.a.push_back(pa[i]);
                r.b.push_back(pb[i]);
                r.w.push_back(0.5 * o[2][3]);
            }
            break;
        }
        default: throw Error(VEMD_ERR_ARG, "nq2d must be one of 1, 3, 4, 7, 9, 12");
    }
    return r;
}

// ------------------------------------------------------------------- basis
static inline double ipow(double x, int e) {
    double r = 1.0;
    for (int i = 0; i < e; ++i) r *= x;
    return r;
}

static void mono_index(int idx, int& a, int& b) {
    int d = 0, pre = 0;
    while (idx >= pre + d + 1) {
        pre += d + 1;
        ++d;
    }
    a = d - (idx - pre);
    b = idx - pre;
}

// H(div) basis: gradients of scaled monomials of degree 1..k+1, then the
// complement (-(b+1)/a X^(a-1) Y^(b+1), X^a Y^b).
static void hdiv_eval(const Shape& S, int nk, double xm, double ym, double h, double& vx, double& vy) {
    if (nk < S.grad) {
        int a, b;
        mono_index(nk + 1, a, b);
        vx = a > 0 ? a * ipow(xm, a - 1) * ipow(ym, b) / h : 0.0;
        vy = b > 0 ? b * ipow(xm, a) * ipow(ym, b - 1) / h : 0.0;
        return;
    }
    int ci = nk - S.grad, deg = 1;
    while (ci >= deg) {
        ci -= deg;
        ++deg;
    }
    int yd = ci, xd = deg - yd;
    vx = -(double)(yd + 1) / xd * ipow(xm, xd - 1) * ipow(ym, yd + 1);
    vy = ipow(xm, xd) * ipow(ym, yd);
}

static long binom(int n, int k) {
    long r = 1;
    k = std::min(k, n - k);
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// ------------------------------------------------------------------- dense
static void inverse(double* A, int n, double* out, int cell) {
    double a[2 * 32 * 32];
    int m = 2 * n;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) a[i * m + j] = A[i * n + j];
        for (int j = 0; j < n; ++j) a[i * m + n + j] = i == j ? 1.0 : 0.0;
    }
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(a[i * m + k]) > std::fabs(a[piv * m + k])) piv = i;
        if (piv != k)
            for (int j = 0; j < m; ++j) std::swap(a[k * m + j], a[piv * m + j]);
        if (std::fabs(a[k * m + k]) < 1e-14)
            throw Error(VEMD_ERR_SINGULAR, "singular element matrix in cell " + std::to_string(cell));
        for (int i = k + 1; i < n; ++i) {
            double f = a[i * m + k] / a[k * m + k];
            for (int j = k; j < m; ++j) a[i * m + j] -= f * a[k * m + j];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double d = a[k * m + k];
        for (int j = k; j < m; ++j) a[k * m + j] /= d;
        for (int i = 0; i < k; ++i) {
            double f = a[i * m + k];
            for (int j = k; j < m; ++j) a[i * m + j] -= f * a[k * m + j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[i * n + j] = a[i * m + n + j];
}

// ------------------------------------------------------------- geometry/data
struct Tri {
    double v[6];
};

// Ear clipping in the reference's vertex order with diameter-relative
// tolerances (the reference uses absolute 1e-8 / 1e-6, which reject every ear
// once cell areas approach 1e-6).
static void triangulate(const Mesh& m, int c, std::vector<Tri>& out, double& total_area) {
    int n = m.npc(c), s = m.off[c];
    double h2 = m.diam[c] * m.diam[c];
    double tol_cross = 1e-8 * h2, tol_area = 1e-6 * h2;
    std::vector<double> p(2 * n);
    for (int i = 0; i < n; ++i) {
        p[2 * i] = m.xy[2 * m.nodes[s + i]];
        p[2 * i + 1] = m.xy[2 * m.nodes[s + i] + 1];
    }
    out.clear();
    auto push = [&](int a, int b, int cc) {
        Tri t = {{p[2 * a], p[2 * a + 1], p[2 * b], p[2 * b + 1], p[2 * cc], p[2 * cc + 1]}};
        out.push_back(t);
    };
    auto inside = [&](int q, int a, int b, int cc) {
        auto ar = [&](int i, int j, int k) {
            return 0.5 * std::fabs((p[2 * j] - p[2 * i]) * (p[2 * k + 1] - p[2 * i + 1]) -
                                   (p[2 * k] - p[2 * i]) * (p[2 * j + 1] - p[2 * i + 1]));
        };
        return std::fabs(ar(a, b, q) + ar(b, cc, q) + ar(cc, a, q) - ar(a, b, cc)) < tol_area;
    };
    int rem = n;
    while (rem > 3) {
        bool found = false;
        for (int i = 0; i < rem && !found; ++i) {
            int pr = (i - 1 + rem) % rem, nx = (i + 1) % rem;
            double cr = (p[2 * nx] - p[2 * i]) * (p[2 * pr + 1] - p[2 * i + 1]) -
                        (p[2 * pr] - p[2 * i]) * (p[2 * nx + 1] - p[2 * i + 1]);
            if (cr <= tol_cross) continue;
            bool ear = true;
            for (int j = 0; j < rem && ear; ++j)
                if (j != i && j != pr && j != nx && inside(j, pr, i, nx)) ear = false;
            if (!ear) continue;
            push(pr, i, nx);
            p.erase(p.begin() + 2 * i, p.begin() + 2 * i + 2);
            --rem;
            found = true;
        }
        if (!found) throw Error(VEMD_ERR_MESH, "no valid ear in cell " + std::to_string(c));
    }
    push(0, 1, 2);
    total_area = 0.0;
    for (auto& t : out)
        total_area += 0.5 * std::fabs((t.v[2] - t.v[0]) * (t.v[5] - t.v[1]) - (t.v[4] - t.v[0]) * (t.v[3] - t.v[1]));
}

struct Chord {
    double ax, ay, bx, by, R;
    int ia, ib;  // local vertex indices of the chord end points
};

struct Geo {
    double X, Y, j00, j01, j10, j11, det;
};

// Wachspress coordinates of a convex polygon and their gradients, in the
// product form l_i = A_i P_i / D (P_i the product of the edge areas not
// adjacent to vertex i), which stays finite on the polygon's edges.
static void wachspress(const double* vx, const double* vy, int n, double x, double y, double* lam, double* gl) {
    double a[16], ga[16][2], P[16], gP[16][2], Ac[16];
    for (int i = 0; i < n; ++i) {
        int j = (i + 1) % n;
        a[i] = 0.5 * ((vx[i] - x) * (vy[j] - y) - (vy[i] - y) * (vx[j] - x));
        ga[i][0] = 0.5 * (vy[i] - vy[j]);
        ga[i][1] = 0.5 * (vx[j] - vx[i]);
    }
    double D = 0, gD0 = 0, gD1 = 0;
    for (int i = 0; i < n; ++i) {
        int p = (i - 1 + n) % n, q = (i + 1) % n;
        Ac[i] = 0.5 * ((vx[i] - vx[p]) * (vy[q] - vy[p]) - (vy[i] - vy[p]) * (vx[q] - vx[p]));
        P[i] = 1.0;
        gP[i][0] = gP[i][1] = 0.0;
        for (int m = 0; m < n; ++m) {
            if (m == p || m == i) continue;
            gP[i][0] = gP[i][0] * a[m] + P[i] * ga[m][0];
            gP[i][1] = gP[i][1] * a[m] + P[i] * ga[m][1];
            P[i] *= a[m];
        }
        D += Ac[i] * P[i];
        gD0 += Ac[i] * gP[i][0];
        gD1 += Ac[i] * gP[i][1];
    }
    for (int i = 0; i < n; ++i) {
        lam[i] = Ac[i] * P[i] / D;
        gl[2 * i] = (Ac[i] * gP[i][0] - lam[i] * gD0) / D;
        gl[2 * i + 1] = (Ac[i] * gP[i][1] - lam[i] * gD1) / D;
    }
}

// Chord-to-arc map of one boundary cell: F = xi + sum over its chords of
// S^2 delta(s), S = l_A + l_B, s = l_B / S, delta(s) = arc(s) - chord(s), with
// Wachspress coordinates l.  F is smooth inside the cell, maps each chord onto
// its arc and is the identity on the straight edges.
static void chord_blend(const Chord& ch, const double* lam, const double* gl, double* F, double* J) {
    double lA = lam[ch.ia], lB = lam[ch.ib], S = lA + lB;
    if (S <= 1e-300) return;
    const double *gA = &gl[2 * ch.ia], *gB = &gl[2 * ch.ib];
    double s = lB / S;
    double tA = std::atan2(ch.ay, ch.ax), dt = std::atan2(ch.by, ch.bx) - tA;
    if (dt > M_PI) dt -= 2 * M_PI;
    if (dt <= -M_PI) dt += 2 * M_PI;
    double th = tA + s * dt, ct = std::cos(th), st = std::sin(th);
    double del0 = ch.R * ct - (ch.ax + s * (ch.bx - ch.ax)), del1 = ch.R * st - (ch.ay + s * (ch.by - ch.ay));
    double dd0 = -ch.R * dt * st - (ch.bx - ch.ax), dd1 = ch.R * dt * ct - (ch.by - ch.ay);
    double gS0 = gA[0] + gB[0], gS1 = gA[1] + gB[1];
    double w0 = lA * gB[0] - lB * gA[0], w1 = lA * gB[1] - lB * gA[1];
    F[0] += S * S * del0;
    F[1] += S * S * del1;
    J[0] += 2 * S * del0 * gS0 + dd0 * w0;
    J[1] += 2 * S * del0 * gS1 + dd0 * w1;
    J[2] += 2 * S * del1 * gS0 + dd1 * w0;
    J[3] += 2 * S * del1 * gS1 + dd1 * w1;
}

struct CellGeom {
    std::vector<Chord> chords;
    double vx[16], vy[16];
    int n = 0;
    Geo eval(const vemd_problem_desc& d, const Tri& t, double x, double y) const {
        (void)t;
        Geo g;
        if (d.map_kind == VEMD_MAP_SIN) {
            double e = d.eps;
            g.X = x + e * std::sin(2 * M_PI * y);
            g.Y = y + e * std::sin(2 * M_PI * x);
            g.j00 = 1.0;
            g.j01 = e * 2 * M_PI * std::cos(2 * M_PI * y);
            g.j10 = e * 2 * M_PI * std::cos(2 * M_PI * x);
            g.j11 = 1.0;
        } else {
            double F[2] = {x, y}, J[4] = {1, 0, 0, 1};
            if (!chords.empty()) {
                double lam[16], gl[32];
                wachspress(vx, vy, n, x, y, lam, gl);
                for (const Chord& ch : chords) chord_blend(ch, lam, gl, F, J);
            }
            g.X = F[0];
            g.Y = F[1];
            g.j00 = J[0];
            g.j01 = J[1];
            g.j10 = J[2];
            g.j11 = J[3];
        }
        if (d.funcs && d.funcs->map) {  // the caller's map replaces the built-in one
            double xy[2], J[4];
            d.funcs->map(d.funcs->ctx, x, y, xy, J);
            g.X = xy[0];
            g.Y = xy[1];
            g.j00 = J[0];
            g.j01 = J[1];
            g.j10 = J[2];
            g.j11 = J[3];
        }
        g.det = g.j00 * g.j11 - g.j01 * g.j10;
        return g;
    }
};

static CellGeom cell_geom(const Mesh& m, const vemd_problem_desc& d, int c) {
    CellGeom cg;
    int ne = m.npc(c), s0 = m.off[c];
    if (d.map_kind != VEMD_MAP_ARCS || m.arcs.empty() || (d.funcs && d.funcs->map)) return cg;
    for (int j = 0; j < ne; ++j) {
        int a0 = m.nodes[s0 + j], a1 = m.nodes[s0 + (j + 1) % ne];
        auto it = m.arcs.find(Mesh::key(a0, a1));
        if (it != m.arcs.end())
            cg.chords.push_back({m.xy[2 * a0], m.xy[2 * a0 + 1], m.xy[2 * a1], m.xy[2 * a1 + 1], it->second, j, (j + 1) % ne});
    }
    if (!cg.chords.empty()) {
        if (ne > 16) throw Error(VEMD_ERR_MESH, "curved cell with more than 16 edges");
        cg.n = ne;
        for (int j = 0; j < ne; ++j) {
            cg.vx[j] = m.xy[2 * m.nodes[s0 + j]];
            cg.vy[j] = m.xy[2 * m.nodes[s0 + j] + 1];
        }
        // the Wachspress blend needs a strictly convex cell
        for (int j = 0; j < ne; ++j) {
            int p = (j - 1 + ne) % ne, q = (j + 1) % ne;
            double cr = (cg.vx[j] - cg.vx[p]) * (cg.vy[q] - cg.vy[p]) - (cg.vy[j] - cg.vy[p]) * (cg.vx[q] - cg.vx[p]);
            if (!(cr > 0.0))
                throw Error(VEMD_ERR_MESH, "curved boundary cell " + std::to_string(c) + " is not strictly convex");
        }
    }
    return cg;
}

static void exact(int sol, double x, double y, double& p, double& ux, double& uy, double& g) {
    double sx = std::sin(M_PI * x), cx = std::cos(M_PI * x), sy = std::sin(M_PI * y), cy = std::cos(M_PI * y);
    if (sol == VEMD_SOL_SINCOS) {
        p = sx * cy;
        ux = -M_PI * cx * cy;
        uy = M_PI * sx * sy;
    } else {
        p = sx * sy;
        ux = -M_PI * cx * sy;
        uy = -M_PI * sx * cy;
    }
    g = 2 * M_PI * M_PI * p;
}

// Problem data at a physical point -- pressure, flux and source -- from the
// caller's callbacks where given (the reference's DarcyPdeData fields
// solution_p, solution_u, source_g), else the manufactured solution.
static void pde_data(const vemd_problem_desc& d, double x, double y, double& p, double& ux, double& uy, double& g) {
    exact(d.solution, x, y, p, ux, uy, g);
    if (const vemd_problem_funcs* f = d.funcs) {
        if (f->solution_p) p = f->solution_p(f->ctx, x, y);
        if (f->solution_u) {
            double u[2] = {0.0, 0.0};
            f->solution_u(f->ctx, x, y, u);
            ux = u[0];
            uy = u[1];
        }
        if (f->source_g) g = f->source_g(f->ctx, x, y);
    }
}

// T = K_hat^{-1} = J^T K^{-1} J / det J at a volume point (the inverse of the
// reference's computational_kappa, det J J^{-1} K J^{-T},
// CPU/examples/darcy_problem.cpp:15-29,72-83): K = I unless the caller gives
// kappa, which must be symmetric positive definite.  t = [t00, t01, t11].
static void kinv_hat(const vemd_problem_desc& d, const Geo& g, int cell, double* t) {
    double i00 = 1.0, i01 = 0.0, i11 = 1.0;
    if (d.funcs && d.funcs->kappa) {
        double K[4] = {1.0, 0.0, 0.0, 1.0};
        d.funcs->kappa(d.funcs->ctx, g.X, g.Y, K);
        const double sc = std::fabs(K[0]) + std::fabs(K[3]);
        const double det = K[0] * K[3] - K[1] * K[2];
        if (!(std::fabs(K[1] - K[2]) <= 1e-12 * sc) || !(K[0] > 0.0) || !(det > 0.0) || !std::isfinite(det))
            throw Error(VEMD_ERR_ARG, "permeability tensor not symmetric positive definite in cell " + std::to_string(cell));
        const double k01 = 0.5 * (K[1] + K[2]);
        const double dd = K[0] * K[3] - k01 * k01;
        i00 = K[3] / dd;
        i01 = -k01 / dd;
        i11 = K[0] / dd;
    }
    // (K^{-1} J), then J^T (K^{-1} J) / det J
    const double a00 = i00 * g.j00 + i01 * g.j10, a01 = i00 * g.j01 + i01 * g.j11;
    const double a10 = i01 * g.j00 + i11 * g.j10, a11 = i01 * g.j01 + i11 * g.j11;
    t[0] = (g.j00 * a00 + g.j10 * a10) / g.det;
    t[1] = (g.j00 * a01 + g.j10 * a11) / g.det;
    t[2] = (g.j01 * a01 + g.j11 * a11) / g.det;
}

// --------------------------------------------------------------- one cell
struct Work {
    std::vector<Tri> tris;
    std::vector<double> px, py, pw;  // volume points
    std::vector<int> pt;             // triangle of each volume point
    std::vector<Geo> geo;
    std::vector<double> gx, gy, mono;  // basis tables (point-major)
    std::vector<double> D, G, GT, W, H, Hs, B, P, Pd, A, tmp, Ginv, Hinv;
};

struct CellOut {
    double* A;   // packed n(n+1)/2
    double* W;   // q x n
    double* L;   // q
    double* f;   // q
    double* ru;  // n natural-bc flux rhs
    double* ub;  // n essential boundary values
    char* isb;   // n
    double* P;   // p x n projector (for flux reconstruction)
    double* alpha;
};

static void build_cell(const Mesh& m, const vemd_problem_desc& d, const Shape& S, const Rule& rule,
                       const std::vector<double>& glx, const std::vector<double>& glw, int c, Work& w,
                       const CellOut& o) {
    const int ne = m.npc(c), k = S.k, p = S.p, q = S.q, gr = S.grad, n = S.nflux(ne), edof = (k + 1) * ne;
    const int nm = (k + 2) * (k + 3) / 2;  // monomials up to degree k+1
    const double xc = m.cx[c], yc = m.cy[c], h = m.diam[c];
    const int s0 = m.off[c];
    CellGeom cg = cell_geom(m, d, c);
    double tri_area;
    triangulate(m, c, w.tris, tri_area);
    // volume points
    w.px.clear();
    w.py.clear();
    w.pw.clear();
    w.pt.clear();
    for (size_t t = 0; t < w.tris.size(); ++t) {
        const double* v = w.tris[t].v;
        double J = std::fabs((v[2] - v[0]) * (v[5] - v[1]) - (v[4] - v[0]) * (v[3] - v[1]));
        for (size_t i = 0; i < rule.w.size(); ++i) {
            w.px.push_back(v[0] + (v[2] - v[0]) * rule.a[i] + (v[4] - v[0]) * rule.b[i]);
            w.py.push_back(v[1] + (v[3] - v[1]) * rule.a[i] + (v[5] - v[1]) * rule.b[i]);
            w.pw.push_back(rule.w[i] * J);
            w.pt.push_back((int)t);
        }
    }
    const int np = (int)w.pw.size();
    w.geo.resize(np);
    w.gx.resize((size_t)np * p);
    w.gy.resize((size_t)np * p);
    w.mono.resize((size_t)np * nm);
    for (int l = 0; l < np; ++l) {
        w.geo[l] = cg.eval(d, w.tris[w.pt[l]], w.px[l], w.py[l]);
        if (!(w.geo[l].det > 0.0))
            throw Error(VEMD_ERR_MESH, "non-positive map Jacobian in cell " + std::to_string(c));
        double xm = (w.px[l] - xc) / h, ym = (w.py[l] - yc) / h;
        for (int a = 0; a < p; ++a) hdiv_eval(S, a, xm, ym, h, w.gx[l * p + a], w.gy[l * p + a]);
        for (int r = 0; r < nm; ++r) {
            int ea, eb;
            mono_index(r, ea, eb);
            w.mono[l * nm + r] = ipow(xm, ea) * ipow(ym, eb);
        }
    }
    // G, G_T (K = I weighted by J^T J / det J), alpha, H, H*
    w.G.assign(p * p, 0.0);
    w.GT.assign(p * p, 0.0);
    w.H.assign(q * q, 0.0);
    w.Hs.assign(gr * q, 0.0);
    double Ti[4] = {0, 0, 0, 0};
    for (int l = 0; l < np; ++l) {
        const Geo& g = w.geo[l];
        double tk[3];
        kinv_hat(d, g, c, tk);
        const double t00 = tk[0], t01 = tk[1], t11 = tk[2];
        double wl = w.pw[l];
        Ti[0] += t00 * wl;
        Ti[1] += t01 * wl;
        Ti[3] += t11 * wl;
        const double* ax = &w.gx[l * p];
        const double* ay = &w.gy[l * p];
        for (int a = 0; a < p; ++a)
            for (int b = 0; b < p; ++b) {
                w.G[a * p + b] += (ax[a] * ax[b] + ay[a] * ay[b]) * wl;
                w.GT[a * p + b] += (ax[a] * (t00 * ax[b] + t01 * ay[b]) + ay[a] * (t01 * ax[b] + t11 * ay[b])) * wl;
            }
        const double* mo = &w.mono[l * nm];
        for (int r = 0; r < q; ++r)
            for (int s = 0; s < q; ++s) w.H[r * q + s] += mo[r] * mo[s] * wl;
        for (int a = 0; a < gr; ++a)
            for (int s = 0; s < q; ++s) w.Hs[a * q + s] += mo[a + 1] * mo[s] * wl;
    }
    Ti[2] = Ti[1];
    if (!(tri_area > 0.0)) throw Error(VEMD_ERR_MESH, "cell " + std::to_string(c) + " has no area");
    *o.alpha = std::sqrt((Ti[0] * Ti[0] + Ti[1] * Ti[1] + Ti[2] * Ti[2] + Ti[3] * Ti[3])) / tri_area;
    // D: edge moments int_e (g.n) t^r, internal moments int g . g_mom
    w.D.assign(n * p, 0.0);
    w.W.assign(q * n, 0.0);
    std::vector<double> ex(glx.size()), ey(glx.size()), ew(glx.size()), et(glx.size());
    w.B.assign(p * n, 0.0);
    for (int j = 0; j < ne; ++j) {
        int a0 = m.nodes[s0 + j], a1 = m.nodes[s0 + (j + 1) % ne];
        double x1 = m.xy[2 * a0], y1 = m.xy[2 * a0 + 1], x2 = m.xy[2 * a1], y2 = m.xy[2 * a1 + 1];
        double dx = x2 - x1, dy = y2 - y1, len = std::sqrt(dx * dx + dy * dy);
        if (!(len > 1e-14 * h)) throw Error(VEMD_ERR_MESH, "degenerate edge in cell " + std::to_string(c));
        double nx = dy / len, ny = -dx / len, tx = dx / len, ty = dy / len;
        double mx = 0.5 * (x1 + x2), my = 0.5 * (y1 + y2);
        for (size_t i = 0; i < glx.size(); ++i) {
            ex[i] = x1 + dx * (glx[i] + 1) / 2;
            ey[i] = y1 + dy * (glx[i] + 1) / 2;
            ew[i] = glw[i] * len / 2;
            et[i] = ((ex[i] - mx) * tx + (ey[i] - my) * ty) / len;
        }
        for (int r = 0; r <= k; ++r)
            for (int a = 0; a < p; ++a) {
                double sacc = 0;
                for (size_t i = 0; i < glx.size(); ++i) {
                    double vx, vy;
                    hdiv_eval(S, a, (ex[i] - xc) / h, (ey[i] - yc) / h, h, vx, vy);
                    sacc += ew[i] * (vx * nx + vy * ny) * ipow(et[i], r);
                }
                w.D[(r * ne + j) * p + a] = sacc;
            }
        // W edge columns: coefficients of m_r on the edge in t (exact)
        for (int r = 0; r < q; ++r) {
            int ea, eb;
            mono_index(r, ea, eb);
            double co[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            double ox = mx - xc, oy = my - yc;
            for (int i = 0; i <= ea; ++i)
                for (int jj = 0; jj <= eb; ++jj)
                    co[i + jj] += binom(ea, i) * binom(eb, jj) * ipow(ox, ea - i) * ipow(oy, eb - jj) * ipow(tx, i) *
                                  ipow(ty, jj);
            for (int dg = 0; dg <= ea + eb; ++dg) co[dg] *= ipow(len, dg) / ipow(h, ea + eb);
            for (int dg = 0; dg <= k; ++dg) w.W[r * n + dg * ne + j] += co[dg];
        }
        // B_2 contribution: H_e* H_e^{-1} W_e on this edge's columns
        int kk = k + 1;
        double He[16], Hei[16], Hes[32 * 4];
        for (int r = 0; r < kk; ++r)
            for (int s = 0; s < kk; ++s) {
                double v = 0;
                for (size_t i = 0; i < glx.size(); ++i) v += ipow(et[i], r) * ipow(et[i], s) * ew[i];
                He[r * kk + s] = v;
            }
        for (int a = 0; a < gr; ++a)
            for (int cc = 0; cc < kk; ++cc) {
                double v = 0;
                for (size_t i = 0; i < glx.size(); ++i) {
                    int ea, eb;
                    mono_index(a + 1, ea, eb);
                    v += ipow((ex[i] - xc) / h, ea) * ipow((ey[i] - yc) / h, eb) * ipow(et[i], cc) * ew[i];
                }
                Hes[a * kk + cc] = v;
            }
        inverse(He, kk, Hei, c);
        // W_e(r, s) = [s mod ne == r] (CPU/solver/HdivMatrix.cpp:275-282)
        for (int a = 0; a < gr; ++a)
            for (int cc = 0; cc < kk; ++cc) {
                double v = 0;
                for (int t = 0; t < kk; ++t)
                    for (int s = 0; s < kk; ++s)
                        if (cc % ne == s) v += Hes[a * kk + t] * Hei[t * kk + s];
                w.B[a * n + cc * ne + j] += v;
            }
        // boundary data on this edge
        int ge = m.cell_edges[s0 + j];
        for (int r = 0; r <= k; ++r) {
            o.ru[r * ne + j] = 0.0;
            o.ub[r * ne + j] = 0.0;
            o.isb[r * ne + j] = 0;
        }
        if (m.edge_occ[ge] == 1) {
            const Tri* tt = &w.tris[0];
            for (auto& t : w.tris) {
                int hit = 0;
                for (int z = 0; z < 3; ++z)
                    if ((t.v[2 * z] == x1 && t.v[2 * z + 1] == y1) || (t.v[2 * z] == x2 && t.v[2 * z + 1] == y2)) ++hit;
                if (hit == 2) {
                    tt = &t;
                    break;
                }
            }
            double mom_p[4] = {0, 0, 0, 0};
            const vemd_problem_funcs* uf = d.funcs;
            for (size_t i = 0; i < glx.size(); ++i) {
                Geo g = cg.eval(d, *tt, ex[i], ey[i]);
                double pe, ux, uy, gg;
                pde_data(d, g.X, g.Y, pe, ux, uy, gg);
                if (uf && uf->dirichlet_p) pe = uf->dirichlet_p(uf->ctx, g.X, g.Y);
                // u_hat . n_hat with u_hat = det J J^{-1} u, i.e. det J u . (J^{-T} n_hat)
                double uhn;
                if (uf && uf->neumann_un) {
                    // the physical normal is J^{-T} n_hat / |J^{-T} n_hat|
                    const double mx = (g.j11 * nx - g.j10 * ny) / g.det, my = (-g.j01 * nx + g.j00 * ny) / g.det;
                    const double mn = std::sqrt(mx * mx + my * my);
                    uhn = g.det * mn * uf->neumann_un(uf->ctx, g.X, g.Y, mx / mn, my / mn);
                } else {
                    const double uhx = g.j11 * ux - g.j01 * uy, uhy = -g.j10 * ux + g.j00 * uy;
                    uhn = uhx * nx + uhy * ny;
                }
                for (int r = 0; r <= k; ++r) {
                    double tr = ipow(et[i], r);
                    o.ub[r * ne + j] += uhn * tr * ew[i];
                    mom_p[r] += pe * tr * ew[i];
                }
            }
            if (d.bc == VEMD_BC_FLUX_MEANZERO) {
                for (int r = 0; r <= k; ++r) o.isb[r * ne + j] = 1;
            } else {
                // natural term -int p_D (phi . n): phi.n = sum_s (H_e^{-1})_{s r} t^s
                for (int r = 0; r <= k; ++r) {
                    double v = 0;
                    for (int s = 0; s < kk; ++s) v += Hei[s * kk + r] * mom_p[s];
                    o.ru[r * ne + j] = -v;
                    o.ub[r * ne + j] = 0.0;
                }
            }
        }
    }
    // internal rows of D and W
    for (int i = edof; i < n; ++i) {
        int mom = i < edof + (q - 1) ? (i - edof) : (i - edof - (q - 1)) + gr;
        for (int a = 0; a < p; ++a) {
            double sacc = 0;
            for (int l = 0; l < np; ++l)
                sacc += (w.gx[l * p + a] * w.gx[l * p + mom] + w.gy[l * p + a] * w.gy[l * p + mom]) * w.pw[l];
            w.D[i * p + a] = sacc;
        }
        int gd = i - edof;
        if (gd < q - 1) w.W[(gd + 1) * n + i] += -1.0;
    }
    // B_1 = -H* H^{-1} W, B_curl
    w.Hinv.resize(q * q);
    inverse(w.H.data(), q, w.Hinv.data(), c);
    w.tmp.assign(q * n, 0.0);
    for (int r = 0; r < q; ++r)
        for (int s = 0; s < q; ++s)
            for (int i = 0; i < n; ++i) w.tmp[r * n + i] += w.Hinv[r * q + s] * w.W[s * n + i];
    for (int a = 0; a < gr; ++a)
        for (int r = 0; r < q; ++r)
            for (int i = 0; i < n; ++i) w.B[a * n + i] -= w.Hs[a * q + r] * w.tmp[r * n + i];
    for (int r = 0; r < S.curl; ++r) w.B[(gr + r) * n + edof + (q - 1) + r] = 1.0;
    // P = G^{-1} B, Pd = D P
    w.Ginv.resize(p * p);
    inverse(w.G.data(), p, w.Ginv.data(), c);
    w.P.assign(p * n, 0.0);
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int i = 0; i < n; ++i) w.P[a * n + i] += w.Ginv[a * p + b] * w.B[b * n + i];
    w.Pd.assign(n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < p; ++a)
            for (int j = 0; j < n; ++j) w.Pd[i * n + j] += w.D[i * p + a] * w.P[a * n + j];
    // A = P^T G_T P + alpha (I - Pd)^T (I - Pd)
    w.tmp.assign(p * n, 0.0);
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int i = 0; i < n; ++i) w.tmp[a * n + i] += w.GT[a * p + b] * w.P[b * n + i];
    w.A.assign(n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < p; ++a)
            for (int j = 0; j < n; ++j) w.A[i * n + j] += w.P[a * n + i] * w.tmp[a * n + j];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) w.Pd[i * n + j] = (i == j ? 1.0 : 0.0) - w.Pd[i * n + j];
    for (int l = 0; l < n; ++l)
        for (int i = 0; i < n; ++i) {
            double a = *o.alpha * w.Pd[l * n + i];
            for (int j = 0; j < n; ++j) w.A[i * n + j] += a * w.Pd[l * n + j];
        }
    // multiplier row and source
    for (int r = 0; r < q; ++r) {
        double sl = 0, sf = 0;
        for (int l = 0; l < np; ++l) {
            const Geo& g = w.geo[l];
            double pe, ux, uy, gg;
            pde_data(d, g.X, g.Y, pe, ux, uy, gg);
            double mr = w.mono[l * nm + r];
            sl += mr * g.det * w.pw[l];
            sf += g.det * gg * mr * w.pw[l];
        }
        o.L[r] = sl;
        o.f[r] = sf;
    }
    // direction signs (CPU/solver/DarcySolver.cpp:229-320) applied to rows/columns
    double sg[64];
    for (int i = 0; i < n; ++i) sg[i] = 1.0;
    for (int r = 0; r <= k; r += 2)
        for (int j = 0; j < ne; ++j) {
            int ge = m.cell_edges[s0 + j];
            if (m.edge_occ[ge] == 2 && m.nodes[s0 + j] == m.edge_nodes[2 * ge + 1]) sg[r * ne + j] = -1.0;
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) o.A[i * (i + 1) / 2 + j] = w.A[i * n + j] * sg[i] * sg[j];
    for (int r = 0; r < q; ++r)
        for (int i = 0; i < n; ++i) o.W[r * n + i] = w.W[r * n + i] * sg[i];
    for (int i = 0; i < n; ++i) o.ru[i] *= sg[i];
    std::memcpy(o.P, w.P.data(), sizeof(double) * p * n);
}

QuadRules quad_rules(int nq2d, int nq1d) {
    QuadRules r;
    Rule t = tri_rule(nq2d);
    r.ta = t.a;
    r.tb = t.b;
    r.tw = t.w;
    gauss_legendre(nq1d, r.glx, r.glw);
    return r;
}

Setup* setup_prepare(const Mesh& m, const vemd_problem_desc& d) {
    if (d.k < 0 || d.k > 3) throw Error(VEMD_ERR_ARG, "k must be between 0 and 3");
    if (d.k == 3 && d.nq2d != 12) throw Error(VEMD_ERR_ARG, "k=3 requires 12 triangle points");
    if (d.map_kind != VEMD_MAP_SIN && d.map_kind != VEMD_MAP_ARCS) throw Error(VEMD_ERR_ARG, "map_kind");
    if (d.bc != VEMD_BC_FLUX_MEANZERO && d.bc != VEMD_BC_DIRICHLET_P) throw Error(VEMD_ERR_ARG, "bc");
    (void)tri_rule(d.nq2d);  // validates nq2d
    if (d.theta)
        for (int c = 0; c < m.nc; ++c)
            if (!std::isfinite(d.theta[c])) throw Error(VEMD_ERR_ARG, "theta must be finite");
    std::vector<double> glx, glw;
    gauss_legendre(d.nq1d, glx, glw);
    Setup* S = new Setup();
    S->sh = Shape(d.k);
    S->desc = d;
    S->desc.theta = nullptr;
    if (d.funcs) {  // a copy the library owns (the caller's struct may go away)
        S->funcs = *d.funcs;
        S->desc.funcs = &S->funcs;
    }
    const Shape& sh = S->sh;
    int nc = m.nc;
    auto& O = S->ofs;
    O.a_off.resize(nc + 1);
    O.w_off.resize(nc + 1);
    O.v_off.resize(nc + 1);
    S->p_off.resize(nc + 1);
    O.a_off[0] = O.w_off[0] = O.v_off[0] = S->p_off[0] = 0;
    for (int c = 0; c < nc; ++c) {
        int n = sh.nflux(m.npc(c));
        if (n > 63) {
            delete S;
            throw Error(VEMD_ERR_MESH, "cell " + std::to_string(c) + " has too many edges");
        }
        O.a_off[c + 1] = O.a_off[c] + n * (n + 1) / 2;
        O.w_off[c + 1] = O.w_off[c] + (int64_t)sh.q * n;
        O.v_off[c + 1] = O.v_off[c] + n;
        S->p_off[c + 1] = S->p_off[c] + (int64_t)sh.p * n;
    }
    S->A.resize(O.a_off[nc]);
    S->W.resize(O.w_off[nc]);
    S->P.resize(S->p_off[nc]);
    S->L.resize((size_t)sh.q * nc);
    S->f.resize((size_t)sh.q * nc);
    S->ru.resize(O.v_off[nc]);
    S->ub.resize(O.v_off[nc]);
    S->isb.resize(O.v_off[nc]);
    S->alpha.resize(nc);
    S->theta.assign(nc, 0.0);
    if (d.theta) S->theta.assign(d.theta, d.theta + nc);
    return S;
}

void setup_finish(const Mesh& m, Setup* S) {
    const vemd_problem_desc& d = S->desc;
    const Shape& sh = S->sh;
    const int nc = m.nc;
    const auto& O = S->ofs;
    // reference global dof of each local flux dof, and its sign
    S->gdof.resize(O.v_off[nc]);
    S->sgn.resize(O.v_off[nc]);
    for (int c = 0; c < nc; ++c) {
        int ne = m.npc(c), s0 = m.off[c];
        int64_t v = O.v_off[c];
        for (int r = 0; r <= d.k; ++r)
            for (int j = 0; j < ne; ++j) {
                int ge = m.cell_edges[s0 + j];
                S->gdof[v + r * ne + j] = (int64_t)r * m.ne + ge;
                bool rev = m.edge_occ[ge] == 2 && m.nodes[s0 + j] == m.edge_nodes[2 * ge + 1];
                S->sgn[v + r * ne + j] = (rev && r % 2 == 0) ? -1.0 : 1.0;
            }
        for (int i = 0; i < sh.ni; ++i) {
            S->gdof[v + (d.k + 1) * ne + i] = (int64_t)m.ne * (d.k + 1) + (int64_t)c * sh.ni + i;
            S->sgn[v + (d.k + 1) * ne + i] = 1.0;
        }
    }
    S->nflux_total = (int64_t)m.ne * (d.k + 1) + (int64_t)nc * sh.ni;
    S->npres_total = (int64_t)nc * sh.q;
    // one mean-zero multiplier per batch instance (the last unknowns, in instance order)
    S->ntot = S->nflux_total + S->npres_total + (d.bc == VEMD_BC_FLUX_MEANZERO ? m.ninst() : 0);
}

Setup* setup_build(const Mesh& m, const vemd_problem_desc& d) {
    std::unique_ptr<Setup> S(setup_prepare(m, d));
    const Shape& sh = S->sh;
    const auto& O = S->ofs;
    Rule rule = tri_rule(d.nq2d);
    std::vector<double> glx, glw;
    gauss_legendre(d.nq1d, glx, glw);
    Setup* Sp = S.get();
    parallel_for(m.nc, [&](int64_t b, int64_t e) {
        Work w;
        for (int64_t c = b; c < e; ++c) {
            CellOut o{&Sp->A[O.a_off[c]], &Sp->W[O.w_off[c]], &Sp->L[c * sh.q], &Sp->f[c * sh.q],
                      &Sp->ru[O.v_off[c]], &Sp->ub[O.v_off[c]], &Sp->isb[O.v_off[c]], &Sp->P[Sp->p_off[c]],
                      &Sp->alpha[c]};
            build_cell(m, Sp->desc, sh, rule, glx, glw, (int)c, w, o);
        }
    });
    setup_finish(m, Sp);
    return S.release();
}

// L2 errors against the manufactured solution (CPU/solver/DarcySolver.cpp:942-1131):
// pressure and projected flux in the computational measure, then in the
// physical measure (det J weight, forward Piola u = J u_hat / det J), then the
// physical divergence error and the mass int p_h det J
// (GPU/src/equation_gpu.cu:385-479): div_x u_h = (H^{-1} W u) . m / det J with
// H the monomial mass matrix of the computational cell.
void errors_host(const struct Solver& s, const double* sol, int nq, double* out);

}  // namespace vemd

#include "solver.h"

namespace vemd {

void errors_host(const Solver& s, const double* sol, int nq, double* out) {
    const Mesh& m = *s.mesh;
    const Setup& S = *s.setup;
    const Shape& sh = S.sh;
    const vemd_problem_desc& d = S.desc;
    Rule rule = tri_rule(nq);
    std::vector<double> e(6 * (size_t)m.nc, 0.0);
    parallel_for(m.nc, [&](int64_t b, int64_t en) {
        std::vector<Tri> tris;
        std::vector<double> u, coef(sh.p), pc(sh.q), wu(sh.q), dc(sh.q), H(sh.q * sh.q), mr(sh.q);
        for (int64_t c = b; c < en; ++c) {
            int ne = m.npc((int)c), n = sh.nflux(ne);
            int64_t v0 = S.ofs.v_off[c];
            u.resize(n);
            for (int i = 0; i < n; ++i) u[i] = S.sgn[v0 + i] * sol[S.gdof[v0 + i]];
            for (int r = 0; r < sh.q; ++r) pc[r] = sol[S.nflux_total + c * sh.q + r];
            const double* P = &S.P[S.p_off[c]];
            for (int a = 0; a < sh.p; ++a) {
                double v = 0;
                for (int i = 0; i < n; ++i) v += P[a * n + i] * u[i];
                coef[a] = v;
            }
            CellGeom cg = cell_geom(m, d, (int)c);
            double ta;
            triangulate(m, (int)c, tris, ta);
            double xc = m.cx[c], yc = m.cy[c], h = m.diam[c];
            // divergence coefficients: H dc = W u (W stored with the direction
            // signs in its columns, so W sol[gdof] = W_unsigned u)
            const double* Wc = &S.W[S.ofs.w_off[c]];
            for (int r = 0; r < sh.q; ++r) {
                double v = 0;
                for (int i = 0; i < n; ++i) v += Wc[r * n + i] * sol[S.gdof[v0 + i]];
                wu[r] = v;
            }
            std::fill(H.begin(), H.end(), 0.0);
            for (auto& t : tris) {
                const double* v = t.v;
                double J = std::fabs((v[2] - v[0]) * (v[5] - v[1]) - (v[4] - v[0]) * (v[3] - v[1]));
                for (size_t i = 0; i < rule.w.size(); ++i) {
                    double x = v[0] + (v[2] - v[0]) * rule.a[i] + (v[4] - v[0]) * rule.b[i];
                    double y = v[1] + (v[3] - v[1]) * rule.a[i] + (v[5] - v[1]) * rule.b[i];
                    double xm = (x - xc) / h, ym = (y - yc) / h;
                    for (int r = 0; r < sh.q; ++r) {
                        int ea, eb;
                        mono_index(r, ea, eb);
                        mr[r] = ipow(xm, ea) * ipow(ym, eb);
                    }
                    for (int r = 0; r < sh.q; ++r)
                        for (int t2 = 0; t2 < sh.q; ++t2) H[r * sh.q + t2] += mr[r] * mr[t2] * rule.w[i] * J;
                }
            }
            {
                // H / |E| is O(1) for the scaled monomials
                const double s0 = H[0] > 0 ? 1.0 / H[0] : 1.0;
                std::vector<double> Hs(H.size()), Hi(H.size());
                for (size_t i = 0; i < H.size(); ++i) Hs[i] = H[i] * s0;
                inverse(Hs.data(), sh.q, Hi.data(), (int)c);
                for (int r = 0; r < sh.q; ++r) {
                    double v = 0;
                    for (int t2 = 0; t2 < sh.q; ++t2) v += Hi[r * sh.q + t2] * wu[t2] * s0;
                    dc[r] = v;
                }
            }
            double ep = 0, eu = 0, epp = 0, eup = 0, ed = 0, ms = 0;
            for (auto& t : tris) {
                const double* v = t.v;
                double J = std::fabs((v[2] - v[0]) * (v[5] - v[1]) - (v[4] - v[0]) * (v[3] - v[1]));
                for (size_t i = 0; i < rule.w.size(); ++i) {
                    double x = v[0] + (v[2] - v[0]) * rule.a[i] + (v[4] - v[0]) * rule.b[i];
                    double y = v[1] + (v[3] - v[1]) * rule.a[i] + (v[5] - v[1]) * rule.b[i];
                    double w = rule.w[i] * J;
                    double xm = (x - xc) / h, ym = (y - yc) / h;
                    double ph = 0, vx = 0, vy = 0, dv = 0;
                    for (int r = 0; r < sh.q; ++r) {
                        int ea, eb;
                        mono_index(r, ea, eb);
                        const double mv = ipow(xm, ea) * ipow(ym, eb);
                        ph += pc[r] * mv;
                        dv += dc[r] * mv;
                    }
                    for (int a = 0; a < sh.p; ++a) {
                        double bx, by;
                        hdiv_eval(sh, a, xm, ym, h, bx, by);
                        vx += coef[a] * bx;
                        vy += coef[a] * by;
                    }
                    Geo g = cg.eval(d, t, x, y);
                    double pe, ux, uy, gg;
                    pde_data(d, g.X, g.Y, pe, ux, uy, gg);
                    double uhx = g.j11 * ux - g.j01 * uy, uhy = -g.j10 * ux + g.j00 * uy;
                    double dp = ph - pe, dx = vx - uhx, dy = vy - uhy;
                    ep += dp * dp * w;
                    eu += (dx * dx + dy * dy) * w;
                    epp += dp * dp * w * g.det;
                    double fx = (g.j00 * dx + g.j01 * dy) / g.det, fy = (g.j10 * dx + g.j11 * dy) / g.det;
                    eup += (fx * fx + fy * fy) * w * g.det;
                    const double dd = dv / g.det - gg;  // div_x u_h - div u (= g)
                    ed += dd * dd * w * g.det;
                    ms += ph * w * g.det;
                }
            }
            double* ec = &e[6 * c];
            ec[0] = ep;
            ec[1] = eu;
            ec[2] = epp;
            ec[3] = eup;
            ec[4] = ed;
            ec[5] = ms;
        }
    });
    double t[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < m.nc; ++c)
        for (int j = 0; j < 6; ++j) t[j] += e[6 * (size_t)c + j];
    for (int j = 0; j < 5; ++j) out[j] = std::sqrt(t[j]);
    out[5] = t[5];
}

}  // namespace vemd
