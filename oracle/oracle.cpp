// CPU oracle for the mixed virtual element Darcy solve.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++ without PETSc,
// the reference project's CPU algorithm (vem/CurvedVEM_Darcy) so that tests can
// check the HIP product path against it.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Every function cites the reference file:line it follows.  Paths are relative
// to the reference root; CPU/ = vem/CurvedVEM_Darcy/.
//
// Extensions beyond the reference (each marked "EXTENSION" below):
//   * per-cell curved maps that push a straight boundary chord onto a circular
//     arc (the reference has only one global map F(xi,eta));
//   * a per-cell scalar log-permeability theta_E (K = exp(theta_E) I);
//   * the manufactured pressure sin(pi x) sin(pi y);
//   * a Dirichlet-pressure (natural) boundary condition, which the reference
//     declares but never defines (CPU/solver/DarcySolver.h:88);
//   * the caller-defined data (OrcFuncs) also feed the reference's declared
//     but unused dirichlet_p and neumann_un fields into the boundary
//     conditions; kappa, solution_p/u, source_g and the map enter as in the
//     reference (computational_kappa etc., CPU/examples/darcy_problem.cpp:37-83).
//
// Parity status of each part is recorded in the test files that pin it.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace orc {

static thread_local std::string g_err;

// ---------------------------------------------------------------------------
// Quadrature.  CPU/core/gauss_quadrature.cpp.
// ---------------------------------------------------------------------------

// Gauss-Legendre nodes/weights on [-1,1], ascending, as in
// CPU/core/gauss_quadrature.cpp:68-187.  Computed here by Newton iteration on
// the Legendre recurrence instead of tabulated constants.
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    if (n < 1 || n > 10) throw std::invalid_argument("Gauss-Legendre order must be between 1 and 10");
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        long double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        long double pp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= n; ++j) {
                long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0L);
            long double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs((double)(z - z1)) < 1e-19) break;
        }
        long double p1 = 1.0L, p2 = 0.0L;
        for (int j = 1; j <= n; ++j) {
            long double p3 = p2;
            p2 = p1;
            p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1.0L);
        double wi = (double)(2.0L / ((1.0L - z * z) * pp * pp));
        double zi = (double)z;
        if (2 * i + 1 == n) zi = 0.0;
        x[n - 1 - i] = zi;
        x[i] = -zi;
        w[n - 1 - i] = wi;
        w[i] = wi;
    }
}

struct TriRule {
    std::vector<double> xi, eta, w;  // on the reference triangle (0,0),(1,0),(0,1); sum w = 1/2
};

// Collapsed (Duffy) Gauss product rule: x = (1+a)/2, y = (1-x)(1+b)/2,
// weight wa*wb*(1-x)/4.  The reference's 4- and 9-point rules are exactly the
// 2x2 and 3x3 instances (CPU/core/gauss_quadrature.cpp:266-273,289-301).
static void collapsed_gauss(int m, TriRule& r) {
    std::vector<double> x, w;
    gauss_legendre(m, x, w);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double px = (1.0 + x[i]) / 2.0;
            double py = (1.0 - px) * (1.0 + x[j]) / 2.0;
            r.xi.push_back(px);
            r.eta.push_back(py);
            r.w.push_back(w[i] * w[j] * (1.0 - px) / 4.0);
        }
}

// Triangle rules with 1, 3, 4, 7, 9, 12 points (CPU/core/gauss_quadrature.cpp:252-322).
static TriRule triangle_rule(int n) {
    TriRule r;
    switch (n) {
        case 1:
            r.xi = {1.0 / 3.0}; r.eta = {1.0 / 3.0}; r.w = {0.5};
            break;
        case 3:
            r.xi = {0.5, 0.5, 0.0}; r.eta = {0.0, 0.5, 0.5}; r.w = {1.0 / 6, 1.0 / 6, 1.0 / 6};
            break;
        case 4: collapsed_gauss(2, r); break;
        case 7: {
            // Radon's degree-5 rule in closed form.
            double s15 = std::sqrt(15.0);
            double a = (9.0 + 2.0 * s15) / 21.0, b = (6.0 - s15) / 21.0;
            double c = (9.0 - 2.0 * s15) / 21.0, d = (6.0 + s15) / 21.0;
            double wa = (155.0 - s15) / 2400.0, wc = (155.0 + s15) / 2400.0;
            r.xi = {1.0 / 3.0, a, b, b, c, d, d};
            r.eta = {1.0 / 3.0, b, a, b, d, c, d};
            r.w = {9.0 / 80.0, wa, wa, wa, wc, wc, wc};
            break;
        }
        case 9: collapsed_gauss(3, r); break;
        case 12: {
            // Tabulated 12-point degree-6 rule used by the reference for k=3
            // (CPU/core/gauss_quadrature.cpp:302-316): three orbits.
            const double a1 = 0.5611400349004341, b1 = 0.2194299825497830, w1 = 0.1713331241529811;
            const double a2 = 0.0397240717755698, b2 = 0.4801379641122151, w2 = 0.0807310895930309;
            const double a3 = 0.0193717243612408, b3 = 0.1416190159239682, c3 = 0.8390092597147911,
                         w3 = 0.0406345597936607;
            double px[12] = {a1, b1, b1, a2, b2, b2, a3, b3, c3, b3, a3, c3};
            double py[12] = {b1, b1, a1, b2, b2, a2, b3, c3, a3, a3, c3, b3};
            double pw[12] = {w1, w1, w1, w2, w2, w2, w3, w3, w3, w3, w3, w3};
            for (int i = 0; i < 12; ++i) {
                r.xi.push_back(px[i]);
                r.eta.push_back(py[i]);
                r.w.push_back(0.5 * pw[i]);
            }
            break;
        }
        default:
            throw std::invalid_argument("Triangle quadrature supports: 1, 3, 4, 7, 9, 12 points");
    }
    return r;
}

struct QP {
    double x, y, w;
};

// CPU/core/gauss_quadrature.cpp:342-363 (affine map, weight times |2*area|).
static void map_triangle(const TriRule& r, const double* v, std::vector<QP>& out) {
    double x1 = v[0], y1 = v[1], x2 = v[2], y2 = v[3], x3 = v[4], y3 = v[5];
    double J = std::fabs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
    out.clear();
    for (size_t i = 0; i < r.w.size(); ++i)
        out.push_back({x1 + (x2 - x1) * r.xi[i] + (x3 - x1) * r.eta[i],
                       y1 + (y2 - y1) * r.xi[i] + (y3 - y1) * r.eta[i], r.w[i] * J});
}

// CPU/core/gauss_quadrature.cpp:30-50.
static void line_points(int n, double x1, double y1, double x2, double y2, std::vector<QP>& out) {
    std::vector<double> x, w;
    gauss_legendre(n, x, w);
    double dx = x2 - x1, dy = y2 - y1;
    double len = std::sqrt(dx * dx + dy * dy);
    out.clear();
    for (int i = 0; i < n; ++i)
        out.push_back({x1 + dx * (x[i] + 1.0) / 2.0, y1 + dy * (x[i] + 1.0) / 2.0, w[i] * len / 2.0});
}

// ---------------------------------------------------------------------------
// Polynomial basis.  CPU/lib/polynomial_basis.cpp.
// ---------------------------------------------------------------------------

static double ipow(double x, int e) {
    double r = 1.0;
    for (int i = 0; i < e; ++i) r *= x;
    return r;
}

static long comb(int n, int k) {
    k = std::min(k, n - k);
    long r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

static int dimP(int k) { return k < 0 ? 0 : (k + 1) * (k + 2) / 2; }

// Monomial index -> (x degree, y degree): total degree first, then increasing
// y power (CPU/lib/polynomial_basis.cpp:309-325).
static void mindex(int idx, int& a, int& b) {
    int d = 0, pre = 0;
    while (idx >= pre + d + 1) {
        pre += d + 1;
        ++d;
    }
    int off = idx - pre;
    a = d - off;
    b = off;
}

struct Basis {
    int k, q, grad, curl, p;
    explicit Basis(int kk) : k(kk) {
        q = dimP(k);
        grad = dimP(k + 1) - 1;
        curl = k == 0 ? 0 : dimP(k - 1);
        p = grad + curl;
    }
    // CPU/lib/polynomial_basis.cpp:92-101.
    double mono(int nk, double xc, double yc, double h, double x, double y) const {
        int a, b;
        mindex(nk, a, b);
        return ipow((x - xc) / h, a) * ipow((y - yc) / h, b);
    }
    // CPU/lib/polynomial_basis.cpp:245-290.
    void hdiv(int nk, double xc, double yc, double h, double x, double y, double& vx, double& vy) const {
        double xm = (x - xc) / h, ym = (y - yc) / h;
        if (nk < grad) {
            int a, b;
            mindex(nk + 1, a, b);
            vx = a > 0 ? a * ipow(xm, a - 1) * ipow(ym, b) / h : 0.0;
            vy = b > 0 ? b * ipow(xm, a) * ipow(ym, b - 1) / h : 0.0;
            return;
        }
        int ci = nk - grad, deg = 1;
        while (ci >= deg) {
            ci -= deg;
            ++deg;
        }
        int yd = ci, xd = deg - yd;
        double f = static_cast<double>(yd + 1) / xd;
        vx = -f * ipow(xm, xd - 1) * ipow(ym, yd + 1);
        vy = ipow(xm, xd) * ipow(ym, yd);
    }
    // Coefficients of m_nk restricted to edge p1->p2 in t in [-1/2,1/2]
    // (CPU/lib/polynomial_basis.cpp:163-203).
    void edge_coeffs(int nk, double xc, double yc, double h, double x1, double y1, double x2, double y2,
                     std::vector<double>& c) const {
        double ex = x2 - x1, ey = y2 - y1;
        double len = std::sqrt(ex * ex + ey * ey);
        int a, b;
        mindex(nk, a, b);
        int tot = a + b;
        c.assign(std::max(k + 1, tot + 1), 0.0);
        double mx = 0.5 * (x1 + x2), my = 0.5 * (y1 + y2);
        double tx = ex / len, ty = ey / len, ox = mx - xc, oy = my - yc;
        for (int i = 0; i <= a; ++i)
            for (int j = 0; j <= b; ++j)
                c[i + j] += comb(a, i) * comb(b, j) * ipow(ox, a - i) * ipow(oy, b - j) * ipow(tx, i) * ipow(ty, j);
        for (int d = 0; d <= tot; ++d) c[d] *= ipow(len, d) / ipow(h, tot);
        c.resize(k + 1);
    }
    // CPU/lib/polynomial_basis.cpp:205-226.
    static double edge_mono(int nk, double x0, double y0, double x1, double y1, double x, double y) {
        double ex = x1 - x0, ey = y1 - y0;
        double len = std::sqrt(ex * ex + ey * ey);
        double mx = 0.5 * (x0 + x1), my = 0.5 * (y0 + y1);
        double t = ((x - mx) * ex / len + (y - my) * ey / len) / len;
        return ipow(t, nk);
    }
};

// ---------------------------------------------------------------------------
// Mesh.  CPU/mesh/straight_mesh.{h,cpp}.
// ---------------------------------------------------------------------------

struct Mesh {
    int nn = 0, nc = 0, ne = 0;
    std::vector<double> xy;
    std::vector<int> off, nodes, npc;        // CCW cell nodes with offsets (size nc+1)
    std::vector<double> cx, cy, area, diam;  // CPU/mesh/straight_mesh.cpp:504-576
    std::vector<int> cell_edges;             // global edge of local edge j (same offsets)
    std::vector<int> edge_nodes;             // [min,max] per edge
    std::vector<int> edge_occ;               // 1 boundary, 2 interior
    // EXTENSION: circular arcs (centre at the origin) keyed by sorted node pair.
    std::map<std::pair<int, int>, double> arcs;
};

// CPU/mesh/straight_mesh.cpp:470-482: reverse if the signed area is negative.
static void ccw(std::vector<int>& v, const std::vector<double>& xy) {
    int n = (int)v.size();
    double a = 0.0;
    for (int i = 0; i < n; ++i) {
        int j = (i + 1) % n;
        a += xy[2 * v[i]] * xy[2 * v[j] + 1] - xy[2 * v[j]] * xy[2 * v[i] + 1];
    }
    if (a / 2.0 < 0) std::reverse(v.begin(), v.end());
}

// CPU/mesh/straight_mesh.cpp:504-576 and :584-660.
static void finalize(Mesh& m) {
    m.cx.assign(m.nc, 0.0);
    m.cy.assign(m.nc, 0.0);
    m.area.assign(m.nc, 0.0);
    m.diam.assign(m.nc, 0.0);
    for (int c = 0; c < m.nc; ++c) {
        int s = m.off[c], n = m.npc[c];
        std::vector<double> x(n), y(n);
        for (int i = 0; i < n; ++i) {
            x[i] = m.xy[2 * m.nodes[s + i]];
            y[i] = m.xy[2 * m.nodes[s + i] + 1];
        }
        if (n == 4) {
            double x0 = *std::min_element(x.begin(), x.end()), x1 = *std::max_element(x.begin(), x.end());
            double y0 = *std::min_element(y.begin(), y.end()), y1 = *std::max_element(y.begin(), y.end());
            m.cx[c] = (x0 + x1) / 2.0;
            m.cy[c] = (y0 + y1) / 2.0;
            m.area[c] = (x1 - x0) * (y1 - y0);
        } else {
            double sx = 0, sy = 0, a = 0;
            for (int i = 0; i < n; ++i) {
                sx += x[i];
                sy += y[i];
            }
            m.cx[c] = sx / n;
            m.cy[c] = sy / n;
            for (int i = 0; i < n; ++i) {
                int j = (i + 1) % n;
                a += x[i] * y[j] - x[j] * y[i];
            }
            m.area[c] = std::fabs(a) / 2.0;
        }
        double md = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                double dx = x[i] - x[j], dy = y[i] - y[j];
                md = std::max(md, std::sqrt(dx * dx + dy * dy));
            }
        m.diam[c] = md;
    }
    // Edges: sort (min,max) pairs, number uniquely in sorted order.
    std::vector<std::pair<std::pair<int, int>, int>> all;
    all.reserve(m.nodes.size());
    for (int c = 0; c < m.nc; ++c) {
        int s = m.off[c], n = m.npc[c];
        for (int j = 0; j < n; ++j) {
            int a = m.nodes[s + j], b = m.nodes[s + (j + 1) % n];
            all.push_back({{std::min(a, b), std::max(a, b)}, s + j});
        }
    }
    std::sort(all.begin(), all.end());
    m.cell_edges.assign(m.nodes.size(), -1);
    m.edge_nodes.clear();
    m.edge_occ.clear();
    int ne = 0;
    for (size_t i = 0; i < all.size(); ++i) {
        if (i == 0 || all[i].first != all[i - 1].first) {
            m.edge_nodes.push_back(all[i].first.first);
            m.edge_nodes.push_back(all[i].first.second);
            m.edge_occ.push_back(0);
            ++ne;
        }
        m.cell_edges[all[i].second] = ne - 1;
        m.edge_occ[ne - 1]++;
    }
    m.ne = ne;
}

static bool add_cell(Mesh& m, std::vector<int> v) {
    ccw(v, m.xy);
    m.npc.push_back((int)v.size());
    m.nodes.insert(m.nodes.end(), v.begin(), v.end());
    m.off.push_back((int)m.nodes.size());
    return true;
}

// Gmsh 4.1 reader: CPU/mesh/straight_mesh.cpp:215-451 (element types 2/3/11/12
// give 3/4/5/6 vertices; other types are skipped).
static Mesh* read_msh(const char* path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    Mesh* m = new Mesh();
    std::string line;
    while (std::getline(f, line))
        if (line.find("$Nodes") != std::string::npos) break;
    std::getline(f, line);
    int nblocks, nn;
    {
        std::istringstream hs(line);
        hs >> nblocks >> nn;
    }
    m->nn = nn;
    m->xy.assign(2 * nn, 0.0);
    for (int b = 0; b < nblocks; ++b) {
        std::getline(f, line);
        std::istringstream bs(line);
        int dim, tag, par, cnt;
        bs >> dim >> tag >> par >> cnt;
        std::vector<int> tags;
        while ((int)tags.size() < cnt) {
            std::getline(f, line);
            std::istringstream ts(line);
            int t;
            while ((int)tags.size() < cnt && ts >> t) tags.push_back(t);
        }
        for (int i = 0; i < cnt; ++i) {
            std::getline(f, line);
            std::istringstream cs(line);
            double x, y, z;
            if (!(cs >> x >> y >> z)) continue;
            int id = tags[i] - 1;
            if (id >= 0 && id < nn) {
                m->xy[2 * id] = x;
                m->xy[2 * id + 1] = y;
            }
        }
    }
    while (std::getline(f, line))
        if (line.find("$Elements") != std::string::npos) break;
    std::getline(f, line);
    {
        std::istringstream hs(line);
        hs >> nblocks;
    }
    m->off.push_back(0);
    for (int b = 0; b < nblocks; ++b) {
        std::getline(f, line);
        std::istringstream bs(line);
        int dim, tag, type, cnt;
        bs >> dim >> tag >> type >> cnt;
        int nv = type == 2 ? 3 : type == 3 ? 4 : type == 11 ? 5 : type == 12 ? 6 : -1;
        for (int i = 0; i < cnt; ++i) {
            std::getline(f, line);
            if (nv < 3) continue;
            std::istringstream es(line);
            int etag;
            es >> etag;
            std::vector<int> v(nv);
            for (int j = 0; j < nv; ++j) {
                es >> v[j];
                v[j] -= 1;
            }
            add_cell(*m, v);
        }
    }
    m->nc = (int)m->npc.size();
    if (m->nc == 0) {
        delete m;
        throw std::runtime_error("no polygon elements");
    }
    finalize(*m);
    return m;
}

// Legacy VTK polygon reader: CPU/mesh/straight_mesh.cpp:92-214.
static Mesh* read_vtk(const char* path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    auto seek = [&](const char* key) {
        std::string t;
        while (f >> t)
            if (t == key) return true;
        return false;
    };
    Mesh* m = new Mesh();
    if (!seek("POINTS")) throw std::runtime_error("POINTS section not found");
    std::string dt;
    f >> m->nn >> dt;
    m->xy.assign(2 * m->nn, 0.0);
    for (int i = 0; i < m->nn; ++i) {
        double z;
        f >> m->xy[2 * i] >> m->xy[2 * i + 1] >> z;
    }
    if (!seek("CELLS")) throw std::runtime_error("CELLS section not found");
    int nc, tot;
    f >> nc >> tot;
    m->off.push_back(0);
    for (int c = 0; c < nc; ++c) {
        int nv;
        f >> nv;
        if (nv < 3) throw std::runtime_error("cell with fewer than 3 vertices");
        std::vector<int> v(nv);
        for (int i = 0; i < nv; ++i) {
            f >> v[i];
            if (v[i] < 0 || v[i] >= m->nn) throw std::runtime_error("node index out of range");
        }
        add_cell(*m, v);
    }
    m->nc = nc;
    if (!seek("CELL_TYPES")) throw std::runtime_error("CELL_TYPES section not found");
    int nt;
    f >> nt;
    if (nt != nc) throw std::runtime_error("CELL_TYPES count mismatch");
    for (int c = 0; c < nc; ++c) {
        int t;
        f >> t;
        if (t != 5 && t != 7 && t != 9) throw std::runtime_error("unsupported VTK cell type");
    }
    finalize(*m);
    return m;
}

// ---------------------------------------------------------------------------
// Ear clipping.  CPU/mesh/polygon_triangulator.cpp:50-145.  The reference tests
// the area sum against an absolute 1e-6 and the convexity cross product against
// an absolute 1e-8; both thresholds are scaled here by the squared cell
// diameter (h^2 = 1 reproduces the reference).  On the reference's meshes the
// decisions are identical (pinned against the reference's own triangulator);
// on fine Voronoi meshes the absolute form rejects every ear.
// ---------------------------------------------------------------------------

struct Tri {
    int n[3];
    double v[6];
};

static bool in_tri(const std::vector<double>& p, int pi, int a, int b, int c, double h2) {
    double px = p[2 * pi], py = p[2 * pi + 1], ax = p[2 * a], ay = p[2 * a + 1];
    double bx = p[2 * b], by = p[2 * b + 1], cx = p[2 * c], cy = p[2 * c + 1];
    double at = 0.5 * std::fabs((bx - ax) * (cy - ay) - (cx - ax) * (by - ay));
    double a1 = 0.5 * std::fabs((ax - px) * (by - py) - (bx - px) * (ay - py));
    double a2 = 0.5 * std::fabs((bx - px) * (cy - py) - (cx - px) * (by - py));
    double a3 = 0.5 * std::fabs((cx - px) * (ay - py) - (ax - px) * (cy - py));
    return std::fabs((a1 + a2 + a3) - at) < 1e-6 * h2;
}

static bool is_ear(int i, const std::vector<double>& p, int n, double h2) {
    int pr = (i - 1 + n) % n, nx = (i + 1) % n;
    double xi = p[2 * i], yi = p[2 * i + 1];
    double xp = p[2 * pr], yp = p[2 * pr + 1], xn = p[2 * nx], yn = p[2 * nx + 1];
    double cr = (xn - xi) * (yp - yi) - (xp - xi) * (yn - yi);
    if (cr <= 1e-8 * h2) return false;
    for (int j = 0; j < n; ++j) {
        if (j == i || j == pr || j == nx) continue;
        if (in_tri(p, j, pr, i, nx, h2)) return false;
    }
    return true;
}

static std::vector<Tri> ear_clip(const Mesh& m, int c, double* total_area) {
    int n = m.npc[c], s = m.off[c];
    std::vector<int> idx(n);
    std::vector<double> p(2 * n);
    for (int i = 0; i < n; ++i) {
        idx[i] = m.nodes[s + i];
        p[2 * i] = m.xy[2 * idx[i]];
        p[2 * i + 1] = m.xy[2 * idx[i] + 1];
    }
    std::vector<Tri> out;
    auto push = [&](int a, int b, int cc) {
        Tri t;
        t.n[0] = idx[a];
        t.n[1] = idx[b];
        t.n[2] = idx[cc];
        double v[6] = {p[2 * a], p[2 * a + 1], p[2 * b], p[2 * b + 1], p[2 * cc], p[2 * cc + 1]};
        std::memcpy(t.v, v, sizeof v);
        out.push_back(t);
    };
    double h2 = m.diam[c] * m.diam[c];
    int rem = n;
    while (rem > 3) {
        bool found = false;
        for (int i = 0; i < rem; ++i) {
            if (is_ear(i, p, rem, h2)) {
                push((i - 1 + rem) % rem, i, (i + 1) % rem);
                p.erase(p.begin() + 2 * i, p.begin() + 2 * (i + 1));
                idx.erase(idx.begin() + i);
                --rem;
                found = true;
                break;
            }
        }
        if (!found) throw std::runtime_error("No valid ear found for polygon");
    }
    push(0, 1, 2);
    double ta = 0.0;
    for (const Tri& t : out)
        ta += 0.5 * std::fabs((t.v[2] - t.v[0]) * (t.v[5] - t.v[1]) - (t.v[4] - t.v[0]) * (t.v[3] - t.v[1]));
    if (total_area) *total_area = ta;
    return out;
}

// ---------------------------------------------------------------------------
// Problem data.  CPU/examples/darcy_problem.{h,cpp}.
// ---------------------------------------------------------------------------

enum { MAP_SIN = 0, MAP_ARCS = 1 };
enum { SOL_SINCOS = 0, SOL_SINSIN = 1 };
enum { BC_FLUX_MEANZERO = 0, BC_DIRICHLET_P = 1 };

// Caller-defined data: the DarcyPdeData fields (CPU/examples/darcy_problem.h:
// 162-180) and the mapping's physical_coords + jacobian (:39-63) as C
// callbacks; a null member keeps the built-in choice.
struct OrcFuncs {
    void* ctx;
    void (*kappa)(void*, double, double, double*);           // [xx, xy, yx, yy]
    double (*solution_p)(void*, double, double);
    void (*solution_u)(void*, double, double, double*);
    double (*source_g)(void*, double, double);
    double (*dirichlet_p)(void*, double, double);
    double (*neumann_un)(void*, double, double, double, double);
    void (*map)(void*, double, double, double*, double*);   // F, J = [dx/dxi, dx/deta, dy/dxi, dy/deta]
};

struct Problem {
    int map_kind = MAP_SIN;
    double eps = 0.05;  // SinPerturbationMapping amplitude (darcy_problem.h:55-96)
    int sol = SOL_SINCOS;
    int bc = BC_FLUX_MEANZERO;
    std::vector<double> theta;  // EXTENSION: per-cell log-permeability
    OrcFuncs f{};               // caller-defined data (all null: built-in)
};

struct Geo {  // map value at a computational point
    double x, y;                  // physical point F(xi)
    double j00, j01, j10, j11;    // J = dF/dxi (row-major, j01 = dx/deta)
    double det;
};

// SinPerturbationMapping (CPU/examples/darcy_problem.h:55-96).
static Geo sin_map(double eps, double xi, double eta) {
    Geo g;
    g.x = xi + eps * std::sin(2.0 * M_PI * eta);
    g.y = eta + eps * std::sin(2.0 * M_PI * xi);
    g.j00 = 1.0;
    g.j01 = eps * 2.0 * M_PI * std::cos(2.0 * M_PI * eta);
    g.j10 = eps * 2.0 * M_PI * std::cos(2.0 * M_PI * xi);
    g.j11 = 1.0;
    g.det = g.j00 * g.j11 - g.j01 * g.j10;
    return g;
}

// EXTENSION: per-cell chord-to-arc map.  For a boundary chord A->B (local
// vertices ia, ib of the cell) of a circle of radius R centred at the origin:
//   F(xi) = xi + S^2 delta(s),  S = l_A + l_B,  s = l_B / S,
//   delta(s) = R (cos th(s), sin th(s)) - (A + s (B - A)),  th linear in s,
// with l the Wachspress coordinates of the (convex) cell.  On AB, S = 1 and
// s is the chord parameter, so F maps the chord onto the arc; on every other
// edge F is the identity; F is smooth inside the cell (a map that is only
// piecewise smooth, e.g. per sub-triangle, costs half an order of flux accuracy).
struct Chord {
    double ax, ay, bx, by, R;
    int ia, ib;
};

// Wachspress coordinates and gradients in product form (no division by the
// edge areas, so edge points are fine): l_i = A_i P_i / D.
static void wachspress(const std::vector<double>& vx, const std::vector<double>& vy, double x, double y, double* lam,
                       double* gl) {
    int n = (int)vx.size();
    double a[64], ga[64][2], Ac[64], P[64], gP[64][2];
    for (int i = 0; i < n; ++i) {
        int j = (i + 1) % n;
        a[i] = 0.5 * ((vx[i] - x) * (vy[j] - y) - (vy[i] - y) * (vx[j] - x));
        ga[i][0] = 0.5 * (vy[i] - vy[j]);
        ga[i][1] = 0.5 * (vx[j] - vx[i]);
    }
    double D = 0.0, gD[2] = {0.0, 0.0};
    for (int i = 0; i < n; ++i) {
        int p = (i - 1 + n) % n, q = (i + 1) % n;
        Ac[i] = 0.5 * ((vx[i] - vx[p]) * (vy[q] - vy[p]) - (vy[i] - vy[p]) * (vx[q] - vx[p]));
        P[i] = 1.0;
        gP[i][0] = gP[i][1] = 0.0;
        for (int m = 0; m < n; ++m) {
            if (m == p || m == i) continue;
            // product rule: d(P * a_m) = dP a_m + P da_m
            gP[i][0] = gP[i][0] * a[m] + P[i] * ga[m][0];
            gP[i][1] = gP[i][1] * a[m] + P[i] * ga[m][1];
            P[i] *= a[m];
        }
        D += Ac[i] * P[i];
        gD[0] += Ac[i] * gP[i][0];
        gD[1] += Ac[i] * gP[i][1];
    }
    for (int i = 0; i < n; ++i) {
        lam[i] = Ac[i] * P[i] / D;
        gl[2 * i] = (Ac[i] * gP[i][0] - lam[i] * gD[0]) / D;
        gl[2 * i + 1] = (Ac[i] * gP[i][1] - lam[i] * gD[1]) / D;
    }
}

static void arc_term(const Chord& ch, const double* lam, const double* gl, double* F, double* J) {
    double lA = lam[ch.ia], lB = lam[ch.ib];
    double S = lA + lB;
    if (S <= 1e-300) return;
    const double* gA = &gl[2 * ch.ia];
    const double* gB = &gl[2 * ch.ib];
    double s = lB / S;
    double tA = std::atan2(ch.ay, ch.ax), tB = std::atan2(ch.by, ch.bx);
    double dt = tB - tA;
    while (dt > M_PI) dt -= 2 * M_PI;
    while (dt <= -M_PI) dt += 2 * M_PI;
    double th = tA + s * dt;
    double del[2] = {ch.R * std::cos(th) - (ch.ax + s * (ch.bx - ch.ax)), ch.R * std::sin(th) - (ch.ay + s * (ch.by - ch.ay))};
    double ddel[2] = {-ch.R * dt * std::sin(th) - (ch.bx - ch.ax), ch.R * dt * std::cos(th) - (ch.by - ch.ay)};
    // grad (S^2 delta(s)) = 2 S delta grad S + delta'(s) (l_A grad l_B - l_B grad l_A)
    double gS[2] = {gA[0] + gB[0], gA[1] + gB[1]};
    double w[2] = {lA * gB[0] - lB * gA[0], lA * gB[1] - lB * gA[1]};
    F[0] += S * S * del[0];
    F[1] += S * S * del[1];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) J[2 * i + j] += 2.0 * S * del[i] * gS[j] + ddel[i] * w[j];
}

struct Ctx;
static Geo eval_map(const Ctx& X, int cell, const Tri* tri, double xi, double eta);

// ---------------------------------------------------------------------------
// Small dense linear algebra (row-major).
// ---------------------------------------------------------------------------

typedef std::vector<double> Mat;

// Gauss-Jordan inverse with partial pivoting, singular below 1e-14
// (CPU/core/petsc_utils.cpp:64-138).
static Mat inverse(const Mat& A, int n) {
    std::vector<double> a(n * 2 * n, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) a[i * 2 * n + j] = A[i * n + j];
        a[i * 2 * n + n + i] = 1.0;
    }
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(a[i * 2 * n + k]) > std::fabs(a[piv * 2 * n + k])) piv = i;
        if (piv != k)
            for (int j = 0; j < 2 * n; ++j) std::swap(a[k * 2 * n + j], a[piv * 2 * n + j]);
        if (std::fabs(a[k * 2 * n + k]) < 1e-14) throw std::runtime_error("Matrix is singular, cannot invert");
        for (int i = k + 1; i < n; ++i) {
            double f = a[i * 2 * n + k] / a[k * 2 * n + k];
            for (int j = k; j < 2 * n; ++j) a[i * 2 * n + j] -= f * a[k * 2 * n + j];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double dv = a[k * 2 * n + k];
        for (int j = k; j < 2 * n; ++j) a[k * 2 * n + j] /= dv;
        for (int i = 0; i < k; ++i) {
            double f = a[i * 2 * n + k];
            for (int j = k; j < 2 * n; ++j) a[i * 2 * n + j] -= f * a[k * 2 * n + j];
        }
    }
    Mat r(n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) r[i * n + j] = a[i * 2 * n + n + j];
    return r;
}

static Mat matmul(const Mat& A, const Mat& B, int m, int k, int n, double s = 1.0) {
    Mat C(m * n, 0.0);
    for (int i = 0; i < m; ++i)
        for (int l = 0; l < k; ++l) {
            double a = A[i * k + l];
            for (int j = 0; j < n; ++j) C[i * n + j] += a * B[l * n + j];
        }
    if (s != 1.0)
        for (double& v : C) v *= s;
    return C;
}

static Mat transpose(const Mat& A, int m, int n) {
    Mat T(n * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) T[j * m + i] = A[i * n + j];
    return T;
}

// ---------------------------------------------------------------------------
// Context: mesh + problem + discretisation parameters.
// ---------------------------------------------------------------------------

struct Csr {
    int n = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
};

struct Ctx {
    const Mesh* m;
    Problem pb;
    int k, nq2, nq1;
    Basis basis;
    TriRule rule;
    int internal, q;
    std::vector<int> nflux;  // per cell local flux dofs
    int Nf, Np, Ntot;        // totals; Ntot includes lambda for BC_FLUX_MEANZERO
    std::vector<std::vector<Chord>> chords;  // per cell
    // assembled system
    Csr A;
    std::vector<double> rhs, ub;  // full rhs, boundary values
    std::vector<char> isb;        // essential boundary dof
    std::vector<int> free_idx;    // free dofs (reduced system)
    Csr Ar;
    std::vector<double> rr;       // reduced rhs
    std::vector<double> sol;      // full solution
    bool have_sys = false, have_sol = false;
    Ctx(const Mesh* mm, const Problem& p, int kk, int a, int b)
        : m(mm), pb(p), k(kk), nq2(a), nq1(b), basis(kk), rule(triangle_rule(a)) {
        if (k < 0 || k > 4) throw std::invalid_argument("polynomial degree k must be between 0 and 4");
        if (k == 3 && nq2 != 12) throw std::invalid_argument("k=3 requires 12 triangle quadrature points");
        std::vector<double> x, w;
        gauss_legendre(nq1, x, w);
        // CPU/solver/HdivMatrix.cpp:58-88
        internal = (dimP(k) - 1) + (dimP(k) - (k + 1));
        q = dimP(k);
        nflux.resize(m->nc);
        for (int c = 0; c < m->nc; ++c) nflux[c] = (k + 1) * m->npc[c] + internal;
        Nf = m->ne * (k + 1) + m->nc * internal;
        Np = q * m->nc;
        Ntot = Nf + Np + (pb.bc == BC_FLUX_MEANZERO ? 1 : 0);
        if (!pb.theta.empty() && (int)pb.theta.size() != m->nc) throw std::invalid_argument("theta size");
        chords.resize(m->nc);
        if (pb.map_kind == MAP_ARCS) {
            for (int c = 0; c < m->nc; ++c) {
                int s = m->off[c], n = m->npc[c];
                for (int j = 0; j < n; ++j) {
                    int a0 = m->nodes[s + j], a1 = m->nodes[s + (j + 1) % n];
                    auto it = m->arcs.find({std::min(a0, a1), std::max(a0, a1)});
                    if (it == m->arcs.end()) continue;
                    chords[c].push_back({m->xy[2 * a0], m->xy[2 * a0 + 1], m->xy[2 * a1], m->xy[2 * a1 + 1], it->second, j, (j + 1) % n});
                }
            }
        }
    }
    double kappa_scale(int c) const { return pb.theta.empty() ? 1.0 : std::exp(pb.theta[c]); }
};

static Geo eval_map(const Ctx& X, int cell, const Tri* tri, double xi, double eta) {
    if (X.pb.f.map) {  // IsoparametricMapping::physical_coords and jacobian
        double F[2], J[4];
        X.pb.f.map(X.pb.f.ctx, xi, eta, F, J);
        Geo g;
        g.x = F[0];
        g.y = F[1];
        g.j00 = J[0];
        g.j01 = J[1];
        g.j10 = J[2];
        g.j11 = J[3];
        g.det = J[0] * J[3] - J[1] * J[2];
        return g;
    }
    if (X.pb.map_kind == MAP_SIN) return sin_map(X.pb.eps, xi, eta);
    (void)tri;
    Geo g;
    double F[2] = {xi, eta}, J[4] = {1.0, 0.0, 0.0, 1.0};
    if (!X.chords[cell].empty()) {
        const Mesh& m = *X.m;
        int n = m.npc[cell], s0 = m.off[cell];
        std::vector<double> vx(n), vy(n);
        for (int i = 0; i < n; ++i) {
            vx[i] = m.xy[2 * m.nodes[s0 + i]];
            vy[i] = m.xy[2 * m.nodes[s0 + i] + 1];
        }
        double lam[64], gl[128];
        wachspress(vx, vy, xi, eta, lam, gl);
        for (const Chord& ch : X.chords[cell]) arc_term(ch, lam, gl, F, J);
    }
    g.x = F[0];
    g.y = F[1];
    g.j00 = J[0];
    g.j01 = J[1];
    g.j10 = J[2];
    g.j11 = J[3];
    g.det = J[0] * J[3] - J[1] * J[2];
    return g;
}

// Physical exact solution (CPU/examples/darcy_problem.cpp:116-131; the
// sin*sin variant is the manufactured solution of the benchmark workloads).
static void exact(int sol, double x, double y, double* p, double* ux, double* uy, double* g) {
    if (sol == SOL_SINCOS) {
        *p = std::sin(M_PI * x) * std::cos(M_PI * y);
        double dpdx = M_PI * std::cos(M_PI * x) * std::cos(M_PI * y);
        double dpdy = -M_PI * std::sin(M_PI * x) * std::sin(M_PI * y);
        *ux = -dpdx;
        *uy = -dpdy;
        *g = 2.0 * M_PI * M_PI * std::sin(M_PI * x) * std::cos(M_PI * y);
    } else {
        *p = std::sin(M_PI * x) * std::sin(M_PI * y);
        double dpdx = M_PI * std::cos(M_PI * x) * std::sin(M_PI * y);
        double dpdy = M_PI * std::sin(M_PI * x) * std::cos(M_PI * y);
        *ux = -dpdx;
        *uy = -dpdy;
        *g = 2.0 * M_PI * M_PI * (*p);
    }
}

// solution_p, solution_u, source_g at a physical point: the caller's where
// given, else the built-in solution.
static void pde_at(const Ctx& X, double x, double y, double* p, double* ux, double* uy, double* g) {
    exact(X.pb.sol, x, y, p, ux, uy, g);
    const OrcFuncs& f = X.pb.f;
    if (f.solution_p) *p = f.solution_p(f.ctx, x, y);
    if (f.solution_u) {
        double u[2];
        f.solution_u(f.ctx, x, y, u);
        *ux = u[0];
        *uy = u[1];
    }
    if (f.source_g) *g = f.source_g(f.ctx, x, y);
}

// Pulled-back data (CPU/examples/darcy_problem.cpp:37-83).  The data (source,
// boundary pressure, boundary normal flux) are those of the K = I manufactured
// solution whatever theta is: with a cellwise permeability the problem stays
// well posed but has no closed-form solution.
static void comp_data(const Ctx& X, int cell, const Geo& G, double* ph, double* uhx, double* uhy, double* gh) {
    (void)cell;
    double p, ux, uy, g;
    pde_at(X, G.x, G.y, &p, &ux, &uy, &g);
    *ph = p;
    // u_hat = det J * J^{-1} u
    double inv00 = G.j11 / G.det, inv01 = -G.j01 / G.det, inv10 = -G.j10 / G.det, inv11 = G.j00 / G.det;
    *uhx = G.det * (inv00 * ux + inv01 * uy);
    *uhy = G.det * (inv10 * ux + inv11 * uy);
    *gh = G.det * g;
}

// K_hat^{-1} = (det J J^{-1} K J^{-T})^{-1} = J^T K^{-1} J / det J  (K = e^theta I).
static void kinv_hat(const Ctx& X, int cell, const Geo& G, double* T) {
    // follow the reference literally: build K_hat then invert it
    // (CPU/examples/darcy_problem.cpp:15-29,72-83; CPU/solver/DarcySolver.cpp:12-25)
    double ks = X.kappa_scale(cell);
    // physical tensor: the caller's kappa(x, y) (times exp(theta_E)) or e^theta I
    double K[4] = {1.0, 0.0, 0.0, 1.0};
    if (X.pb.f.kappa) X.pb.f.kappa(X.pb.f.ctx, G.x, G.y, K);
    for (double& v : K) v *= ks;
    double i00 = G.j11 / G.det, i01 = -G.j01 / G.det, i10 = -G.j10 / G.det, i11 = G.j00 / G.det;
    // left = J^{-1} K, then det J * left * J^{-T} (transform_tensor)
    double lxx = i00 * K[0] + i01 * K[2], lxy = i00 * K[1] + i01 * K[3];
    double lyx = i10 * K[0] + i11 * K[2], lyy = i10 * K[1] + i11 * K[3];
    double kxx = G.det * (lxx * i00 + lxy * i01), kxy = G.det * (lxx * i10 + lxy * i11);
    double kyx = G.det * (lyx * i00 + lyy * i01), kyy = G.det * (lyx * i10 + lyy * i11);
    double d = kxx * kyy - kxy * kyx;
    if (!std::isfinite(d) || std::fabs(d) < 1e-14) throw std::runtime_error("computational permeability tensor is singular");
    T[0] = kyy / d;
    T[1] = -kxy / d;
    T[2] = -kyx / d;
    T[3] = kxx / d;
}

// ---------------------------------------------------------------------------
// Element matrices.  CPU/solver/HdivMatrix.cpp and CPU/solver/DarcySolver.cpp.
// ---------------------------------------------------------------------------

struct Elem {
    int n;                   // local flux dofs
    int ne;                  // edges
    Mat D, G, W, H, Hs, B, P, Pd, GK, A;
    std::vector<double> L, f;  // Lagrange vector, pressure rhs
    double alpha, tri_area;
    std::vector<Tri> tris;
};

static void cell_xy(const Mesh& m, int c, int j, double& x, double& y) {
    int v = m.nodes[m.off[c] + j];
    x = m.xy[2 * v];
    y = m.xy[2 * v + 1];
}

static Elem element(const Ctx& X, int c) {
    const Mesh& m = *X.m;
    const Basis& bs = X.basis;
    Basis bk1(X.k + 1);
    Elem e;
    int ne = m.npc[c], k = X.k, p = bs.p, q = bs.q, gr = bs.grad;
    int n = X.nflux[c], edof = (k + 1) * ne;
    e.n = n;
    e.ne = ne;
    double xc = m.cx[c], yc = m.cy[c], h = m.diam[c];
    e.tris = ear_clip(m, c, &e.tri_area);
    std::vector<QP> vol, tmp;
    std::vector<int> vol_tri;
    for (size_t t = 0; t < e.tris.size(); ++t) {
        map_triangle(X.rule, e.tris[t].v, tmp);
        for (auto& qp : tmp) {
            vol.push_back(qp);
            vol_tri.push_back((int)t);
        }
    }
    // D (CPU/solver/HdivMatrix.cpp:94-172,396-412)
    e.D.assign(n * p, 0.0);
    for (int j = 0; j < ne; ++j) {
        double x1, y1, x2, y2;
        cell_xy(m, c, j, x1, y1);
        cell_xy(m, c, (j + 1) % ne, x2, y2);
        double dx = x2 - x1, dy = y2 - y1, len = std::sqrt(dx * dx + dy * dy);
        if (!(len > 1e-14)) throw std::invalid_argument("degenerate edge");
        double nx = dy / len, ny = -dx / len;  // CPU/solver/HdivMatrix.cpp:38-56
        std::vector<QP> lp;
        line_points(X.nq1, x1, y1, x2, y2, lp);
        for (int r = 0; r <= k; ++r) {
            int row = r * ne + j;
            for (int a = 0; a < p; ++a) {
                double s = 0.0;
                for (auto& qp : lp) {
                    double vx, vy;
                    bs.hdiv(a, xc, yc, h, qp.x, qp.y, vx, vy);
                    s += qp.w * (vx * nx + vy * ny) * Basis::edge_mono(r, x1, y1, x2, y2, qp.x, qp.y);
                }
                e.D[row * p + a] = s;
            }
        }
    }
    for (int i = edof; i < n; ++i) {
        int mom = i < edof + (q - 1) ? (i - edof) : (i - edof - (q - 1)) + gr;
        for (int a = 0; a < p; ++a) {
            double s = 0.0;
            for (auto& qp : vol) {
                double vx, vy, mx, my;
                bs.hdiv(a, xc, yc, h, qp.x, qp.y, vx, vy);
                bs.hdiv(mom, xc, yc, h, qp.x, qp.y, mx, my);
                s += (vx * mx + vy * my) * qp.w;
            }
            e.D[i * p + a] = s;
        }
    }
    // G (CPU/solver/HdivMatrix.cpp:174-200)
    e.G.assign(p * p, 0.0);
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b) {
            double s = 0.0;
            for (auto& qp : vol) {
                double ax, ay, bx, by;
                bs.hdiv(a, xc, yc, h, qp.x, qp.y, ax, ay);
                bs.hdiv(b, xc, yc, h, qp.x, qp.y, bx, by);
                s += (ax * bx + ay * by) * qp.w;
            }
            e.G[a * p + b] = s;
        }
    // W (CPU/solver/HdivMatrix.cpp:202-246,430-451): exact, no quadrature
    e.W.assign(q * n, 0.0);
    for (int r = 0; r < q; ++r)
        for (int i = 0; i < n; ++i) {
            double v = 0.0;
            int gd = i - edof;
            if (gd >= 0 && gd == r - 1) v = -1.0;
            if (i < edof) {
                int j = i % ne, deg = i / ne;
                double x1, y1, x2, y2;
                cell_xy(m, c, j, x1, y1);
                cell_xy(m, c, (j + 1) % ne, x2, y2);
                std::vector<double> co;
                bs.edge_coeffs(r, xc, yc, h, x1, y1, x2, y2, co);
                v += co[deg];
            }
            e.W[r * n + i] = v;
        }
    // H, H* (CPU/solver/HdivMatrix.cpp:248-273,453-485)
    e.H.assign(q * q, 0.0);
    e.Hs.assign(gr * q, 0.0);
    for (int r = 0; r < q; ++r)
        for (int s2 = 0; s2 < q; ++s2) {
            double s = 0.0;
            for (auto& qp : vol) s += bk1.mono(r, xc, yc, h, qp.x, qp.y) * bk1.mono(s2, xc, yc, h, qp.x, qp.y) * qp.w;
            e.H[r * q + s2] = s;
        }
    for (int a = 0; a < gr; ++a)
        for (int s2 = 0; s2 < q; ++s2) {
            double s = 0.0;
            for (auto& qp : vol)
                s += bk1.mono(a + 1, xc, yc, h, qp.x, qp.y) * bk1.mono(s2, xc, yc, h, qp.x, qp.y) * qp.w;
            e.Hs[a * q + s2] = s;
        }
    // B_1 = -H* H^{-1} W (CPU/solver/HdivMatrix.cpp:579-585)
    Mat Hi = inverse(e.H, q);
    Mat B1 = matmul(e.Hs, matmul(Hi, e.W, q, q, n), gr, q, n, -1.0);
    // B_2 = sum_e H_e* H_e^{-1} W_e scattered to edge columns (CPU/solver/HdivMatrix.cpp:545-577)
    Mat B2(gr * n, 0.0);
    int kk = k + 1;
    for (int j = 0; j < ne; ++j) {
        double x1, y1, x2, y2;
        cell_xy(m, c, j, x1, y1);
        cell_xy(m, c, (j + 1) % ne, x2, y2);
        std::vector<QP> lp;
        line_points(X.nq1, x1, y1, x2, y2, lp);
        Mat He(kk * kk, 0.0), Hes(gr * kk, 0.0), We(kk * kk, 0.0);
        for (int r = 0; r < kk; ++r)
            for (int s2 = 0; s2 < kk; ++s2) {
                double s = 0.0;
                for (auto& qp : lp)
                    s += Basis::edge_mono(r, x1, y1, x2, y2, qp.x, qp.y) * Basis::edge_mono(s2, x1, y1, x2, y2, qp.x, qp.y) * qp.w;
                He[r * kk + s2] = s;
                We[r * kk + s2] = (s2 % ne) == r ? 1.0 : 0.0;  // CPU/solver/HdivMatrix.cpp:275-282,507-522
            }
        for (int a = 0; a < gr; ++a)
            for (int cc = 0; cc < kk; ++cc) {
                double s = 0.0;
                for (auto& qp : lp)
                    s += bk1.mono(a + 1, xc, yc, h, qp.x, qp.y) * Basis::edge_mono(cc, x1, y1, x2, y2, qp.x, qp.y) * qp.w;
                Hes[a * kk + cc] = s;
            }
        Mat R = matmul(Hes, matmul(inverse(He, kk), We, kk, kk, kk), gr, kk, kk);
        for (int cc = 0; cc < kk; ++cc)
            for (int a = 0; a < gr; ++a) B2[a * n + cc * ne + j] += R[a * kk + cc];
    }
    // B = [B_1 + B_2 ; B_curl] (CPU/solver/HdivMatrix.cpp:587-639)
    e.B.assign(p * n, 0.0);
    for (int a = 0; a < gr; ++a)
        for (int i = 0; i < n; ++i) e.B[a * n + i] = B1[a * n + i] + B2[a * n + i];
    for (int r = 0; r < bs.curl; ++r)
        for (int i = 0; i < n; ++i) {
            int cd = i - edof - (q - 1);
            e.B[(gr + r) * n + i] = (cd >= 0 && cd == r) ? 1.0 : 0.0;
        }
    // Projections (CPU/solver/HdivMatrix.cpp:641-652)
    e.P = matmul(inverse(e.G, p), e.B, p, p, n);
    e.Pd = matmul(e.D, e.P, n, p, n);
    // G_Kappa and alpha (CPU/solver/DarcySolver.cpp:1133-1226,1454-1474)
    e.GK.assign(p * p, 0.0);
    double Ti[4] = {0, 0, 0, 0};
    std::vector<Geo> geos(vol.size());
    for (size_t l = 0; l < vol.size(); ++l) geos[l] = eval_map(X, c, &e.tris[vol_tri[l]], vol[l].x, vol[l].y);
    for (size_t l = 0; l < vol.size(); ++l) {
        double T[4];
        kinv_hat(X, c, geos[l], T);
        for (int z = 0; z < 4; ++z) Ti[z] += T[z] * vol[l].w;
    }
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b) {
            double s = 0.0;
            for (size_t l = 0; l < vol.size(); ++l) {
                double T[4], ax, ay, bx, by;
                kinv_hat(X, c, geos[l], T);
                bs.hdiv(a, xc, yc, h, vol[l].x, vol[l].y, ax, ay);
                bs.hdiv(b, xc, yc, h, vol[l].x, vol[l].y, bx, by);
                s += (ax * (T[0] * bx + T[1] * by) + ay * (T[2] * bx + T[3] * by)) * vol[l].w;
            }
            e.GK[a * p + b] = s;
        }
    if (!(e.tri_area > 0.0)) throw std::runtime_error("element area must be positive and finite");
    double av[4] = {Ti[0] / e.tri_area, Ti[1] / e.tri_area, Ti[2] / e.tri_area, Ti[3] / e.tri_area};
    e.alpha = std::sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2] + av[3] * av[3]);
    // A = P^T G_K P + alpha (I - Pd)^T (I - Pd) (CPU/solver/DarcySolver.cpp:1228-1269,1476-1509)
    Mat Pt = transpose(e.P, p, n);
    Mat Kac = matmul(Pt, matmul(e.GK, e.P, p, p, n), n, p, n);
    Mat Imp(n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Imp[i * n + j] = (i == j ? 1.0 : 0.0) - e.Pd[i * n + j];
    Mat Kas = matmul(transpose(Imp, n, n), Imp, n, n, n, e.alpha);
    e.A.resize(n * n);
    for (int i = 0; i < n * n; ++i) e.A[i] = Kac[i] + Kas[i];
    // Lagrange vector and source rhs (CPU/solver/DarcySolver.cpp:1271-1452)
    e.L.assign(q, 0.0);
    e.f.assign(q, 0.0);
    for (int r = 0; r < q; ++r) {
        double sl = 0.0, sf = 0.0;
        for (size_t l = 0; l < vol.size(); ++l) {
            double mr = bs.mono(r, xc, yc, h, vol[l].x, vol[l].y);
            double ph, ux, uy, gh;
            comp_data(X, c, geos[l], &ph, &ux, &uy, &gh);
            sf += gh * mr * vol[l].w;
            sl += mr * geos[l].det * vol[l].w;
        }
        e.L[r] = sl;
        e.f[r] = sf;
    }
    return e;
}

// Local-to-global map (CPU/solver/DarcySolver.cpp:182-227) and direction signs
// (CPU/solver/DarcySolver.cpp:229-320).
static void dofmap(const Ctx& X, int c, std::vector<int>& g, std::vector<double>& dir) {
    const Mesh& m = *X.m;
    int ne = m.npc[c], k = X.k, n = X.nflux[c], q = X.q;
    g.assign(n + q + 1, -1);
    dir.assign(n + q + 1, 1.0);
    int s = m.off[c];
    for (int d = 0; d <= k; ++d)
        for (int j = 0; j < ne; ++j) {
            int ge = m.cell_edges[s + j];
            g[d * ne + j] = d * m.ne + ge;
            if (m.edge_occ[ge] == 1) continue;
            int l0 = m.nodes[s + j], l1 = m.nodes[s + (j + 1) % ne];
            bool rev = l0 == m.edge_nodes[2 * ge + 1] && l1 == m.edge_nodes[2 * ge];
            if (rev && d % 2 == 0) dir[d * ne + j] = -1.0;
        }
    int li = (k + 1) * ne;
    int is = m.ne * (k + 1) + c * X.internal;
    for (int i = 0; i < X.internal; ++i) g[li++] = is + i;
    int ps = X.Nf + c * q;
    for (int i = 0; i < q; ++i) g[li++] = ps + i;
    g[li] = X.Nf + X.Np;  // lambda (unused for Dirichlet p)
}

static void coo_to_csr(int n, std::vector<std::pair<long long, double>>& coo, Csr& A) {
    std::sort(coo.begin(), coo.end(), [](const std::pair<long long, double>& a, const std::pair<long long, double>& b) {
        return a.first < b.first;
    });
    A.n = n;
    A.rp.assign(n + 1, 0);
    A.ci.clear();
    A.v.clear();
    long long last = -1;
    for (auto& e : coo) {
        if (e.first != last) {
            A.ci.push_back((int)(e.first % n));
            A.v.push_back(e.second);
            A.rp[e.first / n + 1]++;
            last = e.first;
        } else {
            A.v.back() += e.second;
        }
    }
    for (int i = 0; i < n; ++i) A.rp[i + 1] += A.rp[i];
}

static void csr_mv(const Csr& A, const double* x, double* y) {
    for (int i = 0; i < A.n; ++i) {
        double s = 0.0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) s += A.v[p] * x[A.ci[p]];
        y[i] = s;
    }
}

// Global assembly (CPU/solver/DarcySolver.cpp:402-500), boundary treatment
// (CPU/solver/DarcySolver.cpp:502-663) and, for the Dirichlet-pressure
// EXTENSION, the natural boundary term -int p_D (phi . n) on boundary edges.
static void assemble(Ctx& X) {
    const Mesh& m = *X.m;
    int N = X.Ntot;
    bool lam = X.pb.bc == BC_FLUX_MEANZERO;
    std::vector<std::pair<long long, double>> coo;
    for (int i = 0; i < N; ++i) coo.push_back({(long long)i * N + i, 0.0});
    X.rhs.assign(N, 0.0);
    for (int c = 0; c < m.nc; ++c) {
        Elem e = element(X, c);
        int n = e.n, q = X.q, nl = n + q + (lam ? 1 : 0);
        std::vector<int> g;
        std::vector<double> dir;
        dofmap(X, c, g, dir);
        Mat loc(nl * nl, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) loc[i * nl + j] = e.A[i * n + j];
        for (int r = 0; r < q; ++r)
            for (int i = 0; i < n; ++i) {
                loc[i * nl + n + r] = -e.W[r * n + i];
                loc[(n + r) * nl + i] = e.W[r * n + i];
            }
        if (lam)
            for (int r = 0; r < q; ++r) {
                loc[(n + r) * nl + n + q] = e.L[r];
                loc[(n + q) * nl + n + r] = e.L[r];
            }
        for (int i = 0; i < nl; ++i)
            for (int j = 0; j < nl; ++j)
                coo.push_back({(long long)g[i] * N + g[j], loc[i * nl + j] * dir[i] * dir[j]});
        for (int r = 0; r < q; ++r) X.rhs[g[n + r]] += e.f[r] * dir[n + r];
        if (X.pb.bc == BC_DIRICHLET_P) {
            int s = m.off[c], ne = m.npc[c], kk = X.k + 1;
            for (int j = 0; j < ne; ++j) {
                int ge = m.cell_edges[s + j];
                if (m.edge_occ[ge] != 1) continue;
                double x1, y1, x2, y2;
                cell_xy(m, c, j, x1, y1);
                cell_xy(m, c, (j + 1) % ne, x2, y2);
                std::vector<QP> lp;
                line_points(X.nq1, x1, y1, x2, y2, lp);
                Mat He(kk * kk, 0.0);
                std::vector<double> mom(kk, 0.0);
                for (int r = 0; r < kk; ++r) {
                    for (int s2 = 0; s2 < kk; ++s2) {
                        double v = 0.0;
                        for (auto& qp : lp)
                            v += Basis::edge_mono(r, x1, y1, x2, y2, qp.x, qp.y) * Basis::edge_mono(s2, x1, y1, x2, y2, qp.x, qp.y) * qp.w;
                        He[r * kk + s2] = v;
                    }
                    for (auto& qp : lp) {
                        Geo G = eval_map(X, c, &e.tris[0], qp.x, qp.y);
                        // the chord point maps onto the arc whatever sub-triangle is used:
                        // on the chord the blend weight is 1 for its own triangle; use the
                        // triangle that holds this edge
                        for (auto& t : e.tris) {
                            int hit = 0;
                            for (int z = 0; z < 3; ++z)
                                if ((t.v[2 * z] == x1 && t.v[2 * z + 1] == y1) || (t.v[2 * z] == x2 && t.v[2 * z + 1] == y2)) ++hit;
                            if (hit == 2) {
                                G = eval_map(X, c, &t, qp.x, qp.y);
                                break;
                            }
                        }
                        double p, ux, uy, gg;
                        pde_at(X, G.x, G.y, &p, &ux, &uy, &gg);
                        if (X.pb.f.dirichlet_p) p = X.pb.f.dirichlet_p(X.pb.f.ctx, G.x, G.y);
                        mom[r] += p * Basis::edge_mono(r, x1, y1, x2, y2, qp.x, qp.y) * qp.w;
                    }
                }
                Mat Hi = inverse(He, kk);
                for (int r = 0; r < kk; ++r) {
                    double v = 0.0;
                    for (int s2 = 0; s2 < kk; ++s2) v += Hi[s2 * kk + r] * mom[s2];
                    X.rhs[g[r * ne + j]] += -v * dir[r * ne + j];
                }
            }
        }
    }
    coo_to_csr(N, coo, X.A);
    // essential normal flux (reference boundary condition)
    X.ub.assign(N, 0.0);
    X.isb.assign(N, 0);
    if (lam) {
        for (int c = 0; c < m.nc; ++c) {
            int s = m.off[c], ne = m.npc[c];
            Elem dummy;
            std::vector<Tri> tris = ear_clip(m, c, nullptr);
            for (int j = 0; j < ne; ++j) {
                int ge = m.cell_edges[s + j];
                if (m.edge_occ[ge] != 1) continue;
                double x1, y1, x2, y2;
                cell_xy(m, c, j, x1, y1);
                cell_xy(m, c, (j + 1) % ne, x2, y2);
                double dx = x2 - x1, dy = y2 - y1, len = std::sqrt(dx * dx + dy * dy);
                double nx = dy / len, ny = -dx / len;
                const Tri* tt = &tris[0];
                for (auto& t : tris) {
                    int hit = 0;
                    for (int z = 0; z < 3; ++z)
                        if ((t.v[2 * z] == x1 && t.v[2 * z + 1] == y1) || (t.v[2 * z] == x2 && t.v[2 * z + 1] == y2)) ++hit;
                    if (hit == 2) {
                        tt = &t;
                        break;
                    }
                }
                std::vector<QP> lp;
                line_points(X.nq1, x1, y1, x2, y2, lp);
                for (int d = 0; d <= X.k; ++d) {
                    double mo = 0.0;
                    for (auto& qp : lp) {
                        Geo G = eval_map(X, c, tt, qp.x, qp.y);
                        double ph, ux, uy, gh;
                        comp_data(X, c, G, &ph, &ux, &uy, &gh);
                        double un = ux * nx + uy * ny;
                        if (X.pb.f.neumann_un) {
                            // u_hat . n_hat = det J |J^{-T} n_hat| (u . n), n the physical unit normal
                            double mx = (G.j11 * nx - G.j10 * ny) / G.det, my = (-G.j01 * nx + G.j00 * ny) / G.det;
                            double mn = std::sqrt(mx * mx + my * my);
                            un = G.det * mn * X.pb.f.neumann_un(X.pb.f.ctx, G.x, G.y, mx / mn, my / mn);
                        }
                        mo += un * Basis::edge_mono(d, x1, y1, x2, y2, qp.x, qp.y) * qp.w;
                    }
                    int gd = d * m.ne + ge;
                    X.ub[gd] = mo;
                    X.isb[gd] = 1;
                }
            }
        }
    }
    // reduced system on the free dofs: F - K U_b restricted (CPU/solver/DarcySolver.cpp:600-661)
    std::vector<double> kub(N);
    csr_mv(X.A, X.ub.data(), kub.data());
    std::vector<int> pos(N, -1);
    X.free_idx.clear();
    for (int i = 0; i < N; ++i)
        if (!X.isb[i]) {
            pos[i] = (int)X.free_idx.size();
            X.free_idx.push_back(i);
        }
    int nf = (int)X.free_idx.size();
    X.rr.resize(nf);
    for (int a = 0; a < nf; ++a) X.rr[a] = X.rhs[X.free_idx[a]] - kub[X.free_idx[a]];
    X.Ar.n = nf;
    X.Ar.rp.assign(nf + 1, 0);
    X.Ar.ci.clear();
    X.Ar.v.clear();
    for (int a = 0; a < nf; ++a) {
        int i = X.free_idx[a];
        for (int p = X.A.rp[i]; p < X.A.rp[i + 1]; ++p) {
            int j = X.A.ci[p];
            if (pos[j] < 0) continue;
            X.Ar.ci.push_back(pos[j]);
            X.Ar.v.push_back(X.A.v[p]);
        }
        X.Ar.rp[a + 1] = (int)X.Ar.ci.size();
    }
    X.have_sys = true;
}

// ILU(0) on the stored pattern, natural ordering, no shift (the PETSc PCILU
// defaults used at CPU/core/petsc_utils.cpp:164-171).
struct Ilu {
    Csr F;
    std::vector<int> diag;
};

static Ilu ilu0(const Csr& A) {
    Ilu I;
    I.F = A;
    int n = A.n;
    I.diag.assign(n, -1);
    for (int i = 0; i < n; ++i)
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p)
            if (A.ci[p] == i) I.diag[i] = p;
    std::vector<int> where(n, -1);
    auto& F = I.F;
    for (int i = 0; i < n; ++i) {
        for (int p = F.rp[i]; p < F.rp[i + 1]; ++p) where[F.ci[p]] = p;
        for (int p = F.rp[i]; p < F.rp[i + 1]; ++p) {
            int kcol = F.ci[p];
            if (kcol >= i) continue;
            double piv = F.v[I.diag[kcol]];
            if (std::fabs(piv) < 1e-300) throw std::runtime_error("zero pivot in ILU(0)");
            double lik = F.v[p] / piv;
            F.v[p] = lik;
            for (int r = F.rp[kcol]; r < F.rp[kcol + 1]; ++r) {
                int j = F.ci[r];
                if (j <= kcol) continue;
                int w = where[j];
                if (w >= 0) F.v[w] -= lik * F.v[r];
            }
        }
        for (int p = F.rp[i]; p < F.rp[i + 1]; ++p) where[F.ci[p]] = -1;
        if (I.diag[i] < 0 || F.v[I.diag[i]] == 0.0) throw std::runtime_error("zero pivot in ILU(0)");
    }
    return I;
}

static void ilu_solve(const Ilu& I, const double* b, double* x) {
    const Csr& F = I.F;
    int n = F.n;
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int p = F.rp[i]; p < F.rp[i + 1]; ++p)
            if (F.ci[p] < i) s -= F.v[p] * x[F.ci[p]];
        x[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int p = F.rp[i]; p < F.rp[i + 1]; ++p)
            if (F.ci[p] > i) s -= F.v[p] * x[F.ci[p]];
        x[i] = s / F.v[I.diag[i]];
    }
}

// GMRES(restart), left preconditioned, classical Gram-Schmidt, convergence on
// the preconditioned residual norm relative to ||M^{-1} b|| (PETSc KSPGMRES
// defaults as configured at CPU/core/petsc_utils.cpp:144-211).
static int gmres(const Csr& A, const Ilu& M, const std::vector<double>& b, std::vector<double>& x, double rtol,
                 int restart, int maxit, double* final_rel) {
    int n = A.n;
    x.assign(n, 0.0);
    std::vector<double> r(n), w(n), tmp(n);
    std::vector<std::vector<double>> V(restart + 1, std::vector<double>(n));
    std::vector<double> H((restart + 1) * restart), cs(restart), sn(restart), g(restart + 1), y(restart);
    ilu_solve(M, b.data(), r.data());
    double r0 = 0.0;
    for (double v : r) r0 += v * v;
    r0 = std::sqrt(r0);
    if (r0 == 0.0) {
        if (final_rel) *final_rel = 0.0;
        return 0;
    }
    int it = 0;
    double rel = 1.0;
    while (it < maxit) {
        csr_mv(A, x.data(), tmp.data());
        for (int i = 0; i < n; ++i) tmp[i] = b[i] - tmp[i];
        ilu_solve(M, tmp.data(), r.data());
        double beta = 0.0;
        for (double v : r) beta += v * v;
        beta = std::sqrt(beta);
        rel = beta / r0;
        if (rel <= rtol) break;
        for (int i = 0; i < n; ++i) V[0][i] = r[i] / beta;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < restart && it < maxit; ++j, ++it) {
            csr_mv(A, V[j].data(), tmp.data());
            ilu_solve(M, tmp.data(), w.data());
            for (int i = 0; i <= j; ++i) {  // classical Gram-Schmidt
                double h = 0.0;
                for (int l = 0; l < n; ++l) h += V[i][l] * w[l];
                H[i * restart + j] = h;
            }
            for (int i = 0; i <= j; ++i)
                for (int l = 0; l < n; ++l) w[l] -= H[i * restart + j] * V[i][l];
            double hn = 0.0;
            for (double v : w) hn += v * v;
            hn = std::sqrt(hn);
            H[(j + 1) * restart + j] = hn;
            if (hn > 0)
                for (int l = 0; l < n; ++l) V[j + 1][l] = w[l] / hn;
            for (int i = 0; i < j; ++i) {
                double a = H[i * restart + j], bb = H[(i + 1) * restart + j];
                H[i * restart + j] = cs[i] * a + sn[i] * bb;
                H[(i + 1) * restart + j] = -sn[i] * a + cs[i] * bb;
            }
            double a = H[j * restart + j], bb = H[(j + 1) * restart + j];
            double den = std::sqrt(a * a + bb * bb);
            cs[j] = a / den;
            sn[j] = bb / den;
            H[j * restart + j] = den;
            H[(j + 1) * restart + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            rel = std::fabs(g[j + 1]) / r0;
            if (rel <= rtol || hn == 0.0) {
                ++j;
                ++it;
                break;
            }
        }
        for (int i = j - 1; i >= 0; --i) {
            double s = g[i];
            for (int l = i + 1; l < j; ++l) s -= H[i * restart + l] * y[l];
            y[i] = s / H[i * restart + i];
        }
        for (int i = 0; i < j; ++i)
            for (int l = 0; l < n; ++l) x[l] += y[i] * V[i][l];
        if (rel <= rtol) break;
    }
    if (final_rel) *final_rel = rel;
    return it;
}

// L2 errors in the computational measure (CPU/solver/DarcySolver.cpp:942-1131)
// plus physical-domain variants (det J weight, forward Piola u = J u_hat / det J).
// out: pressure and flux errors in the computational measure, the same in the
// physical measure, the physical divergence error and the physical mass
// int p_h (GPU/src/equation_gpu.cu:385-479: the divergence is the
// polynomial H^{-1} W u of the computational cell, divided by det J).
// cell_sq (nullable, 2 per cell): each cell's squared pressure and flux errors
// in the computational measure, so that tests can split the error by region.
static void errors(const Ctx& X, int nq, double* out, double* cell_sq = nullptr) {
    const Mesh& m = *X.m;
    const Basis& bs = X.basis;
    TriRule rule = triangle_rule(nq);
    double ep = 0, eu = 0, epp = 0, eup = 0, ed = 0, mass = 0;
    for (int c = 0; c < m.nc; ++c) {
        Elem e = element(X, c);
        std::vector<int> g;
        std::vector<double> dir;
        dofmap(X, c, g, dir);
        int n = e.n, p = bs.p, q = X.q;
        std::vector<double> uf(n), coef(p, 0.0), pc(q), wu(q, 0.0), dc(q, 0.0);
        for (int i = 0; i < n; ++i) uf[i] = dir[i] * X.sol[g[i]];
        for (int r = 0; r < q; ++r) pc[r] = X.sol[g[n + r]];
        for (int a = 0; a < p; ++a)
            for (int i = 0; i < n; ++i) coef[a] += e.P[a * n + i] * uf[i];
        for (int r = 0; r < q; ++r)
            for (int i = 0; i < n; ++i) wu[r] += e.W[r * n + i] * uf[i];
        Mat Hi = inverse(e.H, q);
        for (int r = 0; r < q; ++r)
            for (int t = 0; t < q; ++t) dc[r] += Hi[r * q + t] * wu[t];
        double xc = m.cx[c], yc = m.cy[c], h = m.diam[c];
        const double ep0 = ep, eu0 = eu;
        std::vector<QP> pts;
        for (auto& t : e.tris) {
            map_triangle(rule, t.v, pts);
            for (auto& qp : pts) {
                double ph = 0.0, vx = 0.0, vy = 0.0, dv = 0.0;
                for (int r = 0; r < q; ++r) {
                    const double mr = bs.mono(r, xc, yc, h, qp.x, qp.y);
                    ph += pc[r] * mr;
                    dv += dc[r] * mr;
                }
                for (int a = 0; a < p; ++a) {
                    double bx, by;
                    bs.hdiv(a, xc, yc, h, qp.x, qp.y, bx, by);
                    vx += coef[a] * bx;
                    vy += coef[a] * by;
                }
                Geo G = eval_map(X, c, &t, qp.x, qp.y);
                double pe, ux, uy, gh;
                comp_data(X, c, G, &pe, &ux, &uy, &gh);
                double dp = ph - pe, dx = vx - ux, dy = vy - uy;
                ep += dp * dp * qp.w;
                eu += (dx * dx + dy * dy) * qp.w;
                epp += dp * dp * qp.w * G.det;
                double fx = (G.j00 * dx + G.j01 * dy) / G.det, fy = (G.j10 * dx + G.j11 * dy) / G.det;
                eup += (fx * fx + fy * fy) * qp.w * G.det;
                const double dd = (dv - gh) / G.det;  // div_x u_h - g, with g_hat = det J g
                ed += dd * dd * qp.w * G.det;
                mass += ph * qp.w * G.det;
            }
        }
        if (cell_sq) {
            cell_sq[2 * c] = ep - ep0;
            cell_sq[2 * c + 1] = eu - eu0;
        }
    }
    out[0] = std::sqrt(ep);
    out[1] = std::sqrt(eu);
    out[2] = std::sqrt(epp);
    out[3] = std::sqrt(eup);
    out[4] = std::sqrt(ed);
    out[5] = mass;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI for the tests (ctypes).
// ---------------------------------------------------------------------------

using namespace orc;

#define ORC_TRY(...)                     \
    try {                                \
        __VA_ARGS__                      \
    } catch (const std::exception& ex) { \
        orc::g_err = ex.what();          \
        return -1;                       \
    }

extern "C" {

const char* orc_last_error() { return orc::g_err.c_str(); }

int orc_mesh_read(const char* path, int vtk, void** out) {
    ORC_TRY(*out = vtk ? (void*)read_vtk(path) : (void*)read_msh(path); return 0;)
}

// Cells are given as offsets (nc+1) into a node list; arcs as (n0, n1, R) per
// curved boundary edge.
int orc_mesh_from_arrays(int nn, const double* xy, int nc, const int* off, const int* nodes, int narc,
                         const int* arc_nodes, const double* arc_r, void** out) {
    ORC_TRY({
        Mesh* m = new Mesh();
        m->nn = nn;
        m->xy.assign(xy, xy + 2 * nn);
        m->off.push_back(0);
        for (int c = 0; c < nc; ++c) add_cell(*m, std::vector<int>(nodes + off[c], nodes + off[c + 1]));
        m->nc = nc;
        finalize(*m);
        for (int a = 0; a < narc; ++a) {
            int n0 = arc_nodes[2 * a], n1 = arc_nodes[2 * a + 1];
            m->arcs[{std::min(n0, n1), std::max(n0, n1)}] = arc_r[a];
        }
        *out = m;
        return 0;
    })
}

void orc_mesh_free(void* m) { delete (Mesh*)m; }

int orc_mesh_counts(void* mp, int* out) {
    Mesh* m = (Mesh*)mp;
    out[0] = m->nn;
    out[1] = m->nc;
    out[2] = m->ne;
    out[3] = (int)m->nodes.size();
    return 0;
}

int orc_mesh_get(void* mp, double* cx, double* cy, double* area, double* diam, int* off, int* nodes, int* cell_edges,
                 int* edge_nodes, int* edge_occ) {
    Mesh* m = (Mesh*)mp;
    std::copy(m->cx.begin(), m->cx.end(), cx);
    std::copy(m->cy.begin(), m->cy.end(), cy);
    std::copy(m->area.begin(), m->area.end(), area);
    std::copy(m->diam.begin(), m->diam.end(), diam);
    std::copy(m->off.begin(), m->off.end(), off);
    std::copy(m->nodes.begin(), m->nodes.end(), nodes);
    std::copy(m->cell_edges.begin(), m->cell_edges.end(), cell_edges);
    std::copy(m->edge_nodes.begin(), m->edge_nodes.end(), edge_nodes);
    std::copy(m->edge_occ.begin(), m->edge_occ.end(), edge_occ);
    return 0;
}

// Triangulation of one cell: returns the number of triangles; verts gets 6 per triangle.
int orc_triangulate(void* mp, int cell, double* verts, int cap) {
    ORC_TRY({
        std::vector<Tri> t = ear_clip(*(Mesh*)mp, cell, nullptr);
        for (size_t i = 0; i < t.size() && (int)i < cap; ++i) std::memcpy(verts + 6 * i, t[i].v, sizeof t[i].v);
        return (int)t.size();
    })
}

int orc_triangle_rule(int n, double* xi, double* eta, double* w) {
    ORC_TRY({
        TriRule r = triangle_rule(n);
        for (size_t i = 0; i < r.w.size(); ++i) {
            xi[i] = r.xi[i];
            eta[i] = r.eta[i];
            w[i] = r.w[i];
        }
        return (int)r.w.size();
    })
}

int orc_gauss_legendre(int n, double* x, double* w) {
    ORC_TRY({
        std::vector<double> a, b;
        gauss_legendre(n, a, b);
        std::copy(a.begin(), a.end(), x);
        std::copy(b.begin(), b.end(), w);
        return n;
    })
}

int orc_basis_eval(int k, int kind, int nk, double xc, double yc, double h, double x, double y, double* out) {
    ORC_TRY({
        Basis b(k);
        if (kind == 0)
            out[0] = b.mono(nk, xc, yc, h, x, y);
        else
            b.hdiv(nk, xc, yc, h, x, y, out[0], out[1]);
        return 0;
    })
}

int orc_edge_coeffs(int k, int nk, double xc, double yc, double h, double x1, double y1, double x2, double y2,
                    double* out) {
    ORC_TRY({
        Basis b(k);
        std::vector<double> c;
        b.edge_coeffs(nk, xc, yc, h, x1, y1, x2, y2, c);
        std::copy(c.begin(), c.end(), out);
        return 0;
    })
}

// Problem: map_kind 0 = sinusoidal global map with amplitude eps (reference
// examples 1 and 2), 1 = per-cell arc maps.  sol 0 = sin(pi x)cos(pi y),
// 1 = sin(pi x)sin(pi y).  bc 0 = essential normal flux + mean-zero pressure
// multiplier, 1 = Dirichlet pressure.  theta may be NULL (K = I).
// funcs (may be NULL): caller-defined data, the DarcyPdeData fields and the
// map, which replace the built-in choices member by member.
int orc_problem_new2(void* mp, int k, int nq2, int nq1, int map_kind, double eps, int sol, int bc,
                     const double* theta, const OrcFuncs* funcs, void** out) {
    ORC_TRY({
        Problem p;
        p.map_kind = map_kind;
        p.eps = eps;
        p.sol = sol;
        p.bc = bc;
        if (funcs) p.f = *funcs;
        Mesh* m = (Mesh*)mp;
        if (theta) p.theta.assign(theta, theta + m->nc);
        *out = new Ctx(m, p, k, nq2, nq1);
        return 0;
    })
}

int orc_problem_new(void* mp, int k, int nq2, int nq1, int map_kind, double eps, int sol, int bc,
                    const double* theta, void** out) {
    return orc_problem_new2(mp, k, nq2, nq1, map_kind, eps, sol, bc, theta, nullptr, out);
}

void orc_problem_free(void* c) { delete (Ctx*)c; }

// sizes: [Nf, Np, Ntot, Nfree, nnz(A), nnz(A_free)]
int orc_sizes(void* cp, long long* out) {
    Ctx* X = (Ctx*)cp;
    out[0] = X->Nf;
    out[1] = X->Np;
    out[2] = X->Ntot;
    out[3] = X->have_sys ? (long long)X->free_idx.size() : -1;
    out[4] = X->have_sys ? (long long)X->A.ci.size() : -1;
    out[5] = X->have_sys ? (long long)X->Ar.ci.size() : -1;
    return 0;
}

// Element matrices of one cell.  what: 0 D(n x p) 1 G(p x p) 2 W(q x n)
// 3 B(p x n) 4 P(p x n) 5 Pd(n x n) 6 GK(p x p) 7 A(n x n) 8 L(q) 9 f(q)
// 10 [alpha, tri_area, n, p, q] 11 H(q x q) 12 local-to-global map (n+q+1, as doubles)
// 13 direction signs (n+q+1)
int orc_element(void* cp, int cell, int what, double* out) {
    ORC_TRY({
        Ctx* X = (Ctx*)cp;
        if (cell < 0 || cell >= X->m->nc) throw std::out_of_range("cell");
        if (what >= 12) {
            std::vector<int> g;
            std::vector<double> dir;
            dofmap(*X, cell, g, dir);
            for (size_t i = 0; i < g.size(); ++i) out[i] = what == 12 ? (double)g[i] : dir[i];
            return (int)g.size();
        }
        Elem e = element(*X, cell);
        const Mat* src = nullptr;
        switch (what) {
            case 0: src = &e.D; break;
            case 1: src = &e.G; break;
            case 2: src = &e.W; break;
            case 3: src = &e.B; break;
            case 4: src = &e.P; break;
            case 5: src = &e.Pd; break;
            case 6: src = &e.GK; break;
            case 7: src = &e.A; break;
            case 8: src = &e.L; break;
            case 9: src = &e.f; break;
            case 11: src = &e.H; break;
            case 10: {
                out[0] = e.alpha;
                out[1] = e.tri_area;
                out[2] = e.n;
                out[3] = X->basis.p;
                out[4] = X->q;
                return 5;
            }
            default: throw std::invalid_argument("what");
        }
        std::copy(src->begin(), src->end(), out);
        return (int)src->size();
    })
}

int orc_assemble(void* cp) { ORC_TRY(assemble(*(Ctx*)cp); return 0;) }

// CSR of the full (reduced=0) or the free-dof (reduced=1) system, with its rhs.
int orc_csr(void* cp, int reduced, int* rp, int* ci, double* v, double* rhs) {
    Ctx* X = (Ctx*)cp;
    const Csr& A = reduced ? X->Ar : X->A;
    std::copy(A.rp.begin(), A.rp.end(), rp);
    std::copy(A.ci.begin(), A.ci.end(), ci);
    std::copy(A.v.begin(), A.v.end(), v);
    const std::vector<double>& r = reduced ? X->rr : X->rhs;
    std::copy(r.begin(), r.end(), rhs);
    return 0;
}

// free_idx (Nfree) and the full boundary-value vector (Ntot).
int orc_free(void* cp, int* free_idx, double* ub) {
    Ctx* X = (Ctx*)cp;
    std::copy(X->free_idx.begin(), X->free_idx.end(), free_idx);
    std::copy(X->ub.begin(), X->ub.end(), ub);
    return 0;
}

// Reduced solve with GMRES(restart)+ILU(0); scatters into the full solution.
int orc_solve_gmres_ilu(void* cp, double rtol, int restart, int maxit, int* iters, double* rel) {
    ORC_TRY({
        Ctx* X = (Ctx*)cp;
        if (!X->have_sys) assemble(*X);
        Ilu M = ilu0(X->Ar);
        std::vector<double> xf;
        int it = gmres(X->Ar, M, X->rr, xf, rtol, restart, maxit, rel);
        X->sol = X->ub;
        for (size_t a = 0; a < X->free_idx.size(); ++a) X->sol[X->free_idx[a]] = xf[a];
        X->have_sol = true;
        if (iters) *iters = it;
        return 0;
    })
}

// Install a full solution vector (e.g. from a direct solve of the CSR system).
int orc_set_solution(void* cp, const double* x) {
    Ctx* X = (Ctx*)cp;
    X->sol.assign(x, x + X->Ntot);
    X->have_sol = true;
    return 0;
}

int orc_get_solution(void* cp, double* x) {
    Ctx* X = (Ctx*)cp;
    if (!X->have_sol) return -1;
    std::copy(X->sol.begin(), X->sol.end(), x);
    return 0;
}

// out: [p_comp, u_comp, p_phys, u_phys]
int orc_errors(void* cp, int nq, double* out) {
    ORC_TRY({
        Ctx* X = (Ctx*)cp;
        if (!X->have_sol) throw std::runtime_error("errors require a solution");
        errors(*X, nq, out);
        return 0;
    })
}

// The six error values of orc_errors plus each cell's squared pressure and
// flux errors in the computational measure (2 per cell).
int orc_errors_cells(void* cp, int nq, double* out, double* cell_sq) {
    ORC_TRY({
        Ctx* X = (Ctx*)cp;
        if (!X->have_sol) throw std::runtime_error("errors require a solution");
        errors(*X, nq, out, cell_sq);
        return 0;
    })
}

// Pulled-back problem data at a computational point of a cell, in the order
// of Darcy::DarcyProblem's computational_* accessors
// (CPU/examples/darcy_problem.cpp:37-83): [p_hat, u_hat (2), g_hat, K_hat
// (xx, xy, yx, yy), det J].  K_hat is the inverse of the tensor the element
// integrals use (kinv_hat), so this also checks that inversion.
int orc_problem_data(void* cp, int cell, double xi, double eta, double* out) {
    ORC_TRY({
        Ctx* X = (Ctx*)cp;
        if (cell < 0 || cell >= X->m->nc) throw std::out_of_range("cell out of range");
        Geo G = eval_map(*X, cell, nullptr, xi, eta);
        comp_data(*X, cell, G, &out[0], &out[1], &out[2], &out[3]);
        double T[4];
        kinv_hat(*X, cell, G, T);
        const double d = T[0] * T[3] - T[1] * T[2];
        out[4] = T[3] / d;
        out[5] = -T[1] / d;
        out[6] = -T[2] / d;
        out[7] = T[0] / d;
        out[8] = G.det;
        return 0;
    })
}

// Map evaluation at a computational point of a given cell (uses the cell's
// triangulation to pick the sub-triangle holding the point).
int orc_map(void* cp, int cell, double xi, double eta, double* out) {
    ORC_TRY({
        Ctx* X = (Ctx*)cp;
        std::vector<Tri> tris = ear_clip(*X->m, cell, nullptr);
        const Tri* tt = &tris[0];
        for (auto& t : tris) {
            double d = (t.v[2] - t.v[0]) * (t.v[5] - t.v[1]) - (t.v[4] - t.v[0]) * (t.v[3] - t.v[1]);
            double l1 = ((xi - t.v[0]) * (t.v[5] - t.v[1]) - (t.v[4] - t.v[0]) * (eta - t.v[1])) / d;
            double l2 = ((t.v[2] - t.v[0]) * (eta - t.v[1]) - (xi - t.v[0]) * (t.v[3] - t.v[1])) / d;
            if (l1 >= -1e-12 && l2 >= -1e-12 && l1 + l2 <= 1 + 1e-12) {
                tt = &t;
                break;
            }
        }
        Geo g = eval_map(*X, cell, tt, xi, eta);
        out[0] = g.x;
        out[1] = g.y;
        out[2] = g.j00;
        out[3] = g.j01;
        out[4] = g.j10;
        out[5] = g.j11;
        out[6] = g.det;
        return 0;
    })
}

}  // extern "C"
