// Seeded Lloyd-relaxed Voronoi meshes of a disk or an annulus.
//
// Sites are drawn uniformly from the domain with a seeded generator, relaxed by
// Lloyd iterations (each site moves to the centroid of its cell), and the final
// Voronoi diagram becomes the polygon mesh.  Cells at the boundary are closed
// by mirror sites reflected radially across the circles, so the Voronoi
// vertices shared with a mirror cell lie next to the circle; they are projected
// onto it and every chord between two of them carries its exact arc.
//
// The Delaunay triangulation is incremental (Hilbert-ordered insertion,
// walking point location, Lawson flips) with exact predicates: coordinates are
// quantised to multiples of 2^-26 of the outer radius, orientation is exact in
// 64-bit integers and the in-circle test in 128-bit integers.
#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>

#include "common.h"

namespace vemd {

namespace {

// splitmix64: seeded uniform doubles in [0, 1).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

constexpr double QSCALE = 67108864.0;  // 2^26

struct Delaunay {
    std::vector<int64_t> X, Y;
    std::vector<std::array<int, 3>> V, N;
    std::vector<int> vtri;  // one incident triangle per point
    int last = 0;
    std::vector<std::pair<int, int>> stack;

    static int64_t orient(int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t cx, int64_t cy) {
        return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
    }
    int64_t orient(int a, int b, int c) const { return orient(X[a], Y[a], X[b], Y[b], X[c], Y[c]); }
    // > 0 when d lies inside the circumcircle of the CCW triangle (a, b, c)
    bool incircle(int a, int b, int c, int d) const {
        int64_t adx = X[a] - X[d], ady = Y[a] - Y[d], bdx = X[b] - X[d], bdy = Y[b] - Y[d];
        int64_t cdx = X[c] - X[d], cdy = Y[c] - Y[d];
        __int128 al = (__int128)(adx * adx + ady * ady), bl = (__int128)(bdx * bdx + bdy * bdy),
                 cl = (__int128)(cdx * cdx + cdy * cdy);
        __int128 det = al * (__int128)(bdx * cdy - cdx * bdy) + bl * (__int128)(cdx * ady - adx * cdy) +
                       cl * (__int128)(adx * bdy - bdx * ady);
        return det > 0;
    }
    void set_nb(int t, int old, int nw) {
        if (t < 0) return;
        for (int k = 0; k < 3; ++k)
            if (N[t][k] == old) {
                N[t][k] = nw;
                return;
            }
    }
    int make(int a, int b, int c, int na, int nb, int nc) {
        V.push_back({a, b, c});
        N.push_back({na, nb, nc});
        return (int)V.size() - 1;
    }
    void assign(int t, int a, int b, int c, int na, int nb, int nc) {
        V[t] = {a, b, c};
        N[t] = {na, nb, nc};
        vtri[a] = vtri[b] = vtri[c] = t;
    }
    void legalize() {
        while (!stack.empty()) {
            auto [t, i] = stack.back();
            stack.pop_back();
            int p = V[t][i], a = V[t][(i + 1) % 3], b = V[t][(i + 2) % 3];
            int u = N[t][i];
            if (u < 0) continue;
            int j = 0;
            while (N[u][j] != t) ++j;
            int d = V[u][j];
            if (!incircle(p, a, b, d)) continue;
            int X1 = N[u][(j + 1) % 3], X2 = N[u][(j + 2) % 3];
            int Y1 = N[t][(i + 1) % 3], Y2 = N[t][(i + 2) % 3];
            assign(t, p, a, d, X1, u, Y2);
            assign(u, p, d, b, X2, Y1, t);
            set_nb(X1, u, t);
            set_nb(Y1, t, u);
            stack.push_back({t, 0});
            stack.push_back({u, 0});
        }
    }
    // Walk to the triangle holding p; returns it and the index of the edge p
    // lies on (-1 if strictly inside).
    int locate(int p, int& on_edge) {
        int t = last;
        uint32_t r = (uint32_t)p * 2654435761u;
        for (int steps = 0;; ++steps) {
            int s = (int)(r % 3);
            r = r * 1664525u + 1013904223u;
            bool moved = false;
            on_edge = -1;
            for (int kk = 0; kk < 3; ++kk) {
                int k = (s + kk) % 3;
                int a = V[t][(k + 1) % 3], b = V[t][(k + 2) % 3];
                int64_t o = orient(a, b, p);
                if (o < 0) {
                    t = N[t][k];
                    moved = true;
                    break;
                }
                if (o == 0) on_edge = k;
            }
            if (!moved) return t;
            if (t < 0 || steps > 100000000) throw Error(VEMD_ERR_INTERNAL, "Delaunay point location failed");
        }
    }
    bool insert(int p) {
        int e;
        int t = locate(p, e);
        for (int k = 0; k < 3; ++k)
            if (X[V[t][k]] == X[p] && Y[V[t][k]] == Y[p]) return false;  // duplicate site
        if (e < 0) {
            int a = V[t][0], b = V[t][1], c = V[t][2];
            int na = N[t][0], nb = N[t][1], nc = N[t][2];
            int t1 = make(b, c, p, na, -1, -1), t2 = make(c, a, p, nb, -1, -1);
            assign(t, a, b, p, t1, t2, nc);
            N[t1] = {t2, t, na};
            N[t2] = {t, t1, nb};
            V[t1] = {b, c, p};
            V[t2] = {c, a, p};
            // t = (a, b, p): opp a = (b, p) -> t1, opp b = (p, a) -> t2, opp p = (a, b) -> nc
            N[t] = {t1, t2, nc};
            N[t1] = {t2, t, na};
            N[t2] = {t, t1, nb};
            vtri[b] = vtri[c] = vtri[p] = t1;
            vtri[a] = t;
            set_nb(na, t, t1);
            set_nb(nb, t, t2);
            stack.push_back({t, 2});
            stack.push_back({t1, 2});
            stack.push_back({t2, 2});
            last = t;
        } else {
            int i = e, c = V[t][i], a = V[t][(i + 1) % 3], b = V[t][(i + 2) % 3];
            int u = N[t][i];
            if (u < 0) throw Error(VEMD_ERR_INTERNAL, "site on the hull");
            int j = 0;
            while (N[u][j] != t) ++j;
            int d = V[u][j];
            int Nb = N[t][(i + 1) % 3], Na = N[t][(i + 2) % 3];
            int Ma = N[u][(j + 1) % 3], Mb = N[u][(j + 2) % 3];
            int t1 = make(c, p, b, -1, -1, -1), u1 = make(d, p, a, -1, -1, -1);
            assign(t, c, a, p, u1, t1, Na);
            assign(t1, c, p, b, u, Nb, t);
            assign(u, d, b, p, t1, u1, Mb);
            assign(u1, d, p, a, t, Ma, u);
            set_nb(Nb, t, t1);
            set_nb(Ma, u, u1);
            stack.push_back({t, 2});
            stack.push_back({t1, 1});
            stack.push_back({u, 2});
            stack.push_back({u1, 1});
            last = t;
        }
        legalize();
        return true;
    }
};

struct Domain {
    int kind;
    double ri, ro;
    bool inside(double x, double y) const {
        double r = std::hypot(x, y);
        return r <= ro && (kind == VEMD_DOMAIN_DISK || r >= ri);
    }
    double area() const { return M_PI * (ro * ro - (kind == VEMD_DOMAIN_DISK ? 0.0 : ri * ri)); }
};

static uint64_t hkey(double x, double y, double lo, double sc) {
    uint32_t ix = (uint32_t)std::min(65535.0, std::max(0.0, (x - lo) * sc));
    uint32_t iy = (uint32_t)std::min(65535.0, std::max(0.0, (y - lo) * sc));
    uint64_t d = 0;
    for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
        uint32_t rx = (ix & s) > 0, ry = (iy & s) > 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                ix = s - 1 - ix;
                iy = s - 1 - iy;
            }
            std::swap(ix, iy);
        }
    }
    return d;
}

// Triangulates sites (first ns points are real) plus mirrors and corners.
struct Diagram {
    Delaunay dt;
    int ns = 0, npts = 0;
    std::vector<double> cx, cy;  // circumcentres per triangle
};

static void build_diagram(const std::vector<double>& sx, const std::vector<double>& sy, const Domain& D, double h,
                          Diagram& G) {
    int ns = (int)sx.size();
    std::vector<double> px(sx), py(sy);
    // mirrors of sites within 3h of a boundary circle
    double band = 3.0 * h;
    for (int i = 0; i < ns; ++i) {
        double r = std::hypot(sx[i], sy[i]);
        if (r > D.ro - band && r > 0) {
            double f = (2 * D.ro - r) / r;
            px.push_back(sx[i] * f);
            py.push_back(sy[i] * f);
        }
        if (D.kind == VEMD_DOMAIN_ANNULUS && r < D.ri + band) {
            double f = (2 * D.ri - r) / r;
            px.push_back(sx[i] * f);
            py.push_back(sy[i] * f);
        }
    }
    int np = (int)px.size();
    Delaunay& T = G.dt;
    T = Delaunay();
    T.X.resize(np + 4);
    T.Y.resize(np + 4);
    double qs = QSCALE / D.ro;
    for (int i = 0; i < np; ++i) {
        T.X[i] = (int64_t)std::llround(px[i] * qs);
        T.Y[i] = (int64_t)std::llround(py[i] * qs);
    }
    int64_t big = (int64_t)std::llround(2.5 * QSCALE);
    int c0 = np, c1 = np + 1, c2 = np + 2, c3 = np + 3;
    T.X[c0] = -big, T.Y[c0] = -big;
    T.X[c1] = big, T.Y[c1] = -big;
    T.X[c2] = big, T.Y[c2] = big;
    T.X[c3] = -big, T.Y[c3] = big;
    T.vtri.assign(np + 4, -1);
    T.V.reserve(2 * (size_t)np + 16);
    T.N.reserve(2 * (size_t)np + 16);
    // two corner triangles: (c0, c1, c2) and (c0, c2, c3)
    T.make(c0, c1, c2, -1, 1, -1);
    T.make(c0, c2, c3, -1, -1, 0);
    T.vtri[c0] = 0, T.vtri[c1] = 0, T.vtri[c2] = 0, T.vtri[c3] = 1;
    // Hilbert order of insertion
    std::vector<uint64_t> key(np);
    double lo = -2.0 * D.ro, sc = 65535.0 / (4.0 * D.ro);
    for (int i = 0; i < np; ++i) key[i] = hkey(px[i], py[i], lo, sc);
    std::vector<int> ord(np);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
    for (int i : ord)
        if (!T.insert(i) && i < ns) throw Error(VEMD_ERR_MESH, "two Voronoi sites coincide");
    G.ns = ns;
    G.npts = np;
    size_t nt = T.V.size();
    G.cx.resize(nt);
    G.cy.resize(nt);
    double inv = 1.0 / qs;
    for (size_t t = 0; t < nt; ++t) {
        double ax = T.X[T.V[t][0]] * inv, ay = T.Y[T.V[t][0]] * inv;
        double bx = T.X[T.V[t][1]] * inv - ax, by = T.Y[T.V[t][1]] * inv - ay;
        double cxx = T.X[T.V[t][2]] * inv - ax, cyy = T.Y[T.V[t][2]] * inv - ay;
        double d = 2.0 * (bx * cyy - by * cxx);
        double b2 = bx * bx + by * by, c2 = cxx * cxx + cyy * cyy;
        G.cx[t] = ax + (cyy * b2 - by * c2) / d;
        G.cy[t] = ay + (bx * c2 - cxx * b2) / d;
    }
}

// Triangles around real site s in CCW order.
static void ring(const Delaunay& T, int s, std::vector<int>& out) {
    out.clear();
    int t0 = T.vtri[s], t = t0;
    do {
        out.push_back(t);
        int i = 0;
        while (T.V[t][i] != s) ++i;
        t = T.N[t][(i + 1) % 3];
        if (t < 0) throw Error(VEMD_ERR_INTERNAL, "open Voronoi cell");
    } while (t != t0 && out.size() < 64);
    if (t != t0) throw Error(VEMD_ERR_MESH, "Voronoi cell with more than 64 vertices");
}

struct UF {
    std::vector<int> p;
    explicit UF(int n) : p(n) { std::iota(p.begin(), p.end(), 0); }
    int find(int a) {
        while (p[a] != a) a = p[a] = p[p[a]];
        return a;
    }
    void join(int a, int b) {
        a = find(a), b = find(b);
        if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
};

}  // namespace

Mesh* mesh_generate(const vemd_mesh_opts& o) {
    if (o.domain != VEMD_DOMAIN_DISK && o.domain != VEMD_DOMAIN_ANNULUS) throw Error(VEMD_ERR_ARG, "domain");
    if (!(o.r_outer > 0) || (o.domain == VEMD_DOMAIN_ANNULUS && !(o.r_inner > 0 && o.r_inner < o.r_outer)))
        throw Error(VEMD_ERR_ARG, "radii");
    if (o.num_cells < 3 || o.num_cells > 50000000) throw Error(VEMD_ERR_ARG, "num_cells out of range");
    if (o.lloyd_iters < 0) throw Error(VEMD_ERR_ARG, "lloyd_iters");
    Domain D{o.domain, o.r_inner, o.r_outer};
    const int ns = o.num_cells;
    const double h = std::sqrt(D.area() / ns);
    if (o.domain == VEMD_DOMAIN_ANNULUS && 3.0 * h >= o.r_inner) throw Error(VEMD_ERR_ARG, "too few cells for the annulus");
    // seeded uniform sites
    Rng rng(o.seed);
    std::vector<double> sx(ns), sy(ns);
    double r0 = o.domain == VEMD_DOMAIN_DISK ? 0.0 : o.r_inner;
    for (int i = 0; i < ns; ++i) {
        double r = std::sqrt(r0 * r0 + rng.uniform() * (D.ro * D.ro - r0 * r0));
        double th = 2.0 * M_PI * rng.uniform();
        sx[i] = r * std::cos(th);
        sy[i] = r * std::sin(th);
    }
    Diagram G;
    std::vector<int> rg;
    for (int it = 0; it <= o.lloyd_iters; ++it) {
        build_diagram(sx, sy, D, h, G);
        if (it == o.lloyd_iters) break;
        // Lloyd step: centroid of each cell polygon
        std::vector<double> nx(ns), ny(ns);
        parallel_for(ns, [&](int64_t b, int64_t e) {
            std::vector<int> r;
            for (int64_t s = b; s < e; ++s) {
                ring(G.dt, (int)s, r);
                double a = 0, mx = 0, my = 0;
                int n = (int)r.size();
                for (int k = 0; k < n; ++k) {
                    double x0 = G.cx[r[k]], y0 = G.cy[r[k]], x1 = G.cx[r[(k + 1) % n]], y1 = G.cy[r[(k + 1) % n]];
                    double cr = x0 * y1 - x1 * y0;
                    a += cr;
                    mx += (x0 + x1) * cr;
                    my += (y0 + y1) * cr;
                }
                double cxs = a != 0 ? mx / (3 * a) : sx[s], cys = a != 0 ? my / (3 * a) : sy[s];
                if (!D.inside(cxs, cys)) {  // keep sites strictly inside
                    double rr = std::hypot(cxs, cys);
                    double target = std::min(std::max(rr, D.kind == VEMD_DOMAIN_DISK ? 0.0 : D.ri + 1e-3 * h), D.ro - 1e-3 * h);
                    cxs *= target / rr;
                    cys *= target / rr;
                }
                nx[s] = cxs;
                ny[s] = cys;
            }
        });
        sx.swap(nx);
        sy.swap(ny);
    }
    // Voronoi vertices used by real cells
    const Delaunay& T = G.dt;
    size_t nt = T.V.size();
    std::vector<int> vid(nt, -1);
    std::vector<char> bnd(nt, 0);
    int nv = 0;
    std::vector<double> vx, vy;
    std::vector<int> cell_off(1, 0), cell_vs;
    for (int s = 0; s < ns; ++s) {
        ring(T, s, rg);
        for (int t : rg) {
            if (vid[t] < 0) {
                vid[t] = nv++;
                vx.push_back(G.cx[t]);
                vy.push_back(G.cy[t]);
                bnd[t] = (T.V[t][0] >= ns || T.V[t][1] >= ns || T.V[t][2] >= ns);
            }
            cell_vs.push_back(vid[t]);
        }
        cell_off.push_back((int)cell_vs.size());
    }
    std::vector<char> vb(nv, 0);
    for (size_t t = 0; t < nt; ++t)
        if (vid[t] >= 0) vb[vid[t]] = bnd[t];
    // project boundary vertices onto their circle; check interior vertices
    std::vector<double> vr(nv, 0.0);  // circle radius of boundary vertices
    for (int v = 0; v < nv; ++v) {
        double r = std::hypot(vx[v], vy[v]);
        if (vb[v]) {
            double R = D.ro;
            if (D.kind == VEMD_DOMAIN_ANNULUS && std::fabs(r - D.ri) < std::fabs(r - D.ro)) R = D.ri;
            vx[v] *= R / r;
            vy[v] *= R / r;
            vr[v] = R;
        } else if (!(r < D.ro) || (D.kind == VEMD_DOMAIN_ANNULUS && !(r > D.ri))) {
            throw Error(VEMD_ERR_MESH, "interior Voronoi vertex outside the domain");
        }
    }
    // merge vertices joined by (near-)zero or short edges
    UF uf(nv);
    double tol = std::max(1e-9 * h, o.min_edge_frac * h);
    for (int s = 0; s < ns; ++s) {
        int b = cell_off[s], n = cell_off[s + 1] - b;
        for (int k = 0; k < n; ++k) {
            int a = cell_vs[b + k], c = cell_vs[b + (k + 1) % n];
            if (a != c && std::hypot(vx[a] - vx[c], vy[a] - vy[c]) < tol) {
                if (vb[a] && vb[c] && vr[a] != vr[c]) continue;
                uf.join(a, c);
            }
        }
    }
    // representative position: a boundary vertex wins
    std::vector<int> rep(nv), newid(nv, -1);
    for (int v = 0; v < nv; ++v) rep[v] = uf.find(v);
    for (int v = 0; v < nv; ++v)
        if (vb[v] && !vb[rep[v]]) {
            vx[rep[v]] = vx[v];
            vy[rep[v]] = vy[v];
            vb[rep[v]] = 1;
            vr[rep[v]] = vr[v];
        }
    Mesh* m = new Mesh();
    int nn = 0;
    for (int v = 0; v < nv; ++v)
        if (rep[v] == v) {
            newid[v] = nn++;
            m->xy.push_back(vx[v]);
            m->xy.push_back(vy[v]);
        }
    m->nn = nn;
    m->off.push_back(0);
    std::vector<int> poly;
    for (int s = 0; s < ns; ++s) {
        poly.clear();
        for (int j = cell_off[s]; j < cell_off[s + 1]; ++j) {
            int v = newid[rep[cell_vs[j]]];
            if (poly.empty() || poly.back() != v) poly.push_back(v);
        }
        while (poly.size() > 1 && poly.front() == poly.back()) poly.pop_back();
        if (poly.size() < 3) {
            delete m;
            throw Error(VEMD_ERR_MESH, "a Voronoi cell collapsed below three vertices");
        }
        m->nodes.insert(m->nodes.end(), poly.begin(), poly.end());
        m->off.push_back((int)m->nodes.size());
    }
    mesh_finalize(*m);
    // arcs: boundary edges whose end points lie on the same circle
    std::vector<double> nr(nn, 0.0);
    for (int v = 0; v < nv; ++v)
        if (rep[v] == v && vb[v]) nr[newid[v]] = vr[v];
    for (int e = 0; e < m->ne; ++e) {
        if (m->edge_occ[e] != 1) continue;
        int a = m->edge_nodes[2 * e], b = m->edge_nodes[2 * e + 1];
        if (nr[a] == 0.0 || nr[a] != nr[b]) {
            delete m;
            throw Error(VEMD_ERR_MESH, "boundary edge not on a boundary circle");
        }
        m->arcs[Mesh::key(a, b)] = nr[a];
    }
    return m;
}

}  // namespace vemd
