// FP64 element setup on the GPU: projector, stabilisation, divergence, data
// and boundary terms of the mixed VEM (k = 0, 1) for every cell, one thread
// per cell, cells of one valence side by side in a wavefront.
//
// Each thread follows the host setup's arithmetic step for step (the same
// quadrature, ear clipping decisions, basis, small Gauss-Jordan inverses and
// summation orders), so that the two agree to rounding of the math library:
//   quadrature      CPU/core/gauss_quadrature.cpp:30-50,252-363
//   basis           CPU/lib/polynomial_basis.cpp:92-101,163-226,245-325
//   ear clipping    CPU/mesh/polygon_triangulator.cpp:50-145
//   D,G,W,H,B,P     CPU/solver/HdivMatrix.cpp:94-652
//   G_K, alpha, A   CPU/solver/DarcySolver.cpp:1133-1269,1476-1509
//   rhs, multiplier CPU/solver/DarcySolver.cpp:1271-1452
//   boundary        CPU/solver/DarcySolver.cpp:502-663
// The per-thread matrices live in private memory; the quadrature points are
// visited once and every volume integral is accumulated in that pass.
//
// The same cell function, in its error mode, evaluates the six error
// functionals of a device-resident solution (the reference GPU port's
// error_kernel / reduce_errors, GPU/src/equation_gpu.cu:385-479): the
// projector is rebuilt per cell instead of being stored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>

#include "precond.h"
#include "solver.h"
#include "store.h"

namespace vemd {

namespace {

constexpr int NE = MAX_NE;  // polygons of up to 10 edges
constexpr int MAXQ2 = 12, MAXQ1 = 10;

struct SetupIn {
    const double* xy;
    const int *off, *nodes, *cell_edges, *edge_nodes, *edge_occ;
    const double *cx, *cy, *diam;
    const double* edge_r;  // arc radius of each edge, 0 for a straight edge
    const int* order;      // cells in processing order
    const int64_t *a_off, *w_off, *v_off, *p_off;
    int nc;
    int map_kind, solution, bc, arcs;
    double eps;
    int nt, ng;
    double ta[MAXQ2], tb[MAXQ2], tw[MAXQ2], gx[MAXQ1], gw[MAXQ1];
};

struct SetupOut {
    double *A, *W, *L, *f, *ru, *ub, *P, *alpha;
    char* isb;
    int* err;  // first failure: code, cell
};

// Error functionals of a solution (the error mode of the cell function): the
// solution in the solver's internal layout on the device, where each cell
// finds its dofs, the error quadrature rule, and six values per cell.
struct ErrIn {
    const double* x;           // internal-layout solution
    const int* edge_int;       // internal number of every mesh edge
    const int* gpos;           // store position of every cell
    int64_t cbase, cpad;
    int nt;
    double ta[MAXQ2], tb[MAXQ2], tw[MAXQ2];
    double* out;               // 6 per cell: the squared errors of errors_host and the mass
};

enum { E_JAC = 1, E_EAR, E_EDGE, E_SINGULAR, E_CONVEX, E_AREA, E_NE };

template <int K>
struct Sh {
    static constexpr int q = (K + 1) * (K + 2) / 2;
    static constexpr int grad = (K + 2) * (K + 3) / 2 - 1;
    static constexpr int curl = K * (K + 1) / 2;
    static constexpr int p = grad + curl;
    static constexpr int ni = (q - 1) + (q - (K + 1));
    static constexpr int nm = (K + 2) * (K + 3) / 2;
    static constexpr int kk = K + 1;
    static constexpr int nmax = kk * NE + ni;
};

__host__ __device__ inline double dpow(double x, int e) {
    double r = 1.0;
    for (int i = 0; i < e; ++i) r *= x;
    return r;
}

__host__ __device__ inline void dmono(int idx, int& a, int& b) {
    int d = 0, pre = 0;
    while (idx >= pre + d + 1) {
        pre += d + 1;
        ++d;
    }
    a = d - (idx - pre);
    b = idx - pre;
}

__host__ __device__ inline long dbinom(int n, int k) {
    long r = 1;
    k = k < n - k ? k : n - k;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// H(div) basis function nk: gradients of scaled monomials of degree 1..k+1,
// then the complement (-(b+1)/a X^(a-1) Y^(b+1), X^a Y^b).
template <int K>
__host__ __device__ inline void dhdiv(int nk, double xm, double ym, double h, double& vx, double& vy) {
    if (nk < Sh<K>::grad) {
        int a, b;
        dmono(nk + 1, a, b);
        vx = a > 0 ? a * dpow(xm, a - 1) * dpow(ym, b) / h : 0.0;
        vy = b > 0 ? b * dpow(xm, a) * dpow(ym, b - 1) / h : 0.0;
        return;
    }
    int ci = nk - Sh<K>::grad, deg = 1;
    while (ci >= deg) {
        ci -= deg;
        ++deg;
    }
    int yd = ci, xd = deg - yd;
    vx = -(double)(yd + 1) / xd * dpow(xm, xd - 1) * dpow(ym, yd + 1);
    vy = dpow(xm, xd) * dpow(ym, yd);
}

// Gauss-Jordan inverse with partial pivoting of an n x n matrix (n <= 6).
__host__ __device__ inline bool dinverse(const double* A, int n, double* out) {
    double a[6 * 12];
    const int m = 2 * n;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) a[i * m + j] = A[i * n + j];
        for (int j = 0; j < n; ++j) a[i * m + n + j] = i == j ? 1.0 : 0.0;
    }
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (fabs(a[i * m + k]) > fabs(a[piv * m + k])) piv = i;
        if (piv != k)
            for (int j = 0; j < m; ++j) {
                const double t = a[k * m + j];
                a[k * m + j] = a[piv * m + j];
                a[piv * m + j] = t;
            }
        if (fabs(a[k * m + k]) < 1e-14) return false;
        for (int i = k + 1; i < n; ++i) {
            const double f = a[i * m + k] / a[k * m + k];
            for (int j = k; j < m; ++j) a[i * m + j] -= f * a[k * m + j];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        const double d = a[k * m + k];
        for (int j = k; j < m; ++j) a[k * m + j] /= d;
        for (int i = 0; i < k; ++i) {
            const double f = a[i * m + k];
            for (int j = k; j < m; ++j) a[i * m + j] -= f * a[k * m + j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[i * n + j] = a[i * m + n + j];
    return true;
}

struct DGeo {
    double X, Y, j00, j01, j10, j11, det;
};

// Wachspress coordinates of a convex polygon and their gradients (product form).
__host__ __device__ inline void dwachspress(const double* vx, const double* vy, int n, double x, double y, double* lam,
                                   double* gl) {
    double a[NE], ga0[NE], ga1[NE], P[NE], gP0[NE], gP1[NE], Ac[NE];
    for (int i = 0; i < n; ++i) {
        const int j = (i + 1) % n;
        a[i] = 0.5 * ((vx[i] - x) * (vy[j] - y) - (vy[i] - y) * (vx[j] - x));
        ga0[i] = 0.5 * (vy[i] - vy[j]);
        ga1[i] = 0.5 * (vx[j] - vx[i]);
    }
    double D = 0, gD0 = 0, gD1 = 0;
    for (int i = 0; i < n; ++i) {
        const int p = (i - 1 + n) % n, q = (i + 1) % n;
        Ac[i] = 0.5 * ((vx[i] - vx[p]) * (vy[q] - vy[p]) - (vy[i] - vy[p]) * (vx[q] - vx[p]));
        P[i] = 1.0;
        gP0[i] = gP1[i] = 0.0;
        for (int m = 0; m < n; ++m) {
            if (m == p || m == i) continue;
            gP0[i] = gP0[i] * a[m] + P[i] * ga0[m];
            gP1[i] = gP1[i] * a[m] + P[i] * ga1[m];
            P[i] *= a[m];
        }
        D += Ac[i] * P[i];
        gD0 += Ac[i] * gP0[i];
        gD1 += Ac[i] * gP1[i];
    }
    for (int i = 0; i < n; ++i) {
        lam[i] = Ac[i] * P[i] / D;
        gl[2 * i] = (Ac[i] * gP0[i] - lam[i] * gD0) / D;
        gl[2 * i + 1] = (Ac[i] * gP1[i] - lam[i] * gD1) / D;
    }
}

// Chord-to-arc blend of one curved edge (ia -> ib, radius R about the origin).
__host__ __device__ inline void dchord(double ax, double ay, double bx, double by, double R, int ia, int ib, const double* lam,
                              const double* gl, double* F, double* J) {
    const double lA = lam[ia], lB = lam[ib], S = lA + lB;
    if (S <= 1e-300) return;
    const double *gA = &gl[2 * ia], *gB = &gl[2 * ib];
    const double s = lB / S;
    const double tA = atan2(ay, ax);
    double dt = atan2(by, bx) - tA;
    if (dt > M_PI) dt -= 2 * M_PI;
    if (dt <= -M_PI) dt += 2 * M_PI;
    const double th = tA + s * dt, ct = cos(th), st = sin(th);
    const double del0 = R * ct - (ax + s * (bx - ax)), del1 = R * st - (ay + s * (by - ay));
    const double dd0 = -R * dt * st - (bx - ax), dd1 = R * dt * ct - (by - ay);
    const double gS0 = gA[0] + gB[0], gS1 = gA[1] + gB[1];
    const double w0 = lA * gB[0] - lB * gA[0], w1 = lA * gB[1] - lB * gA[1];
    F[0] += S * S * del0;
    F[1] += S * S * del1;
    J[0] += 2 * S * del0 * gS0 + dd0 * w0;
    J[1] += 2 * S * del0 * gS1 + dd0 * w1;
    J[2] += 2 * S * del1 * gS0 + dd1 * w0;
    J[3] += 2 * S * del1 * gS1 + dd1 * w1;
}

// The cell's map at (x, y): sinusoidal perturbation, identity, or the
// chord-to-arc blend of its curved edges (chord_mask bit j: edge j -> j + 1).
__host__ __device__ inline DGeo dmap(const SetupIn& in, const double* vx, const double* vy, int ne, const double* er,
                            unsigned chord_mask, double x, double y) {
    DGeo g;
    if (in.map_kind == VEMD_MAP_SIN) {
        const double e = in.eps;
        g.X = x + e * sin(2 * M_PI * y);
        g.Y = y + e * sin(2 * M_PI * x);
        g.j00 = 1.0;
        g.j01 = e * 2 * M_PI * cos(2 * M_PI * y);
        g.j10 = e * 2 * M_PI * cos(2 * M_PI * x);
        g.j11 = 1.0;
    } else {
        double F[2] = {x, y}, J[4] = {1, 0, 0, 1};
        if (chord_mask) {
            double lam[NE], gl[2 * NE];
            dwachspress(vx, vy, ne, x, y, lam, gl);
            for (int j = 0; j < ne; ++j)
                if (chord_mask >> j & 1u) {
                    const int j1 = (j + 1) % ne;
                    dchord(vx[j], vy[j], vx[j1], vy[j1], er[j], j, j1, lam, gl, F, J);
                }
        }
        g.X = F[0];
        g.Y = F[1];
        g.j00 = J[0];
        g.j01 = J[1];
        g.j10 = J[2];
        g.j11 = J[3];
    }
    g.det = g.j00 * g.j11 - g.j01 * g.j10;
    return g;
}

__host__ __device__ inline void dexact(int sol, double x, double y, double& p, double& ux, double& uy, double& g) {
    const double sx = sin(M_PI * x), cx = cos(M_PI * x), sy = sin(M_PI * y), cy = cos(M_PI * y);
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

// Element data of cell c; returns 0 or the failure code.  ERR: the same
// projector construction, then the cell's error functionals of the solution
// in *ei instead of the element outputs.
template <int K, bool ERR = false>
__host__ __device__ int setup_cell(const SetupIn& in, const SetupOut& o, int c, const ErrIn* ei = nullptr) {
    using S = Sh<K>;
    constexpr int q = S::q, p = S::p, gr = S::grad, nm = S::nm, kk = S::kk, NM = S::nmax;
    const int s0 = in.off[c], ne = in.off[c + 1] - s0;
    if (ne > NE || ne < 3) return E_NE;
    const int n = kk * ne + S::ni, edof = kk * ne;
    const double xc = in.cx[c], yc = in.cy[c], h = in.diam[c];
    double vx[NE], vy[NE], er[NE];
    unsigned chord = 0;
    for (int j = 0; j < ne; ++j) {
        const int a = in.nodes[s0 + j];
        vx[j] = in.xy[2 * a];
        vy[j] = in.xy[2 * a + 1];
        er[j] = in.arcs ? in.edge_r[in.cell_edges[s0 + j]] : 0.0;
        if (er[j] != 0.0) chord |= 1u << j;
    }
    if (chord)
        for (int j = 0; j < ne; ++j) {  // the Wachspress blend needs a strictly convex cell
            const int pp = (j - 1 + ne) % ne, qq = (j + 1) % ne;
            const double cr = (vx[j] - vx[pp]) * (vy[qq] - vy[pp]) - (vy[j] - vy[pp]) * (vx[qq] - vx[pp]);
            if (!(cr > 0.0)) {
                return E_CONVEX;
            }
        }
    // ear clipping with diameter-relative tolerances, in the host's order
    unsigned char tri[NE - 2][3];
    double tri_area = 0.0;
    {
        const double h2 = h * h, tol_cross = 1e-8 * h2, tol_area = 1e-6 * h2;
        int ids[NE];
        for (int i = 0; i < ne; ++i) ids[i] = i;
        int rem = ne, nt = 0;
        auto ar = [&](int i, int j, int k) {
            return 0.5 * fabs((vx[ids[j]] - vx[ids[i]]) * (vy[ids[k]] - vy[ids[i]]) -
                              (vx[ids[k]] - vx[ids[i]]) * (vy[ids[j]] - vy[ids[i]]));
        };
        while (rem > 3) {
            bool found = false;
            for (int i = 0; i < rem && !found; ++i) {
                const int pr = (i - 1 + rem) % rem, nx = (i + 1) % rem;
                const double cr = (vx[ids[nx]] - vx[ids[i]]) * (vy[ids[pr]] - vy[ids[i]]) -
                                  (vx[ids[pr]] - vx[ids[i]]) * (vy[ids[nx]] - vy[ids[i]]);
                if (cr <= tol_cross) continue;
                bool ear = true;
                for (int j = 0; j < rem && ear; ++j)
                    if (j != i && j != pr && j != nx &&
                        fabs(ar(pr, i, j) + ar(i, nx, j) + ar(nx, pr, j) - ar(pr, i, nx)) < tol_area)
                        ear = false;
                if (!ear) continue;
                tri[nt][0] = (unsigned char)ids[pr];
                tri[nt][1] = (unsigned char)ids[i];
                tri[nt][2] = (unsigned char)ids[nx];
                ++nt;
                for (int j = i; j + 1 < rem; ++j) ids[j] = ids[j + 1];
                --rem;
                found = true;
            }
            if (!found) {
                return E_EAR;
            }
        }
        tri[nt][0] = (unsigned char)ids[0];
        tri[nt][1] = (unsigned char)ids[1];
        tri[nt][2] = (unsigned char)ids[2];
        for (int k = 0; k < ne - 2; ++k) {
            const int a = tri[k][0], b = tri[k][1], d = tri[k][2];
            tri_area += 0.5 * fabs((vx[b] - vx[a]) * (vy[d] - vy[a]) - (vx[d] - vx[a]) * (vy[b] - vy[a]));
        }
    }
    // volume quadrature: G, G_T, H, H*, alpha, internal rows of D, L, f
    double G[p * p] = {}, GT[p * p] = {}, H[q * q] = {}, Hs[gr * q + 1] = {};
    double D[NM * p] = {}, W[q * NM] = {}, B[p * NM] = {};
    double Lr[q] = {}, fr[q] = {};
    double Ti0 = 0, Ti1 = 0, Ti3 = 0;
    for (int k = 0; k < ne - 2; ++k) {
        const double v0 = vx[tri[k][0]], v1 = vy[tri[k][0]], v2 = vx[tri[k][1]], v3 = vy[tri[k][1]],
                     v4 = vx[tri[k][2]], v5 = vy[tri[k][2]];
        const double J = fabs((v2 - v0) * (v5 - v1) - (v4 - v0) * (v3 - v1));
        for (int i = 0; i < in.nt; ++i) {
            const double x = v0 + (v2 - v0) * in.ta[i] + (v4 - v0) * in.tb[i];
            const double y = v1 + (v3 - v1) * in.ta[i] + (v5 - v1) * in.tb[i];
            const double wl = in.tw[i] * J;
            const DGeo g = dmap(in, vx, vy, ne, er, chord, x, y);
            if (!(g.det > 0.0)) {
                return E_JAC;
            }
            const double xm = (x - xc) / h, ym = (y - yc) / h;
            double ax[p], ay[p], mo[nm];
#pragma unroll
            for (int a = 0; a < p; ++a) dhdiv<K>(a, xm, ym, h, ax[a], ay[a]);
#pragma unroll
            for (int r = 0; r < nm; ++r) {
                int ea, eb;
                dmono(r, ea, eb);
                mo[r] = dpow(xm, ea) * dpow(ym, eb);
            }
            const double t00 = (g.j00 * g.j00 + g.j10 * g.j10) / g.det,
                         t01 = (g.j00 * g.j01 + g.j10 * g.j11) / g.det,
                         t11 = (g.j01 * g.j01 + g.j11 * g.j11) / g.det;
            Ti0 += t00 * wl;
            Ti1 += t01 * wl;
            Ti3 += t11 * wl;
#pragma unroll
            for (int a = 0; a < p; ++a)
#pragma unroll
                for (int b = 0; b < p; ++b) {
                    G[a * p + b] += (ax[a] * ax[b] + ay[a] * ay[b]) * wl;
                    GT[a * p + b] += (ax[a] * (t00 * ax[b] + t01 * ay[b]) + ay[a] * (t01 * ax[b] + t11 * ay[b])) * wl;
                }
#pragma unroll
            for (int r = 0; r < q; ++r)
#pragma unroll
                for (int s = 0; s < q; ++s) H[r * q + s] += mo[r] * mo[s] * wl;
#pragma unroll
            for (int a = 0; a < gr; ++a)
#pragma unroll
                for (int s = 0; s < q; ++s) Hs[a * q + s] += mo[a + 1] * mo[s] * wl;
#pragma unroll
            for (int ii = 0; ii < S::ni; ++ii) {
                const int mom = ii < q - 1 ? ii : (ii - (q - 1)) + gr;
#pragma unroll
                for (int a = 0; a < p; ++a)
                    D[(edof + ii) * p + a] += (ax[a] * ax[mom] + ay[a] * ay[mom]) * wl;
            }
            double pe, ux, uy, gg;
            dexact(in.solution, g.X, g.Y, pe, ux, uy, gg);
#pragma unroll
            for (int r = 0; r < q; ++r) {
                Lr[r] += mo[r] * g.det * wl;
                fr[r] += g.det * gg * mo[r] * wl;
            }
        }
    }
    if (!(tri_area > 0.0)) {
        return E_AREA;
    }
    const double alpha = sqrt(Ti0 * Ti0 + Ti1 * Ti1 + Ti1 * Ti1 + Ti3 * Ti3) / tri_area;
    // edges: D edge rows, exact W edge columns, B_2, boundary data
    const int64_t v0 = in.v_off[c];
    for (int j = 0; j < ne; ++j) {
        const int j1 = (j + 1) % ne;
        const double x1 = vx[j], y1 = vy[j], x2 = vx[j1], y2 = vy[j1];
        const double dx = x2 - x1, dy = y2 - y1, len = sqrt(dx * dx + dy * dy);
        if (!(len > 1e-14 * h)) {
            return E_EDGE;
        }
        const double nx = dy / len, ny = -dx / len, tx = dx / len, ty = dy / len;
        const double mx = 0.5 * (x1 + x2), my = 0.5 * (y1 + y2);
        double ex[MAXQ1], ey[MAXQ1], ew[MAXQ1], et[MAXQ1];
        for (int i = 0; i < in.ng; ++i) {
            ex[i] = x1 + dx * (in.gx[i] + 1) / 2;
            ey[i] = y1 + dy * (in.gx[i] + 1) / 2;
            ew[i] = in.gw[i] * len / 2;
            et[i] = ((ex[i] - mx) * tx + (ey[i] - my) * ty) / len;
        }
        double De[kk * p] = {};
        for (int i = 0; i < in.ng; ++i) {
            double bx[p], by[p];
#pragma unroll
            for (int a = 0; a < p; ++a) dhdiv<K>(a, (ex[i] - xc) / h, (ey[i] - yc) / h, h, bx[a], by[a]);
#pragma unroll
            for (int r = 0; r < kk; ++r)
#pragma unroll
                for (int a = 0; a < p; ++a) De[r * p + a] += ew[i] * (bx[a] * nx + by[a] * ny) * dpow(et[i], r);
        }
#pragma unroll
        for (int r = 0; r < kk; ++r)
#pragma unroll
            for (int a = 0; a < p; ++a) D[(r * ne + j) * p + a] = De[r * p + a];
        // W edge columns: coefficients of m_r on the edge in t (exact)
#pragma unroll
        for (int r = 0; r < q; ++r) {
            int ea, eb;
            dmono(r, ea, eb);
            double co[K + 3] = {};
            const double ox = mx - xc, oy = my - yc;
            for (int i = 0; i <= ea; ++i)
                for (int jj = 0; jj <= eb; ++jj)
                    co[i + jj] += dbinom(ea, i) * dbinom(eb, jj) * dpow(ox, ea - i) * dpow(oy, eb - jj) *
                                  dpow(tx, i) * dpow(ty, jj);
            for (int dg = 0; dg <= ea + eb; ++dg) co[dg] *= dpow(len, dg) / dpow(h, ea + eb);
            for (int dg = 0; dg <= K; ++dg) W[r * n + dg * ne + j] += co[dg];
        }
        // B_2: H_e* H_e^{-1} W_e on this edge's columns
        double He[kk * kk], Hei[kk * kk], Hes[gr * kk];
#pragma unroll
        for (int r = 0; r < kk; ++r)
#pragma unroll
            for (int s = 0; s < kk; ++s) {
                double v = 0;
                for (int i = 0; i < in.ng; ++i) v += dpow(et[i], r) * dpow(et[i], s) * ew[i];
                He[r * kk + s] = v;
            }
#pragma unroll
        for (int a = 0; a < gr; ++a)
#pragma unroll
            for (int cc = 0; cc < kk; ++cc) {
                double v = 0;
                int ea, eb;
                dmono(a + 1, ea, eb);
                for (int i = 0; i < in.ng; ++i)
                    v += dpow((ex[i] - xc) / h, ea) * dpow((ey[i] - yc) / h, eb) * dpow(et[i], cc) * ew[i];
                Hes[a * kk + cc] = v;
            }
        if (!dinverse(He, kk, Hei)) {
            return E_SINGULAR;
        }
#pragma unroll
        for (int a = 0; a < gr; ++a)
#pragma unroll
            for (int cc = 0; cc < kk; ++cc) {
                double v = 0;
                for (int t2 = 0; t2 < kk; ++t2)
                    for (int s = 0; s < kk; ++s)
                        if (cc % ne == s) v += Hes[a * kk + t2] * Hei[t2 * kk + s];
                B[a * n + cc * ne + j] += v;
            }
        if constexpr (ERR) continue;  // no boundary data in the error mode
        // boundary data on this edge
        double rub[kk] = {}, ubb[kk] = {};
        char isbb = 0;
        if (in.edge_occ[in.cell_edges[s0 + j]] == 1) {
            double mom_p[kk] = {};
            for (int i = 0; i < in.ng; ++i) {
                const DGeo g = dmap(in, vx, vy, ne, er, chord, ex[i], ey[i]);
                double pe, ux, uy, gg;
                dexact(in.solution, g.X, g.Y, pe, ux, uy, gg);
                const double uhx = g.j11 * ux - g.j01 * uy, uhy = -g.j10 * ux + g.j00 * uy;
#pragma unroll
                for (int r = 0; r < kk; ++r) {
                    const double tr = dpow(et[i], r);
                    ubb[r] += (uhx * nx + uhy * ny) * tr * ew[i];
                    mom_p[r] += pe * tr * ew[i];
                }
            }
            if (in.bc == VEMD_BC_FLUX_MEANZERO) {
                isbb = 1;
            } else {
#pragma unroll
                for (int r = 0; r < kk; ++r) {
                    double v = 0;
                    for (int s = 0; s < kk; ++s) v += Hei[s * kk + r] * mom_p[s];
                    rub[r] = -v;
                    ubb[r] = 0.0;
                }
            }
        }
        // direction sign of the even moments of a reversed interior edge
        const int ge = in.cell_edges[s0 + j];
        const bool rev = in.edge_occ[ge] == 2 && in.nodes[s0 + j] == in.edge_nodes[2 * ge + 1];
#pragma unroll
        for (int r = 0; r < kk; ++r) {
            const double sg = (rev && r % 2 == 0) ? -1.0 : 1.0;
            o.ru[v0 + r * ne + j] = rub[r] * sg;
            o.ub[v0 + r * ne + j] = ubb[r];
            o.isb[v0 + r * ne + j] = isbb;
        }
    }
    for (int i = edof; i < n; ++i) {
        const int gd = i - edof;
        if (gd < q - 1) W[(gd + 1) * n + i] += -1.0;
        if constexpr (!ERR) {
            o.ru[v0 + i] = 0.0;  // internal dofs carry no boundary data
            o.ub[v0 + i] = 0.0;
            o.isb[v0 + i] = 0;
        }
    }
    // B_1 = -H* H^{-1} W, B_curl
    double Hinv[q * q];
    if (!dinverse(H, q, Hinv)) {
        return E_SINGULAR;
    }
    {
        double tmp[q * NM] = {};
        for (int r = 0; r < q; ++r)
            for (int s = 0; s < q; ++s)
                for (int i = 0; i < n; ++i) tmp[r * n + i] += Hinv[r * q + s] * W[s * n + i];
        for (int a = 0; a < gr; ++a)
            for (int r = 0; r < q; ++r)
                for (int i = 0; i < n; ++i) B[a * n + i] -= Hs[a * q + r] * tmp[r * n + i];
    }
    for (int r = 0; r < S::curl; ++r) B[(gr + r) * n + edof + (q - 1) + r] = 1.0;
    // P = G^{-1} B
    double Ginv[p * p], P[p * NM] = {};
    if (!dinverse(G, p, Ginv)) {
        return E_SINGULAR;
    }
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int i = 0; i < n; ++i) P[a * n + i] += Ginv[a * p + b] * B[b * n + i];
    if constexpr (ERR) {
        // errors_host on one cell: flux coefficients P u, divergence
        // coefficients H^-1 W u (H of the error rule), then the integrands at
        // the error rule's points of every sub-triangle
        const ErrIn& E = *ei;
        const int64_t gp = E.gpos[c];
        double u[NM], coef[p], wu[q], pc[q];
        for (int i = 0; i < n; ++i) {
            int64_t idx;
            double sgn = 1.0;
            if (i < edof) {
                const int r = i / ne, j = i % ne, ge = in.cell_edges[s0 + j];
                idx = (int64_t)kk * E.edge_int[ge] + r;
                if (r % 2 == 0 && in.edge_occ[ge] == 2 && in.nodes[s0 + j] == in.edge_nodes[2 * ge + 1]) sgn = -1.0;
            } else {
                idx = E.cbase + (int64_t)(i - edof) * E.cpad + gp;
            }
            u[i] = sgn * E.x[idx];
        }
#pragma unroll
        for (int r = 0; r < q; ++r) pc[r] = E.x[E.cbase + (int64_t)(S::ni + r) * E.cpad + gp];
#pragma unroll
        for (int a = 0; a < p; ++a) {
            double v = 0.0;
            for (int i = 0; i < n; ++i) v += P[a * n + i] * u[i];
            coef[a] = v;
        }
#pragma unroll
        for (int r = 0; r < q; ++r) {
            double v = 0.0;
            for (int i = 0; i < n; ++i) v += W[r * n + i] * u[i];
            wu[r] = v;
        }
        double He[q * q] = {};
        for (int k = 0; k < ne - 2; ++k) {
            const double v0x = vx[tri[k][0]], v0y = vy[tri[k][0]], v1x = vx[tri[k][1]], v1y = vy[tri[k][1]],
                         v2x = vx[tri[k][2]], v2y = vy[tri[k][2]];
            const double J = fabs((v1x - v0x) * (v2y - v0y) - (v2x - v0x) * (v1y - v0y));
            for (int i = 0; i < E.nt; ++i) {
                const double x = v0x + (v1x - v0x) * E.ta[i] + (v2x - v0x) * E.tb[i];
                const double y = v0y + (v1y - v0y) * E.ta[i] + (v2y - v0y) * E.tb[i];
                const double xm = (x - xc) / h, ym = (y - yc) / h;
                double mr[q];
#pragma unroll
                for (int r = 0; r < q; ++r) {
                    int ea, eb;
                    dmono(r, ea, eb);
                    mr[r] = dpow(xm, ea) * dpow(ym, eb);
                }
#pragma unroll
                for (int r = 0; r < q; ++r)
#pragma unroll
                    for (int t2 = 0; t2 < q; ++t2) He[r * q + t2] += mr[r] * mr[t2] * E.tw[i] * J;
            }
        }
        double dcf[q];
        {
            // H / |E| is O(1) for the scaled monomials
            const double sc0 = He[0] > 0 ? 1.0 / He[0] : 1.0;
            double Hsc[q * q], Hi[q * q];
#pragma unroll
            for (int i = 0; i < q * q; ++i) Hsc[i] = He[i] * sc0;
            if (!dinverse(Hsc, q, Hi)) return E_SINGULAR;
#pragma unroll
            for (int r = 0; r < q; ++r) {
                double v = 0.0;
#pragma unroll
                for (int t2 = 0; t2 < q; ++t2) v += Hi[r * q + t2] * wu[t2] * sc0;
                dcf[r] = v;
            }
        }
        double ep = 0, eu = 0, epp = 0, eup = 0, ed = 0, ms = 0;
        for (int k = 0; k < ne - 2; ++k) {
            const double v0x = vx[tri[k][0]], v0y = vy[tri[k][0]], v1x = vx[tri[k][1]], v1y = vy[tri[k][1]],
                         v2x = vx[tri[k][2]], v2y = vy[tri[k][2]];
            const double J = fabs((v1x - v0x) * (v2y - v0y) - (v2x - v0x) * (v1y - v0y));
            for (int i = 0; i < E.nt; ++i) {
                const double x = v0x + (v1x - v0x) * E.ta[i] + (v2x - v0x) * E.tb[i];
                const double y = v0y + (v1y - v0y) * E.ta[i] + (v2y - v0y) * E.tb[i];
                const double w = E.tw[i] * J;
                const double xm = (x - xc) / h, ym = (y - yc) / h;
                double ph = 0, vxh = 0, vyh = 0, dv = 0;
#pragma unroll
                for (int r = 0; r < q; ++r) {
                    int ea, eb;
                    dmono(r, ea, eb);
                    const double mv = dpow(xm, ea) * dpow(ym, eb);
                    ph += pc[r] * mv;
                    dv += dcf[r] * mv;
                }
#pragma unroll
                for (int a = 0; a < p; ++a) {
                    double bx, by;
                    dhdiv<K>(a, xm, ym, h, bx, by);
                    vxh += coef[a] * bx;
                    vyh += coef[a] * by;
                }
                const DGeo g = dmap(in, vx, vy, ne, er, chord, x, y);
                double pe, ux, uy, gg;
                dexact(in.solution, g.X, g.Y, pe, ux, uy, gg);
                const double uhx = g.j11 * ux - g.j01 * uy, uhy = -g.j10 * ux + g.j00 * uy;
                const double dp = ph - pe, dx = vxh - uhx, dy = vyh - uhy;
                ep += dp * dp * w;
                eu += (dx * dx + dy * dy) * w;
                epp += dp * dp * w * g.det;
                const double fx = (g.j00 * dx + g.j01 * dy) / g.det, fy = (g.j10 * dx + g.j11 * dy) / g.det;
                eup += (fx * fx + fy * fy) * w * g.det;
                const double dd = dv / g.det - gg;
                ed += dd * dd * w * g.det;
                ms += ph * w * g.det;
            }
        }
        double* eo = E.out + 6 * (int64_t)c;
        eo[0] = ep;
        eo[1] = eu;
        eo[2] = epp;
        eo[3] = eup;
        eo[4] = ed;
        eo[5] = ms;
        return 0;
    }
    // A = P^T G_T P + alpha (I - D P)^T (I - D P), lower triangle
    double T[p * NM] = {};
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b)
            for (int i = 0; i < n; ++i) T[a * n + i] += GT[a * p + b] * P[b * n + i];
    double Pd[NM * NM];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double v = 0.0;
            for (int a = 0; a < p; ++a) v += D[i * p + a] * P[a * n + j];
            Pd[i * n + j] = (i == j ? 1.0 : 0.0) - v;
        }
    double sg[NM];
    for (int i = 0; i < n; ++i) sg[i] = 1.0;
    for (int r = 0; r <= K; r += 2)
        for (int j = 0; j < ne; ++j) {
            const int ge = in.cell_edges[s0 + j];
            if (in.edge_occ[ge] == 2 && in.nodes[s0 + j] == in.edge_nodes[2 * ge + 1]) sg[r * ne + j] = -1.0;
        }
    double* Ao = o.A + in.a_off[c];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double v = 0.0;
            for (int a = 0; a < p; ++a) v += P[a * n + i] * T[a * n + j];
            for (int l = 0; l < n; ++l) v += alpha * Pd[l * n + i] * Pd[l * n + j];
            Ao[i * (i + 1) / 2 + j] = v * sg[i] * sg[j];
        }
    double* Wo = o.W + in.w_off[c];
    for (int r = 0; r < q; ++r)
        for (int i = 0; i < n; ++i) Wo[r * n + i] = W[r * n + i] * sg[i];
    double* Po = o.P + in.p_off[c];
    for (int a = 0; a < p; ++a)
        for (int i = 0; i < n; ++i) Po[a * n + i] = P[a * n + i];
    for (int r = 0; r < q; ++r) {
        o.L[(int64_t)c * q + r] = Lr[r];
        o.f[(int64_t)c * q + r] = fr[r];
    }
    o.alpha[c] = alpha;
    return 0;
}

template <int K>
__global__ __launch_bounds__(64) void k_setup(SetupIn in, SetupOut o) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= in.nc) return;
    const int c = in.order[t];
    const int e = setup_cell<K>(in, o, c);
    if (e && atomicCAS(&o.err[0], 0, e) == 0) atomicExch(&o.err[1], c);
}

template <int K>
__global__ __launch_bounds__(64) void k_errors(SetupIn in, ErrIn ei, int* err) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= in.nc) return;
    const int c = in.order[t];
    const int e = setup_cell<K, true>(in, SetupOut{}, c, &ei);
    if (e && atomicCAS(&err[0], 0, e) == 0) atomicExch(&err[1], c);
}

// Fixed-order sums of the six per-cell values: stage 1 sums 2048 cells per
// block (8 per thread, then a shuffle tree and the four wavefronts in order),
// stage 2 the block partials with one block.
__device__ __forceinline__ void sum6_block(double (&v)[6], double* out) {
    __shared__ double sh[6][4];
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] += __shfl_xor(v[k], o, 64);
    const int w = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) sh[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) out[threadIdx.x] = (sh[threadIdx.x][0] + sh[threadIdx.x][1]) + (sh[threadIdx.x][2] + sh[threadIdx.x][3]);
}

__global__ __launch_bounds__(256) void k_sum6(const double* in, int64_t n, double* out) {
    double v[6] = {0, 0, 0, 0, 0, 0};
    const int64_t b0 = (int64_t)blockIdx.x * 2048;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int64_t i = b0 + threadIdx.x + 256 * u;
        if (i < n)
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] += in[6 * i + k];
    }
    sum6_block(v, out + 6 * (int64_t)blockIdx.x);
}

// Device buffers of one setup call, released on every path.
struct Bufs {
    std::vector<void*> p;
    ~Bufs() {
        for (void* x : p) (void)hipFree(x);
    }
    template <typename T>
    T* alloc(size_t n) {
        void* x = nullptr;
        HIPCHK(hipMalloc(&x, std::max<size_t>(n * sizeof(T), 256)));
        p.push_back(x);
        return (T*)x;
    }
    template <typename T>
    T* up(const std::vector<T>& v) {
        T* x = alloc<T>(v.size());
        if (!v.empty()) HIPCHK(hipMemcpy(x, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return x;
    }
};

template <typename T>
void down(std::vector<T>& v, const T* d) {
    if (!v.empty()) HIPCHK(hipMemcpy(v.data(), d, v.size() * sizeof(T), hipMemcpyDeviceToHost));
}

}  // namespace

// Device copies of the mesh and numbering the error mode reads, kept per
// solver after the first device error evaluation.
struct ErrDev {
    Bufs b;
    SetupIn in{};
    const int* edge_int = nullptr;
    const int* gpos = nullptr;
    double* cell6 = nullptr;
    double* part = nullptr;
    int* err = nullptr;
    int device = -1;
};

void errdev_free(ErrDev* e) { delete e; }

// Cells in processing order: grouped by valence so that a wavefront's
// threads run the same loops.
static std::vector<int> valence_order(const Mesh& m) {
    std::vector<int> order(m.nc), cnt(65, 0), start(65, 0);
    for (int c = 0; c < m.nc; ++c) ++cnt[std::min(m.npc(c), 64)];
    for (int v = 1; v < 65; ++v) start[v] = start[v - 1] + cnt[v - 1];
    for (int c = 0; c < m.nc; ++c) order[start[std::min(m.npc(c), 64)]++] = c;
    return order;
}

// The setup's inputs that do not depend on a quadrature rule: mesh arrays on
// the device, problem kind, arcs.
static void setup_in_mesh(SetupIn& in, Bufs& b, const Mesh& m, const vemd_problem_desc& d) {
    in.nc = m.nc;
    in.map_kind = d.map_kind;
    in.solution = d.solution;
    in.bc = d.bc;
    in.eps = d.eps;
    in.arcs = d.map_kind == VEMD_MAP_ARCS && !m.arcs.empty();
    // arc radius of every edge (the chords of the curved cells)
    std::vector<double> er(in.arcs ? m.ne : 0, 0.0);
    for (int e = 0; e < (int)er.size(); ++e) {
        auto it = m.arcs.find(Mesh::key(m.edge_nodes[2 * e], m.edge_nodes[2 * e + 1]));
        if (it != m.arcs.end()) er[e] = it->second;
    }
    in.xy = b.up(m.xy);
    in.off = b.up(m.off);
    in.nodes = b.up(m.nodes);
    in.cell_edges = b.up(m.cell_edges);
    in.edge_nodes = b.up(m.edge_nodes);
    in.edge_occ = b.up(m.edge_occ);
    in.cx = b.up(m.cx);
    in.cy = b.up(m.cy);
    in.diam = b.up(m.diam);
    in.edge_r = b.up(er);
    in.order = b.up(valence_order(m));
}

static void setup_rules(SetupIn& in, const QuadRules& qr) {
    in.nt = (int)qr.tw.size();
    in.ng = (int)qr.glx.size();
    if (in.nt > MAXQ2 || in.ng > MAXQ1) throw Error(VEMD_ERR_ARG, "quadrature too large for the device setup");
    for (int i = 0; i < in.nt; ++i) {
        in.ta[i] = qr.ta[i];
        in.tb[i] = qr.tb[i];
        in.tw[i] = qr.tw[i];
    }
    for (int i = 0; i < in.ng; ++i) {
        in.gx[i] = qr.glx[i];
        in.gw[i] = qr.glw[i];
    }
}

static void throw_cell_error(const int (&err)[2]) {
    const std::string cell = std::to_string(err[1]);
    switch (err[0]) {
        case E_JAC: throw Error(VEMD_ERR_MESH, "non-positive map Jacobian in cell " + cell);
        case E_EAR: throw Error(VEMD_ERR_MESH, "no valid ear in cell " + cell);
        case E_EDGE: throw Error(VEMD_ERR_MESH, "degenerate edge in cell " + cell);
        case E_SINGULAR: throw Error(VEMD_ERR_SINGULAR, "singular element matrix in cell " + cell);
        case E_CONVEX: throw Error(VEMD_ERR_MESH, "curved boundary cell " + cell + " is not strictly convex");
        case E_AREA: throw Error(VEMD_ERR_MESH, "cell " + cell + " has no area");
        default:
            throw Error(VEMD_ERR_ARG, "cell " + cell + " has more than " + std::to_string(NE) +
                                          " edges (the device setup's limit)");
    }
}

void errors_device(Solver& s, const double* x_dev, int nq, void* stream, double* out6) {
    const hipStream_t st = (hipStream_t)stream;
    const Mesh& m = *s.mesh;
    const Setup& S = *s.setup;
    const vemd_problem_desc& d = S.desc;
    if (S.sh.k > 1) throw Error(VEMD_ERR_ARG, "device error functionals: k = 0 and k = 1");
    if (d.funcs) throw Error(VEMD_ERR_ARG, "problem callbacks run on the host: use vemd_errors_full for them");
    for (int c = 0; c < m.nc; ++c)
        if (m.npc(c) > NE) throw Error(VEMD_ERR_ARG, "device error functionals: polygons of up to 10 edges");
    if (!s.errdev || s.errdev->device != s.device) {
        errdev_free(s.errdev);
        s.errdev = nullptr;
        std::unique_ptr<ErrDev> e(new ErrDev());
        e->device = s.device;
        setup_in_mesh(e->in, e->b, m, d);
        setup_rules(e->in, quad_rules(d.nq2d, d.nq1d));
        e->in.v_off = e->b.up(S.ofs.v_off);
        e->edge_int = e->b.up(s.layout.edge_int);
        e->gpos = e->b.up(s.layout.gpos_of_cell);
        e->cell6 = e->b.alloc<double>(6 * (size_t)std::max(m.nc, 1));
        e->part = e->b.alloc<double>(6 * (size_t)((m.nc + 2047) / 2048 + 1));
        e->err = e->b.alloc<int>(2);
        s.errdev = e.release();
    }
    ErrDev& E = *s.errdev;
    ErrIn ei{};
    ei.x = x_dev;
    ei.edge_int = E.edge_int;
    ei.gpos = E.gpos;
    ei.cbase = s.layout.cbase;
    ei.cpad = s.layout.cpad;
    const QuadRules er = quad_rules(nq, 2);
    ei.nt = (int)er.tw.size();
    if (ei.nt > MAXQ2) throw Error(VEMD_ERR_ARG, "error quadrature too large for the device");
    for (int i = 0; i < ei.nt; ++i) {
        ei.ta[i] = er.ta[i];
        ei.tb[i] = er.tb[i];
        ei.tw[i] = er.tw[i];
    }
    ei.out = E.cell6;
    HIPCHK(hipMemsetAsync(E.err, 0, 2 * sizeof(int), st));
    if (m.nc > 0) {
        const dim3 grid((unsigned)((m.nc + 63) / 64)), block(64);
        if (S.sh.k == 0)
            hipLaunchKernelGGL(k_errors<0>, grid, block, 0, st, E.in, ei, E.err);
        else
            hipLaunchKernelGGL(k_errors<1>, grid, block, 0, st, E.in, ei, E.err);
        HIPCHK(hipGetLastError());
    }
    // fixed-order sums: cells -> blocks of 2048 -> ... -> one
    int64_t n = m.nc;
    const double* src = E.cell6;
    double* dst = E.part;
    double* spare = E.cell6;
    do {
        const int64_t nb = std::max<int64_t>(1, (n + 2047) / 2048);
        hipLaunchKernelGGL(k_sum6, dim3((unsigned)nb), dim3(256), 0, st, src, n, dst);
        HIPCHK(hipGetLastError());
        n = nb;
        src = dst;
        std::swap(dst, spare);
    } while (n > 1);
    double t[6];
    int err[2];
    HIPCHK(hipMemcpyAsync(t, src, sizeof(t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(err, E.err, sizeof(err), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (err[0]) throw_cell_error(err);
    for (int j = 0; j < 5; ++j) out6[j] = std::sqrt(t[j]);
    out6[5] = t[5];
}

Setup* setup_build_device(const Mesh& m, const vemd_problem_desc& d, int device) {
    if (d.k > 2) throw Error(VEMD_ERR_ARG, "the device element setup supports k = 0, 1 and 2");
    if (d.funcs)
        throw Error(VEMD_ERR_ARG, "problem callbacks run on the host: use vemd_solver_create for them");
    for (int c = 0; c < m.nc; ++c)
        if (m.npc(c) > NE)
            throw Error(VEMD_ERR_ARG, "cell " + std::to_string(c) + " has more than " + std::to_string(NE) +
                                          " edges (the device path's limit)");
    std::unique_ptr<Setup> S(setup_prepare(m, d));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(VEMD_ERR_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) throw Error(VEMD_ERR_ARG, "device out of range");
    HIPCHK(hipSetDevice(device));
    SetupIn in{};
    setup_rules(in, quad_rules(d.nq2d, d.nq1d));
    Bufs b;
    setup_in_mesh(in, b, m, d);
    in.a_off = b.up(S->ofs.a_off);
    in.w_off = b.up(S->ofs.w_off);
    in.v_off = b.up(S->ofs.v_off);
    in.p_off = b.up(S->p_off);
    SetupOut o{};
    o.A = b.alloc<double>(S->A.size());
    o.W = b.alloc<double>(S->W.size());
    o.L = b.alloc<double>(S->L.size());
    o.f = b.alloc<double>(S->f.size());
    o.ru = b.alloc<double>(S->ru.size());
    o.ub = b.alloc<double>(S->ub.size());
    o.P = b.alloc<double>(S->P.size());
    o.alpha = b.alloc<double>(S->alpha.size());
    o.isb = b.alloc<char>(S->isb.size());
    o.err = b.alloc<int>(2);
    HIPCHK(hipMemset(o.err, 0, 2 * sizeof(int)));
    const dim3 grid((unsigned)((m.nc + 63) / 64)), block(64);
    if (m.nc > 0) {
        if (d.k == 0)
            hipLaunchKernelGGL(k_setup<0>, grid, block, 0, 0, in, o);
        else if (d.k == 1)
            hipLaunchKernelGGL(k_setup<1>, grid, block, 0, 0, in, o);
        else
            hipLaunchKernelGGL(k_setup<2>, grid, block, 0, 0, in, o);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipDeviceSynchronize());
    int err[2] = {0, 0};
    HIPCHK(hipMemcpy(err, o.err, sizeof(err), hipMemcpyDeviceToHost));
    if (err[0]) throw_cell_error(err);
    down(S->A, o.A);
    down(S->W, o.W);
    down(S->L, o.L);
    down(S->f, o.f);
    down(S->ru, o.ru);
    down(S->ub, o.ub);
    down(S->P, o.P);
    down(S->alpha, o.alpha);
    down(S->isb, (const char*)o.isb);
    setup_finish(m, S.get());
    return S.release();
}

}  // namespace vemd
