// Test-only C driver over the reference project's PETSc-free modules
// (quadrature, polynomial basis, mesh reader, triangulator, Darcy problem
// data).  It is compiled together with those reference sources, where they lie
// under the reference tree, into oracle/_ref/libref.so (see oracle/Makefile) so
// that tests can pin the oracle's restatement against the reference itself.
// Nothing here is product code.

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "core/gauss_quadrature.h"
#include "examples/darcy_problem.h"
#include "lib/polynomial_basis.h"
#include "mesh/polygon_triangulator.h"
#include "mesh/straight_mesh.h"

static std::unique_ptr<vem::StraightMeshReader> g_reader;

extern "C" {

// Reads a mesh (vtk != 0: legacy VTK); returns [nn, nc, ne, total cell nodes].
int ref_read(const char* path, int vtk, int* counts) {
    g_reader.reset(new vem::StraightMeshReader());
    bool ok = vtk ? g_reader->read_mesh_vtk(path) : g_reader->read_mesh(path);
    if (!ok) return -1;
    const auto& m = g_reader->get_mesh_data();
    counts[0] = m.num_nodes;
    counts[1] = m.num_cells;
    counts[2] = m.num_edges;
    counts[3] = (int)m.cell_nodes.size();
    return 0;
}

int ref_mesh_get(double* cx, double* cy, double* area, double* diam, int* off, int* nodes, int* cell_edges,
                 int* edge_nodes, int* edge_occ) {
    const auto& m = g_reader->get_mesh_data();
    for (int c = 0; c < m.num_cells; ++c) {
        cx[c] = m.cell_centroid_x[c];
        cy[c] = m.cell_centroid_y[c];
        area[c] = m.cell_area[c];
        diam[c] = m.cell_diameter[c];
    }
    for (size_t i = 0; i < m.cell_node_indices.size(); ++i) off[i] = m.cell_node_indices[i];
    for (size_t i = 0; i < m.cell_nodes.size(); ++i) nodes[i] = m.cell_nodes[i];
    for (size_t i = 0; i < m.global_edges.size(); ++i) cell_edges[i] = m.global_edges[i];
    for (size_t i = 0; i < m.edge_endpoints.size(); ++i) edge_nodes[i] = m.edge_endpoints[i];
    for (size_t i = 0; i < m.edge_occurrence.size(); ++i) edge_occ[i] = m.edge_occurrence[i];
    return 0;
}

int ref_triangulate(int cell, double* verts, int cap) {
    vem::PolygonTriangulator t;
    vem::TriangulatedElement e = t.triangulateElement(g_reader->get_mesh_data(), cell);
    int n = (int)e.triangles.size();
    for (int i = 0; i < n && i < cap; ++i)
        for (int j = 0; j < 6; ++j) verts[6 * i + j] = e.triangles[i].vertices[j];
    return n;
}

int ref_triangle_points(int npts, const double* v, double* out) {
    vem::GaussQuadrature q;
    std::vector<double> vv(v, v + 6);
    auto pts = q.get_triangle_points(vv, npts);
    for (size_t i = 0; i < pts.size(); ++i) {
        out[3 * i] = pts[i].x[0];
        out[3 * i + 1] = pts[i].x[1];
        out[3 * i + 2] = pts[i].w;
    }
    return (int)pts.size();
}

int ref_line_points(int n, double x1, double y1, double x2, double y2, double* out) {
    vem::GaussQuadrature q;
    auto pts = q.get_line_2D_points(vem::QuadratureType::GAUSS_LEGENDRE, n, x1, y1, x2, y2);
    for (size_t i = 0; i < pts.size(); ++i) {
        out[3 * i] = pts[i].x[0];
        out[3 * i + 1] = pts[i].x[1];
        out[3 * i + 2] = pts[i].w;
    }
    return (int)pts.size();
}

int ref_basis_eval(int k, int kind, int nk, double xc, double yc, double h, double x, double y, double* out) {
    vem::basis::BasisFunctionPloy b(k);
    vem::basis::Point2D c(xc, yc), p(x, y);
    if (kind == 0) {
        out[0] = b.evalMonomial2D(nk, c, h, p);
    } else {
        auto v = b.evalHdivBasis(nk, c, h, p);
        out[0] = v.x;
        out[1] = v.y;
    }
    return 0;
}

int ref_edge_coeffs(int k, int nk, double xc, double yc, double h, double x1, double y1, double x2, double y2,
                    double* out) {
    vem::basis::BasisFunctionPloy b(k);
    std::vector<double> c;
    b.getEdgeBasisCoeffs(nk, vem::basis::Point2D(xc, yc), h, vem::basis::Point2D(x1, y1),
                         vem::basis::Point2D(x2, y2), c);
    for (size_t i = 0; i < c.size(); ++i) out[i] = c[i];
    return (int)c.size();
}

double ref_edge_mono(int k, int nk, double x0, double y0, double x1, double y1, double x, double y) {
    vem::basis::BasisFunctionPloy b(k);
    return b.evalEdgeMonomial(nk, vem::basis::Point2D(x0, y0), vem::basis::Point2D(x1, y1),
                              vem::basis::Point2D(x, y));
}

// Problem data on the computational domain: [p, ux, uy, g, Kxx, Kxy, Kyx, Kyy, detJ].
int ref_problem_data(int index, double xi, double eta, double* out) {
    Darcy::DarcyProblem p;
    p.set_problem_index(index);
    if (!p.init_problem()) return -1;
    out[0] = p.computational_solution_p(xi, eta);
    auto u = p.computational_solution_u(xi, eta);
    out[1] = u.x;
    out[2] = u.y;
    out[3] = p.computational_source_g(xi, eta);
    auto K = p.computational_kappa(xi, eta);
    out[4] = K.xx;
    out[5] = K.xy;
    out[6] = K.yx;
    out[7] = K.yy;
    out[8] = p.get_mapping().jacobian_det(xi, eta);
    return 0;
}

}  // extern "C"
