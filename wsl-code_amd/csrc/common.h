// Shared host-side definitions of libvemd: status handling, the mesh, and the
// per-cell operator data produced by the FP64 setup.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/vemd.h"

namespace vemd {

// Error carrying a C-ABI status code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

// Number of worker threads for host setup (bounded by the CPU share that a
// machine gives one GPU).
int host_threads();

// Runs f(begin, end) over [0, n) on host_threads() threads.
template <class F>
void parallel_for(int64_t n, F f);

// ---------------------------------------------------------------------------
// Mesh (the reference's StraightMeshData conventions, CPU/mesh/straight_mesh.h:18-77)
// ---------------------------------------------------------------------------
struct Mesh {
    int nn = 0, nc = 0, ne = 0;
    std::vector<double> xy;             // node coordinates (x, y)
    std::vector<int> off, nodes;        // CCW cell nodes, offsets of size nc + 1
    std::vector<double> cx, cy, area, diam;
    std::vector<int> cell_edges;        // global edge of each local edge (same offsets)
    std::vector<int> edge_nodes;        // [min, max] node of each edge
    std::vector<int> edge_occ;          // 1 = boundary, 2 = interior
    std::vector<int> edge_cells;        // two cells per edge, second = -1 on the boundary
    // Circular boundary arcs centred at the origin: key (min << 32 | max) -> radius.
    std::unordered_map<uint64_t, double> arcs;
    // Disjoint instances of a batch: cells [inst_cell0[i], inst_cell0[i + 1])
    // (empty: one instance).
    std::vector<int> inst_cell0;
    int ninst() const { return inst_cell0.empty() ? 1 : (int)inst_cell0.size() - 1; }

    int npc(int c) const { return off[c + 1] - off[c]; }
    static uint64_t key(int a, int b) {
        uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
        return ((uint64_t)lo << 32) | hi;
    }
};

// Reorders cells CCW and computes the cell properties and edge numbering.
void mesh_finalize(Mesh& m);
Mesh* mesh_read_msh(const char* path);
Mesh* mesh_read_vtk(const char* path);
Mesh* mesh_from_arrays(int nn, const double* xy, int nc, const int* off, const int* nodes, int narc,
                       const int* arc_nodes, const double* arc_r);
Mesh* mesh_generate(const vemd_mesh_opts& o);
// Disjoint union of meshes (node and cell numbering concatenated in order).
Mesh* mesh_concat(const std::vector<const Mesh*>& parts);

// ---------------------------------------------------------------------------
// Element setup output (FP64, reference DOF conventions, direction signs
// already applied to rows and columns).
// ---------------------------------------------------------------------------
struct Shape {
    int k, q, p, grad, curl, ni;  // ni = internal flux dofs per cell
    explicit Shape(int kk = 1);
    int nflux(int ne) const { return (k + 1) * ne + ni; }
};

struct CellOps {
    // Offsets into the flat arrays of Setup for each cell.
    std::vector<int64_t> a_off, w_off, v_off;  // packed A, W (q x n), per-flux-dof vectors
};

struct Setup {
    Shape sh;
    vemd_problem_desc desc;   // desc.funcs points to funcs (or is null)
    vemd_problem_funcs funcs{};
    int64_t nflux_total = 0, npres_total = 0, ntot = 0;
    // Per cell, reference ordering of local flux dofs; A packed lower triangle
    // (i*(i+1)/2 + j, j <= i), K = I part (the cell scale exp(-theta) applied
    // separately), direction signs applied.  Essential boundary dofs are
    // decoupled only when the device store is packed.
    CellOps ofs;
    std::vector<double> A, W;        // packed A^(I), W (q x n)
    std::vector<double> L, f;        // per cell q: mean-zero multiplier row, pressure source
    std::vector<double> ru, ub;      // per flux dof: natural-bc rhs and essential boundary values
    std::vector<char> isb;           // per flux dof: essential boundary dof
    std::vector<int64_t> gdof;       // per flux dof: reference global dof
    std::vector<double> sgn;         // per flux dof: direction sign
    std::vector<double> alpha;       // per cell alpha^(I) (diagnostics)
    std::vector<double> theta;       // per cell log-permeability
    std::vector<double> P;           // per cell projector P (p x n), unsigned, for flux reconstruction
    std::vector<int64_t> p_off;
};

// Host (FP64, all cores) element setup.
Setup* setup_build(const Mesh& m, const vemd_problem_desc& d);
// The same element data computed on a GPU (k = 0 and k = 1, polygons of up to
// 10 edges) and copied back into the host arrays.
Setup* setup_build_device(const Mesh& m, const vemd_problem_desc& d, int device);
// Pieces shared by both: validation, offsets and sized arrays; the global dof
// map and signs; the quadrature rules.
Setup* setup_prepare(const Mesh& m, const vemd_problem_desc& d);
void setup_finish(const Mesh& m, Setup* S);
struct QuadRules {
    std::vector<double> ta, tb, tw;  // triangle rule on (0,0), (1,0), (0,1)
    std::vector<double> glx, glw;    // Gauss-Legendre on [-1, 1]
};
QuadRules quad_rules(int nq2d, int nq1d);

}  // namespace vemd

#include "parallel.h"
