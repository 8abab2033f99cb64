// Smoothed-aggregation algebraic multigrid for the pressure block of the
// block-diagonal MINRES preconditioner: host construction (FP64) of the
// hierarchy for S = W diag(A)^-1 W^T, the pressure Schur complement with the
// flux block replaced by its assembled diagonal.
#pragma once

#include <cstdint>
#include <vector>

#include "store.h"

namespace vemd {

// Compressed sparse rows with 32-bit column indices.
struct Csr {
    int64_t nrows = 0, ncols = 0;
    std::vector<int64_t> rp;  // nrows + 1
    std::vector<int> ci;
    std::vector<double> v;
    int64_t nnz() const { return (int64_t)ci.size(); }
};

struct AmgOptions {
    int max_levels = 12;
    int coarse_size = 1500;    // the coarsest operator is inverted densely
    double strength = 0.08;    // strength-of-connection threshold on the finest level (halved per level)
    int degree0 = 2;           // Chebyshev steps before and after the coarse correction, finest level
    int degree = 2;            // the same on the coarser levels
    double cheb_ratio = 7.0;   // smoothing interval [lmax / ratio, lmax] of D^-1 A
};

struct AmgLevel {
    Csr A;                     // operator of this level
    std::vector<double> dinv;  // inverse of its diagonal
    double lmax = 0.0;         // upper estimate of the spectral radius of D^-1 A
    Csr P, R;                  // prolongation from the next coarser level and R = P^T (not on the coarsest)
    Csr M;                     // R A: the next level's right-hand side R (b - A x) = R b - M x (not on the coarsest)
    Csr Pt;                    // tentative prolongator (the aggregates), kept for a rebuild of the values
    int bs = 1;                // rows per node of this level (level 0: set by the caller)
    std::vector<double> ev;    // last iterate of the spectral-radius estimate (start of the next)
};

struct Amg {
    AmgOptions opt;
    int bs0 = 1;                      // block size of the finest level (pressure moments per cell)
    std::vector<AmgLevel> lv;         // lv[0] finest
    std::vector<double> coarse_inv;   // dense inverse of the coarsest operator, row-major
    int64_t ncoarse = 0;
    double rho = 0.0;                 // weight of the rank-one coarse term (0: none)
    int nvec0 = 1;                    // near-null vectors of the first coarsening (level 1 has that many rows per node)
};

// Pressure Schur approximation on the cells' pressure moments, cell-major:
// row gpos * q + r is moment r of the cell at store position gpos (q x q
// blocks per pair of cells sharing an edge); padding positions get unit rows.  dflux is the assembled flux diagonal (internal
// order, exp(-theta) included); essential boundary flux dofs are left out.
Csr schur_pressure(const Mesh& m, const Setup& S, const Layout& L, const std::vector<double>& dflux);

// The same in two parts: the theta-independent sparsity and per-dof slots
// (built once), then the values for a flux diagonal.
struct SchurPlan {
    Csr pattern;                      // structure of S (values empty)
    std::vector<int> self_slot;       // per cell: its own block among its row's blocks
    std::vector<int> nbk;             // per store position: blocks in its rows
    std::vector<int> other_slot;      // per local flux dof: the other cell's block (-1: none)
    std::vector<int> other_i;         // per local flux dof: its index in the other cell
    std::vector<int64_t> tflux;       // per local flux dof: internal slot (-1: essential dof)
};
SchurPlan schur_plan(const Mesh& m, const Setup& S, const Layout& L);
Csr schur_values(const SchurPlan& P, const Mesh& m, const Setup& S, const Layout& L, const std::vector<double>& dflux);

// Builds the hierarchy.  Rows are grouped into nodes (node_of_row, -1 = row
// left to the smoother), b0 is the near-null vector (one value per row) and
// node_order the order in which aggregates are grown.  With l0 (one value per
// row) the operator may be singular: the coarsest operator becomes A_c + rho
// l_c l_c^T, l_c = R ... R l0 (the restrictions of the levels), so that the
// V-cycle is symmetric positive definite and approximates (A + rho l0 l0^T)^-1
// (a null vector of A with a component along l0 is moved off zero).
// With reuse, the aggregates and depth of that hierarchy (built on an
// operator of the same sparsity) are kept and only the values are rebuilt:
// the smoothed prolongators, Galerkin operators, spectral bounds and the
// coarsest inverse (node_of_row, b0 and node_order are then unused).
// With row_inst (one instance per row of A0, -1 for none; aggregates never
// join rows of different instances) the rank-one term is one per instance:
// rho l_c l_c^T is added only where both coarse rows belong to one instance.
// With nv (further near-null vectors, each one value per row of A0) the
// first coarsening carries b0 and nv: per aggregate the vectors restricted
// to its rows, orthonormalised in order, so that level 1 has 1 + nv.size()
// rows per aggregate; the later coarsenings carry b0's coarse image only.
Amg amg_build(Csr A0, const std::vector<int>& node_of_row, int nnodes, const std::vector<double>& b0,
              const std::vector<int>& node_order, const AmgOptions& o, const std::vector<double>* l0 = nullptr,
              double rho = 0.0, const Amg* reuse = nullptr, const std::vector<int>* row_inst = nullptr,
              const std::vector<std::vector<double>>* nv = nullptr);

// In-place inverse of a dense row-major n x n matrix (Gauss-Jordan with
// partial pivoting; the coarsest operator).
void dense_inverse(std::vector<double>& a, int64_t n);

// Helpers exposed for tests: y = A x, C = A B, transpose.
void csr_spmv(const Csr& A, const double* x, double* y);
Csr csr_mult(const Csr& A, const Csr& B);
Csr csr_transpose(const Csr& A);

}  // namespace vemd
