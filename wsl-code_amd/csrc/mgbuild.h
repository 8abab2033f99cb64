// Device rebuild of the multigrid hierarchy's values after a theta change.
//
// A theta change alters the flux diagonal and with it every value of the
// hierarchy, but not its aggregates or any sparsity pattern (the host build
// with reuse keeps them).  The rebuild on the GPU recomputes, level by level
// and in FP64, the values the host build computes: the pressure Schur
// approximation S = W diag(A)^-1 W^T, the inverse diagonal, the spectral
// bound of D^-1 A by the same power iteration (warm-started from the previous
// iterate), the smoothed prolongator P = (I - omega D^-1 A) Pt, R = P^T,
// M = R A and the Galerkin operator M P; every sum runs in the host's order
// and no multiply-add is fused, so the values are bitwise those of the host
// build.  Only the coarsest operator (a few thousand values) and the
// restricted multiplier row go back to the host, which forms and inverts the
// dense coarsest matrix.  The uploaded cycles are then refilled in place on
// the device from these values (precision conversions, the diagonal scaling
// and half-precision normalisation of the FP32 finest level).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "precond.h"

namespace vemd {

struct Solver;

// Device copy of a host CSR pattern.
struct CsrPat {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    int64_t* rp = nullptr;
    int* ci = nullptr;
};

struct MgLevelBuild {
    CsrPat A, Pt, P, R, M;  // P, R, M, Pt: not on the coarsest level
    double *Av = nullptr, *Ptv = nullptr, *Pv = nullptr, *Rv = nullptr, *Mv = nullptr;
    int64_t* diag = nullptr;   // slot of every row's diagonal entry (-1: none)
    int* rperm = nullptr;      // entry e of R is entry rperm[e] of P
    double* dinv = nullptr;    // inverse diagonal (1 where it is zero or missing)
    double* ev = nullptr;      // power-iteration vector (kept between rebuilds)
    double* y = nullptr;       // power-iteration scratch
    double* l = nullptr;       // multiplier row restricted to this level (mean-zero problems)
    double lmax = 0.0;         // 1.1 x the spectral estimate of D^-1 A (host copy after a rebuild)
};

struct MgBuild {
    std::vector<MgLevelBuild> lv;
    int q = 1;
    int64_t cpad = 0, nlive = 0;
    double ln0 = 0.0;          // sum of the multiplier row over the first pressure moments
    bool meanzero = false;
    // level-0 operator (the Schur approximation), per store position, cell
    // and local flux dof
    const int* cell_of_gpos = nullptr;  // borrowed from the solver's device state
    int* nbk = nullptr;
    int* self_slot = nullptr;
    int* n1 = nullptr;
    int64_t *v0 = nullptr, *w0 = nullptr;
    int64_t* tflux = nullptr;
    int* oslot = nullptr;
    int64_t* ow = nullptr;
    int* on2 = nullptr;
    double* W = nullptr;
    double* scal = nullptr;    // per level [lam, 1/|y|, stop]; the diagonal sum; storage statistics
    double* part = nullptr;    // per-block partials of the statistics (npart pairs)
    double* gpart = nullptr;   // per-64-row sums of the power iteration's norms (two arrays)
    int npart = 0;
    double rho = 0.0;                  // rank-one weight of the last rebuild
    std::vector<double> cinv;          // its coarsest inverse
    std::vector<void*> allocs;
    int64_t bytes = 0;
    ~MgBuild();
};

// Index data and warm-start vectors from the host hierarchy s.amg (built with
// the solver's current aggregates) and the solver's setup; cell_of_gpos is
// the device copy of the layout's table.
MgBuild* mg_build_create(const Solver& s, const int* cell_of_gpos_dev);

// Rebuilds the values on the stream for the flux-diagonal inverse minv
// (internal order, FP64, in device memory), then refills both uploaded
// cycles.  Updates the host hierarchy's spectral bounds, rank-one weight and
// coarsest inverse and s.amg_lam_minv; the host hierarchy's other values are
// then stale (Solver::amg_host_stale).  Returns false when a value is not
// finite or the storage formats of the uploaded cycles would change; the
// FP64 cycle's operators may then hold new values already, and the caller
// rebuilds on the host and refills both cycles.
bool mg_build_run(MgBuild& B, Solver& s, const double* minv_dev, AmgDev<double>& D64, AmgDev<float>& D32,
                  hipStream_t st);

// Compares the device values of the last rebuild with the host hierarchy
// s.amg (rebuilt from the same state on the host): the number of values that
// differ, with the largest relative difference in *maxrel.
int64_t mg_build_compare(const MgBuild& B, const Solver& s, double* maxrel);

}  // namespace vemd
