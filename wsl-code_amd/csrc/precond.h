// Device side of the multigrid pressure preconditioner: the hierarchy in
// sliced-ELLPACK form (one wavefront per slice of 64 rows) and the V-cycle
// (Chebyshev smoothing with the Jacobi diagonal, Galerkin coarse levels, dense
// coarsest solve).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "amg.h"

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw ::vemd::Error(VEMD_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));  \
    } while (0)

namespace vemd {

// Krylov scalar state kept in device memory (MINRES recurrences, stop test);
// the host reads it only at the periodic convergence check.
struct MState {
    double beta1, beta, oldb, alfa, dbar, epsln, phibar, cs, sn, rho;
    double c_az, c_r1, c_r2;
    double u_v, u_w1, u_w2, u_x;
    double tol;
    int itn, maxit, done, pending;
};

// Block sliced ELLPACK with square blocks of bs rows: block row s * 64 + l
// has width[s] block entries; entry j sits at e = off[s] + j * 64 + l with
// block column col[e]; value (r, c) of its block is
// val[(off[s] + j * 64) * bs * bs + (r * bs + c) * 64 + l], so that the 64
// rows of a slice read every block component as 64 consecutive values.
// Padding entries carry zero blocks and column 0.  bs = 1 is sliced ELLPACK.
// Values are in val, or in half precision in hval (val null).
template <typename T>
struct Sell {
    int64_t nrows = 0;        // block rows
    int bs = 1;
    int nslices = 0;
    const int64_t* off = nullptr;
    const int* width = nullptr;
    const int* col = nullptr;
    const T* val = nullptr;
    const _Float16* hval = nullptr;
    float hscale = 1.0f;  // hval holds the values divided by hscale
    int lanes = 1;  // lanes per row in the single-column kernels of a scalar level (1 or 8)
};

// Compressed sparse rows (restriction: few long rows).
template <typename T>
struct CsrDev {
    int64_t nrows = 0;
    const int64_t* rp = nullptr;
    const int* ci = nullptr;
    const T* v = nullptr;
    const _Float16* hv = nullptr;  // values in half precision (divided by hscale) instead of v
    float hscale = 1.0f;
};

// Rows of 1 x bs blocks (bs = 1: plain compressed rows): row i holds the
// entries [rp[i], rp[i + 1]); entry e has block column bc[e] (columns
// bc[e] * bs .. bc[e] * bs + bs - 1) and values v[e * bs + c].
template <typename T>
struct BcsrDev {
    int64_t nrows = 0;
    int bs = 1;
    const int64_t* rp = nullptr;
    const int* bc = nullptr;
    const T* v = nullptr;
    const _Float16* hv = nullptr;  // values in half precision (divided by hscale) instead of v
    float hscale = 1.0f;
};

// Where the values of a host CSR matrix went in a device matrix: entry p of
// the CSR arrays sits at index at[p] of the device value array (of n values;
// at empty: the CSR order itself).  Kept from the upload so that new values
// of the same sparsity refill the device arrays without rebuilding them.
struct ValMap {
    std::vector<uint32_t> at;
    size_t n = 0;
    uint32_t* dev = nullptr;  // device copy of at (made by a device refill; owned by the cycle)
};

template <typename T>
struct AmgLevelDev {
    ValMap vA, vP, vR, vM, vRb;  // value placement of A, P, R, M, Rb
    int64_t n = 0;
    Sell<T> A, P;             // P: prolongation from the next level
    CsrDev<T> R;              // R = P^T
    // Restriction of the residual in one pass (single column): the next
    // level's right-hand side Rb b - M x with M = R A (scaled level: M carries
    // the scaling, M = R D^1/2 (D^-1/2 A D^-1/2), and Rb the plain R, which
    // acts on the unscaled b), both in rows of 1 x bs blocks.  M.nrows = 0:
    // residual and restriction apart.
    BcsrDev<T> M, Rb;
    T* dinv = nullptr;        // null: unit diagonal
    // Scaled level (sc non-null): the operator is D^-1/2 A D^-1/2 (unit
    // diagonal, entries in [-1, 1], stored in half precision) acting on
    // D^1/2 x; sc = D^-1/2 scales the level's right-hand side on entry and its
    // solution on exit, and P, R of the level carry D^1/2.
    T* sc = nullptr;
    T *b = nullptr, *x = nullptr, *r = nullptr, *d0 = nullptr, *d1 = nullptr;       // one column
    T *mb = nullptr, *mx = nullptr, *mr = nullptr, *md0 = nullptr, *md1 = nullptr;  // MCOLS interleaved columns
    double centre = 0.0, half = 0.0;  // Chebyshev interval of D^-1 A
    int degree = 2;
};

constexpr int MCOLS = 32;  // columns of the multi-vector cycle (right-hand sides)

template <typename T>
struct AmgDev {
    std::vector<AmgLevelDev<T>> lv;
    bool multi = false;       // level buffers for MCOLS columns allocated
    T* cinv = nullptr;        // dense inverse of the coarsest operator
    int64_t nc = 0;
    int64_t bytes = 0;        // device memory held
    std::vector<void*> allocs;
    ~AmgDev();
};

// Uploads the hierarchy in precision T (level buffers included).
template <typename T>
AmgDev<T>* amg_upload(const Amg& H);

// Krylov side of the cycle's last kernel (single column): instead of lv[0].x
// the finest level's last smoothing step writes the pressure block of the
// MINRES vector z (entry pb + r * cpad + g for moment r of store position g)
// and per-tile (64 positions) partials q1 = b.z and q2 = z.r2 (b the cycle's
// input), selecting z and r2 by the iteration parity of the tile's instance
// as the Krylov sweeps do; tiles of a finished instance are left alone.  init:
// z = Z[0] and no q2.
template <typename T>
struct KOut {
    T* Z[2];
    const T* R[2];
    const MState* st;
    const int* tile_inst;  // instance of every tile (null: one instance)
    int64_t pb, cpad;
    double *q1, *q2;
    int init;
};

// Refills the values of an uploaded hierarchy from a host hierarchy of the
// same sparsity (the aggregates kept by amg_build's reuse): operators,
// transfers, diagonals, Chebyshev intervals and the coarsest inverse.
// Returns false, changing nothing that a fresh upload would not replace,
// when the two do not match in shape or storage format.
template <typename T>
bool amg_refresh(AmgDev<T>& D, const Amg& H);

// Enqueues one V-cycle on the stream: x = B b with b = lv[0].b, x = lv[0].x.
// With done non-null every kernel returns at once while *done is set (work
// skipped inside a captured chunk of iterations after convergence).  With
// pre_done the producer of b has also written lv[0].d0 = b / centre times
// lv[0].sc (scaled level) or lv[0].dinv, the cycle's first update, which
// saves the cycle's first sweep over the level (only if amg_pre_fusable).
// With ko non-null and amg_kout_fusable(D) the cycle ends in the Krylov
// output described by KOut (lv[0].x is not written).
template <typename T>
void amg_vcycle(const AmgDev<T>& D, hipStream_t stream, const int* done, bool pre_done = false,
                const KOut<T>* ko = nullptr);
template <typename T>
bool amg_pre_fusable(const AmgDev<T>& D);
template <typename T>
bool amg_kout_fusable(const AmgDev<T>& D);

// The same cycle applied to MCOLS interleaved columns (entry (row, j) at
// row * MCOLS + j): b = lv[0].mb, x = lv[0].mx.  amg_multi_alloc first.
// pre_done as for amg_vcycle, with the first update in lv[0].md0.
template <typename T>
void amg_multi_alloc(AmgDev<T>& D);
template <typename T>
void amg_vcycle_multi(const AmgDev<T>& D, hipStream_t stream, const int* done, bool pre_done = false);

}  // namespace vemd
