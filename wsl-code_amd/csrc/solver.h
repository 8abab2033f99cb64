// Solver object shared by the C ABI and the device code.
#pragma once

#include <memory>
#include <vector>

#include "amg.h"
#include "store.h"

namespace vemd {

struct DeviceState;
struct ErrDev;  // device copies for the error functionals (setup.hip)

struct Solver {
    std::shared_ptr<Mesh> mesh;
    std::unique_ptr<Setup> setup;
    Layout layout;
    std::vector<double> scale;     // exp(-theta) per store position
    std::vector<double> minv;      // preconditioner inverse, internal order
    std::vector<double> rowscale;  // internal = rowscale * reference-form rows
    std::vector<double> rhs_ref;   // reference-form right-hand side
    bool rhs_changed = true;       // the last theta refresh rebuilt rhs_ref
    std::unique_ptr<DiagPlan> dplan;  // index data of the preconditioner diagonal (theta-independent)
    // multigrid hierarchy of the pressure Schur approximation (with the
    // mean-zero multiplier: of that approximation plus a rank-one term along
    // the multiplier row, whose inverse weight preconditions the multiplier)
    std::unique_ptr<Amg> amg;
    AmgOptions amg_opt;
    std::vector<double> amg_lam_minv;  // per instance: multiplier entry of M^-1 with the multigrid cycle (empty: no multiplier)
    std::unique_ptr<SchurPlan> splan;  // sparsity of the pressure Schur approximation (theta-independent)
    bool amg_values_only = false;  // the last rebuild kept every sparsity pattern of the hierarchy
    // the hierarchy's values were last rebuilt on the device: the host copy
    // keeps its patterns, aggregates, spectral bounds and coarsest inverse, but
    // not its matrix values (rebuilt on the host before another upload)
    bool amg_host_stale = false;
    double amg_setup_s = 0.0;
    DeviceState* dev = nullptr;
    ErrDev* errdev = nullptr;  // allocated by the first device error evaluation
    int device = -1;
    // per-instance results of the last solve (iterations, true relative residual)
    std::vector<int> inst_iterations;
    std::vector<double> inst_rel;
    vemd_solve_info last_info{};
    ~Solver();
    // host-side conversions between reference order and internal order
    void to_internal(const double* ref, double* in, bool scale_rows) const;
    void to_reference(const double* in, double* ref, bool unscale_rows) const;
    // host_amg false: the multigrid values are left to the device rebuild
    void refresh_theta_dependent(bool host_amg = true);
    void build_amg();
};

int device_count();
void dev_upload(Solver& s, int device);
// Cell scales and preconditioner after a theta change; with device_mg the
// multigrid values are rebuilt on the device (on the host if that fails).
void dev_refresh(Solver& s, bool device_mg = false);
// Whether a theta change can rebuild the multigrid values on the device.
bool dev_mg_rebuild_possible(const Solver& s);
void dev_free(Solver& s);
// Apply variant chosen by the autotuner for valences 3..10 (FP64, FP32 pairs; -1: none).
int dev_tuning(const Solver& s, int* out16);
void dev_apply_ref(Solver& s, int precision, const double* x_ref, double* y_ref);
// y1 == nullptr: complete apply; otherwise y1 receives the edges' second sides.
void dev_apply_int(Solver& s, int precision, const void* x, void* y, void* h, void* stream);
// The same complete apply with its side slots and partials in the library's
// static device array (no allocation); the bytes of that array it needs.
void dev_apply_static(Solver& s, int precision, const void* x, void* y, void* stream);
size_t dev_static_scratch_need(const Solver& s, int precision);
void dev_solve(Solver& s, const vemd_solve_opts& o, const double* rhs_ref, double* sol_ref, vemd_solve_info* info);
void dev_solve_int(Solver& s, const vemd_solve_opts& o, const void* b_dev, void* x_dev, vemd_solve_info* info);
// Up to 32 right-hand sides at once (FP32 MFMA iterations, FP64 refinement):
// b and x are ncol column-major device vectors of the internal layout.
void dev_solve_multi(Solver& s, const vemd_solve_opts& o, int ncol, const double* b, double* x, vemd_solve_info* info);
// MFMA apply on 32 interleaved FP32 columns (second sides of edge dofs in y1).
void dev_apply_multi(Solver& s, const void* x, void* y, void* y1, void* stream);
void dev_adjoint(Solver& s, const vemd_solve_opts& o, const double* sol_ref, const double* dldx_ref, double* grad,
                 vemd_solve_info* info);
// The six error functionals of errors_host for an internal-layout solution in
// device memory, evaluated on the device (after the caller's stream's work).
void dev_errors(Solver& s, const void* x_dev, int nq, void* stream, double* out6);
// setup.hip: the evaluation on a stream of the solver's device, and the release
// of its device copies.
void errors_device(Solver& s, const double* x_dev, int nq, void* stream, double* out6);
void errdev_free(ErrDev* e);

}  // namespace vemd
