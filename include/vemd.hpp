// vemd.hpp -- header-only C++ host interface over the C ABI (vemd.h), shaped
// like the reference's classes for this path so that a caller of
//   vem::StraightMeshReader (CPU/mesh/straight_mesh.h:80-143),
//   Darcy::DarcyProblem    (CPU/examples/darcy_problem.h:185-229),
//   DarcySolver            (CPU/solver/DarcySolver.h:17-186)
// keeps its call sequence: read_mesh -> set_problem_index/init_problem ->
// DarcySolver(reader, problem, nq2d, nq1d, k) -> solve() -> getSolution() ->
// computeL2ErrorPressure/Flux().  Errors surface as vemd::Error exceptions
// built from the C status codes; nothing throws across the C ABI itself.
#ifndef VEMD_HPP
#define VEMD_HPP

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vemd.h"

namespace vemd {

struct Error : std::runtime_error {
    vemd_status status;
    Error(vemd_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(vemd_status s) {
    if (s != VEMD_OK) throw Error(s, vemd_last_error());
}

class StraightMeshReader {
public:
    StraightMeshReader() = default;
    ~StraightMeshReader() { reset(); }
    StraightMeshReader(const StraightMeshReader&) = delete;
    StraightMeshReader& operator=(const StraightMeshReader&) = delete;

    bool read_mesh(const std::string& filename) { return load(filename, VEMD_MESH_MSH); }
    bool read_mesh_vtk(const std::string& filename) { return load(filename, VEMD_MESH_VTK); }
    void generate(const vemd_mesh_opts& o) {
        reset();
        check(vemd_mesh_generate(&o, &m_));
    }
    vemd_mesh handle() const { return m_; }

private:
    vemd_mesh m_ = nullptr;
    bool load(const std::string& f, int fmt) {
        reset();
        return vemd_mesh_read(f.c_str(), fmt, &m_) == VEMD_OK;
    }
    void reset() {
        if (m_) vemd_mesh_destroy(m_);
        m_ = nullptr;
    }
};

class DarcyProblem {
public:
    void set_problem_index(int i) { index_ = i; }
    bool init_problem() {
        desc_.map_kind = VEMD_MAP_SIN;
        desc_.solution = VEMD_SOL_SINCOS;
        desc_.bc = VEMD_BC_FLUX_MEANZERO;
        if (index_ == 1) {
            desc_.eps = 0.05;  // sinusoidal curved map
        } else if (index_ == 2) {
            desc_.eps = 0.0;   // straight limit
        } else {
            return false;
        }
        return true;
    }
    // Extensions of the reference problem: per-cell arc maps, the sin*sin
    // solution, Dirichlet pressure and a cellwise log-permeability.
    void set_arcs(bool on) { desc_.map_kind = on ? VEMD_MAP_ARCS : VEMD_MAP_SIN; }
    void set_solution(int s) { desc_.solution = s; }
    void set_bc(int bc) { desc_.bc = bc; }
    void set_theta(std::vector<double> theta) { theta_ = std::move(theta); }
    vemd_problem_desc desc(int k, int nq2d, int nq1d) const {
        vemd_problem_desc d = desc_;
        d.k = k;
        d.nq2d = nq2d;
        d.nq1d = nq1d;
        d.theta = theta_.empty() ? nullptr : theta_.data();
        return d;
    }

private:
    int index_ = 0;
    vemd_problem_desc desc_{1, 9, 9, VEMD_MAP_SIN, 0.05, VEMD_SOL_SINCOS, VEMD_BC_FLUX_MEANZERO, nullptr};
    std::vector<double> theta_;
};

class DarcySolver {
public:
    DarcySolver(const StraightMeshReader& reader, const DarcyProblem& problem, int gauss_point_num = 9,
                int gauss_point_num_1d = 9, int k = 1, int device = 0)
        : device_(device) {
        vemd_problem_desc d = problem.desc(k, gauss_point_num, gauss_point_num_1d);
        check(vemd_solver_create(reader.handle(), &d, &s_));
        int64_t z[6];
        check(vemd_solver_sizes(s_, z));
        nflux_ = z[0];
        npres_ = z[1];
        ntot_ = z[2];
        opts_ = vemd_solve_opts{VEMD_FP64, 1e-10, 1000000, 1e-4, 32, 1, VEMD_PC_AUTO, nullptr};
    }
    ~DarcySolver() {
        if (s_) vemd_solver_destroy(s_);
    }
    DarcySolver(const DarcySolver&) = delete;
    DarcySolver& operator=(const DarcySolver&) = delete;

    int64_t getTotalDofFlux() const { return nflux_; }
    int64_t getTotalDofConc() const { return npres_; }
    int64_t getTotalDof() const { return nflux_ + npres_; }
    vemd_solve_opts& options() { return opts_; }

    // Assembles (HBM store upload on first use) and solves; false when the
    // Krylov solve does not converge.  A missing device throws.
    bool solve() {
        if (!uploaded_) {
            check(vemd_solver_upload(s_, device_));
            uploaded_ = true;
        }
        solution_.assign(ntot_, 0.0);
        vemd_status st = vemd_solve(s_, &opts_, nullptr, solution_.data(), &info_);
        if (st == VEMD_ERR_NOT_CONVERGED) return false;
        check(st);
        return true;
    }
    const std::vector<double>& getSolution() const { return solution_; }
    const vemd_solve_info& info() const { return info_; }
    double computeL2ErrorPressure(int nq = 9) { return errors(nq)[0]; }
    double computeL2ErrorFlux(int nq = 9) { return errors(nq)[1]; }
    // The error fields of the reference GPU driver's report.json
    // (GPU/src/solver_driver.cpp:340-351): scalar_comp, flux_comp, scalar_phys,
    // flux_phys, div_phys, mass.
    struct Report {
        double scalar_comp, flux_comp, scalar_phys, flux_phys, div_phys, mass;
    };
    Report computeErrorReport(int nq = 9) {
        const std::vector<double> e = errors(nq);
        return Report{e[0], e[1], e[2], e[3], e[4], e[5]};
    }

private:
    vemd_solver s_ = nullptr;
    int device_;
    bool uploaded_ = false;
    int64_t nflux_ = 0, npres_ = 0, ntot_ = 0;
    vemd_solve_opts opts_;
    vemd_solve_info info_{};
    std::vector<double> solution_;
    std::vector<double> errors(int nq) {
        if (solution_.empty()) throw std::runtime_error("errors require a completed numerical solution");
        std::vector<double> e(6);
        check(vemd_errors_full(s_, solution_.data(), nq, e.data()));
        return e;
    }
};

}  // namespace vemd

#endif  // VEMD_HPP
