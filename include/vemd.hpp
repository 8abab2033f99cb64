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

#include <exception>
#include <functional>
#include <limits>
#include <mutex>
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

// The reference's data types and its DarcyPdeData fields
// (CPU/examples/darcy_problem.h:17-33,162-180), and its mapping as the pair
// physical_coords / jacobian of IsoparametricMapping (:39-63).
struct Vector2D {
    double x = 0.0, y = 0.0;
};
struct Tensor2D {
    double xx = 1.0, xy = 0.0, yx = 0.0, yy = 1.0;
};
struct DarcyPdeData {
    std::function<Tensor2D(double x, double y)> kappa;
    std::function<double(double x, double y)> dirichlet_p;
    std::function<double(double x, double y, double nx, double ny)> neumann_un;
    std::function<double(double x, double y)> solution_p;
    std::function<Vector2D(double x, double y)> solution_u;
    std::function<double(double x, double y)> source_g;
};
struct IsoparametricMapping {
    std::function<Vector2D(double xi, double eta)> physical_coords;
    std::function<Tensor2D(double xi, double eta)> jacobian;  // xx = dx/dxi, xy = dx/deta, ...
};

class DarcyProblem {
public:
    DarcyProblem() = default;
    DarcyProblem(const DarcyProblem&) = delete;  // the solver keeps a pointer to it, as the reference does
    DarcyProblem& operator=(const DarcyProblem&) = delete;

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
    // Caller-defined data and map (any empty member keeps the built-in one).
    // The problem must outlive the solvers built on it.
    void set_pde_data(DarcyPdeData data) { pde_ = std::move(data); }
    void set_mapping(IsoparametricMapping map) { map_ = std::move(map); }
    const DarcyPdeData& get_pde_data() const { return pde_; }

    vemd_problem_desc desc(int k, int nq2d, int nq1d) const {
        vemd_problem_desc d = desc_;
        d.k = k;
        d.nq2d = nq2d;
        d.nq1d = nq1d;
        d.theta = theta_.empty() ? nullptr : theta_.data();
        funcs_ = vemd_problem_funcs{};
        funcs_.ctx = const_cast<DarcyProblem*>(this);
        if (pde_.kappa) funcs_.kappa = &DarcyProblem::t_kappa;
        if (pde_.solution_p) funcs_.solution_p = &DarcyProblem::t_solution_p;
        if (pde_.solution_u) funcs_.solution_u = &DarcyProblem::t_solution_u;
        if (pde_.source_g) funcs_.source_g = &DarcyProblem::t_source_g;
        if (pde_.dirichlet_p) funcs_.dirichlet_p = &DarcyProblem::t_dirichlet_p;
        if (pde_.neumann_un) funcs_.neumann_un = &DarcyProblem::t_neumann_un;
        if (map_.physical_coords && map_.jacobian) funcs_.map = &DarcyProblem::t_map;
        const bool any = funcs_.kappa || funcs_.solution_p || funcs_.solution_u || funcs_.source_g ||
                         funcs_.dirichlet_p || funcs_.neumann_un || funcs_.map;
        d.funcs = any ? &funcs_ : nullptr;
        return d;
    }
    // The first exception thrown by a callback since the last call (the
    // library sees NaN; the exception must not cross the C ABI).
    void rethrow_pending() const {
        std::exception_ptr e;
        {
            std::lock_guard<std::mutex> l(mu_);
            std::swap(e, err_);
        }
        if (e) std::rethrow_exception(e);
    }

private:
    int index_ = 0;
    vemd_problem_desc desc_{1, 9, 9, VEMD_MAP_SIN, 0.05, VEMD_SOL_SINCOS, VEMD_BC_FLUX_MEANZERO, nullptr, nullptr};
    std::vector<double> theta_;
    DarcyPdeData pde_;
    IsoparametricMapping map_;
    mutable vemd_problem_funcs funcs_{};
    mutable std::mutex mu_;
    mutable std::exception_ptr err_;

    static constexpr double nan_() { return std::numeric_limits<double>::quiet_NaN(); }
    template <class F>
    static void guarded(void* ctx, F&& f) {
        const DarcyProblem* p = static_cast<const DarcyProblem*>(ctx);
        try {
            f(*p);
        } catch (...) {
            std::lock_guard<std::mutex> l(p->mu_);
            if (!p->err_) p->err_ = std::current_exception();
        }
    }
    static void t_kappa(void* c, double x, double y, double* k) {
        k[0] = k[1] = k[2] = k[3] = nan_();
        guarded(c, [&](const DarcyProblem& p) {
            const Tensor2D t = p.pde_.kappa(x, y);
            k[0] = t.xx;
            k[1] = t.xy;
            k[2] = t.yx;
            k[3] = t.yy;
        });
    }
    static double t_solution_p(void* c, double x, double y) {
        double v = nan_();
        guarded(c, [&](const DarcyProblem& p) { v = p.pde_.solution_p(x, y); });
        return v;
    }
    static void t_solution_u(void* c, double x, double y, double* u) {
        u[0] = u[1] = nan_();
        guarded(c, [&](const DarcyProblem& p) {
            const Vector2D v = p.pde_.solution_u(x, y);
            u[0] = v.x;
            u[1] = v.y;
        });
    }
    static double t_source_g(void* c, double x, double y) {
        double v = nan_();
        guarded(c, [&](const DarcyProblem& p) { v = p.pde_.source_g(x, y); });
        return v;
    }
    static double t_dirichlet_p(void* c, double x, double y) {
        double v = nan_();
        guarded(c, [&](const DarcyProblem& p) { v = p.pde_.dirichlet_p(x, y); });
        return v;
    }
    static double t_neumann_un(void* c, double x, double y, double nx, double ny) {
        double v = nan_();
        guarded(c, [&](const DarcyProblem& p) { v = p.pde_.neumann_un(x, y, nx, ny); });
        return v;
    }
    static void t_map(void* c, double xi, double eta, double* xy, double* j) {
        xy[0] = xy[1] = j[0] = j[1] = j[2] = j[3] = nan_();
        guarded(c, [&](const DarcyProblem& p) {
            const Vector2D f = p.map_.physical_coords(xi, eta);
            const Tensor2D t = p.map_.jacobian(xi, eta);
            xy[0] = f.x;
            xy[1] = f.y;
            j[0] = t.xx;
            j[1] = t.xy;
            j[2] = t.yx;
            j[3] = t.yy;
        });
    }
};

class DarcySolver {
public:
    DarcySolver(const StraightMeshReader& reader, const DarcyProblem& problem, int gauss_point_num = 9,
                int gauss_point_num_1d = 9, int k = 1, int device = 0)
        : problem_(&problem), device_(device) {
        vemd_problem_desc d = problem.desc(k, gauss_point_num, gauss_point_num_1d);
        const vemd_status st = vemd_solver_create(reader.handle(), &d, &s_);
        problem.rethrow_pending();
        check(st);
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
    const DarcyProblem* problem_;
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
        const vemd_status st = vemd_errors_full(s_, solution_.data(), nq, e.data());
        problem_->rethrow_pending();
        check(st);
        return e;
    }
};

}  // namespace vemd

#endif  // VEMD_HPP
