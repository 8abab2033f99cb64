// C++ caller of the drop-in wrapper with its own problem data (test program):
// prints the element operator of cell 3 of an anisotropic problem, then
// checks that an exception thrown by a data field reaches the caller.
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "vemd.hpp"

static vemd::DarcyPdeData aniso() {
    const double pi = M_PI;
    vemd::DarcyPdeData d;
    d.kappa = [](double x, double y) {
        const double a = 0.3 + 0.2 * y;
        return vemd::Tensor2D{2.0 + x, a, a, 1.0 + y * y};
    };
    d.solution_p = [pi](double x, double y) { return std::sin(pi * x) * std::cos(pi * y); };
    d.solution_u = [pi](double x, double y) {
        const double a = 0.3 + 0.2 * y;
        const double px = pi * std::cos(pi * x) * std::cos(pi * y), py = -pi * std::sin(pi * x) * std::sin(pi * y);
        return vemd::Vector2D{-((2.0 + x) * px + a * py), -(a * px + (1.0 + y * y) * py)};
    };
    return d;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    vemd::StraightMeshReader reader;
    if (!reader.read_mesh(argv[1])) return 3;
    vemd::DarcyProblem problem;
    problem.set_problem_index(1);
    problem.init_problem();
    problem.set_pde_data(aniso());
    {
        vemd::DarcySolver solver(reader, problem, 9, 9, 1);
        vemd_solver s = nullptr;
        (void)s;
        int n = 0;
        double A[64 * 64];
        // the wrapper keeps its C handle private: build the same solver through the C ABI
        vemd_problem_desc d = problem.desc(1, 9, 9);
        vemd_solver h;
        vemd::check(vemd_solver_create(reader.handle(), &d, &h));
        vemd::check(vemd_element_matrices(h, 3, &n, A, nullptr, nullptr, nullptr, nullptr, nullptr));
        for (int i = 0; i < n * n; ++i) std::printf("%.17g%c", A[i], i + 1 == n * n ? '\n' : ' ');
        vemd_solver_destroy(h);
        if (solver.getTotalDof() <= 0) return 4;
    }
    vemd::DarcyProblem bad;
    bad.set_problem_index(1);
    bad.init_problem();
    vemd::DarcyPdeData d = aniso();
    d.kappa = [](double x, double) -> vemd::Tensor2D {
        if (x > -1.0) throw std::out_of_range("kappa out of range");
        return vemd::Tensor2D{};
    };
    bad.set_pde_data(d);
    try {
        vemd::DarcySolver solver(reader, bad, 9, 9, 1);
        std::printf("not caught\n");
    } catch (const std::out_of_range& e) {
        std::printf("caught: %s\n", e.what());
    }
    return 0;
}
