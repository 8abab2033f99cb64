// vemd_darcy: convergence driver over the C++ host interface (include/vemd.hpp),
// following the reference's CPU/main/Darcy_curved_main.cpp:48-93 (read mesh,
// set up the problem, construct the solver, solve, report L2 errors and orders).
//
//   vemd_darcy [--example 1|2] [--k K] [--nq2 N] [--nq1 N] [--report] mesh1.msh [mesh2.msh ...]
//
// Prints one line per mesh: cells, dofs, pressure and flux L2 errors (in the
// computational measure, as the reference), observed orders, iterations and
// solve seconds. With --report, one more line per mesh with the error fields
// of the reference GPU driver's report (physical-domain pressure, flux and
// divergence errors, mass).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vemd.hpp"

int main(int argc, char** argv) {
    int example = 1, k = 1, nq2 = 9, nq1 = 9;
    bool report = false;
    std::vector<std::string> meshes;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--example" && i + 1 < argc)
            example = std::atoi(argv[++i]);
        else if (a == "--k" && i + 1 < argc)
            k = std::atoi(argv[++i]);
        else if (a == "--nq2" && i + 1 < argc)
            nq2 = std::atoi(argv[++i]);
        else if (a == "--nq1" && i + 1 < argc)
            nq1 = std::atoi(argv[++i]);
        else if (a == "--report")
            report = true;
        else
            meshes.push_back(a);
    }
    if (meshes.empty()) {
        std::fprintf(stderr, "usage: vemd_darcy [--example 1|2] [--k K] [--nq2 N] [--nq1 N] [--report] mesh.msh ...\n");
        return 2;
    }
    try {
        double prev_h = 0, prev_p = 0, prev_u = 0;
        std::printf("%-28s %8s %10s %16s %7s %16s %7s %7s %9s\n", "mesh", "cells", "dofs", "pressure_L2", "order",
                    "flux_L2", "order", "its", "solve_s");
        for (const std::string& path : meshes) {
            vemd::StraightMeshReader reader;
            if (!reader.read_mesh(path)) throw std::runtime_error("cannot read mesh " + path);
            int64_t counts[5];
            vemd::check(vemd_mesh_counts(reader.handle(), counts));
            vemd::DarcyProblem problem;
            problem.set_problem_index(example);
            if (!problem.init_problem()) throw std::runtime_error("unknown example");
            vemd::DarcySolver solver(reader, problem, nq2, nq1, k);
            solver.options().rtol = 1e-12;
            if (!solver.solve()) throw std::runtime_error("solve did not converge");
            double ep = solver.computeL2ErrorPressure(nq2), eu = solver.computeL2ErrorFlux(nq2);
            double h = 1.0 / std::sqrt((double)counts[1]);  // as CPU/main/Darcy_curved_main.cpp:77
            std::string op = "-", ou = "-";
            if (prev_h > 0) {
                char b[32];
                std::snprintf(b, sizeof b, "%.4f", std::log(prev_p / ep) / std::log(prev_h / h));
                op = b;
                std::snprintf(b, sizeof b, "%.4f", std::log(prev_u / eu) / std::log(prev_h / h));
                ou = b;
            }
            std::printf("%-28s %8lld %10lld %16.10e %7s %16.10e %7s %7d %9.4f\n", path.substr(path.rfind('/') + 1).c_str(),
                        (long long)counts[1], (long long)(solver.getTotalDof() + 1), ep, op.c_str(), eu, ou.c_str(),
                        solver.info().iterations, solver.info().seconds);
            if (report) {
                const vemd::DarcySolver::Report r = solver.computeErrorReport(nq2);
                std::printf("report scalar_comp %.16e flux_comp %.16e scalar_phys %.16e flux_phys %.16e div_phys %.16e "
                            "mass %.3e\n",
                            r.scalar_comp, r.flux_comp, r.scalar_phys, r.flux_phys, r.div_phys, r.mass);
            }
            prev_h = h;
            prev_p = ep;
            prev_u = eu;
        }
    } catch (const vemd::Error& e) {
        std::fprintf(stderr, "vemd_darcy: %s\n", e.what());
        return e.status == VEMD_ERR_NO_DEVICE ? 3 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "vemd_darcy: %s\n", e.what());
        return 1;
    }
    return 0;
}
