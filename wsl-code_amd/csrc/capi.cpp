// extern "C" entry points of libvemd (declared in include/vemd.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <fstream>
#include <memory>
#include <thread>

#include "solver.h"

namespace vemd {

static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }

int host_threads() {
    static int n = [] {
        const char* e = std::getenv("VEMD_THREADS");
        int v = e ? std::atoi(e) : (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(v, 16));
    }();
    return n;
}

Solver::~Solver() {
    try {
        dev_free(*this);
    } catch (...) {
    }
}

void Solver::to_internal(const double* ref, double* in, bool scale_rows) const {
    std::fill(in, in + layout.nvec, 0.0);
    for (int64_t g = 0; g < setup->ntot; ++g) {
        int64_t i = layout.int_of_ref[g];
        in[i] = scale_rows ? ref[g] * rowscale[i] : ref[g];
    }
}

void Solver::to_reference(const double* in, double* ref, bool unscale_rows) const {
    for (int64_t g = 0; g < setup->ntot; ++g) {
        int64_t i = layout.int_of_ref[g];
        ref[g] = unscale_rows ? in[i] / rowscale[i] : in[i];
    }
}

void Solver::refresh_theta_dependent(bool host_amg) {
    auto t0 = std::chrono::steady_clock::now();
    if (!dplan) dplan.reset(new DiagPlan(diag_plan(*mesh, *setup, layout)));
    pack_scale(*setup, layout, scale);
    build_minv(*mesh, *setup, layout, *dplan, minv);
    // the row scales depend on theta only through essential flux rows
    if (rowscale.empty() || dplan->essential) build_row_scale(*mesh, *setup, layout, rowscale);
    // without essential dofs the right-hand side has no theta-dependent lift
    rhs_changed = rhs_ref.empty() || dplan->essential;
    if (rhs_changed) build_rhs_ref(*mesh, *setup, rhs_ref);
    auto t1 = std::chrono::steady_clock::now();
    // the multigrid cycle: k = 0, 1 and 2 (q = 1, 3 or 6 pressure moments per cell)
    if (setup->sh.k <= 2 && host_amg) build_amg();
    if (std::getenv("VEMD_VERBOSE"))
        std::fprintf(stderr, "[vemd] theta-dependent host data %.3fs (scales, diagonal, rhs %.3fs)\n",
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(),
                     std::chrono::duration<double>(t1 - t0).count());
}

void Solver::build_amg() {
    // a rebuild after a theta change keeps the aggregates (and with them every
    // sparsity pattern of the hierarchy) and recomputes the values only;
    // VEMD_MG_REAGGREGATE=1 builds it afresh
    std::unique_ptr<Amg> prev = std::move(amg);
    const char* ra = std::getenv("VEMD_MG_REAGGREGATE");
    if (ra && ra[0] == '1') prev.reset();
    amg_values_only = false;
    amg_lam_minv.clear();
    auto t0 = std::chrono::steady_clock::now();
    // tuning overrides: smoothing degrees, Chebyshev interval ratio, strength threshold
    if (const char* e = std::getenv("VEMD_MG_DEGREE0")) amg_opt.degree0 = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("VEMD_MG_DEGREE")) amg_opt.degree = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("VEMD_MG_RATIO")) amg_opt.cheb_ratio = std::max(2.0, std::atof(e));
    if (const char* e = std::getenv("VEMD_MG_STRENGTH")) amg_opt.strength = std::max(0.0, std::atof(e));
    const Layout& L = layout;
    if (!splan) splan.reset(new SchurPlan(schur_plan(*mesh, *setup, L)));
    Csr S0 = schur_values(*splan, *mesh, *setup, L, minv);
    const int64_t n = S0.nrows;
    std::vector<int> node(n, -1);
    std::vector<double> b0(n, 0.0);
    for (int64_t g = 0; g < L.cpad; ++g)
        if (L.cell_of_gpos[g] >= 0)
            for (int r = 0; r < L.q; ++r) {
                node[g * L.q + r] = (int)g;
                b0[g * L.q + r] = r == 0 ? 1.0 : 0.0;
            }
    // Mean-zero multiplier problem: S0 is singular, its null vector the
    // constant pressure n0 (moment 0 of every cell), which the multiplier row
    // L fixes.  The cycle then approximates (S0 + rho L L^T)^-1 and the
    // multiplier row is preconditioned by 1 / rho (exact inverses make this
    // block preconditioner independent of rho).  The cycle sees the rank-one
    // term only on its coarsest level, while its smoothers also act on n0;
    // rho is set so that along n0 the rank-one term, rho (L.n0)^2 / |n0|^2, is
    // a fraction c of S0's mean diagonal: then the coarse correction, not the
    // smoothers, sets the cycle's response to n0 (as the exact inverse does),
    // on any mesh size, and the coarsest matrix stays well conditioned.
    // A batch has one multiplier per instance and S0 one null vector per
    // instance; each gets its own rank-one term (rho_i from the instance's own
    // rows; aggregates never cross instances, so the terms stay in the
    // instances' blocks of the coarsest matrix).
    std::vector<double> l0;
    std::vector<int> row_inst;
    double rho = 0.0;
    if (setup->desc.bc == VEMD_BC_FLUX_MEANZERO) {
        const int ni = L.ninst;
        l0.assign(n, 0.0);
        std::vector<double> sd(ni, 0.0), ln0(ni, 0.0);
        std::vector<int64_t> nlive(ni, 0);
        for (int i = 0; i < ni; ++i)
            for (int64_t g = L.inst_gpos0[i]; g < L.inst_gpos0[i + 1]; ++g) {
                const int c = L.cell_of_gpos[g];
                if (c < 0) continue;
                ++nlive[i];
                for (int r = 0; r < L.q; ++r) {
                    const int64_t row = g * L.q + r;
                    l0[row] = setup->L[(size_t)c * L.q + r];
                    if (r == 0) ln0[i] += l0[row];
                }
            }
        // S0's diagonal over each instance's live rows, in 256 blocks of rows
        // summed in order (one instance: all rows, the order of the device
        // rebuild)
        for (int i = 0; i < ni; ++i) {
            const int64_t r0 = ni == 1 ? 0 : (int64_t)L.q * L.inst_gpos0[i];
            const int64_t r1 = ni == 1 ? n : (int64_t)L.q * L.inst_gpos0[i + 1];
            const int64_t bl = (r1 - r0 + 255) / 256;
            for (int k = 0; k < 256; ++k) {
                double sk = 0.0;
                for (int64_t row = r0 + k * bl; row < std::min(r1, r0 + (k + 1) * bl); ++row) {
                    if (L.cell_of_gpos[row / L.q] < 0) continue;
                    for (int64_t p = S0.rp[row]; p < S0.rp[row + 1]; ++p)
                        if (S0.ci[p] == row) sk += S0.v[p];
                }
                sd[i] += sk;
            }
        }
        double c = 0.03;  // measured: 203, 250, 262 FP64 iterations at 4k, 64k, 256k disk cells
        if (const char* e = std::getenv("VEMD_MG_RHO")) c = std::max(1e-12, std::atof(e));
        std::vector<double> rhoi(ni);
        amg_lam_minv.assign(ni, 0.0);
        for (int i = 0; i < ni; ++i) {
            if (!(sd[i] > 0.0) || !(ln0[i] != 0.0))
                throw Error(VEMD_ERR_INTERNAL, "degenerate pressure Schur approximation");
            rhoi[i] = c * (sd[i] / ((double)nlive[i] * L.q)) * (double)nlive[i] / (ln0[i] * ln0[i]);
            amg_lam_minv[i] = 1.0 / rhoi[i];
        }
        if (ni == 1) {
            rho = rhoi[0];
        } else {
            // rho_i folded into the instance's rows of l0 (rho_i l l^T = (sqrt(rho_i) l)(sqrt(rho_i) l)^T)
            rho = 1.0;
            row_inst.assign(n, -1);
            for (int i = 0; i < ni; ++i) {
                const double f = std::sqrt(rhoi[i]);
                for (int64_t row = (int64_t)L.q * L.inst_gpos0[i]; row < (int64_t)L.q * L.inst_gpos0[i + 1]; ++row) {
                    l0[row] *= f;
                    row_inst[row] = i;
                }
            }
        }
    }
    // VEMD_MG_LINEAR=1: the first coarsening also carries the linear pressure
    // fields x and y (cell value x_c, linear moment h; the pressure
    // coefficients are those of the scaled monomials (x - x_c) / h), so that
    // level 1 has three rows per aggregate
    std::vector<std::vector<double>> nv;
    if (const char* e = std::getenv("VEMD_MG_LINEAR"))
        if (e[0] == '1' && L.q >= 3) {
            nv.assign(2, std::vector<double>(n, 0.0));
            for (int64_t g = 0; g < L.cpad; ++g) {
                const int c = L.cell_of_gpos[g];
                if (c < 0) continue;
                const double h = mesh->diam[c];
                nv[0][g * L.q] = mesh->cx[c];
                nv[0][g * L.q + 1] = h;
                nv[1][g * L.q] = mesh->cy[c];
                nv[1][g * L.q + 2] = h;
            }
        }
    amg.reset(new Amg(amg_build(std::move(S0), node, (int)L.cpad, b0, L.spatial_gpos, amg_opt,
                                l0.empty() ? nullptr : &l0, rho, prev.get(),
                                row_inst.empty() ? nullptr : &row_inst, nv.empty() ? nullptr : &nv)));
    amg->bs0 = L.q;
    amg_values_only = prev != nullptr;
    amg_host_stale = false;
    amg_setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (std::getenv("VEMD_VERBOSE")) {
        std::fprintf(stderr, "[vemd] multigrid setup %.2fs:", amg_setup_s);
        for (const AmgLevel& l : amg->lv)
            std::fprintf(stderr, " %lld rows/%lld nnz (R %lld, RA %lld)", (long long)l.A.nrows, (long long)l.A.nnz(),
                         (long long)l.R.nnz(), (long long)l.M.nnz());
        std::fprintf(stderr, "\n");
    }
}

// L2 errors of a reference-order solution (computational and physical measure)
// against the manufactured solution, following CPU/solver/DarcySolver.cpp:942-1131.
void errors_host(const Solver& s, const double* sol, int nq, double* out);

}  // namespace vemd

using namespace vemd;

struct vemd_mesh_s {
    std::shared_ptr<Mesh> m;
};
struct vemd_solver_s {
    std::unique_ptr<Solver> s;
};

#define API_BEGIN try {
#define API_END                                  \
    return VEMD_OK;                              \
    }                                            \
    catch (const Error& e) {                     \
        set_last_error(e.what());                \
        return (vemd_status)e.code;              \
    }                                            \
    catch (const std::bad_alloc&) {              \
        set_last_error("out of host memory");    \
        return VEMD_ERR_INTERNAL;                \
    }                                            \
    catch (const std::exception& e) {            \
        set_last_error(e.what());                \
        return VEMD_ERR_INTERNAL;                \
    }

#define REQUIRE(c, msg) \
    if (!(c)) throw Error(VEMD_ERR_ARG, msg)

extern "C" {

const char* vemd_last_error(void) { return g_last_error.c_str(); }

int vemd_device_count(void) { return device_count(); }

vemd_status vemd_mesh_read(const char* path, int format, vemd_mesh* out) {
    API_BEGIN
    REQUIRE(path && out, "null argument");
    REQUIRE(format == VEMD_MESH_MSH || format == VEMD_MESH_VTK, "unknown mesh format");
    Mesh* m = format == VEMD_MESH_MSH ? mesh_read_msh(path) : mesh_read_vtk(path);
    *out = new vemd_mesh_s{std::shared_ptr<Mesh>(m)};
    API_END
}

vemd_status vemd_mesh_from_arrays(int num_nodes, const double* xy, int num_cells, const int* cell_off,
                                  const int* cell_nodes, int num_arcs, const int* arc_nodes, const double* arc_radius,
                                  vemd_mesh* out) {
    API_BEGIN
    REQUIRE(out, "null argument");
    REQUIRE(num_arcs == 0 || (arc_nodes && arc_radius), "arc arrays missing");
    Mesh* m = mesh_from_arrays(num_nodes, xy, num_cells, cell_off, cell_nodes, num_arcs, arc_nodes, arc_radius);
    *out = new vemd_mesh_s{std::shared_ptr<Mesh>(m)};
    API_END
}

vemd_status vemd_mesh_generate(const vemd_mesh_opts* opts, vemd_mesh* out) {
    API_BEGIN
    REQUIRE(opts && out, "null argument");
    Mesh* m = mesh_generate(*opts);
    *out = new vemd_mesh_s{std::shared_ptr<Mesh>(m)};
    API_END
}

vemd_status vemd_mesh_concat(int n, const vemd_mesh* meshes, vemd_mesh* out) {
    API_BEGIN
    REQUIRE(n > 0 && meshes && out, "null argument or empty batch");
    std::vector<const Mesh*> parts;
    for (int i = 0; i < n; ++i) {
        REQUIRE(meshes[i], "null mesh in the batch");
        parts.push_back(meshes[i]->m.get());
    }
    Mesh* m = mesh_concat(parts);
    *out = new vemd_mesh_s{std::shared_ptr<Mesh>(m)};
    API_END
}

vemd_status vemd_mesh_counts(vemd_mesh mh, int64_t* c) {
    API_BEGIN
    REQUIRE(mh && c, "null argument");
    const Mesh& m = *mh->m;
    c[0] = m.nn;
    c[1] = m.nc;
    c[2] = m.ne;
    c[3] = (int64_t)m.nodes.size();
    c[4] = (int64_t)m.arcs.size();
    API_END
}

vemd_status vemd_mesh_export(vemd_mesh mh, double* xy, int* cell_off, int* cell_nodes, int* cell_edges,
                             int* edge_nodes, int* edge_occ, double* cx, double* cy, double* area, double* diam,
                             int* arc_nodes, double* arc_radius) {
    API_BEGIN
    REQUIRE(mh, "null mesh");
    const Mesh& m = *mh->m;
    if (xy) std::copy(m.xy.begin(), m.xy.end(), xy);
    if (cell_off) std::copy(m.off.begin(), m.off.end(), cell_off);
    if (cell_nodes) std::copy(m.nodes.begin(), m.nodes.end(), cell_nodes);
    if (cell_edges) std::copy(m.cell_edges.begin(), m.cell_edges.end(), cell_edges);
    if (edge_nodes) std::copy(m.edge_nodes.begin(), m.edge_nodes.end(), edge_nodes);
    if (edge_occ) std::copy(m.edge_occ.begin(), m.edge_occ.end(), edge_occ);
    if (cx) std::copy(m.cx.begin(), m.cx.end(), cx);
    if (cy) std::copy(m.cy.begin(), m.cy.end(), cy);
    if (area) std::copy(m.area.begin(), m.area.end(), area);
    if (diam) std::copy(m.diam.begin(), m.diam.end(), diam);
    if (arc_nodes || arc_radius) {
        std::vector<std::pair<uint64_t, double>> arcs(m.arcs.begin(), m.arcs.end());
        std::sort(arcs.begin(), arcs.end());
        for (size_t i = 0; i < arcs.size(); ++i) {
            if (arc_nodes) {
                arc_nodes[2 * i] = (int)(arcs[i].first >> 32);
                arc_nodes[2 * i + 1] = (int)(arcs[i].first & 0xffffffffu);
            }
            if (arc_radius) arc_radius[i] = arcs[i].second;
        }
    }
    API_END
}

vemd_status vemd_mesh_write_vtk(vemd_mesh mh, const char* path) {
    API_BEGIN
    REQUIRE(mh && path, "null argument");
    const Mesh& m = *mh->m;
    FILE* f = std::fopen(path, "w");
    if (!f) throw Error(VEMD_ERR_IO, std::string("cannot write ") + path);
    std::fprintf(f, "# vtk DataFile Version 3.0\nvemd mesh\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS %d double\n", m.nn);
    for (int i = 0; i < m.nn; ++i) std::fprintf(f, "%.17g %.17g 0\n", m.xy[2 * i], m.xy[2 * i + 1]);
    std::fprintf(f, "CELLS %d %d\n", m.nc, (int)(m.nc + m.nodes.size()));
    for (int c = 0; c < m.nc; ++c) {
        std::fprintf(f, "%d", m.npc(c));
        for (int j = m.off[c]; j < m.off[c + 1]; ++j) std::fprintf(f, " %d", m.nodes[j]);
        std::fprintf(f, "\n");
    }
    std::fprintf(f, "CELL_TYPES %d\n", m.nc);
    for (int c = 0; c < m.nc; ++c) std::fprintf(f, "7\n");
    std::fclose(f);
    API_END
}

void vemd_mesh_destroy(vemd_mesh m) { delete m; }

vemd_status vemd_solver_create(vemd_mesh mh, const vemd_problem_desc* d, vemd_solver* out) {
    API_BEGIN
    REQUIRE(mh && d && out, "null argument");
    std::unique_ptr<Solver> s(new Solver());
    s->mesh = mh->m;
    s->setup.reset(setup_build(*s->mesh, *d));
    if (d->k <= 3) {
        s->layout = layout_build(*s->mesh, *s->setup);
        s->refresh_theta_dependent();
    } else {
        build_rhs_ref(*s->mesh, *s->setup, s->rhs_ref);
    }
    *out = new vemd_solver_s{std::move(s)};
    API_END
}

vemd_status vemd_solver_create_device(vemd_mesh mh, const vemd_problem_desc* d, int device, vemd_solver* out) {
    API_BEGIN
    REQUIRE(mh && d && out, "null argument");
    std::unique_ptr<Solver> s(new Solver());
    s->mesh = mh->m;
    auto t0 = std::chrono::steady_clock::now();
    s->setup.reset(setup_build_device(*s->mesh, *d, device));
    auto t1 = std::chrono::steady_clock::now();
    s->layout = layout_build(*s->mesh, *s->setup);
    auto t2 = std::chrono::steady_clock::now();
    s->refresh_theta_dependent();
    if (std::getenv("VEMD_VERBOSE"))
        std::fprintf(stderr, "[vemd] device element setup %.2fs, store layout %.2fs, theta-dependent data %.2fs\n",
                     std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count(),
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t2).count());
    *out = new vemd_solver_s{std::move(s)};
    API_END
}

vemd_status vemd_solver_sizes(vemd_solver sh, int64_t* z) {
    API_BEGIN
    REQUIRE(sh && z, "null argument");
    const Solver& s = *sh->s;
    z[0] = s.setup->nflux_total;
    z[1] = s.setup->npres_total;
    z[2] = s.setup->ntot;
    z[3] = s.layout.nvec;
    z[4] = s.layout.ngroups;
    z[5] = s.layout.max_ne;
    API_END
}

vemd_status vemd_solver_upload(vemd_solver sh, int device) {
    API_BEGIN
    REQUIRE(sh, "null solver");
    if (sh->s->setup->sh.k > 3) throw Error(VEMD_ERR_ARG, "the device path supports k = 0 to 3");
    dev_upload(*sh->s, device);
    API_END
}

vemd_status vemd_solver_tuning(vemd_solver sh, int* out) {
    API_BEGIN
    REQUIRE(sh && out, "null argument");
    dev_tuning(*sh->s, out);
    API_END
}

vemd_status vemd_set_theta(vemd_solver sh, const double* theta) {
    API_BEGIN
    REQUIRE(sh && theta, "null argument");
    Solver& s = *sh->s;
    for (int c = 0; c < s.mesh->nc; ++c)
        if (!std::isfinite(theta[c])) throw Error(VEMD_ERR_ARG, "theta must be finite");
    s.setup->theta.assign(theta, theta + s.mesh->nc);
    if (s.setup->sh.k <= 3) {
        // an uploaded solver rebuilds the multigrid values on its device
        const bool device_mg = dev_mg_rebuild_possible(s);
        s.refresh_theta_dependent(!device_mg);
        dev_refresh(s, device_mg);
    }
    API_END
}

vemd_status vemd_solve(vemd_solver sh, const vemd_solve_opts* o, const double* rhs, double* sol, vemd_solve_info* info) {
    API_BEGIN
    REQUIRE(sh && o, "null argument");
    REQUIRE(o->rtol > 0 && o->rtol < 1, "rtol must be in (0, 1)");
    REQUIRE(o->precision == VEMD_FP64 || o->precision == VEMD_MIXED, "precision must be VEMD_FP64 or VEMD_MIXED");
    dev_solve(*sh->s, *o, rhs, sol, info);
    API_END
}

vemd_status vemd_solve_device(vemd_solver sh, const vemd_solve_opts* o, const void* b_dev, void* x_dev,
                              vemd_solve_info* info) {
    API_BEGIN
    REQUIRE(sh && o, "null argument");
    REQUIRE(o->rtol > 0 && o->rtol < 1, "rtol must be in (0, 1)");
    REQUIRE(o->precision == VEMD_FP64 || o->precision == VEMD_MIXED, "precision must be VEMD_FP64 or VEMD_MIXED");
    dev_solve_int(*sh->s, *o, b_dev, x_dev, info);
    API_END
}

vemd_status vemd_to_internal(vemd_solver sh, const double* ref, double* internal, int scale_rows) {
    API_BEGIN
    REQUIRE(sh && ref && internal, "null argument");
    sh->s->to_internal(ref, internal, scale_rows != 0);
    API_END
}

vemd_status vemd_to_reference(vemd_solver sh, const double* internal, double* ref, int unscale_rows) {
    API_BEGIN
    REQUIRE(sh && ref && internal, "null argument");
    sh->s->to_reference(internal, ref, unscale_rows != 0);
    API_END
}

vemd_status vemd_apply(vemd_solver sh, int precision, const double* x, double* y) {
    API_BEGIN
    REQUIRE(sh && x && y && x != y, "null or aliased argument");
    REQUIRE(precision == VEMD_FP64 || precision == VEMD_FP32, "precision must be VEMD_FP64 or VEMD_FP32");
    dev_apply_ref(*sh->s, precision, x, y);
    API_END
}

vemd_status vemd_apply_device(vemd_solver sh, int precision, const void* x, void* y, void* stream) {
    API_BEGIN
    REQUIRE(sh && x && y && x != y, "null or aliased argument");
    REQUIRE(precision == VEMD_FP64 || precision == VEMD_FP32, "precision must be VEMD_FP64 or VEMD_FP32");
    dev_apply_int(*sh->s, precision, x, y, nullptr, stream);
    API_END
}

vemd_status vemd_apply_device_static(vemd_solver sh, int precision, const void* x, void* y, void* stream) {
    API_BEGIN
    REQUIRE(sh && x && y && x != y, "null or aliased argument");
    REQUIRE(precision == VEMD_FP64 || precision == VEMD_FP32, "precision must be VEMD_FP64 or VEMD_FP32");
    dev_apply_static(*sh->s, precision, x, y, stream);
    API_END
}

vemd_status vemd_apply_device_split(vemd_solver sh, int precision, const void* x, void* y, void* h, void* stream) {
    API_BEGIN
    REQUIRE(sh && x && y && h && x != y && x != h && y != h, "null or aliased argument");
    REQUIRE(precision == VEMD_FP64 || precision == VEMD_FP32, "precision must be VEMD_FP64 or VEMD_FP32");
    dev_apply_int(*sh->s, precision, x, y, h, stream);
    API_END
}

vemd_status vemd_edge_sides(vemd_solver sh, int64_t* nside, int32_t* table) {
    API_BEGIN
    REQUIRE(sh && nside, "null argument");
    const Layout& L = sh->s->layout;
    *nside = L.nside;
    if (table) std::copy(L.hside.begin(), L.hside.end(), table);
    API_END
}

vemd_status vemd_solve_multi_device(vemd_solver sh, const vemd_solve_opts* o, int nrhs, const void* b, void* x,
                                    vemd_solve_info* info) {
    API_BEGIN
    REQUIRE(sh && o && b, "null argument");
    REQUIRE(o->rtol > 0 && o->rtol < 1, "rtol must be in (0, 1)");
    dev_solve_multi(*sh->s, *o, nrhs, (const double*)b, (double*)x, info);
    API_END
}

vemd_status vemd_apply_multi_device(vemd_solver sh, const void* x, void* y, void* y1, void* stream) {
    API_BEGIN
    REQUIRE(sh && x && y && y1 && x != y && x != y1 && y != y1, "null or aliased argument");
    dev_apply_multi(*sh->s, x, y, y1, stream);
    API_END
}

vemd_status vemd_rhs(vemd_solver sh, double* rhs) {
    API_BEGIN
    REQUIRE(sh && rhs, "null argument");
    std::copy(sh->s->rhs_ref.begin(), sh->s->rhs_ref.end(), rhs);
    API_END
}

vemd_status vemd_errors(vemd_solver sh, const double* sol, int nq, double* out) {
    API_BEGIN
    REQUIRE(sh && sol && out, "null argument");
    double all[6];
    errors_host(*sh->s, sol, nq, all);
    std::copy(all, all + 4, out);
    API_END
}

vemd_status vemd_errors_full(vemd_solver sh, const double* sol, int nq, double* out) {
    API_BEGIN
    REQUIRE(sh && sol && out, "null argument");
    errors_host(*sh->s, sol, nq, out);
    API_END
}

vemd_status vemd_errors_device(vemd_solver sh, const void* x_dev, int nq, double* out, void* stream) {
    API_BEGIN
    REQUIRE(sh && x_dev && out, "null argument");
    dev_errors(*sh->s, x_dev, nq, stream, out);
    API_END
}

vemd_status vemd_adjoint_theta(vemd_solver sh, const vemd_solve_opts* o, const double* sol, const double* dldx,
                               double* grad, vemd_solve_info* info) {
    API_BEGIN
    REQUIRE(sh && o && sol && dldx && grad, "null argument");
    if (sh->s->setup->desc.bc != VEMD_BC_DIRICHLET_P)
        throw Error(VEMD_ERR_ARG, "the adjoint gradient is defined for the Dirichlet-pressure condition");
    dev_adjoint(*sh->s, *o, sol, dldx, grad, info);
    API_END
}

vemd_status vemd_element_matrices(vemd_solver sh, int cell, int* n_out, double* A, double* W, double* L,
                                  double* f, double* sign, int64_t* gdof) {
    API_BEGIN
    REQUIRE(sh && n_out, "null argument");
    const Solver& s = *sh->s;
    REQUIRE(cell >= 0 && cell < s.mesh->nc, "cell out of range");
    const Setup& S = *s.setup;
    int n = S.sh.nflux(s.mesh->npc(cell)), q = S.sh.q;
    *n_out = n;
    const double* Ap = &S.A[S.ofs.a_off[cell]];
    if (A)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) A[i * n + j] = Ap[i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i];
    if (W) std::copy(&S.W[S.ofs.w_off[cell]], &S.W[S.ofs.w_off[cell]] + (size_t)q * n, W);
    if (L) std::copy(&S.L[(size_t)cell * q], &S.L[(size_t)cell * q] + q, L);
    if (f) std::copy(&S.f[(size_t)cell * q], &S.f[(size_t)cell * q] + q, f);
    if (sign) std::copy(&S.sgn[S.ofs.v_off[cell]], &S.sgn[S.ofs.v_off[cell]] + n, sign);
    if (gdof) std::copy(&S.gdof[S.ofs.v_off[cell]], &S.gdof[S.ofs.v_off[cell]] + n, gdof);
    API_END
}

vemd_status vemd_batch_info(vemd_solver sh, int n, vemd_solve_info* out) {
    API_BEGIN
    REQUIRE(sh && out, "null argument");
    const Solver& s = *sh->s;
    REQUIRE(n == s.mesh->ninst(), "n must be the number of instances");
    REQUIRE((int)s.inst_iterations.size() == n, "no solve has run");
    for (int i = 0; i < n; ++i) {
        out[i] = s.last_info;
        out[i].iterations = s.inst_iterations[i];
        out[i].rel_residual = s.inst_rel[i];
    }
    API_END
}

void vemd_solver_destroy(vemd_solver s) { delete s; }

}  // extern "C"
