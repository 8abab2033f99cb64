/*
 * vemd.h -- C ABI of libvemd, an MI355X-native matrix-free mixed virtual
 * element Darcy solver (lowest and first order, curved polygonal meshes).
 *
 * The reference exposes this path only through C++ classes; each entry point
 * below names the reference interface it replaces (paths relative to the
 * reference root; CPU/ = vem/CurvedVEM_Darcy/, GPU/ = vem/CurvedVEM_Darcy_gpu/).
 *
 * Conventions:
 *   - every function returns a vemd_status; on failure vemd_last_error()
 *     gives a message for the calling thread.  No C++ exception crosses the
 *     ABI (the reference's rule at GPU/src/krylov.cu:8-29);
 *   - all buffers are caller-owned; host pointers unless a name ends in _dev;
 *   - vectors of unknowns use the reference's global DOF order
 *     (CPU/solver/DarcySolver.cpp:182-227): edge moments degree-major
 *     (r * num_edges + edge), then cell-internal flux dofs, then q pressure
 *     coefficients per cell, then (essential-flux condition only) the
 *     mean-zero multiplier;
 *   - a solver object is used by one thread at a time.
 */
#ifndef VEMD_H
#define VEMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    VEMD_OK = 0,
    VEMD_ERR_ARG = 1,           /* invalid argument */
    VEMD_ERR_IO = 2,            /* file missing or malformed */
    VEMD_ERR_MESH = 3,          /* degenerate or unsupported mesh */
    VEMD_ERR_SINGULAR = 4,      /* singular element matrix (first bad cell in the message) */
    VEMD_ERR_NO_DEVICE = 5,     /* no HIP device: the solve path has no CPU fallback */
    VEMD_ERR_DEVICE = 6,        /* HIP runtime error */
    VEMD_ERR_NOT_CONVERGED = 7, /* Krylov iteration limit reached (info holds the residual) */
    VEMD_ERR_STATE = 8,         /* call out of order (e.g. solve before upload) */
    VEMD_ERR_INTERNAL = 9
} vemd_status;

typedef struct vemd_mesh_s* vemd_mesh;
typedef struct vemd_solver_s* vemd_solver;

/* Message of the last failure on this thread (never NULL). */
const char* vemd_last_error(void);

/* ---------------------------------------------------------------- meshes */

enum { VEMD_MESH_MSH = 0, VEMD_MESH_VTK = 1 };

/* Reads a Gmsh 4.1 (.msh) or legacy VTK polygon mesh with the reference's
 * conventions (CCW cells, quad centroid = bounding-box centre, others = vertex
 * average, diameter = max vertex distance, edges numbered by sorted node pair).
 * Replaces vem::StraightMeshReader::read_mesh / read_mesh_vtk
 * (CPU/mesh/straight_mesh.cpp:10-41,64-90).  Cells with a non-finite vertex,
 * a repeated vertex or zero area are rejected with VEMD_ERR_MESH (the
 * reference's reader does not check for them). */
vemd_status vemd_mesh_read(const char* path, int format, vemd_mesh* out);

/* Builds a mesh from arrays: cell c has nodes cell_nodes[cell_off[c] ..
 * cell_off[c+1]); cell_off[0] = 0, offsets do not decrease, and cell_nodes
 * holds cell_off[num_cells] entries.  Cells are checked as in vemd_mesh_read
 * and clockwise cells reordered.  Optional circular boundary arcs (centre at the origin):
 * arc_nodes holds (n0, n1) per arc, arc_radius its radius; the straight edge
 * n0-n1 is the chord.  Replaces the same reader's post-processing chain
 * (CPU/mesh/straight_mesh.cpp:504-660) for meshes made in memory. */
vemd_status vemd_mesh_from_arrays(int num_nodes, const double* xy, int num_cells, const int* cell_off,
                                  const int* cell_nodes, int num_arcs, const int* arc_nodes,
                                  const double* arc_radius, vemd_mesh* out);

enum { VEMD_DOMAIN_DISK = 0, VEMD_DOMAIN_ANNULUS = 1 };

typedef struct {
    int domain;          /* VEMD_DOMAIN_DISK (r <= r_outer) or VEMD_DOMAIN_ANNULUS */
    double r_inner;      /* annulus inner radius */
    double r_outer;
    int num_cells;       /* number of Voronoi sites */
    int lloyd_iters;     /* Lloyd relaxation steps */
    uint64_t seed;
    double min_edge_frac;/* edges shorter than this fraction of the local spacing are collapsed */
} vemd_mesh_opts;

/* Seeded Lloyd-relaxed Voronoi mesh of a disk or annulus whose boundary edges
 * carry exact circular arcs.  No reference counterpart (the reference's Lloyd
 * generator is a Python script on the unit square, vem/gmesh_py/get_gmesh.py:15-76). */
vemd_status vemd_mesh_generate(const vemd_mesh_opts* opts, vemd_mesh* out);

/* Batch of independent problems: the disjoint union of n meshes (node and
 * cell numbering concatenated in order).  A solver created on it solves every
 * instance at once -- one operator store, one captured Krylov loop -- with
 * separate Krylov recurrences, stopping tests and refinement per instance
 * (Dirichlet-pressure condition).  desc->theta then covers all cells of the
 * union.  No reference counterpart. */
vemd_status vemd_mesh_concat(int n, const vemd_mesh* meshes, vemd_mesh* out);

/* counts: [num_nodes, num_cells, num_edges, total cell-node entries, num_arcs] */
vemd_status vemd_mesh_counts(vemd_mesh m, int64_t* counts);

/* Copies the mesh out (any pointer may be NULL).  xy: 2*num_nodes; cell_off:
 * num_cells+1; cell_nodes and cell_edges: total entries; edge_nodes: 2*num_edges;
 * edge_occ: num_edges; centroid/area/diameter: num_cells each; arc_nodes:
 * 2*num_arcs; arc_radius: num_arcs. */
vemd_status vemd_mesh_export(vemd_mesh m, double* xy, int* cell_off, int* cell_nodes, int* cell_edges,
                             int* edge_nodes, int* edge_occ, double* centroid_x, double* centroid_y,
                             double* area, double* diameter, int* arc_nodes, double* arc_radius);

/* Writes the mesh as a legacy VTK polygon file readable by vemd_mesh_read and
 * by the reference's read_mesh_vtk. */
vemd_status vemd_mesh_write_vtk(vemd_mesh m, const char* path);

void vemd_mesh_destroy(vemd_mesh m);

/* ---------------------------------------------------------------- problem */

enum { VEMD_MAP_SIN = 0, VEMD_MAP_ARCS = 1 };
enum { VEMD_SOL_SINCOS = 0, VEMD_SOL_SINSIN = 1 };
enum { VEMD_BC_FLUX_MEANZERO = 0, VEMD_BC_DIRICHLET_P = 1 };

/* Problem data as callbacks: the fields of the reference's DarcyPdeData and
 * its IsoparametricMapping (CPU/examples/darcy_problem.h:39-63,162-180), so a
 * caller of DarcyProblem can pass its own permeability, solution, source,
 * boundary data and map.  Points are physical (x, y) except for map, which
 * takes computational (xi, eta).  The host element setup and the error
 * functionals call them, from several library threads at once (they must be
 * thread-safe); the library keeps a copy of this struct, the caller keeps ctx
 * valid while the solver exists.  A NULL member keeps the built-in choice of
 * vemd_problem_desc (K = I, the manufactured solution `solution`, the map
 * `map_kind`).  Only vemd_solver_create evaluates callbacks; the GPU element
 * setup rejects them (VEMD_ERR_ARG). */
typedef struct {
    void* ctx;
    /* permeability tensor K(x, y), row-major [xx, xy, yx, yy], symmetric
       positive definite (the reference's kappa); per cell, exp(theta_E)
       multiplies it */
    void (*kappa)(void* ctx, double x, double y, double* k4);
    /* exact pressure (the errors; Dirichlet data unless dirichlet_p is set) */
    double (*solution_p)(void* ctx, double x, double y);
    /* exact flux u = -K grad p as [ux, uy] (the errors; the essential
       normal-flux data unless neumann_un is set) */
    void (*solution_u)(void* ctx, double x, double y, double* u2);
    /* source g = div u */
    double (*source_g)(void* ctx, double x, double y);
    /* boundary pressure (VEMD_BC_DIRICHLET_P) */
    double (*dirichlet_p)(void* ctx, double x, double y);
    /* boundary normal flux u . n for the outward unit normal (nx, ny)
       (VEMD_BC_FLUX_MEANZERO) */
    double (*neumann_un)(void* ctx, double x, double y, double nx, double ny);
    /* computational -> physical map: xy = F(xi, eta) and its Jacobian
       j4 = [dx/dxi, dx/deta, dy/dxi, dy/deta] (det > 0) */
    void (*map)(void* ctx, double xi, double eta, double* xy, double* j4);
} vemd_problem_funcs;

typedef struct {
    int k;              /* polynomial degree: 0 or 1 on the device (0..3 in setup) */
    int nq2d;           /* triangle rule points: 1, 3, 4, 7, 9, 12 */
    int nq1d;           /* Gauss-Legendre points on edges: 1..10 */
    int map_kind;       /* VEMD_MAP_SIN: global x + eps sin(2 pi eta), y + eps sin(2 pi xi)
                           (reference examples 1 and 2); VEMD_MAP_ARCS: per-cell maps that
                           push boundary chords onto their arcs */
    double eps;         /* VEMD_MAP_SIN amplitude */
    int solution;       /* manufactured pressure: VEMD_SOL_SINCOS (reference) or SINSIN */
    int bc;             /* VEMD_BC_FLUX_MEANZERO (reference: essential normal flux and a
                           mean-zero pressure multiplier) or VEMD_BC_DIRICHLET_P */
    const double* theta;/* per-cell log-permeability (K = exp(theta) I), or NULL for K = I */
    const vemd_problem_funcs* funcs; /* callbacks replacing the built-in data, or NULL */
} vemd_problem_desc;

/* Builds every element operator in FP64 on the host (projector, stabilisation,
 * divergence, right-hand side, boundary data).  Replaces the constructor
 * DarcySolver(reader, problem, nq2d, nq1d, k) and its per-element assembly
 * (CPU/solver/DarcySolver.h:21-25, CPU/solver/DarcySolver.cpp:402-500). */
vemd_status vemd_solver_create(vemd_mesh m, const vemd_problem_desc* desc, vemd_solver* out);

/* The same solver with the element operators built on a GPU (one thread per
 * cell, FP64; k = 0, 1 or 2, polygons of up to 10 edges) and copied back into
 * the host arrays; agrees with vemd_solver_create to rounding.  Replaces the
 * same constructor and the reference GPU port's per-element integration from
 * host-packed quadrature samples (GPU/src/hdiv_gpu.cu:45-135,
 * GPU/src/mesh_pack.cpp:31-142).  Does not upload the solver. */
vemd_status vemd_solver_create_device(vemd_mesh m, const vemd_problem_desc* desc, int device, vemd_solver* out);

/* sizes: [total flux dofs, total pressure dofs, total unknowns (incl. multiplier),
 *         internal device vector length, number of valence groups, max valence].
 * Replaces HdivMatrix::getTotalDofFlux / getTotalDofConc (CPU/solver/HdivMatrix.h:69-71). */
vemd_status vemd_solver_sizes(vemd_solver s, int64_t* sizes);

/* Uploads the valence-grouped, spatially ordered operator store (FP64 and FP32)
 * to HBM of the given device.  No reference counterpart as a separate call
 * (GPU/src/hdiv_gpu.cu:139-192 does it inside the constructor). */
vemd_status vemd_solver_upload(vemd_solver s, int device);

/* The apply variant (operator-stream chunk shape of the cell kernel) the
 * autotuner picked at upload for every polygon valence: out[2 * (ne - 3)] for
 * FP64 and out[2 * (ne - 3) + 1] for FP32, ne = 3..10, -1 where the mesh has
 * no such cells (16 ints). No counterpart in the reference, which has one
 * fixed kernel; VEMD_AUTOTUNE=0 turns the tuner off. */
vemd_status vemd_solver_tuning(vemd_solver s, int* out);

/* Replaces the per-cell log-permeability; updates the cell scales, the
 * preconditioner and (essential-flux condition) the lifted right-hand side.
 * Replaces MixedGpu::update (GPU/include/equation_gpu.cuh:76-84). */
vemd_status vemd_set_theta(vemd_solver s, const double* theta);

enum { VEMD_FP64 = 0, VEMD_MIXED = 1, VEMD_FP32 = 2 };
enum { VEMD_PC_AUTO = 0, VEMD_PC_JACOBI = 1, VEMD_PC_MULTIGRID = 2 };

typedef struct {
    int precision;      /* VEMD_FP64, or VEMD_MIXED (FP32 operator and preconditioner inside
                           FP64 iterative refinement) */
    double rtol;        /* target true relative residual ||b - K x||_2 / ||b||_2 */
    int max_iters;      /* total Krylov iteration limit */
    double inner_rtol;  /* VEMD_MIXED: inner FP32 MINRES tolerance per refinement step */
    int check_every;    /* iterations between host convergence checks */
    int use_graph;      /* capture the iteration chunk in a hipGraph */
    int precond;        /* VEMD_PC_JACOBI: block-diagonal with the assembled flux diagonal and
                           diagonal pressure/multiplier Schur approximations (the reference's
                           choice, GPU/src/equation_gpu.cu:282-315); VEMD_PC_MULTIGRID: the same
                           flux block and one smoothed-aggregation V-cycle on the pressure Schur
                           approximation S = W diag(A)^-1 W^T (with the mean-zero multiplier, on
                           S + rho L L^T, the multiplier row then preconditioned by 1 / rho);
                           VEMD_PC_AUTO (0): multigrid where built (k = 0, 1, 2), else Jacobi */
    void* stream;       /* the caller's HIP stream (hipStream_t; NULL = the null stream).  The
                           device-resident solves (vemd_solve_device, vemd_solve_multi_device)
                           wait for work queued on it before they read b_dev, and work queued
                           on it after the call is ordered after their writes of x_dev.  The
                           solver's own stream runs the iteration (the reference synchronises
                           the device at every borrow and restore instead,
                           GPU/include/petsc_gpu.h:62-214). */
} vemd_solve_opts;

typedef struct {
    int iterations;     /* total Krylov iterations */
    int refinements;    /* outer FP64 refinement steps (VEMD_MIXED) */
    double rel_residual;/* final true relative residual (FP64) */
    double seconds;     /* solve wall time, device-resident data */
    double iter_ms;     /* mean wall time per Krylov iteration (seconds / iterations, in ms) */
} vemd_solve_info;

/* Solves with preconditioned MINRES on the symmetric form of the saddle-point
 * system (pressure rows negated).  rhs: optional right-hand side in reference
 * order (NULL = the problem's own); solution: reference order, boundary values
 * included.  Replaces DarcySolver::solve + getSolution (CPU/solver/DarcySolver.cpp:
 * 665-799) and Krylov::solve (GPU/src/krylov.cu:73-175). */
vemd_status vemd_solve(vemd_solver s, const vemd_solve_opts* opts, const double* rhs, double* solution,
                       vemd_solve_info* info);

/* Device-resident solve in the internal layout: b_dev holds the symmetric-form
 * right-hand side (NULL = the problem's own, kept in HBM since upload) and
 * x_dev receives the solution (NULL = keep it on the device only); both are
 * vectors of sizes[3] doubles.  info->seconds covers the whole call.  Used by
 * benchmarks so that no host transfer is timed. */
vemd_status vemd_solve_device(vemd_solver s, const vemd_solve_opts* opts, const void* b_dev, void* x_dev,
                              vemd_solve_info* info);

/* Host conversions between the reference DOF order and the internal device
 * layout; scale_rows applies the row scaling from the reference-form system to
 * the internal symmetric one (pressure and multiplier rows negated, essential
 * boundary rows times exp(-theta_E)) and unscale_rows undoes it. */
vemd_status vemd_to_internal(vemd_solver s, const double* ref, double* internal, int scale_rows);
vemd_status vemd_to_reference(vemd_solver s, const double* internal, double* ref, int unscale_rows);

/* y = K x for the reference-form operator (rows as in CPU/solver/DarcySolver.cpp:
 * 445-476 with direction signs; essential boundary rows are identity).
 * precision VEMD_FP64 or VEMD_FP32.  Replaces MixedGpu::apply
 * (GPU/src/equation_gpu.cu:580-598). */
vemd_status vemd_apply(vemd_solver s, int precision, const double* x, double* y);

/* Device-resident apply in the internal (store-ordered) layout, on the given
 * HIP stream (NULL = default); x_dev and y_dev must not alias.  Used by
 * benchmarks; vector length is sizes[3], element type double or float.
 * Asynchronous: it returns once the kernels are queued.  The solver's device
 * scratch (side slots, partial sums) is shared by every apply and solve of
 * one solver, so each such call is ordered after the previous one, whatever
 * stream either ran on. */
vemd_status vemd_apply_device(vemd_solver s, int precision, const void* x_dev, void* y_dev, void* stream);

/* The apply of vemd_apply_device, allocating nothing: its cell-side slots and
 * multiplier-row partials live in a static device array of the library
 * (64 MiB on each GPU it runs on, held until the process ends) instead of the
 * solver's scratch, and the call only launches kernels on the given stream.
 * No counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  The mesh must fit: (k + 1) * nside side values (nside from
 * vemd_edge_sides) of the precision plus 8 bytes per 64-cell tile, at most
 * 64 MiB (about 650k k = 1 cells in FP64, twice that in FP32); a larger one
 * fails with VEMD_ERR_ARG.  Every call on a GPU shares that GPU's array and
 * leaves its contents behind: calls that overlap on one GPU race, and the
 * caller orders them (one stream, or events).  Same results as
 * vemd_apply_device, bit for bit. */
vemd_status vemd_apply_device_static(vemd_solver s, int precision, const void* x_dev, void* y_dev, void* stream);

/* The same apply without its final edge sum, for consumers that fuse it: y_dev
 * receives all cell-owned and multiplier rows (its leading edge-dof block is
 * not written), h_dev (nside * (k + 1) values, nside from vemd_edge_sides)
 * the edge results of every cell side: side slot p holds values
 * [p * (k + 1), (p + 1) * (k + 1)).  K x at edge dof (k + 1) * e + r (internal
 * edge e) is h[t[2e] * (k + 1) + r] + h[t[2e + 1] * (k + 1) + r] with t the
 * table of vemd_edge_sides.  The last slot (second side of boundary edges) is
 * never written: zero h_dev once before the first call.  This is the form the
 * solver's own iteration uses (one kernel launch, whose stores cover whole
 * cache lines, plus, with the multiplier, one reduction).  No reference
 * counterpart. */
vemd_status vemd_apply_device_split(vemd_solver s, int precision, const void* x_dev, void* y_dev, void* h_dev,
                                    void* stream);

/* The cell sides of the edges for vemd_apply_device_split: *nside receives
 * the number of side slots, and table (if not NULL; 2 * number of mesh edges
 * int32 values) the slots of side 0 and side 1 of every internal edge (edge e
 * owns internal dofs (k + 1) * e .. (k + 1) * e + k).  A boundary edge's
 * side 1 is the zero slot nside - 1.  No reference counterpart. */
vemd_status vemd_edge_sides(vemd_solver s, int64_t* nside, int32_t* table);

/* Simultaneous solves with 1..32 right-hand sides: b_dev and x_dev hold nrhs
 * columns of sizes[3] doubles each (column-major, internal layout, as
 * vemd_solve_device).  The Krylov iterations run in FP32 on 32 interleaved
 * columns with an MFMA element apply (precision must be VEMD_MIXED), inside
 * FP64 iterative refinement per column to opts->rtol; info receives nrhs
 * entries.  Dirichlet-pressure problems on one mesh.  No reference
 * counterpart. */
vemd_status vemd_solve_multi_device(vemd_solver s, const vemd_solve_opts* opts, int nrhs, const void* b_dev,
                                    void* x_dev, vemd_solve_info* info);

/* The MFMA element apply of the multi-right-hand-side solves on 32
 * interleaved FP32 columns (entry i of column j at i * 32 + j, sizes[3] * 32
 * floats): y_dev receives the first sides of the edge dofs and all other
 * rows, y1_dev (also sizes[3] * 32 floats, leading edge-dof block used) the
 * second sides; zero y1_dev once.  On the given stream. */
vemd_status vemd_apply_multi_device(vemd_solver s, const void* x_dev, void* y_dev, void* y1_dev, void* stream);

/* The problem's right-hand side in reference order (ntot values), after the
 * essential-boundary lift. */
vemd_status vemd_rhs(vemd_solver s, double* rhs);

/* L2 errors of a solution against the manufactured one: [pressure, flux] in the
 * computational measure (the reference's metric, CPU/solver/DarcySolver.cpp:
 * 942-1131) and [pressure, flux] in the physical measure.  Host FP64. */
vemd_status vemd_errors(vemd_solver s, const double* solution, int nq2d, double* out4);

/* The four errors of vemd_errors, then the physical-domain divergence error
 * ||div u_h - div u|| (div u_h = (H^-1 W u) . m / det J, H the monomial mass
 * matrix of the computational cell) and the mass int p_h det J: the fields
 * scalar_comp, flux_comp, scalar_phys, flux_phys, div_phys and mass of the
 * reference GPU driver's report.  Replaces error_kernel / reduce_errors
 * (GPU/src/equation_gpu.cu:385-479).  Host FP64. */
vemd_status vemd_errors_full(vemd_solver s, const double* solution, int nq2d, double* out6);

/* The six values of vemd_errors_full for a solution already in device
 * memory: x_dev holds sizes[3] doubles in the internal layout (as
 * vemd_solve_device writes it).  Evaluated on the GPU, one thread per cell
 * rebuilding the cell's projector as the device element setup does, with
 * fixed-order sums (deterministic); it runs after the work queued on the
 * caller's HIP stream (NULL: the null stream) and returns the values.
 * Built-in problem data, k = 0 or 1, polygons of up to 10 edges; problems
 * with callbacks use vemd_errors_full.  Replaces error_kernel /
 * reduce_errors (GPU/src/equation_gpu.cu:385-479). */
vemd_status vemd_errors_device(vemd_solver s, const void* x_dev, int nq2d, double* out6, void* stream);

/* Adjoint gradient of a loss L(x) with respect to theta: solves K lambda =
 * dL/dx (dldx in reference order) and returns dL/dtheta_E =
 * lambda_E^T (exp(-theta_E) A_E) x_E per cell.  No reference counterpart. */
vemd_status vemd_adjoint_theta(vemd_solver s, const vemd_solve_opts* opts, const double* solution,
                               const double* dldx, double* grad_theta, vemd_solve_info* info);

/* FP64 element operator of one cell (host setup output): n local flux dofs;
 * A (n x n, K = I part, direction signs applied to rows and columns; the cell
 * operator is exp(-theta_E) A), W (q x n, signed columns; row r is int m_r div
 * phi_i), L (q, int m_r det J), f (q, pressure source), sign (n), gdof (n,
 * reference global dof of each local flux dof).  Any output may be NULL.
 * Replaces HdivMatrix::getMatrixW and DarcySolver::getMatrixK_ac/K_as,
 * getVecLagrange, getLocalRhs (CPU/solver/HdivMatrix.h:36-55,
 * CPU/solver/DarcySolver.h:54-73). */
vemd_status vemd_element_matrices(vemd_solver s, int cell, int* n, double* A, double* W, double* L, double* f,
                                  double* sign, int64_t* gdof);

/* Per-instance results of the last solve on a batch (n = number of
 * instances, 1 for a plain mesh): iterations and true relative residual;
 * the other fields are those of the whole solve. */
vemd_status vemd_batch_info(vemd_solver s, int n, vemd_solve_info* per_instance);

void vemd_solver_destroy(vemd_solver s);

/* Number of visible HIP devices (0 on a machine without a GPU). */
int vemd_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* VEMD_H */
