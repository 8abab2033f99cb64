"""ctypes declarations of libvemd's C ABI (include/vemd.h)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libvemd.so")

MAP_SIN, MAP_ARCS = 0, 1
SOL_SINCOS, SOL_SINSIN = 0, 1
BC_FLUX_MEANZERO, BC_DIRICHLET_P = 0, 1
FP64, MIXED, FP32 = 0, 1, 2
PC_AUTO, PC_JACOBI, PC_MULTIGRID = 0, 1, 2
DOMAIN_DISK, DOMAIN_ANNULUS = 0, 1

STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_IO", 3: "ERR_MESH", 4: "ERR_SINGULAR", 5: "ERR_NO_DEVICE",
          6: "ERR_DEVICE", 7: "ERR_NOT_CONVERGED", 8: "ERR_STATE", 9: "ERR_INTERNAL"}

c_dp = ctypes.POINTER(ctypes.c_double)
c_ip = ctypes.POINTER(ctypes.c_int)
c_lp = ctypes.POINTER(ctypes.c_int64)
c_vp = ctypes.c_void_p


class VemdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code
        self.status = STATUS.get(code, str(code))


class MeshOpts(ctypes.Structure):
    _fields_ = [("domain", ctypes.c_int), ("r_inner", ctypes.c_double), ("r_outer", ctypes.c_double),
                ("num_cells", ctypes.c_int), ("lloyd_iters", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("min_edge_frac", ctypes.c_double)]


# callback types of vemd_problem_funcs
KAPPA_FN = ctypes.CFUNCTYPE(None, c_vp, ctypes.c_double, ctypes.c_double, c_dp)
SCALAR_FN = ctypes.CFUNCTYPE(ctypes.c_double, c_vp, ctypes.c_double, ctypes.c_double)
VECTOR_FN = ctypes.CFUNCTYPE(None, c_vp, ctypes.c_double, ctypes.c_double, c_dp)
NORMAL_FN = ctypes.CFUNCTYPE(ctypes.c_double, c_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double)
MAP_FN = ctypes.CFUNCTYPE(None, c_vp, ctypes.c_double, ctypes.c_double, c_dp, c_dp)


class ProblemFuncs(ctypes.Structure):
    _fields_ = [("ctx", c_vp), ("kappa", KAPPA_FN), ("solution_p", SCALAR_FN), ("solution_u", VECTOR_FN),
                ("source_g", SCALAR_FN), ("dirichlet_p", SCALAR_FN), ("neumann_un", NORMAL_FN), ("map", MAP_FN)]


class ProblemDesc(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int), ("nq2d", ctypes.c_int), ("nq1d", ctypes.c_int), ("map_kind", ctypes.c_int),
                ("eps", ctypes.c_double), ("solution", ctypes.c_int), ("bc", ctypes.c_int), ("theta", c_dp),
                ("funcs", ctypes.POINTER(ProblemFuncs))]


class SolveOptsC(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("rtol", ctypes.c_double), ("max_iters", ctypes.c_int),
                ("inner_rtol", ctypes.c_double), ("check_every", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("precond", ctypes.c_int), ("stream", ctypes.c_void_p)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("refinements", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("seconds", ctypes.c_double), ("iter_ms", ctypes.c_double)]


# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "vemd_last_error": [],
    "vemd_device_count": [],
    "vemd_mesh_read": [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(c_vp)],
    "vemd_mesh_from_arrays": [ctypes.c_int, c_dp, ctypes.c_int, c_ip, c_ip, ctypes.c_int, c_ip, c_dp,
                              ctypes.POINTER(c_vp)],
    "vemd_mesh_generate": [ctypes.POINTER(MeshOpts), ctypes.POINTER(c_vp)],
    "vemd_mesh_counts": [c_vp, c_lp],
    "vemd_mesh_export": [c_vp, c_dp, c_ip, c_ip, c_ip, c_ip, c_ip, c_dp, c_dp, c_dp, c_dp, c_ip, c_dp],
    "vemd_mesh_write_vtk": [c_vp, ctypes.c_char_p],
    "vemd_mesh_destroy": [c_vp],
    "vemd_solver_create": [c_vp, ctypes.POINTER(ProblemDesc), ctypes.POINTER(c_vp)],
    "vemd_solver_create_device": [c_vp, ctypes.POINTER(ProblemDesc), ctypes.c_int, ctypes.POINTER(c_vp)],
    "vemd_solver_tuning": [c_vp, c_ip],
    "vemd_solver_sizes": [c_vp, c_lp],
    "vemd_solver_upload": [c_vp, ctypes.c_int],
    "vemd_set_theta": [c_vp, c_dp],
    "vemd_solve": [c_vp, ctypes.POINTER(SolveOptsC), c_dp, c_dp, ctypes.POINTER(SolveInfo)],
    "vemd_solve_device": [c_vp, ctypes.POINTER(SolveOptsC), c_vp, c_vp, ctypes.POINTER(SolveInfo)],
    "vemd_to_internal": [c_vp, c_dp, c_dp, ctypes.c_int],
    "vemd_to_reference": [c_vp, c_dp, c_dp, ctypes.c_int],
    "vemd_apply": [c_vp, ctypes.c_int, c_dp, c_dp],
    "vemd_apply_device": [c_vp, ctypes.c_int, c_vp, c_vp, c_vp],
    "vemd_apply_device_split": [c_vp, ctypes.c_int, c_vp, c_vp, c_vp, c_vp],
    "vemd_apply_device_static": [c_vp, ctypes.c_int, c_vp, c_vp, c_vp],
    "vemd_edge_sides": [c_vp, c_lp, c_ip],
    "vemd_errors_device": [c_vp, c_vp, ctypes.c_int, c_dp, c_vp],
    "vemd_solve_multi_device": [c_vp, ctypes.POINTER(SolveOptsC), ctypes.c_int, c_vp, c_vp, ctypes.POINTER(SolveInfo)],
    "vemd_apply_multi_device": [c_vp, c_vp, c_vp, c_vp, c_vp],
    "vemd_rhs": [c_vp, c_dp],
    "vemd_errors": [c_vp, c_dp, ctypes.c_int, c_dp],
    "vemd_errors_full": [c_vp, c_dp, ctypes.c_int, c_dp],
    "vemd_adjoint_theta": [c_vp, ctypes.POINTER(SolveOptsC), c_dp, c_dp, c_dp, ctypes.POINTER(SolveInfo)],
    "vemd_element_matrices": [c_vp, ctypes.c_int, c_ip, c_dp, c_dp, c_dp, c_dp, c_dp, c_lp],
    "vemd_mesh_concat": [ctypes.c_int, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp)],
    "vemd_batch_info": [c_vp, ctypes.c_int, ctypes.POINTER(SolveInfo)],
    "vemd_solver_destroy": [c_vp],
}
_RESTYPE = {"vemd_last_error": ctypes.c_char_p, "vemd_mesh_destroy": None, "vemd_solver_destroy": None}

_LIB = None


def lib():
    """Loads libvemd.so from the package directory (built by __graft_entry__.build())."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libvemd.so is not built; run __graft_entry__.build()")
        # torch wheels bundle their own HIP runtime under the same soname; the
        # first one loaded serves the whole process, and torch's HIP fails on
        # another build's runtime.  Load torch's first when it is installed so
        # that device tensors and libvemd share one runtime.
        if os.environ.get("VEMD_NO_TORCH") != "1":
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
        L = ctypes.CDLL(LIBPATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, ctypes.c_int)
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise VemdError(rc, lib().vemd_last_error().decode())


def device_count():
    return lib().vemd_device_count()
