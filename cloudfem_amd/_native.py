"""ctypes binding of libcloudfem_hip.so (declared in include/cloudfem_hip.h).

The library is the only compute path: if it is not built, or no GPU is
visible, every entry point raises instead of falling back to the CPU.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcloudfem_hip.so")

# status codes
CF_OK = 0
CF_ERR_ARG = 1
CF_ERR_HIP = 2
CF_ERR_WORKSPACE = 3
CF_ERR_DEGENERATE = 4
CF_ERR_NOT_CONVERGED = 5
CF_ERR_BREAKDOWN = 6

# matrix operators
OP_P1_STIFF = 1
OP_P1_MASS = 2
OP_T10_STIFF = 3
OP_T10_MASS = 4
OP_T10_ELAST = 5
OP_P1_ELAST = 6
OP_P1_VMASS = 7

# rhs operators
RHS_P1_SOURCE = 1
RHS_T10_SOURCE = 2
RHS_P1_BODY = 3
RHS_T10_BODY = 4
RHS_P1_THERMAL = 5
RHS_T10_THERMAL = 6

# face operators
FACE_P1_FLUX = 1
FACE_P1_ROBIN_K = 2
FACE_T10_FLUX = 3
FACE_T10_ROBIN_K = 4
FACE_P1_TRACTION = 5
FACE_T10_TRACTION = 6

FMT_CSR = 0
FMT_NB3 = 1
PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED16 = 2
PC_JACOBI = 0
PC_BLOCK3 = 1

EXPORTED = (
    "cf_version", "cf_last_error", "cf_pattern_workspace_bytes",
    "cf_pattern_sort", "cf_pattern_finalize", "cf_incidence_workspace_bytes",
    "cf_incidence_build", "cf_tet_geometry", "cf_assemble_matrix",
    "cf_assemble_rhs", "cf_face_rhs", "cf_face_matrix",
    "cf_csr_add_subpattern", "cf_expand_pattern3", "cf_spmv",
    "cf_pcg_workspace_bytes", "cf_pcg_solve", "cf_dirichlet_workspace_bytes",
    "cf_dirichlet_count", "cf_dirichlet_fill", "cf_axpby", "cf_copy_masked",
    "cf_element_stress", "cf_spmv_profile", "cf_adjoint_dk", "cf_mf_t10_workspace_bytes",
    "cf_mf_t10_apply", "cf_pcg_solve_mf_t10", "cf_invert_permutation",
    "cf_batch_workspace_bytes", "cf_batch_autotune", "cf_batch_theta_run",
    "cf_mf_t10_plan_bytes", "cf_mf_t10_plan_build", "cf_mf_t10_plan_workspace_bytes",
    "cf_mf_t10_plan_apply", "cf_pcg_solve_mf_t10_plan",
    "cf_mesh_workspace_bytes", "cf_tet10_edges", "cf_edge_midpoints", "cf_boundary_faces",
    "cf_mf_t10_mixed_workspace_bytes", "cf_pcg_solve_mf_t10_mixed",
    "cf_mf_t10_twolevel_workspace_bytes", "cf_pcg_solve_mf_t10_twolevel",
    "cf_t10_twolevel_workspace_bytes", "cf_pcg_solve_t10_twolevel", "cf_spmv_abs",
    "cf_pcg_solve_noalloc", "cf_mf_t10_lp_workspace_bytes", "cf_mf_t10_lp_apply",
)


class CFError(RuntimeError):
    """A libcloudfem_hip call failed."""


class SolverNotConverged(CFError):
    """PCG reached maxit before the tolerance."""


class SolverBreakdown(CFError):
    """PCG broke down (NaN/Inf, non-positive curvature or diagonal)."""


class CfMatrix(ctypes.Structure):
    _fields_ = [("format", ctypes.c_int32), ("n_rows", ctypes.c_int32),
                ("nnz", ctypes.c_int64), ("indptr", ctypes.c_void_p),
                ("indices", ctypes.c_void_p), ("values", ctypes.c_void_p)]


class CfPcgStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32),
                ("residual_updates", ctypes.c_int32),
                ("rel_residual", ctypes.c_double), ("b_norm", ctypes.c_double),
                ("restarts", ctypes.c_int32), ("floor_limited", ctypes.c_int32)]


class CfBatchStats(ctypes.Structure):
    _fields_ = [("n_steps", ctypes.c_int32), ("lanes", ctypes.c_int32),
                ("rows_per_tile", ctypes.c_int32), ("step_iterations", ctypes.c_int32),
                ("graph_replays", ctypes.c_int32), ("extra_replays", ctypes.c_int32),
                ("iterations", ctypes.c_int64), ("sweeps", ctypes.c_int64),
                ("worst_step", ctypes.c_int32), ("failed", ctypes.c_int32),
                ("max_rel_residual", ctypes.c_double)]


_lib = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double
_SZ = ctypes.POINTER(ctypes.c_size_t)

_SIGS = {
    "cf_version": ([], _I32),
    "cf_last_error": ([], ctypes.c_char_p),
    "cf_pattern_workspace_bytes": ([_I64, _I32, _I64, _SZ], _I32),
    "cf_pattern_sort": ([_P, _I64, _I32, _I64, _P, ctypes.c_size_t, _P, _P, _P], _I32),
    "cf_pattern_finalize": ([_I64, _I32, _I64, _I64, _P, _P, _P, _P, _P], _I32),
    "cf_incidence_workspace_bytes": ([_I64, _I64, _SZ], _I32),
    "cf_incidence_build": ([_P, _I64, _I64, _P, ctypes.c_size_t, _P, _P, _P], _I32),
    "cf_tet_geometry": ([_P, _P, _P, _P, _I64, _I32, _P, _P, _P], _I32),
    "cf_assemble_matrix": ([_I32, _P, _P, _P, _P, _P, _I64, _I64, _I32, _P, _P, _D, _D,
                            _P, _P], _I32),
    "cf_assemble_rhs": ([_I32, _P, _P, _I64, _I32, _P, _P, _D, _D, _P, _P], _I32),
    "cf_face_rhs": ([_I32, _P, _P, _P, _P, _I64, _I32, _P, _P, _I64, _P, _P, _P], _I32),
    "cf_face_matrix": ([_I32, _P, _P, _P, _P, _I32, _P, _P, _P, _P, _I64, _I64, _P, _P,
                        _P], _I32),
    "cf_csr_add_subpattern": ([_I64, _P, _P, _P, _P, _P, _P, _D, _P, _P], _I32),
    "cf_expand_pattern3": ([_I64, _P, _P, _I32, _P, _P, _P], _I32),
    "cf_spmv": ([ctypes.POINTER(CfMatrix), _P, _P, _P, _P], _I32),
    "cf_pcg_workspace_bytes": ([ctypes.POINTER(CfMatrix), _I32, _I32, _SZ], _I32),
    "cf_dirichlet_workspace_bytes": ([_I64, _SZ], _I32),
    "cf_dirichlet_count": ([_I64, _P, _P, _P, _P, _P, _P, _I64, _P, ctypes.c_size_t, _P, _P,
                            _P, _P], _I32),
    "cf_dirichlet_fill": ([_I64, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P], _I32),
    "cf_axpby": ([_I64, _D, _P, _D, _P, _P], _I32),
    "cf_copy_masked": ([_I64, _P, _P, _P, _P], _I32),
    "cf_mf_t10_workspace_bytes": ([_I64, _I64, _I32, _SZ], _I32),
    "cf_mf_t10_apply": ([_I64, _I64, _P, _P, _D, _D, _P, _P, _P, _P, _P, _P,
                         ctypes.c_size_t, _P], _I32),
    "cf_pcg_solve_mf_t10": ([_I64, _I64, _P, _P, _D, _D, _P, _P, _P, _P, _P, _I32, _D, _I32,
                             _P, ctypes.c_size_t, ctypes.POINTER(CfPcgStats), _P], _I32),
    "cf_invert_permutation": ([_I64, _P, _P, _P], _I32),
    "cf_adjoint_dk": ([_P, _I64, _I32, _P, _P, _P, _P, _P], _I32),
    "cf_spmv_profile": ([_I32, ctypes.POINTER(_D), ctypes.POINTER(_I64), ctypes.POINTER(_D),
                         ctypes.POINTER(_I64)], _I32),
    "cf_element_stress": ([_P, _I64, _I32, _P, _P, _P, _D, _D, _P, _P, _P], _I32),
    "cf_pcg_solve": ([ctypes.POINTER(CfMatrix), _P, _P, _P, _I32, _I32, _D, _I32, _P,
                      ctypes.c_size_t, ctypes.POINTER(CfPcgStats), _P], _I32),
    "cf_batch_workspace_bytes": ([_I32, _P, _I64, _I32, _SZ], _I32),
    "cf_mf_t10_plan_bytes": ([_I64, _SZ], _I32),
    "cf_mesh_workspace_bytes": ([_I64, _SZ], _I32),
    "cf_mf_t10_mixed_workspace_bytes": ([_I64, _I64, _SZ], _I32),
    "cf_pcg_solve_mf_t10_mixed": ([_I64, _I64, _P, _P, _D, _D, _P, _P, _P, _P, _P, _D, _I32,
                                   _P, ctypes.c_size_t, ctypes.POINTER(CfPcgStats), _P], _I32),
    "cf_mf_t10_twolevel_workspace_bytes": ([_I64, _I64, _SZ], _I32),
    "cf_pcg_solve_mf_t10_twolevel": ([_I64, _I64, _P, _P, _D, _D, _P, _P, _P, _P, _P, _P, _P, _P,
                                      _D, _I32, _P, ctypes.c_size_t, ctypes.POINTER(CfPcgStats),
                                      _P], _I32),
    "cf_t10_twolevel_workspace_bytes": ([ctypes.POINTER(CfMatrix), _I64, _I64, _I32, _SZ],
                                        _I32),
    "cf_spmv_abs": ([ctypes.POINTER(CfMatrix), _P, _P, _P, _P], _I32),
    "cf_mf_t10_lp_workspace_bytes": ([_I64, _I64, _I32, _SZ], _I32),
    "cf_mf_t10_lp_apply": ([_I64, _I64, _P, _P, _D, _D, _P, _P, _I32, _P, _P, _P, _I32,
                            ctypes.POINTER(ctypes.c_float), _P, ctypes.c_size_t, _P], _I32),
    "cf_pcg_solve_noalloc": ([ctypes.POINTER(CfMatrix), _P, _P, _P, _D, _I32,
                              ctypes.POINTER(CfPcgStats), _P], _I32),
    "cf_pcg_solve_t10_twolevel": ([ctypes.POINTER(CfMatrix), _I64, _I64, _P, _P, _D, _D, _P, _P, _P, _P, _P, _P, _P,
                                   _P, _I32, _D, _I32, _P, ctypes.c_size_t,
                                   ctypes.POINTER(CfPcgStats), _P], _I32),
    "cf_tet10_edges": ([_I64, _P, _I32, _I64, _P, _P, _P, _P, _P, ctypes.c_size_t, _P], _I32),
    "cf_edge_midpoints": ([_I64, _I64, _P, _P, _P, _P, _P, _P], _I32),
    "cf_boundary_faces": ([_I64, _P, _I32, _I64, _P, _I32, _P, _P, _P, ctypes.c_size_t, _P], _I32),
    "cf_mf_t10_plan_build": ([_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_size_t, _P,
                              _P], _I32),
    "cf_mf_t10_plan_workspace_bytes": ([_I64, _I64, _P, _I32, _SZ], _I32),
    "cf_mf_t10_plan_apply": ([_I64, _I64, _P, _P, _D, _D, _P, _P, _P, _P, ctypes.c_size_t, _P],
                             _I32),
    "cf_pcg_solve_mf_t10_plan": ([_I64, _I64, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P, _P, _I32,
                                  _D, _I32, _P, ctypes.c_size_t, ctypes.POINTER(CfPcgStats), _P],
                                 _I32),
    "cf_batch_autotune": ([_I32, _P, _P, _P, _P, _I64, _I32, ctypes.POINTER(_I32),
                           ctypes.POINTER(_I32), _P, _P, ctypes.c_size_t, _P], _I32),
    "cf_batch_theta_run": ([_I32, _P, _P, _P, _I64, _P, _P, _P, _D, _P, _P, _I32, _I32, _D,
                            _I32, _I32, _I32, _P, ctypes.c_size_t,
                            ctypes.POINTER(CfBatchStats), _P], _I32),
}


def load():
    """Load the shared library (no GPU needed) and declare every entry point."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CFError(f"{LIB_PATH} is not built; run "
                          "`python -m cloudfem_amd.csrc.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def last_error():
    return load().cf_last_error().decode(errors="replace")


def check(rc, what):
    """Raise the exception matching a non-zero status."""
    if rc == CF_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc == CF_ERR_DEGENERATE:
        raise np.linalg.LinAlgError(msg)
    if rc == CF_ERR_NOT_CONVERGED:
        raise SolverNotConverged(msg)
    if rc == CF_ERR_BREAKDOWN:
        raise SolverBreakdown(msg)
    raise CFError(f"{msg} (status {rc})")
