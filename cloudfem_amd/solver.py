"""Solver options and the PCG call on device-resident matrices.

The reference solves with SciPy SuperLU (spsolve / factorized).  Here the
free-DOF SPD system is solved with preconditioned CG on the GPU; the
options below have no reference counterpart (the reference has no solver
configuration), and the defaults are chosen so that solutions agree with a
direct solve to well within the 1e-10 relative bar on the validation
problems:

  precision  "auto" (default), "fp64", "mixed" or "mixed16".  "auto" is
             "mixed" for Tet10 elasticity (with the two-level solver) and
             "fp64" for every other system.  For 3x3-block
             (vector) systems: FP64 flexible CG preconditioned by an FP32
             Chebyshev polynomial of the Jacobi-scaled matrix, whose copy is
             stored in FP32 ("mixed") or FP16 ("mixed16"); for scalar systems
             both mean FP32 inner CG with FP64 residual replacement.  The stop
             test is the FP64 residual in every case, so the error against
             FP64 is bounded by the same tolerance (cf_pcg_solve)
  precond    "auto" (default: "twolevel" for Tet10 elasticity, "jacobi"
             otherwise), "jacobi" or "block3" (3x3 node blocks, vector
             problems, FP64 only); "twolevel" (Tet10 elasticity): a matrix-free
             Chebyshev polynomial, FP32 with precision "mixed" or FP64 with
             "fp64", wrapped in a coarse correction on the rigid-body modes of
             node aggregates (cf_pcg_solve_t10_twolevel); the outer FP64
             products use the assembled matrix (operator "assembled") or the
             matrix-free operator ("matfree")
  tol        relative residual ||b_f - A_ff x_f|| / ||b_f|| (default 1e-12)
  maxit      iteration cap (default 200000)
  operator   "assembled" (default: CSR/NB3 SpMV) or "matfree" (Tet10
             elasticity only: element forces recomputed from the geometry
             record every product, cf_pcg_solve_mf_t10; with precision
             "mixed" the Chebyshev preconditioner's products are matrix-free
             in FP32, cf_pcg_solve_mf_t10_mixed)
"""

import contextlib
import ctypes
import threading

import torch

from . import _native as N
from .device import ptr, stream_handle, workspace

_DEFAULTS = {"precision": "auto", "precond": "auto", "tol": 1e-12, "maxit": 200000,
             "operator": "assembled"}
_local = threading.local()


def options():
    if not hasattr(_local, "opts"):
        _local.opts = dict(_DEFAULTS)
    return _local.opts


@contextlib.contextmanager
def solver_options(**kw):
    """Temporarily override solver options on this thread."""
    bad = set(kw) - set(_DEFAULTS)
    if bad:
        raise KeyError(f"unknown solver option(s) {sorted(bad)}")
    saved = dict(options())
    options().update(kw)
    try:
        yield options()
    finally:
        _local.opts = saved


last_stats = threading.local()

# What "auto" means for the general solvers; Tet10 elasticity resolves it to
# the two-level mixed solver instead (physics.elastic_tet10.solve_device).
_AUTO = {"precision": "fp64", "precond": "jacobi"}


def resolve(name, value):
    """An option value with "auto" replaced by the general default."""
    return _AUTO[name] if value == "auto" else value


def _base_stats(st):
    """Fields every solver reports.  floor_limited: the true FP64 residual
    stopped falling above tol (tol below the attainable FP64 floor of the
    system); rel_residual is then that true residual."""
    return {"iterations": st.iterations, "residual_updates": st.residual_updates,
            "rel_residual": st.rel_residual, "b_norm": st.b_norm,
            "restarts": st.restarts, "floor_limited": bool(st.floor_limited)}


def _pc_code(precond):
    codes = {"jacobi": N.PC_JACOBI, "block3": N.PC_BLOCK3}
    if precond not in codes:
        raise ValueError(f"precond {precond!r} is not available here (jacobi, block3; twolevel "
                         "needs operator='matfree' and precision='mixed')")
    return codes[precond]


class DeviceMatrix:
    """A square matrix in HBM: scalar CSR or NB3 (3x3 node blocks)."""

    def __init__(self, fmt, pattern, values):
        self.fmt, self.pattern, self.values = fmt, pattern, values

    @property
    def n_dofs(self):
        return self.pattern.n_rows * (3 if self.fmt == N.FMT_NB3 else 1)

    def cstruct(self):
        p = self.pattern
        return N.CfMatrix(self.fmt, p.n_rows, p.nnz, p.indptr.data_ptr(),
                          p.indices.data_ptr(), self.values.data_ptr())

    def abs_matvec(self, x, fixed=None):
        """|A| |x| with fixed rows zeroed (cf_spmv_abs)."""
        y = torch.empty_like(x)
        A = self.cstruct()
        N.check(N.load().cf_spmv_abs(ctypes.byref(A), ptr(x), ptr(y), ptr(fixed),
                                     stream_handle()), "cf_spmv_abs")
        return y

    def fp64_floor(self, x, b, fixed=None):
        """eps || |A_ff| |x_f| || / ||b_f||: the smallest relative residual
        FP64 resolves for this system at the solution x."""
        xf = x if fixed is None else torch.where(fixed.bool(), torch.zeros_like(x), x)
        bf = b if fixed is None else torch.where(fixed.bool(), torch.zeros_like(b), b)
        y = self.abs_matvec(xf, fixed)
        return float(torch.finfo(torch.float64).eps * torch.linalg.vector_norm(y)
                     / torch.linalg.vector_norm(bf))

    def matvec(self, x, fixed=None):
        y = torch.empty_like(x)
        A = self.cstruct()
        N.check(N.load().cf_spmv(ctypes.byref(A), ptr(x), ptr(y), ptr(fixed), stream_handle()),
                "cf_spmv")
        return y


def pcg(A, b, x, fixed=None, precision=None, precond=None, tol=None, maxit=None):
    """Solve the free-DOF system in place (x holds fixed values on entry).

    Returns a dict of statistics; raises SolverNotConverged / SolverBreakdown.
    """
    o = options()
    precision = resolve("precision", precision or o["precision"])
    precond = resolve("precond", precond or o["precond"])
    tol = float(tol if tol is not None else o["tol"])
    maxit = int(maxit if maxit is not None else o["maxit"])
    prec = {"fp64": N.PREC_FP64, "mixed": N.PREC_MIXED, "mixed16": N.PREC_MIXED16}[precision]
    pc = _pc_code(precond)
    if pc == N.PC_BLOCK3 and (A.fmt != N.FMT_NB3 or prec != N.PREC_FP64):
        pc = N.PC_JACOBI
    lib = N.load()
    As = A.cstruct()
    nb = ctypes.c_size_t()
    N.check(lib.cf_pcg_workspace_bytes(ctypes.byref(As), prec, pc, ctypes.byref(nb)),
            "cf_pcg_workspace_bytes")
    ws = workspace(nb.value, b.device)
    st = N.CfPcgStats()
    rc = lib.cf_pcg_solve(ctypes.byref(As), ptr(b), ptr(x), ptr(fixed), prec, pc, tol, maxit,
                          ptr(ws), nb.value, ctypes.byref(st), stream_handle())
    stats = {**_base_stats(st),
             "precision": precision,
             "precond": "block3" if pc == N.PC_BLOCK3 else "jacobi", "tol": tol}
    last_stats.value = stats
    N.check(rc, "cf_pcg_solve")
    return stats


def pcg_noalloc(A, b, x, fixed=None, tol=None, maxit=None):
    """cf_pcg_solve_noalloc: FP64 Jacobi PCG on a scalar CSR DeviceMatrix with
    the library's static scratch array as workspace (nothing is allocated;
    at most 262,144 unknowns; overlapping calls on one GPU race)."""
    o = options()
    tol = float(tol if tol is not None else o["tol"])
    maxit = int(maxit if maxit is not None else o["maxit"])
    As = A.cstruct()
    st = N.CfPcgStats()
    rc = N.load().cf_pcg_solve_noalloc(ctypes.byref(As), ptr(b), ptr(x), ptr(fixed), tol, maxit,
                                       ctypes.byref(st), stream_handle())
    stats = {**_base_stats(st), "precision": "fp64", "precond": "jacobi", "tol": tol}
    last_stats.value = stats
    N.check(rc, "cf_pcg_solve_noalloc")
    return stats


def pcg_matfree_t10_mixed(dm, lam, mu, b, x, fixed=None, tol=None, maxit=None):
    """Mixed precision with the matrix-free Tet10 operator: FP64 flexible CG,
    FP32 matrix-free Chebyshev preconditioner (cf_pcg_solve_mf_t10_mixed)."""
    o = options()
    tol = float(tol if tol is not None else o["tol"])
    maxit = int(maxit if maxit is not None else o["maxit"])
    lib = N.load()
    conn, geo, offs, perm = dm.matfree_arrays()
    nb = ctypes.c_size_t()
    N.check(lib.cf_mf_t10_mixed_workspace_bytes(dm.n_nodes, dm.n_elem, ctypes.byref(nb)),
            "cf_mf_t10_mixed_workspace_bytes")
    ws = workspace(nb.value, b.device)
    st = N.CfPcgStats()
    rc = lib.cf_pcg_solve_mf_t10_mixed(dm.n_nodes, dm.n_elem, ptr(conn), ptr(geo), float(lam),
                                       float(mu), ptr(offs), ptr(perm), ptr(b), ptr(x),
                                       ptr(fixed), tol, maxit, ptr(ws), nb.value,
                                       ctypes.byref(st), stream_handle())
    stats = {**_base_stats(st), "precision": "mixed",
             "precond": "jacobi", "tol": tol, "operator": "matfree"}
    last_stats.value = stats
    N.check(rc, "cf_pcg_solve_mf_t10_mixed")
    return stats


def pcg_t10_twolevel(dm, lam, mu, b, x, fixed=None, A=None, precision="mixed", tol=None,
                     maxit=None):
    """Two-level solver for Tet10 elasticity (cf_pcg_solve_t10_twolevel):
    FP64 flexible CG preconditioned by a coarse correction on the rigid-body
    modes of node aggregates around the matrix-free Chebyshev polynomial,
    which runs in FP32 (precision "mixed") or FP64 ("fp64", the same solver
    with every inner product in FP64).  The outer products use the assembled
    NB3 matrix A (a DeviceMatrix) when given, else the FP64 matrix-free
    operator."""
    o = options()
    tol = float(tol if tol is not None else o["tol"])
    maxit = int(maxit if maxit is not None else o["maxit"])
    prec = {"fp64": N.PREC_FP64, "mixed": N.PREC_MIXED}.get(precision)
    if prec is None:
        raise ValueError(f"two-level precision {precision!r} is not available (fp64, mixed)")
    lib = N.load()
    conn, geo, offs, perm = dm.matfree_arrays()
    As = A.cstruct() if A is not None else None
    Ap = ctypes.byref(As) if As is not None else None
    nb = ctypes.c_size_t()
    N.check(lib.cf_t10_twolevel_workspace_bytes(Ap, dm.n_nodes, dm.n_elem, prec,
                                                ctypes.byref(nb)),
            "cf_t10_twolevel_workspace_bytes")
    ws = workspace(nb.value, b.device)
    st = N.CfPcgStats()
    rc = lib.cf_pcg_solve_t10_twolevel(Ap, dm.n_nodes, dm.n_elem, ptr(conn), ptr(geo),
                                       float(lam), float(mu), ptr(offs), ptr(perm), ptr(dm.x),
                                       ptr(dm.y), ptr(dm.z), ptr(b), ptr(x), ptr(fixed), prec, tol,
                                       maxit, ptr(ws), nb.value, ctypes.byref(st),
                                       stream_handle())
    stats = {**_base_stats(st), "precision": precision, "precond": "two-level", "tol": tol,
             "operator": "assembled" if A is not None else "matfree"}
    last_stats.value = stats
    N.check(rc, "cf_pcg_solve_t10_twolevel")
    return stats


def pcg_matfree_t10_twolevel(dm, lam, mu, b, x, fixed=None, tol=None, maxit=None):
    """Two-level mixed precision with the matrix-free Tet10 operator
    outside and FP32 matrix-free products inside (pcg_t10_twolevel)."""
    return pcg_t10_twolevel(dm, lam, mu, b, x, fixed, A=None, precision="mixed", tol=tol,
                            maxit=maxit)


def pcg_matfree_t10(dm, lam, mu, b, x, fixed=None, precond=None, tol=None, maxit=None,
                    chunked=False):
    """PCG with the matrix-free Tet10 elasticity operator (FP64).  Default:
    the run-sum operator (cf_pcg_solve_mf_t10: element forces summed per node
    in LDS per tile of 256 consecutive elements, one partial per (tile, node)
    gathered per node); chunked=True: the Morton-chunked operator
    (cf_pcg_solve_mf_t10_plan)."""
    o = options()
    precond = resolve("precond", precond or o["precond"])
    tol = float(tol if tol is not None else o["tol"])
    maxit = int(maxit if maxit is not None else o["maxit"])
    pc = _pc_code(precond)
    lib = N.load()
    if chunked:
        plan, counts = dm.matfree_plan()
        offs, perm = dm.incidence()
        nb = ctypes.c_size_t()
        N.check(lib.cf_mf_t10_plan_workspace_bytes(dm.n_nodes, dm.n_elem, counts.ctypes.data, pc,
                                                   ctypes.byref(nb)),
                "cf_mf_t10_plan_workspace_bytes")
        ws = workspace(nb.value, b.device)
        st = N.CfPcgStats()
        rc = lib.cf_pcg_solve_mf_t10_plan(dm.n_nodes, dm.n_elem, ptr(dm.conn), ptr(dm.geometry()),
                                          ptr(offs), ptr(perm), ptr(plan), counts.ctypes.data,
                                          float(lam), float(mu), ptr(b), ptr(x), ptr(fixed), pc,
                                          tol, maxit, ptr(ws), nb.value, ctypes.byref(st),
                                          stream_handle())
        stats = {**_base_stats(st), "residual_updates": 0, "precision": "fp64",
                 "precond": precond, "tol": tol, "operator": "matfree"}
        last_stats.value = stats
        N.check(rc, "cf_pcg_solve_mf_t10_plan")
        return stats
    conn, geo, offs, perm = dm.matfree_arrays()
    nb = ctypes.c_size_t()
    N.check(lib.cf_mf_t10_workspace_bytes(dm.n_nodes, dm.n_elem, pc, ctypes.byref(nb)),
            "cf_mf_t10_workspace_bytes")
    ws = workspace(nb.value, b.device)
    st = N.CfPcgStats()
    rc = lib.cf_pcg_solve_mf_t10(dm.n_nodes, dm.n_elem, ptr(conn), ptr(geo),
                                 float(lam), float(mu), ptr(offs), ptr(perm), ptr(b), ptr(x),
                                 ptr(fixed), pc, tol, maxit, ptr(ws), nb.value,
                                 ctypes.byref(st), stream_handle())
    stats = {**_base_stats(st), "residual_updates": 0, "precision": "fp64",
             "precond": precond, "tol": tol, "operator": "matfree"}
    last_stats.value = stats
    N.check(rc, "cf_pcg_solve_mf_t10")
    return stats


def matfree_t10_lp_apply(dm, lam, mu, x, precision="mixed", scale=None, reps=0):
    """y = S K S x with the inner-precision operator of the Chebyshev and
    two-level solvers (cf_mf_t10_lp_apply): x float32 for "mixed", float64
    for "fp64"; scale (float64, nullable) gives S.  Returns (y, avg_ms of
    `reps` timed repetitions, or None)."""
    prec = {"fp64": N.PREC_FP64, "mixed": N.PREC_MIXED}[precision]
    want = torch.float64 if precision == "fp64" else torch.float32
    if x.dtype != want:
        raise TypeError(f"x must be {want} for precision {precision!r}")
    lib = N.load()
    conn, geo, offs, perm = dm.matfree_arrays()
    nb = ctypes.c_size_t()
    N.check(lib.cf_mf_t10_lp_workspace_bytes(dm.n_nodes, dm.n_elem, prec, ctypes.byref(nb)),
            "cf_mf_t10_lp_workspace_bytes")
    ws = workspace(nb.value, x.device)
    y = torch.empty_like(x)
    ms = ctypes.c_float(0.0)
    N.check(lib.cf_mf_t10_lp_apply(dm.n_nodes, dm.n_elem, ptr(conn), ptr(geo), float(lam),
                                   float(mu), ptr(offs), ptr(perm), prec, ptr(scale), ptr(x),
                                   ptr(y), int(reps), ctypes.byref(ms), ptr(ws), nb.value,
                                   stream_handle()), "cf_mf_t10_lp_apply")
    return y, (ms.value if reps > 0 else None)


def matfree_t10_apply(dm, lam, mu, x, fixed=None, chunked=False):
    """y = K x with the matrix-free Tet10 elasticity operator (chunked: the
    Morton-chunked plan, else the run-sum operator)."""
    lib = N.load()
    if chunked:
        plan, counts = dm.matfree_plan()
        nb = ctypes.c_size_t()
        N.check(lib.cf_mf_t10_plan_workspace_bytes(dm.n_nodes, dm.n_elem, counts.ctypes.data,
                                                   N.PC_JACOBI, ctypes.byref(nb)),
                "cf_mf_t10_plan_workspace_bytes")
        ws = workspace(nb.value, x.device)
        y = torch.empty_like(x)
        N.check(lib.cf_mf_t10_plan_apply(dm.n_nodes, dm.n_elem, ptr(plan), counts.ctypes.data,
                                         float(lam), float(mu), ptr(x), ptr(y), ptr(fixed),
                                         ptr(ws), nb.value, stream_handle()),
                "cf_mf_t10_plan_apply")
        return y
    conn, geo, offs, perm = dm.matfree_arrays()
    nb = ctypes.c_size_t()
    N.check(lib.cf_mf_t10_workspace_bytes(dm.n_nodes, dm.n_elem, N.PC_JACOBI, ctypes.byref(nb)),
            "cf_mf_t10_workspace_bytes")
    ws = workspace(nb.value, x.device)
    y = torch.empty_like(x)
    N.check(lib.cf_mf_t10_apply(dm.n_nodes, dm.n_elem, ptr(conn), ptr(geo),
                                float(lam), float(mu), ptr(offs), ptr(perm), ptr(x), ptr(y),
                                ptr(fixed), ptr(ws), nb.value, stream_handle()),
            "cf_mf_t10_apply")
    return y
