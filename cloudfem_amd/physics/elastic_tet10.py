"""Tet10 3D linear elasticity on the GPU.

Drop-in for cloudfem/physics/elastic_tet10.py: assemble_elasticity_tet10
returns the (3N, 3N) scipy csr_matrix (int32 indices, explicit zeros kept)
and the load vector; solve_elasticity_tet10 solves the free-DOF system of
apply_dirichlet_vec3 with PCG instead of SuperLU.

The stiffness lives on the device in NB3 order (the value array of the DOF
CSR, 3x3 node blocks row-interleaved), so the solve never needs DOF-level
column indices.
"""

import numpy as np
import torch

from .. import _native as N
from ..coef import LazyPoints, dirichlet_vec3, eval_components, eval_scalar
from ..device import DeviceMesh, to_dev, vector_csr
from ..fem.tet10 import quadrature, shape_functions
from ..mesh.faces import T10_FACES as FACES10
from ..mesh.faces import boundary_faces10
from .. import solver as solver_mod
from ..solver import DeviceMatrix, pcg
from .elastic3d import constitutive_3d, lame

__all__ = ["boundary_faces10", "assemble_elasticity_tet10", "solve_elasticity_tet10",
           "element_stress_tet10", "FACES10"]

# 3-point degree-2 face rule (elastic_tet10.py:50-53)
_TRI_L = np.array([[2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0],
                   [1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0],
                   [1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0]])


def _qpoints(mesh):
    """Physical Keast points (M, 5, 3) = Lq @ corners (tet10.py:125)."""
    Lq, _ = quadrature()
    p = np.asarray(mesh.nodes)[np.asarray(mesh.elements)[:, :4]]
    return np.einsum("qa,mai->mqi", Lq, p)


def _assemble_device(mesh, E, nu, body=None, traction=None, T=None, alpha=0.0, T_ref=0.0,
                     dm=None, values=True):
    """(DeviceMesh, NB3 stiffness values, rhs (3N,)) on the GPU; with values=False the
    stiffness values are not assembled (None), for the matrix-free solvers."""
    dm = dm or DeviceMesh(mesh)
    if dm.k != 10:
        raise ValueError("assemble_elasticity_tet10 needs a 10-node mesh")
    lam, mu = lame(E, nu)
    vals = dm.matrix_values(N.OP_T10_ELAST, lam=lam, mu=mu) if values else None
    rhs = torch.zeros(3 * dm.n_nodes, dtype=torch.float64, device=dm.dev)
    if body is not None:
        xq = LazyPoints(5 * dm.n_elem, lambda: _qpoints(mesh).reshape(-1, 3))
        bq = eval_components(body, xq)
        if np.any(bq):
            dm.add_rhs(N.RHS_T10_BODY, to_dev(bq.ravel(), torch.float64, dm.dev), rhs)
    if T is not None and alpha != 0.0:
        if callable(T):
            Tq = eval_scalar(T, _qpoints(mesh).reshape(-1, 3)).reshape(-1, 5)
        else:
            Nq, _ = shape_functions(quadrature()[0])          # (5, 10)
            Tn = np.asarray(T, dtype=float)[np.asarray(mesh.elements)]  # (M, 10)
            Tq = Tn @ Nq.T
        s = alpha * (Tq - T_ref)
        dm.add_rhs(N.RHS_T10_THERMAL, to_dev(s.ravel(), torch.float64, dm.dev), rhs,
                   lam=lam, mu=mu)
    for name, comps in (traction or {}).items():
        faces = boundary_faces10(mesh, name)
        if faces.shape[0] == 0:
            continue
        pc = np.asarray(mesh.nodes)[faces[:, :3]]               # (F, 3, 3)
        xq = np.einsum("qa,fai->fqi", _TRI_L, pc).reshape(-1, 3)
        tq = eval_components(comps, xq)                          # (F*3, 3)
        dm.add_face_rhs(N.FACE_T10_TRACTION, faces, tq, rhs)
    return dm, vals, rhs


def assemble_elasticity_tet10(mesh, E, nu, body=None, traction=None, T=None, alpha=0.0,
                              T_ref=0.0):
    """(A csr_matrix (3N, 3N), rhs (3N,)), DOF order ux0, uy0, uz0, ...
    (elastic_tet10.py:135-211)."""
    dm, vals, rhs = _assemble_device(mesh, E, nu, body, traction, T, alpha, T_ref)
    return vector_csr(dm.pattern(), vals, dense3=True), rhs.cpu().numpy()


def solve_device(dm, vals, rhs, mesh, bc, lam=None, mu=None, **pcg_kw):
    """Solve on the device; returns (u (3N,) tensor, stats).

    The default (precision and precond "auto") is the two-level solver with
    FP32 inside and the assembled matrix outside.  precond "twolevel"
    selects the two-level solver (rigid-body coarse
    correction around a matrix-free Chebyshev polynomial, FP32 inside with
    precision "mixed", FP64 with "fp64"), whose outer products use the
    assembled NB3 matrix `vals` or, with solver option operator="matfree",
    the matrix-free operator.  Otherwise operator="matfree" applies the
    stiffness matrix-free from the geometry record (needs lam, mu; precision
    "mixed" adds the FP32 matrix-free Chebyshev preconditioner), and the
    assembled NB3 matrix is used by default.
    """
    n3 = 3 * dm.n_nodes
    if bc:
        mask, fv = dirichlet_vec3(mesh, bc)
    else:
        mask, fv = np.zeros(n3, dtype=bool), np.zeros(n3)
    x = to_dev(np.where(mask, fv, 0.0), torch.float64, dm.dev)
    fixed = to_dev(mask.astype(np.uint8), torch.uint8, dm.dev) if mask.any() else None
    o = solver_mod.options()
    prec = pcg_kw.get("precision", o["precision"])
    precond = pcg_kw.get("precond", o["precond"])
    if lam is not None:
        # Tet10 elasticity: the two-level solver by default, FP32 inside (it
        # solves the assembled matrix to the same FP64 residual bar)
        prec = "mixed" if prec == "auto" else prec
        precond = "twolevel" if precond == "auto" else precond
    prec = solver_mod.resolve("precision", prec)
    precond = solver_mod.resolve("precond", precond)
    pcg_kw = dict(pcg_kw, precision=prec, precond=precond)
    kw = {k: v for k, v in pcg_kw.items() if k in ("tol", "maxit")}
    matfree = o["operator"] == "matfree" and lam is not None
    if precond == "twolevel" and lam is not None:
        A = None
        if not matfree:
            if vals is None:
                vals = dm.matrix_values(N.OP_T10_ELAST, lam=lam, mu=mu)
            A = DeviceMatrix(N.FMT_NB3, dm.pattern(), vals)
        stats = solver_mod.pcg_t10_twolevel(dm, lam, mu, rhs, x, fixed, A=A,
                                            precision="fp64" if prec == "fp64" else "mixed",
                                            **kw)
    elif matfree and prec == "fp64":
        kw = {k: v for k, v in pcg_kw.items() if k != "precision"}
        stats = solver_mod.pcg_matfree_t10(dm, lam, mu, rhs, x, fixed, **kw)
    elif matfree and prec in ("mixed", "mixed16"):
        stats = solver_mod.pcg_matfree_t10_mixed(dm, lam, mu, rhs, x, fixed, **kw)
    else:
        if vals is None:
            vals = dm.matrix_values(N.OP_T10_ELAST, lam=lam, mu=mu)
        A = DeviceMatrix(N.FMT_NB3, dm.pattern(), vals)
        stats = pcg(A, rhs, x, fixed, **pcg_kw)
    return x, stats


def solve_elasticity_tet10(mesh, E, nu, bc=None, body=None, traction=None, T=None, alpha=0.0,
                           T_ref=0.0):
    """(ux, uy, uz) over all Tet10 nodes (elastic_tet10.py:214-229).  With the
    matrix-free operator selected the stiffness matrix is never formed."""
    matfree = solver_mod.options()["operator"] == "matfree"
    dm, vals, rhs = _assemble_device(mesh, E, nu, body, traction, T, alpha, T_ref,
                                     values=not matfree)
    lam, mu = lame(E, nu)
    x, _ = solve_device(dm, vals, rhs, mesh, bc, lam=lam, mu=mu)
    u = x.cpu().numpy()
    return u[0::3].copy(), u[1::3].copy(), u[2::3].copy()


def element_stress_tet10(mesh, u, E, nu, T=None, alpha=0.0, T_ref=0.0):
    """Centroid strain/stress (M, 6) of Tet10 elements (elastic_tet10.py:232-279)."""
    from .stress import element_stress
    return element_stress(mesh, u, E, nu, T=T, alpha=alpha, T_ref=T_ref, quadratic=True)
