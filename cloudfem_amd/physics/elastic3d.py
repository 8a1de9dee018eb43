"""P1 3D linear elasticity on the GPU.  Drop-in for cloudfem/physics/elastic3d.py."""

import numpy as np
import torch

from .. import _native as N
from ..coef import centroid_points, centroids, dirichlet_vec3, eval_components
from ..device import DeviceMesh, to_dev, vector_csr
from ..mesh.faces import boundary_faces as _faces_p1
from ..solver import DeviceMatrix, pcg


def lame(E, nu):
    """(lambda, mu) exactly as constitutive_3d forms them (elastic3d.py:28-29)."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    return lam, mu


def constitutive_3d(E, nu):
    """6x6 isotropic D, Voigt [xx, yy, zz, yz, xz, xy] (elastic3d.py:26-34)."""
    lam, mu = lame(E, nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[0, 0] = D[1, 1] = D[2, 2] = lam + 2.0 * mu
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    return D


def boundary_faces(mesh, name):
    """Exterior faces (F, 3) of a named boundary (elastic3d.py:37-59)."""
    return _faces_p1(mesh, name)


def _assemble_device(mesh, E, nu, body=None, traction=None, T=None, alpha=0.0, T_ref=0.0,
                     dm=None):
    dm = dm or DeviceMesh(mesh)
    if dm.k != 4:
        raise ValueError("assemble_elasticity_3d needs a 4-node mesh")
    lam, mu = lame(E, nu)
    vals = dm.matrix_values(N.OP_P1_ELAST, lam=lam, mu=mu)
    rhs = torch.zeros(3 * dm.n_nodes, dtype=torch.float64, device=dm.dev)
    if body is not None:
        bv = eval_components(body, centroid_points(mesh))
        if np.any(bv):
            dm.add_rhs(N.RHS_P1_BODY, to_dev(bv.ravel(), torch.float64, dm.dev), rhs)
    if T is not None and alpha != 0.0:
        if callable(T):
            from ..coef import eval_scalar
            Tc = eval_scalar(T, centroids(mesh))
        else:
            Tc = np.asarray(T, dtype=float)[np.asarray(mesh.elements)].mean(axis=1)
        dm.add_rhs(N.RHS_P1_THERMAL, to_dev(alpha * (Tc - T_ref), torch.float64, dm.dev), rhs,
                   lam=lam, mu=mu)
    for name, comps in (traction or {}).items():
        faces = boundary_faces(mesh, name)
        if faces.shape[0] == 0:
            continue
        tv = eval_components(comps, np.asarray(mesh.nodes)[faces].mean(axis=1))
        dm.add_face_rhs(N.FACE_P1_TRACTION, faces, tv, rhs)
    return dm, vals, rhs


def assemble_elasticity_3d(mesh, E, nu, body=None, traction=None, T=None, alpha=0.0,
                           T_ref=0.0):
    """(A csr_matrix (3N, 3N), rhs (3N,)) (elastic3d.py:105-189)."""
    dm, vals, rhs = _assemble_device(mesh, E, nu, body, traction, T, alpha, T_ref)
    return vector_csr(dm.pattern(), vals, dense3=True), rhs.cpu().numpy()


def apply_dirichlet_vec3(A, rhs, mesh, bc):
    """Full (3N,) solution of the free-DOF system (elastic3d.py:192-233).

    A (any square sparse matrix, e.g. from assemble_elasticity_3d/_tet10) is
    uploaded as scalar CSR and the free block is solved on the GPU with PCG.
    """
    from ..device import require_gpu, upload_csr
    dev = require_gpu()
    dmat = upload_csr(A, dev)
    n3 = A.shape[0]
    mask, fv = dirichlet_vec3(mesh, bc)
    if mask.shape[0] != n3:
        raise ValueError("matrix size does not match the mesh")
    x = to_dev(np.where(mask, fv, 0.0), torch.float64, dev)
    fixed = to_dev(mask.astype(np.uint8), torch.uint8, dev) if mask.any() else None
    b = to_dev(np.asarray(rhs, dtype=float), torch.float64, dev)
    pcg(dmat, b, x, fixed)
    return x.cpu().numpy()


def solve_device(dm, vals, rhs, mesh, bc, **pcg_kw):
    n3 = 3 * dm.n_nodes
    A = DeviceMatrix(N.FMT_NB3, dm.pattern(), vals)
    mask, fv = dirichlet_vec3(mesh, bc) if bc else (np.zeros(n3, bool), np.zeros(n3))
    x = to_dev(np.where(mask, fv, 0.0), torch.float64, dm.dev)
    fixed = to_dev(mask.astype(np.uint8), torch.uint8, dm.dev) if mask.any() else None
    stats = pcg(A, rhs, x, fixed, **pcg_kw)
    return x, stats


def solve_elasticity_3d(mesh, E, nu, bc=None, body=None, traction=None, T=None, alpha=0.0,
                        T_ref=0.0):
    """(ux, uy, uz) (elastic3d.py:236-249)."""
    dm, vals, rhs = _assemble_device(mesh, E, nu, body, traction, T, alpha, T_ref)
    x, _ = solve_device(dm, vals, rhs, mesh, bc)
    u = x.cpu().numpy()
    return u[0::3].copy(), u[1::3].copy(), u[2::3].copy()


def element_stress_strain_3d(mesh, u, E, nu, T=None, alpha=0.0, T_ref=0.0):
    """Per-element (sig, eps) (M, 6) of P1 elements (elastic3d.py:252-289)."""
    from .stress import element_stress
    return element_stress(mesh, u, E, nu, T=T, alpha=alpha, T_ref=T_ref)


def von_mises_3d(sig):
    """Von Mises stress of Voigt stresses (elastic3d.py:292-302); elementwise
    NumPy on (M, 6) or (6,) host arrays."""
    s = np.asarray(sig, dtype=float)
    return np.sqrt(0.5 * ((s[..., 0] - s[..., 1]) ** 2 + (s[..., 1] - s[..., 2]) ** 2
                          + (s[..., 2] - s[..., 0]) ** 2)
                   + 3.0 * (s[..., 3] ** 2 + s[..., 4] ** 2 + s[..., 5] ** 2))
