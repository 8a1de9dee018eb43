"""P1 capacity matrix, Neumann/Robin face terms and theta-method transient
heat on the GPU.  Drop-in for cloudfem/fem/tet_heat.py."""

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native as N
from ..coef import centroid_points, dirichlet_scalar, eval_scalar
from ..device import DeviceMesh, scalar_csr, to_dev
from ..mesh.faces import boundary_faces as _boundary_faces_np
from . import _transient
from .tet_poisson import _assemble_device, apply_dirichlet, assemble  # noqa: F401

_THETA = _transient.THETA


def _mass_device(dm, mesh, rho_cp):
    rc = to_dev(eval_scalar(rho_cp, centroid_points(mesh)), torch.float64, dm.dev)
    return dm.matrix_values(N.OP_P1_MASS, coef=rc)


def assemble_mass_tet(mesh, rho_cp=1.0):
    """Consistent capacity C = sum_e rho_cp(x_c) V (1 + d_ab)/20
    (tet_heat.py:72-94)."""
    dm = DeviceMesh(mesh)
    return scalar_csr(dm.pattern(), _mass_device(dm, mesh, rho_cp))


def _boundary_faces(mesh, name):
    """Exterior faces (F, 3) of a named boundary (tet_heat.py:97-120)."""
    return _boundary_faces_np(mesh, name)


def _face_centroids(mesh, faces):
    return np.asarray(mesh.nodes)[faces[:, :3]].mean(axis=1)


def _boundary_device(dm, mesh, neumann, robin):
    """(Robin face pattern + values or None, load vector) on the device.

    All Robin faces of all names form one face matrix, as the reference
    gathers them into one COO (tet_heat.py:156-171).
    """
    F = torch.zeros(dm.n_nodes, dtype=torch.float64, device=dm.dev)
    for name, q in (neumann or {}).items():
        faces = _boundary_faces(mesh, name)
        if faces.shape[0]:
            dm.add_face_rhs(N.FACE_P1_FLUX, faces, eval_scalar(q, _face_centroids(mesh, faces)),
                            F)
    all_faces, all_h = [], []
    for name, (h, tinf) in (robin or {}).items():
        faces = _boundary_faces(mesh, name)
        if not faces.shape[0]:
            continue
        cen = _face_centroids(mesh, faces)
        hv = eval_scalar(h, cen)
        dm.add_face_rhs(N.FACE_P1_FLUX, faces, hv * eval_scalar(tinf, cen), F)
        all_faces.append(faces)
        all_h.append(hv)
    robin_term = None
    if all_faces:
        robin_term = dm.face_matrix(N.FACE_P1_ROBIN_K, np.concatenate(all_faces),
                                    np.concatenate(all_h))
    return robin_term, F


def _face_csr(n, term):
    if term is None:
        return sp.coo_matrix((n, n)).tocsr()
    return scalar_csr(*term)


def assemble_boundary_terms(mesh, neumann=None, robin=None):
    """(K_add csr_matrix, F_add (N,)): Neumann q A/3; Robin h A (1+d)/12 and
    h T_inf A/3 (tet_heat.py:136-172)."""
    dm = DeviceMesh(mesh)
    term, F = _boundary_device(dm, mesh, neumann, robin)
    return _face_csr(dm.n_nodes, term), F.cpu().numpy()


def _dirichlet_values(mesh, bc):
    """{node: value} of a bc dict (tet_heat.py:175-190)."""
    mask, vals = dirichlet_scalar(mesh, bc)
    return {int(i): float(vals[i]) for i in np.nonzero(mask)[0]}


def transient_heat_tet(mesh, k=1.0, rho_cp=1.0, f=0.0, bc=None, neumann=None, robin=None,
                       t_end=1.0, dt=0.1, T0=0.0, method="cn"):
    """Theta-method transient heat on a P1 mesh (tet_heat.py:197-274).

    Returns (times (S+1,), T_history (S+1, N)).
    """
    theta = _transient.check_method(method)
    dm, K_vals, F = _assemble_device(mesh, k, f)
    term, Fb = _boundary_device(dm, mesh, neumann, robin)
    if term is not None:
        dm.add_subpattern(K_vals, *term)
    _transient.axpby(1.0, Fb, 1.0, F)
    C_vals = _mass_device(dm, mesh, rho_cp)
    steps, h = _transient.step_count(t_end, dt)
    mask, fvals = dirichlet_scalar(mesh, bc)
    return _transient.theta_run(dm, K_vals, C_vals, F, theta, t_end, steps, h, T0, mask, fvals)
