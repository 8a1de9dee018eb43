"""Tet10 scalar elements on the GPU: stiffness/load, exact mass, 6-node face
terms and theta-method transient heat.  Drop-in for cloudfem/fem/tet10.py.

shape_functions / quadrature return the constant reference-element tables
(the device kernels use the same tables, compiled into constant memory);
element_operators is the single-element helper of the reference API.
"""

import inspect
from math import factorial

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native as N
from ..coef import LazyPoints, centroid_points, dirichlet_scalar, eval_scalar
from ..device import DeviceMesh, scalar_csr, to_dev
from ..mesh.faces import boundary_faces10 as _faces10
from ..mesh.tet10_mesh import EDGES
from . import _transient
from .tet_poisson import _solve_scalar

__all__ = [
    "shape_functions", "quadrature", "element_operators",
    "assemble_tet10", "assemble_mass_tet10", "solve_tet10_poisson",
    "assemble_boundary_terms_tet10", "transient_heat_tet10",
]

_THETA = _transient.THETA


def shape_functions(L):
    """(N (..., 10), dN/dL (..., 10, 4)) at barycentric points L (tet10.py:63-81):
    corners L_i (2 L_i - 1), edges 4 L_i L_j."""
    L = np.asarray(L, dtype=float)
    N_ = np.empty(L.shape[:-1] + (10,))
    dN = np.zeros(L.shape[:-1] + (10, 4))
    for i in range(4):
        N_[..., i] = L[..., i] * (2.0 * L[..., i] - 1.0)
        dN[..., i, i] = 4.0 * L[..., i] - 1.0
    for s, (i, j) in enumerate(EDGES):
        N_[..., 4 + s] = 4.0 * L[..., i] * L[..., j]
        dN[..., 4 + s, i] = 4.0 * L[..., j]
        dN[..., 4 + s, j] = 4.0 * L[..., i]
    return N_, dN


def quadrature():
    """5-point Keast rule (Lq (5, 4), wq (5,)), sum(wq) = 1 (tet10.py:84-100)."""
    a, b = 0.5, 1.0 / 6.0
    Lq = np.array([[0.25, 0.25, 0.25, 0.25], [a, b, b, b], [b, a, b, b], [b, b, a, b],
                   [b, b, b, a]])
    wq = np.array([-4.0 / 5.0] + [9.0 / 20.0] * 4)
    return Lq, wq


def _dN_dxi(Lq):
    _, dL = shape_functions(Lq)
    return np.stack([dL[..., 1] - dL[..., 0], dL[..., 2] - dL[..., 0],
                     dL[..., 3] - dL[..., 0]], axis=-1)


def element_operators(mesh, el):
    """(G (5, 10, 3), N (5, 10), xq (5, 3), vol) of one element (tet10.py:103-126).

    Single-element API helper; assembly computes the inverse Jacobians of
    all elements on the GPU (cf_tet_geometry).
    """
    el = np.asarray(el, dtype=int)
    p = np.asarray(mesh.nodes)[el[:4]]
    jac = np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1)
    vol = abs(float(np.linalg.det(jac))) / 6.0
    Lq, _ = quadrature()
    N_, _ = shape_functions(Lq)
    G = _dN_dxi(Lq) @ np.linalg.inv(jac)
    return G, N_, Lq @ p, vol


def _qpoints(mesh):
    Lq, _ = quadrature()
    p = np.asarray(mesh.nodes)[np.asarray(mesh.elements)[:, :4]]
    return np.einsum("qa,mai->mqi", Lq, p)


def _assemble_device(mesh, k=1.0, f=0.0, dm=None):
    dm = dm or DeviceMesh(mesh)
    if dm.k != 10:
        raise ValueError("Tet10 assembly needs a 10-node mesh")
    xq = LazyPoints(5 * dm.n_elem, lambda: _qpoints(mesh).reshape(-1, 3))
    kv = to_dev(eval_scalar(k, xq), torch.float64, dm.dev)
    fv = to_dev(eval_scalar(f, xq), torch.float64, dm.dev)
    vals = dm.matrix_values(N.OP_T10_STIFF, coef=kv)
    rhs = torch.zeros(dm.n_nodes, dtype=torch.float64, device=dm.dev)
    dm.add_rhs(N.RHS_T10_SOURCE, fv, rhs)
    return dm, vals, rhs


def assemble_tet10(mesh, k=1.0, f=0.0):
    """(A csr_matrix (N, N), rhs (N,)): K_e = V sum_q w_q k_q G_q G_q^T,
    f_e = V sum_q w_q f_q N_q (tet10.py:138-165)."""
    dm, vals, rhs = _assemble_device(mesh, k, f)
    return scalar_csr(dm.pattern(), vals), rhs.cpu().numpy()


def solve_tet10_poisson(mesh, k=1.0, f=0.0, bc=None):
    """Steady Tet10 Poisson with Dirichlet bc (tet10.py:168-178)."""
    dm, vals, rhs = _assemble_device(mesh, k, f)
    return _solve_scalar(dm, vals, rhs, mesh, bc)


def _mass_reference():
    """int N_a N_b dV / V via barycentric moments 6 a!/(|a|+3)! (tet10.py:185-226)."""
    def poly(a):
        if a < 4:
            sq, ln = [0] * 4, [0] * 4
            sq[a], ln[a] = 2, 1
            return [(2.0, sq), (-1.0, ln)]
        e = [0] * 4
        i, j = EDGES[a - 4]
        e[i] += 1
        e[j] += 1
        return [(4.0, e)]

    def moment(al):
        num = 1
        for x in al:
            num *= factorial(x)
        return 6.0 * num / factorial(sum(al) + 3)

    M = np.zeros((10, 10))
    for a in range(10):
        for b in range(10):
            M[a, b] = sum(c1 * c2 * moment([u + v for u, v in zip(e1, e2)])
                          for c1, e1 in poly(a) for c2, e2 in poly(b))
    return M


_MASS_REF = _mass_reference()


def _mass_device(dm, mesh, rho_cp):
    rc = to_dev(eval_scalar(rho_cp, centroid_points(mesh)), torch.float64, dm.dev)
    return dm.matrix_values(N.OP_T10_MASS, coef=rc)


def assemble_mass_tet10(mesh, rho_cp=1.0):
    """Consistent Tet10 capacity rho_cp(x_c) V MASS_REF per element
    (tet10.py:229-249)."""
    dm = DeviceMesh(mesh)
    return scalar_csr(dm.pattern(), _mass_device(dm, mesh, rho_cp))


# 6-point degree-4 triangle rule, the reference's literals (tet10.py:300-308)
_TRI_L6 = np.array([
    [0.445948490915965, 0.445948490915965, 0.108103018168070],
    [0.445948490915965, 0.108103018168070, 0.445948490915965],
    [0.108103018168070, 0.445948490915965, 0.445948490915965],
    [0.091576213509771, 0.091576213509771, 0.816847572980459],
    [0.091576213509771, 0.816847572980459, 0.091576213509771],
    [0.816847572980459, 0.091576213509771, 0.091576213509771]])
_TRI_W6 = np.array([0.223381589678011] * 3 + [0.109951743655322] * 3)


def _boundary_faces10(mesh, name):
    """(F, 6) exterior faces [c0,c1,c2,m01,m12,m20] (tet10.py:329-352)."""
    return _faces10(mesh, name)


def _face_points(mesh, faces):
    pc = np.asarray(mesh.nodes)[faces[:, :3]]
    return np.einsum("qa,fai->fqi", _TRI_L6, pc).reshape(-1, 3)


def _flux_fn(q):
    """(fn(x, y, z, t), time_dependent) of a Neumann spec (tet10.py:366-381)."""
    if not callable(q):
        v = float(q)
        return (lambda x, y, z, t, v=v: v), False
    try:
        n_par = len(inspect.signature(q).parameters)
    except (TypeError, ValueError):
        n_par = 3
    if n_par >= 4:
        return q, True
    return (lambda x, y, z, t, q=q: q(x, y, z)), False


def _neumann_load(dm, mesh, neumann, t, out, only_timedep=None):
    """out += int q N dA over the Neumann faces at time t."""
    for name, q in neumann.items():
        fn, td = _flux_fn(q)
        if only_timedep is not None and td != only_timedep:
            continue
        faces = _boundary_faces10(mesh, name)
        if not faces.shape[0]:
            continue
        xq = _face_points(mesh, faces)
        qv = eval_scalar(lambda x, y, z, fn=fn: fn(x, y, z, t), xq)
        dm.add_face_rhs(N.FACE_T10_FLUX, faces, qv, out)


def _robin_device(dm, mesh, robin, F):
    all_faces, all_h = [], []
    for name, (h, tinf) in robin.items():
        faces = _boundary_faces10(mesh, name)
        if not faces.shape[0]:
            continue
        xq = _face_points(mesh, faces)
        hv = eval_scalar(h, xq)
        dm.add_face_rhs(N.FACE_T10_FLUX, faces, hv * eval_scalar(tinf, xq), F)
        all_faces.append(faces)
        all_h.append(hv)
    if not all_faces:
        return None
    return dm.face_matrix(N.FACE_T10_ROBIN_K, np.concatenate(all_faces), np.concatenate(all_h))


def assemble_boundary_terms_tet10(mesh, neumann=None, robin=None, t=0.0):
    """(K_add, F_add) of Neumann/Robin data on 6-node faces (tet10.py:438-464).

    As in the reference (an empty csr_matrix plus the Robin matrix), K_add
    stores no exact zeros.
    """
    dm = DeviceMesh(mesh)
    n = dm.n_nodes
    F = torch.zeros(n, dtype=torch.float64, device=dm.dev)
    K = sp.csr_matrix((n, n))
    if neumann:
        _neumann_load(dm, mesh, neumann, t, F)
    if robin:
        term = _robin_device(dm, mesh, robin, F)
        if term is not None:
            K = scalar_csr(*term)
            K.eliminate_zeros()
    return K, F.cpu().numpy()


def transient_heat_tet10(mesh, k=1.0, rho_cp=1.0, f=0.0, bc=None, neumann=None, robin=None,
                         t_end=1.0, dt=0.1, T0=0.0, method="cn"):
    """Theta-method transient heat on a Tet10 mesh (tet10.py:467-548)."""
    theta = _transient.check_method(method)
    dm, K_vals, F = _assemble_device(mesh, k, f)
    if robin:
        term = _robin_device(dm, mesh, robin, F)
        if term is not None:
            dm.add_subpattern(K_vals, *term)
    step_load = None
    if neumann:
        if any(_flux_fn(q)[1] for q in neumann.values()):
            steps_, h_ = _transient.step_count(t_end, dt)

            def step_load(s, h=h_):
                extra = torch.zeros_like(F)
                tmp = torch.zeros_like(F)
                _neumann_load(dm, mesh, neumann, s * h, tmp)
                _transient.axpby(1.0 - theta, tmp, 0.0, extra)
                tmp.zero_()
                _neumann_load(dm, mesh, neumann, (s + 1) * h, tmp)
                _transient.axpby(theta, tmp, 1.0, extra)
                return extra
        else:
            _neumann_load(dm, mesh, neumann, 0.0, F)
    C_vals = _mass_device(dm, mesh, rho_cp)
    steps, h = _transient.step_count(t_end, dt)
    mask, fvals = dirichlet_scalar(mesh, bc)
    return _transient.theta_run(dm, K_vals, C_vals, F, theta, t_end, steps, h, T0, mask, fvals,
                                step_load=step_load)
