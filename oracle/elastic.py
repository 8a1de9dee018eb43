"""Oracle for P1/Tet10 linear elasticity and the P1 vector mass.

Restates cloudfem/physics/elastic3d.py:26-302,
cloudfem/physics/elastic_tet10.py:42-279 and
cloudfem/physics/modal3d.py:40-75 (CPU, test-only).
"""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .p1 import tet_gradients
from .tet10 import element_operators, quadrature, shape_functions

VOIGT_I = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])


def constitutive_3d(E, nu):
    """Isotropic D, Voigt [xx,yy,zz,yz,xz,xy] (elastic3d.py:26-34)."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[0, 0] = D[1, 1] = D[2, 2] = lam + 2.0 * mu
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    return D


def strain_displacement(G):
    """B (6, 3n) from gradients G (n, 3), engineering shear
    (elastic3d.py:62-77, elastic_tet10.py:94-109)."""
    n = G.shape[0]
    B = np.zeros((6, 3 * n))
    for a in range(n):
        gx, gy, gz = G[a]
        c = 3 * a
        B[0, c], B[1, c + 1], B[2, c + 2] = gx, gy, gz
        B[3, c + 1], B[3, c + 2] = gz, gy
        B[4, c], B[4, c + 2] = gz, gx
        B[5, c], B[5, c + 1] = gy, gx
    return B


def dof_map(nodes):
    """Interleaved DOFs 3i+c (elastic3d.py:80-85)."""
    nodes = np.asarray(nodes, dtype=int)
    return (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()


def component_fns(vals):
    """3-component spec -> callables, None -> 0 (elastic3d.py:88-102)."""
    out = []
    for v in vals:
        if v is None:
            out.append(lambda x, y, z: 0.0)
        elif callable(v):
            out.append(v)
        else:
            out.append(lambda x, y, z, _v=v: _v)
    return out


def boundary_faces_p1(mesh, name):
    """elastic3d.py:37-59 (same rule as tet_heat._boundary_faces)."""
    from .p1 import boundary_faces
    return boundary_faces(mesh, name)


def assemble_elasticity_3d(mesh, E, nu, body=None, traction=None, T=None,
                           alpha=0.0, T_ref=0.0):
    """P1: K_e = V B^T D B, body b V/4, thermal V B^T D eps_th, traction
    t A/3 per face node (elastic3d.py:105-189)."""
    D = constitutive_3d(E, nu)
    n3 = 3 * mesh.n_nodes
    rows, cols, data = [], [], []
    rhs = np.zeros(n3)
    bf = component_fns(body) if body is not None else [lambda x, y, z: 0.0] * 3
    Tcall = callable(T)
    if T is not None and not Tcall:
        T = np.asarray(T, dtype=float)
    for el in mesh.elements:
        p = mesh.nodes[el]
        G, vol = tet_gradients(p)
        B = strain_displacement(G)
        ke = (B.T @ D @ B) * vol
        dofs = dof_map(el)
        rows.append(np.repeat(dofs, 12))
        cols.append(np.tile(dofs, 12))
        data.append(ke.ravel())
        xc, yc, zc = p.mean(axis=0)
        bv = np.array([fn(xc, yc, zc) for fn in bf])
        rhs[dofs] += np.tile(bv * vol / 4.0, 4)
        if T is not None and alpha != 0.0:
            Tc = T(xc, yc, zc) if Tcall else T[el].mean()
            rhs[dofs] += (B.T @ D @ (alpha * (Tc - T_ref) * VOIGT_I)) * vol
    for name, vals in (traction or {}).items():
        tf = component_fns(vals)
        for tri in boundary_faces_p1(mesh, name):
            p = mesh.nodes[tri]
            xc, yc, zc = p.mean(axis=0)
            tv = np.array([fn(xc, yc, zc) for fn in tf])
            area = 0.5 * float(np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0])))
            for node in tri:
                rhs[3 * int(node):3 * int(node) + 3] += tv * area / 3.0
    A = sp.coo_matrix((np.concatenate(data),
                       (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n3, n3)).tocsr()
    return A, rhs


def fixed_dofs_vec3(mesh, bc):
    """{dof: value} of a vector bc dict (elastic3d.py:204-218)."""
    fixed = {}
    for name, vals in bc.items():
        ids = (np.arange(mesh.n_nodes) if name == "all"
               else np.asarray(mesh.boundaries[name]))
        for comp, v in enumerate(vals):
            if v is None:
                continue
            fn = v if callable(v) else (lambda x, y, z, _v=v: _v)
            for i in ids:
                i = int(i)
                fixed[3 * i + comp] = fn(*mesh.nodes[i])
    return fixed


def apply_dirichlet_vec3(A, rhs, mesh, bc):
    """Reduced free system + spsolve (elastic3d.py:192-233)."""
    rhs = np.asarray(rhs, dtype=float).copy()
    n3 = 3 * mesh.n_nodes
    fixed = fixed_dofs_vec3(mesh, bc)
    u = np.zeros(n3)
    idx = np.array(sorted(fixed), dtype=int)
    if len(idx):
        v = np.array([fixed[i] for i in idx])
        u[idx] = v
        free = np.ones(n3, dtype=bool)
        free[idx] = False
        A = A.tocsr()
        rf = rhs[free] - A[free][:, idx] @ v
        Aff = A[free][:, free].tocsr()
        u[free] = spla.spsolve(Aff, rf)
    else:
        u = np.asarray(spla.spsolve(A.tocsr(), rhs)).ravel()
    return u


def reduced_system_vec3(A, rhs, mesh, bc):
    """(A_ff, rhs_f, free mask, u with fixed values) -- the system that
    apply_dirichlet_vec3 hands to spsolve (elastic3d.py:220-229)."""
    rhs = np.asarray(rhs, dtype=float)
    n3 = 3 * mesh.n_nodes
    fixed = fixed_dofs_vec3(mesh, bc)
    idx = np.array(sorted(fixed), dtype=int)
    u = np.zeros(n3)
    free = np.ones(n3, dtype=bool)
    if len(idx):
        v = np.array([fixed[i] for i in idx])
        u[idx] = v
        free[idx] = False
        A = A.tocsr()
        return A[free][:, free].tocsr(), rhs[free] - A[free][:, idx] @ v, free, u
    return A.tocsr(), rhs.copy(), free, u


def solve_elasticity_3d(mesh, E, nu, bc=None, body=None, traction=None,
                        T=None, alpha=0.0, T_ref=0.0):
    """elastic3d.py:236-249."""
    A, rhs = assemble_elasticity_3d(mesh, E, nu, body, traction, T, alpha, T_ref)
    u = (apply_dirichlet_vec3(A, rhs, mesh, bc) if bc
         else np.asarray(spla.spsolve(A.tocsr(), rhs)).ravel())
    return u[0::3].copy(), u[1::3].copy(), u[2::3].copy()


def _as_vec(u, n):
    if isinstance(u, tuple):
        uv = np.empty(3 * n)
        uv[0::3], uv[1::3], uv[2::3] = u
        return uv
    return np.asarray(u, dtype=float).ravel()


def element_stress_strain_3d(mesh, u, E, nu, T=None, alpha=0.0, T_ref=0.0):
    """P1 centroid strain/stress (elastic3d.py:252-289)."""
    uv = _as_vec(u, mesh.n_nodes)
    D = constitutive_3d(E, nu)
    Tcall = callable(T)
    if T is not None and not Tcall:
        T = np.asarray(T, dtype=float)
    sig = np.zeros((mesh.n_elements, 6))
    eps = np.zeros((mesh.n_elements, 6))
    for e, el in enumerate(mesh.elements):
        p = mesh.nodes[el]
        G, _ = tet_gradients(p)
        ee = strain_displacement(G) @ uv[dof_map(el)]
        eps[e] = ee
        eth = np.zeros(6)
        if T is not None and alpha != 0.0:
            Tc = T(*p.mean(axis=0)) if Tcall else T[el].mean()
            eth = alpha * (Tc - T_ref) * VOIGT_I
        sig[e] = D @ (ee - eth)
    return sig, eps


def von_mises_3d(sig):
    """elastic3d.py:292-302."""
    s = np.asarray(sig, dtype=float)
    return np.sqrt(0.5 * ((s[..., 0] - s[..., 1]) ** 2
                          + (s[..., 1] - s[..., 2]) ** 2
                          + (s[..., 2] - s[..., 0]) ** 2)
                   + 3.0 * (s[..., 3] ** 2 + s[..., 4] ** 2 + s[..., 5] ** 2))


# ---------------------------------------------------------------- Tet10

TRI_L3 = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6],
                   [1 / 6, 1 / 6, 2 / 3]])
TRI_W3 = np.full(3, 1.0 / 3.0)


def boundary_faces10(mesh, name):
    """elastic_tet10.py:68-91 (same rule and order as the fem copy)."""
    from .tet10 import boundary_faces10 as bf10
    return bf10(mesh, name)


def assemble_elasticity_tet10(mesh, E, nu, body=None, traction=None, T=None,
                              alpha=0.0, T_ref=0.0):
    """K_e = V sum_q w_q B_q^T D B_q (30x30); body/thermal at the Keast
    points; traction on 6-node faces with the 3-point rule
    (elastic_tet10.py:135-211)."""
    _, wq = quadrature()
    D = constitutive_3d(E, nu)
    n3 = 3 * mesh.n_nodes
    rows, cols, data = [], [], []
    rhs = np.zeros(n3)
    bf = component_fns(body) if body is not None else [lambda x, y, z: 0.0] * 3
    Tcall = callable(T)
    if T is not None and not Tcall:
        T = np.asarray(T, dtype=float)
    for el in mesh.elements:
        G, N, xq, vol = element_operators(mesh, el)
        dofs = dof_map(el)
        B = np.stack([strain_displacement(G[q]) for q in range(5)])
        DB = np.einsum("ab,qbi->qai", D, B)
        Ke = vol * np.einsum("q,qai,qaj->ij", wq, B, DB)
        rows.append(np.repeat(dofs, 30))
        cols.append(np.tile(dofs, 30))
        data.append(Ke.ravel())
        bv = np.array([[fn(*x) for fn in bf] for x in xq])
        rhs[dofs] += (vol * np.einsum("q,qa,qc->ac", wq, N, bv)).ravel()
        if T is not None and alpha != 0.0:
            Tq = np.array([T(*x) for x in xq]) if Tcall else N @ T[el]
            eth = alpha * (Tq - T_ref)[:, None] * VOIGT_I[None, :]
            rhs[dofs] += vol * np.einsum("q,qai,qa->i", wq, B, eth @ D.T)
    if traction:
        from .tet10 import face_shape6
        Nf = face_shape6(TRI_L3)
        for name, vals in traction.items():
            tf = component_fns(vals)
            for face in boundary_faces10(mesh, name):
                pc = mesh.nodes[face[:3]]
                area = 0.5 * float(np.linalg.norm(np.cross(pc[1] - pc[0],
                                                           pc[2] - pc[0])))
                xq = TRI_L3 @ pc
                tv = np.array([[fn(*x) for fn in tf] for x in xq])
                fe = area * np.einsum("q,qa,qc->ac", TRI_W3, Nf, tv)
                rhs[dof_map(face)] += fe.ravel()
    A = sp.coo_matrix((np.concatenate(data),
                       (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n3, n3)).tocsr()
    return A, rhs


def solve_elasticity_tet10(mesh, E, nu, bc=None, body=None, traction=None,
                           T=None, alpha=0.0, T_ref=0.0):
    """elastic_tet10.py:214-229."""
    A, rhs = assemble_elasticity_tet10(mesh, E, nu, body, traction, T, alpha,
                                       T_ref)
    u = (apply_dirichlet_vec3(A, rhs, mesh, bc) if bc
         else np.asarray(spla.spsolve(A.tocsr(), rhs)).ravel())
    return u[0::3].copy(), u[1::3].copy(), u[2::3].copy()


def element_stress_tet10(mesh, u, E, nu, T=None, alpha=0.0, T_ref=0.0):
    """Centroid strain/stress of Tet10 elements (elastic_tet10.py:232-279)."""
    uv = _as_vec(u, mesh.n_nodes)
    D = constitutive_3d(E, nu)
    Tcall = callable(T)
    if T is not None and not Tcall:
        T = np.asarray(T, dtype=float)
    Nc, dL = shape_functions(np.array([[0.25] * 4]))
    dxi = np.stack([dL[0, :, 1] - dL[0, :, 0], dL[0, :, 2] - dL[0, :, 0],
                    dL[0, :, 3] - dL[0, :, 0]], axis=-1)
    sig = np.zeros((mesh.n_elements, 6))
    eps = np.zeros((mesh.n_elements, 6))
    for e, el in enumerate(mesh.elements):
        p = mesh.nodes[el[:4]]
        jac = np.column_stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]])
        B = strain_displacement(dxi @ np.linalg.inv(jac))
        ee = B @ uv[dof_map(el)]
        eps[e] = ee
        eth = np.zeros(6)
        if T is not None and alpha != 0.0:
            Tc = T(*p.mean(axis=0)) if Tcall else float(Nc[0] @ T[el])
            eth = alpha * (Tc - T_ref) * VOIGT_I
        sig[e] = D @ (ee - eth)
    return sig, eps


# ---------------------------------------------------------------- mass

def assemble_mass_3d(mesh, rho, lumped=False):
    """P1 vector mass: consistent rho V (1+d)/20 per component block, or
    lumped rho V/4 on the diagonal (modal3d.py:40-75)."""
    n3 = 3 * mesh.n_nodes
    vols = mesh.element_volumes()
    if lumped:
        diag = np.zeros(n3)
        for e, el in enumerate(mesh.elements):
            diag[dof_map(el)] += rho * vols[e] / 4.0
        return sp.diags(diag, format="csr")
    base = (np.ones((4, 4)) + np.eye(4)) / 20.0
    rows, cols, data = [], [], []
    for e, el in enumerate(mesh.elements):
        me = rho * vols[e] * base
        dofs = dof_map(el)
        for c in range(3):
            dc = dofs[c::3]
            rows.append(np.repeat(dc, 4))
            cols.append(np.tile(dc, 4))
            data.append(me.ravel())
    return sp.coo_matrix((np.concatenate(data),
                          (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n3, n3)).tocsr()
