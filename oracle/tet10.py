"""Oracle for Tet10 scalar assembly, exact mass, face terms, transient.

Restates cloudfem/fem/tet10.py:63-548 (CPU, test-only).
"""

import inspect
from math import factorial

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .meshes import TET10_EDGES
from .p1 import apply_dirichlet, as_fn, dirichlet_values, THETA


def shape_functions(L):
    """N_i = L_i(2L_i - 1), N_ij = 4 L_i L_j and dN/dL (tet10.py:63-81)."""
    L = np.asarray(L, dtype=float)
    N = np.empty(L.shape[:-1] + (10,))
    dN = np.zeros(L.shape[:-1] + (10, 4))
    for i in range(4):
        N[..., i] = L[..., i] * (2.0 * L[..., i] - 1.0)
        dN[..., i, i] = 4.0 * L[..., i] - 1.0
    for s, (i, j) in enumerate(TET10_EDGES):
        N[..., 4 + s] = 4.0 * L[..., i] * L[..., j]
        dN[..., 4 + s, i] = 4.0 * L[..., j]
        dN[..., 4 + s, j] = 4.0 * L[..., i]
    return N, dN


def quadrature():
    """5-point Keast rule, weights summing to 1 (tet10.py:84-100)."""
    a, b = 0.5, 1.0 / 6.0
    Lq = np.array([[0.25] * 4, [a, b, b, b], [b, a, b, b], [b, b, a, b],
                   [b, b, b, a]])
    wq = np.array([-0.8, 0.45, 0.45, 0.45, 0.45])
    return Lq, wq


def ref_gradients():
    """dN/dxi at the 5 points with xi = (L2, L3, L4) (tet10.py:120-123)."""
    Lq, _ = quadrature()
    N, dL = shape_functions(Lq)
    dxi = np.stack([dL[..., 1] - dL[..., 0], dL[..., 2] - dL[..., 0],
                    dL[..., 3] - dL[..., 0]], axis=-1)
    return N, dxi


def element_operators(mesh, el):
    """(G (5,10,3), N (5,10), xq (5,3), vol) of one element (tet10.py:103-126)."""
    el = np.asarray(el, dtype=int)
    p = mesh.nodes[el[:4]]
    jac = np.column_stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]])
    vol = abs(float(np.linalg.det(jac))) / 6.0
    Lq, _ = quadrature()
    N, dxi = ref_gradients()
    G = dxi @ np.linalg.inv(jac)
    return G, N, Lq @ p, vol


def assemble_tet10(mesh, k=1.0, f=0.0):
    """K_e = V sum_q w_q k_q G_q G_q^T, f_e = V sum_q w_q f_q N_q
    (tet10.py:138-165)."""
    n = mesh.n_nodes
    kf, ff = as_fn(k), as_fn(f)
    _, wq = quadrature()
    rows, cols, data = [], [], []
    rhs = np.zeros(n)
    for el in mesh.elements:
        G, N, xq, vol = element_operators(mesh, el)
        kv = np.array([kf(*x) for x in xq])
        fv = np.array([ff(*x) for x in xq])
        Ke = vol * np.einsum("q,qai,qbi->ab", wq * kv, G, G)
        rhs[el] += vol * ((wq * fv)[:, None] * N).sum(axis=0)
        rows.append(np.repeat(el, 10))
        cols.append(np.tile(el, 10))
        data.append(Ke.ravel())
    K = sp.coo_matrix((np.concatenate(data),
                       (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    return K, rhs


def solve_tet10_poisson(mesh, k=1.0, f=0.0, bc=None):
    """tet10.py:168-178."""
    A, rhs = assemble_tet10(mesh, k, f)
    if bc:
        A, rhs = apply_dirichlet(A, rhs, mesh, bc)
    return np.asarray(spla.spsolve(A, rhs)).ravel()


def mass_reference():
    """int N_a N_b dV / V from barycentric moments 6 a!/(|a|+3)!
    (tet10.py:185-226)."""
    polys = []
    for i in range(4):
        sq = [0] * 4
        sq[i] = 2
        ln = [0] * 4
        ln[i] = 1
        polys.append([(2.0, sq), (-1.0, ln)])
    for i, j in TET10_EDGES:
        e = [0] * 4
        e[i] += 1
        e[j] += 1
        polys.append([(4.0, e)])

    def moment(alpha):
        num = 1
        for a in alpha:
            num *= factorial(a)
        return 6.0 * num / factorial(sum(alpha) + 3)

    M = np.zeros((10, 10))
    for a in range(10):
        for b in range(10):
            M[a, b] = sum(c1 * c2 * moment([x + y for x, y in zip(e1, e2)])
                          for c1, e1 in polys[a] for c2, e2 in polys[b])
    return M


MASS_REF = mass_reference()


def assemble_mass_tet10(mesh, rho_cp=1.0):
    """rho(x_c) V MASS_REF per element, centroid of the corners
    (tet10.py:229-249)."""
    n = mesh.n_nodes
    rf = as_fn(rho_cp)
    vols = mesh.element_volumes()
    rows, cols, data = [], [], []
    for e, el in enumerate(mesh.elements):
        cen = mesh.nodes[el[:4]].mean(axis=0)
        rows.append(np.repeat(el, 10))
        cols.append(np.tile(el, 10))
        data.append((rf(*cen) * vols[e] * MASS_REF).ravel())
    return sp.coo_matrix((np.concatenate(data),
                          (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n)).tocsr()


# (corner triple, midpoint triple) per tet face (tet10.py:287-292)
FACES10 = (((0, 1, 2), (4, 5, 6)), ((0, 3, 1), (7, 8, 4)),
           ((0, 2, 3), (6, 9, 7)), ((1, 3, 2), (8, 9, 5)))
FACE_EDGES = ((0, 1), (1, 2), (2, 0))

# 6-point degree-4 triangle rule (tet10.py:300-308)
TRI_L6 = np.array([
    [0.445948490915965, 0.445948490915965, 0.108103018168070],
    [0.445948490915965, 0.108103018168070, 0.445948490915965],
    [0.108103018168070, 0.445948490915965, 0.445948490915965],
    [0.091576213509771, 0.091576213509771, 0.816847572980459],
    [0.091576213509771, 0.816847572980459, 0.091576213509771],
    [0.816847572980459, 0.091576213509771, 0.091576213509771]])
TRI_W6 = np.array([0.223381589678011] * 3 + [0.109951743655322] * 3)


def face_shape6(l):
    """6-node triangle shape functions [c0,c1,c2,m01,m12,m20] (tet10.py:311-323)."""
    N = np.empty(l.shape[:-1] + (6,))
    for i in range(3):
        N[..., i] = l[..., i] * (2.0 * l[..., i] - 1.0)
    for s, (i, j) in enumerate(FACE_EDGES):
        N[..., 3 + s] = 4.0 * l[..., i] * l[..., j]
    return N


FACE_N6 = face_shape6(TRI_L6)


def boundary_faces10(mesh, name):
    """Exterior 6-node faces, first-seen order (tet10.py:329-352)."""
    members = set(int(i) for i in mesh.boundaries[name])
    seen, count = {}, {}
    for el in mesh.elements:
        for cor, mid in FACES10:
            tri = tuple(int(el[c]) for c in cor)
            key = tuple(sorted(tri))
            count[key] = count.get(key, 0) + 1
            if key not in seen:
                seen[key] = tuple(int(el[c]) for c in cor + mid)
    out = [seen[k] for k, cnt in count.items()
           if cnt == 1 and all(v in members for v in k)]
    return np.array(out, dtype=int).reshape(-1, 6)


def _face_quad(mesh, name):
    out = []
    for face in boundary_faces10(mesh, name):
        pc = mesh.nodes[face[:3]]
        area = 0.5 * float(np.linalg.norm(np.cross(pc[1] - pc[0], pc[2] - pc[0])))
        out.append((face, area, TRI_L6 @ pc))
    return out


def flux_fn(q):
    """(fn(x,y,z,t), time_dependent) from float / f(x,y,z) / f(x,y,z,t)
    (tet10.py:366-381)."""
    if not callable(q):
        v = float(q)
        return (lambda x, y, z, t, _v=v: _v), False
    try:
        npar = len(inspect.signature(q).parameters)
    except (TypeError, ValueError):
        npar = 3
    if npar >= 4:
        return q, True
    return (lambda x, y, z, t, _q=q: _q(x, y, z)), False


def neumann_spec(mesh, neumann):
    return [(fn, _face_quad(mesh, name), td)
            for name, q in neumann.items() for fn, td in [flux_fn(q)]]


def neumann_load(n, spec, t):
    """F += area sum_q w_q q(x_q, t) N_q (tet10.py:393-401)."""
    F = np.zeros(n)
    for fn, faces, _ in spec:
        for face, area, xq in faces:
            qv = np.array([fn(x, y, z, t) for x, y, z in xq])
            F[face] += area * ((TRI_W6 * qv)[:, None] * FACE_N6).sum(axis=0)
    return F


def robin_terms(n, mesh, robin):
    """Face mass h int N N dA and load h T_inf int N dA (tet10.py:404-435)."""
    rows, cols, data = [], [], []
    F = np.zeros(n)
    for name, (h, tinf) in robin.items():
        hf, tf = as_fn(h), as_fn(tinf)
        for face, area, xq in _face_quad(mesh, name):
            hv = np.array([hf(*x) for x in xq])
            tv = np.array([tf(*x) for x in xq])
            Ke = area * (FACE_N6 * (TRI_W6 * hv)[:, None]).T @ FACE_N6
            F[face] += area * ((TRI_W6 * hv * tv)[:, None] * FACE_N6).sum(axis=0)
            rows.append(np.repeat(face, 6))
            cols.append(np.tile(face, 6))
            data.append(Ke.ravel())
    if data:
        K = sp.coo_matrix((np.concatenate(data),
                           (np.concatenate(rows), np.concatenate(cols))),
                          shape=(n, n)).tocsr()
    else:
        K = sp.csr_matrix((n, n))
    return K, F


def assemble_boundary_terms_tet10(mesh, neumann=None, robin=None, t=0.0):
    """tet10.py:438-464."""
    n = mesh.n_nodes
    F = np.zeros(n)
    K = sp.csr_matrix((n, n))
    if neumann:
        F += neumann_load(n, neumann_spec(mesh, neumann), t)
    if robin:
        Kr, Fr = robin_terms(n, mesh, robin)
        K = K + Kr
        F += Fr
    return K, F


def transient_heat_tet10(mesh, k=1.0, rho_cp=1.0, f=0.0, bc=None,
                         neumann=None, robin=None, t_end=1.0, dt=0.1,
                         T0=0.0, method="cn"):
    """Theta method, one LU, per-step time-dependent Neumann
    (tet10.py:467-548)."""
    if method not in THETA:
        raise ValueError(f"unknown method {method!r}")
    th = THETA[method]
    n = mesh.n_nodes
    K, F = assemble_tet10(mesh, k, f)
    if robin:
        Kr, Fr = assemble_boundary_terms_tet10(mesh, robin=robin)
        K = K + Kr
        F = F + Fr
    spec_t = None
    if neumann:
        spec = neumann_spec(mesh, neumann)
        if any(td for _, _, td in spec):
            spec_t = spec
        else:
            F = F + neumann_load(n, spec, 0.0)
    C = assemble_mass_tet10(mesh, rho_cp)
    steps = int(np.ceil(t_end / dt - 1e-12))
    if steps < 1:
        raise ValueError("t_end must be positive")
    h = t_end / steps
    A_sys = C + th * h * K
    R_sys = C - (1.0 - th) * h * K
    fixed = dirichlet_values(mesh, bc)
    if fixed:
        A_bc, _ = apply_dirichlet(A_sys, np.zeros(n), mesh, bc)
        idx = np.array(sorted(fixed), dtype=int)
        val = np.array([fixed[i] for i in idx])
        colsF = A_sys.tocsr()[:, idx]
    else:
        A_bc, idx, val, colsF = A_sys, None, None, None
    solve = spla.factorized(A_bc.tocsc())
    T = (np.full(n, T0, dtype=float) if np.isscalar(T0)
         else np.asarray(T0, dtype=float).copy())
    if idx is not None:
        T[idx] = val
    times, hist = [0.0], [T.copy()]
    for s in range(steps):
        if spec_t is not None:
            Fs = F + ((1.0 - th) * neumann_load(n, spec_t, s * h)
                      + th * neumann_load(n, spec_t, (s + 1) * h))
            rhs = R_sys @ T + h * Fs
        else:
            rhs = R_sys @ T + h * F
        if idx is not None:
            rhs = rhs - colsF @ val
            rhs[idx] = val
        T = np.asarray(solve(rhs)).ravel()
        if idx is not None:
            T[idx] = val
        times.append(times[-1] + h)
        hist.append(T.copy())
    times[-1] = t_end
    return np.array(times), np.array(hist)
