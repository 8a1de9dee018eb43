"""Oracle for P1 scalar assembly, Dirichlet, steady and transient solves.

Restates cloudfem/fem/tet_poisson.py:17-122 and
cloudfem/fem/tet_heat.py:72-274 (CPU, test-only).
"""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_GREF = np.array([[-1.0, -1.0, -1.0],
                  [1.0, 0.0, 0.0],
                  [0.0, 1.0, 0.0],
                  [0.0, 0.0, 1.0]])


def as_fn(v):
    """float-or-callable(x, y, z) -> callable (tet_heat.py:131-133)."""
    return v if callable(v) else (lambda x, y, z, _v=v: _v)


def tet_gradients(p):
    """P1 gradients + volume of one tet (tet_poisson.py:17-32).

    J has columns p1-p0, p2-p0, p3-p0; V = |det J|/6; the gradient rows
    G solve J^T G^T = g_ref^T.
    """
    jac = np.column_stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]])
    vol = abs(np.linalg.det(jac)) / 6.0
    grads = np.linalg.solve(jac.T, _GREF.T).T
    return grads, vol


def assemble(mesh, k=1.0, f=0.0):
    """K = sum_e k(x_c) V G G^T, rhs_i += f(x_c) V / 4 (tet_poisson.py:35-69)."""
    n = mesh.n_nodes
    kf, ff = as_fn(k), as_fn(f)
    ri, ci, vals = [], [], []
    rhs = np.zeros(n)
    for el in mesh.elements:
        p = mesh.nodes[el]
        g, vol = tet_gradients(p)
        xc, yc, zc = p.mean(axis=0)
        ke = kf(xc, yc, zc) * vol * (g @ g.T)
        fe = ff(xc, yc, zc)
        for a in range(4):
            rhs[el[a]] += fe * vol / 4.0
        ri.extend(np.repeat(el, 4))
        ci.extend(np.tile(el, 4))
        vals.extend(ke.ravel())
    K = sp.coo_matrix((vals, (ri, ci)), shape=(n, n)).tocsr()
    return K, rhs


def dirichlet_values(mesh, bc):
    """{node: value} from a bc dict (tet_heat.py:175-190)."""
    fixed = {}
    if not bc:
        return fixed
    for name, value in bc.items():
        ids = np.arange(mesh.n_nodes) if name == "all" else mesh.boundaries[name]
        fn = as_fn(value)
        for i in ids:
            i = int(i)
            fixed[i] = fn(*mesh.nodes[i])
    return fixed


def apply_dirichlet(A, rhs, mesh, bc):
    """Identity rows + column elimination (tet_poisson.py:72-110).

    Fixed rows become e_i; rhs -= A[:, F] v (after the row change);
    A[:, F] = 0 stores an explicit zero in every row of each fixed
    column; then A_ii = 1 for fixed i.
    """
    import warnings
    A = A.tolil()
    rhs = rhs.copy()
    fixed = dirichlet_values(mesh, bc)
    for i, v in fixed.items():
        A.rows[i] = [i]
        A.data[i] = [1.0]
        rhs[i] = v
    A = A.tocsr()
    idx = np.array(sorted(fixed), dtype=int)
    if len(idx):
        v = np.array([fixed[i] for i in idx])
        rhs = rhs - A[:, idx] @ v
        for i, vv in fixed.items():
            rhs[i] = vv
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", sp.SparseEfficiencyWarning)
            A[:, idx] = 0
            A = A.tocsr()
            for i in fixed:
                A[i, i] = 1.0
    return A, rhs


def solve_tet_poisson(mesh, k=1.0, f=0.0, bc=None):
    """assemble -> apply_dirichlet -> spsolve (tet_poisson.py:113-122)."""
    A, rhs = assemble(mesh, k, f)
    if bc:
        A, rhs = apply_dirichlet(A, rhs, mesh, bc)
    return np.asarray(spla.spsolve(A, rhs)).ravel()


def assemble_mass(mesh, rho_cp=1.0):
    """C = sum_e rho_cp(x_c) V (1 + d_ab)/20 (tet_heat.py:72-94)."""
    n = mesh.n_nodes
    rf = as_fn(rho_cp)
    base = (np.ones((4, 4)) + np.eye(4)) / 20.0
    vols = mesh.element_volumes()
    ri, ci, vals = [], [], []
    for e, el in enumerate(mesh.elements):
        xc, yc, zc = mesh.nodes[el].mean(axis=0)
        me = rf(xc, yc, zc) * vols[e] * base
        ri.extend(np.repeat(el, 4))
        ci.extend(np.tile(el, 4))
        vals.extend(me.ravel())
    return sp.coo_matrix((vals, (ri, ci)), shape=(n, n)).tocsr()


def boundary_faces(mesh, name):
    """Exterior faces of a named set, first-seen order (tet_heat.py:97-120).

    Faces (a,b,c), (a,b,d), (a,c,d), (b,c,d) of each element; a face is
    exterior iff its sorted triple occurs once; kept iff all three nodes
    are in the set.  The stored orientation is the first occurrence.
    """
    members = set(int(i) for i in mesh.boundaries[name])
    seen = {}
    count = {}
    for el in mesh.elements:
        a, b, c, d = (int(x) for x in el[:4])
        for tri in ((a, b, c), (a, b, d), (a, c, d), (b, c, d)):
            key = tuple(sorted(tri))
            count[key] = count.get(key, 0) + 1
            seen.setdefault(key, tri)
    out = [seen[k] for k, cnt in count.items()
           if cnt == 1 and all(v in members for v in k)]
    return np.array(out, dtype=int).reshape(-1, 3)


def _face_geom(mesh, tri):
    p = mesh.nodes[tri]
    return p.mean(axis=0), 0.5 * float(np.linalg.norm(np.cross(p[1] - p[0],
                                                             p[2] - p[0])))


def assemble_boundary_terms(mesh, neumann=None, robin=None):
    """Neumann q A/3 and Robin h A (1+d)/12, h T_inf A/3 (tet_heat.py:136-172)."""
    n = mesh.n_nodes
    F = np.zeros(n)
    ri, ci, vals = [], [], []
    fm = (np.ones((3, 3)) + np.eye(3)) / 12.0
    for name, q in (neumann or {}).items():
        qf = as_fn(q)
        for tri in boundary_faces(mesh, name):
            cen, area = _face_geom(mesh, tri)
            F[tri] += qf(*cen) * area / 3.0
    for name, (h, tinf) in (robin or {}).items():
        hf, tf = as_fn(h), as_fn(tinf)
        for tri in boundary_faces(mesh, name):
            cen, area = _face_geom(mesh, tri)
            hv = hf(*cen)
            ke = hv * area * fm
            ri.extend(np.repeat(tri, 3))
            ci.extend(np.tile(tri, 3))
            vals.extend(ke.ravel())
            F[tri] += hv * tf(*cen) * area / 3.0
    K = sp.coo_matrix((vals, (ri, ci)), shape=(n, n)).tocsr()
    return K, F


THETA = {"be": 1.0, "implicit_euler": 1.0, "cn": 0.5, "crank_nicolson": 0.5}


def transient_heat(mesh, k=1.0, rho_cp=1.0, f=0.0, bc=None, neumann=None,
                   robin=None, t_end=1.0, dt=0.1, T0=0.0, method="cn"):
    """Theta method with one LU (tet_heat.py:197-274).

    (C + th h K) T+ = (C - (1-th) h K) T + h F, h = t_end/ceil(t_end/dt),
    Dirichlet by the apply_dirichlet elimination of the system matrix and
    rhs - A_sys[:, F] v, rhs_F = v each step.
    """
    if method not in THETA:
        raise ValueError(f"unknown method {method!r}")
    th = THETA[method]
    n = mesh.n_nodes
    K, F = assemble(mesh, k, f)
    Kb, Fb = assemble_boundary_terms(mesh, neumann, robin)
    K = K + Kb
    F = F + Fb
    C = assemble_mass(mesh, rho_cp)
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
    for _ in range(steps):
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
