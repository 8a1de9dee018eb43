"""Oracle for the P1 steady-heat adjoint sensitivity (CPU, test-only).

No reference counterpart: CloudFEM has no adjoint.  The forward problem is
the reference's steady P1 solve (tet_poisson.py:35-122) with a per-element
conductivity fed through the reference's own convention for element data,
a nearest-centroid lookup (induction3d.py:111-123).  The adjoint is the
textbook one: K lambda = S^T (T_S - T_target) on the free nodes,
lambda_F = 0, dJ/dk_e = -V_e grad(lambda_e) . grad(T_e).
"""

import numpy as np
import scipy.sparse.linalg as spla
from scipy.spatial import cKDTree

from .p1 import assemble, dirichlet_values, tet_gradients


def centroid_lookup(centroids, values):
    """callable(x, y, z) -> value of the nearest centroid (induction3d.py:111-123)."""
    tree = cKDTree(np.asarray(centroids, dtype=float))
    values = np.asarray(values, dtype=float)

    def fn(x, y, z):
        _, i = tree.query([x, y, z])
        return float(values[i])
    return fn


def _reduced_solve(K, rhs, fixed_idx, fixed_val, n):
    free = np.ones(n, dtype=bool)
    free[fixed_idx] = False
    u = np.zeros(n)
    u[fixed_idx] = fixed_val
    K = K.tocsr()
    u[free] = spla.spsolve(K[free][:, free].tocsc(), rhs[free] - K[free][:, fixed_idx] @ fixed_val)
    return u, free


def forward(mesh, kvals, f, bc):
    cen = mesh.nodes[mesh.elements].mean(axis=1)
    K, F = assemble(mesh, k=centroid_lookup(cen, kvals), f=f)
    fixed = dirichlet_values(mesh, bc)
    idx = np.array(sorted(fixed), dtype=int)
    val = np.array([fixed[i] for i in idx])
    T, free = _reduced_solve(K, F, idx, val, mesh.n_nodes)
    return T, K, idx


def objective(T, sensors, target):
    r = T[sensors] - target
    return 0.5 * float(r @ r)


def gradient(mesh, kvals, f, bc, sensors, target):
    """(J, dJ/dk (M,)) by the FP64 adjoint."""
    T, K, idx = forward(mesh, kvals, f, bc)
    g = np.zeros(mesh.n_nodes)
    np.add.at(g, sensors, T[sensors] - target)
    lam, _ = _reduced_solve(K, g, idx, np.zeros(len(idx)), mesh.n_nodes)
    dk = np.empty(mesh.n_elements)
    for e, el in enumerate(mesh.elements):
        G, vol = tet_gradients(mesh.nodes[el])
        dk[e] = -vol * float((lam[el] @ G) @ (T[el] @ G))
    return objective(T, sensors, target), dk


# ---- the same adjoint, vectorised over elements (for the 40^3 workload,
# where the per-element loops above and SuperLU's fill on a 3D mesh take
# minutes).  Same formulas: G_e = g_ref J_e^-1 and V_e = |det J_e| / 6
# (tet_poisson.py:17-32), K_e = k_e V_e G_e G_e^T, rhs_i += f V_e / 4
# (tet_poisson.py:35-69); the two reduced solves use SciPy's Jacobi CG to a
# relative residual of 1e-13 (a library routine independent of the GPU
# solver).  Pinned against `gradient` on small meshes (tests/test_oracle_p1.py).

_G_REF = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def _element_grads(mesh):
    p = np.asarray(mesh.nodes, dtype=float)[np.asarray(mesh.elements)[:, :4]]
    J = np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=2)
    vol = np.abs(np.linalg.det(J)) / 6.0
    G = np.einsum("ab,mbc->mac", _G_REF, np.linalg.inv(J))
    return G, vol


def _reduced_solve_cg(K, rhs, fixed_idx, fixed_val, n):
    import scipy.sparse as sp
    free = np.ones(n, dtype=bool)
    free[fixed_idx] = False
    u = np.zeros(n)
    u[fixed_idx] = fixed_val
    K = K.tocsr()
    Kf = K[free][:, free].tocsr()
    b = rhs[free] - K[free][:, fixed_idx] @ fixed_val
    M = sp.diags(1.0 / Kf.diagonal())
    x, info = spla.cg(Kf, b, rtol=1e-13, atol=0.0, maxiter=100000, M=M)
    if info != 0:
        raise RuntimeError(f"oracle CG did not converge (info {info})")
    u[free] = x
    return u, free


def forward_vec(mesh, kvals, f, bc):
    import scipy.sparse as sp
    els = np.asarray(mesh.elements)[:, :4]
    n = mesh.n_nodes
    G, vol = _element_grads(mesh)
    Ke = (np.asarray(kvals, dtype=float) * vol)[:, None, None] * np.einsum("mac,mbc->mab", G, G)
    rows = np.repeat(els, 4, axis=1).ravel()
    cols = np.tile(els, (1, 4)).ravel()
    K = sp.coo_matrix((Ke.ravel(), (rows, cols)), shape=(n, n)).tocsr()
    F = np.zeros(n)
    np.add.at(F, els.ravel(), np.repeat(f * vol / 4.0, 4))
    fixed = dirichlet_values(mesh, bc)
    idx = np.array(sorted(fixed), dtype=int)
    val = np.array([fixed[i] for i in idx])
    T, _ = _reduced_solve_cg(K, F, idx, val, n)
    return T, K, idx, G, vol


def gradient_vec(mesh, kvals, f, bc, sensors, target):
    """(J, dJ/dk (M,)) by the FP64 adjoint, vectorised over elements."""
    T, K, idx, G, vol = forward_vec(mesh, kvals, f, bc)
    g = np.zeros(mesh.n_nodes)
    np.add.at(g, sensors, T[sensors] - target)
    lam, _ = _reduced_solve_cg(K, g, idx, np.zeros(len(idx)), mesh.n_nodes)
    els = np.asarray(mesh.elements)[:, :4]
    gl = np.einsum("ma,mac->mc", lam[els], G)
    gt = np.einsum("ma,mac->mc", T[els], G)
    return objective(T, sensors, target), -vol * np.einsum("mc,mc->m", gl, gt)
