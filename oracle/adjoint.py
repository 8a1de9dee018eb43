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
