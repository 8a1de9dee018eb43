"""Host evaluation of the reference's float-or-callable coefficients.

The reference calls user callables f(x, y, z) once per centroid, quadrature
point or node (tet_poisson.py:48-56, tet10.py:152-153,
elastic_tet10.py:176-186).  Python callables cannot run on the GPU, so they
are evaluated here into arrays that are then uploaded.  A callable is first
tried once on whole coordinate arrays; if that raises or returns the wrong
shape it is called point by point, exactly as the reference does.
"""

import numpy as np


def eval_scalar(v, pts):
    """Values of a float / callable(x, y, z) at points (P, 3) -> (P,)."""
    pts = np.asarray(pts, dtype=float)
    n = pts.shape[0]
    if not callable(v):
        return np.full(n, float(v))
    if n == 0:
        return np.zeros(0)
    try:
        out = np.asarray(v(pts[:, 0], pts[:, 1], pts[:, 2]), dtype=float)
        if out.shape == ():
            return np.full(n, float(out))
        if out.shape == (n,):
            return out
    except Exception:  # noqa: BLE001 - any failure means "not vectorisable"
        pass
    return np.array([float(v(x, y, z)) for x, y, z in pts], dtype=float)


def eval_components(vals, pts):
    """3-component spec (each float / callable / None) -> (P, 3)
    (elastic3d.py:88-102)."""
    pts = np.asarray(pts, dtype=float)
    out = np.zeros((pts.shape[0], 3))
    for c, v in enumerate(vals):
        if v is not None:
            out[:, c] = eval_scalar(v, pts)
    return out


def centroids(mesh):
    """Corner centroids (M, 3) as p[:, i].mean() (tet_poisson.py:54)."""
    return np.asarray(mesh.nodes)[np.asarray(mesh.elements)[:, :4]].mean(axis=1)


def dirichlet_scalar(mesh, bc):
    """(mask (N,) bool, values (N,)) of a scalar bc dict; later names
    override earlier ones, as the reference's dict does
    (tet_poisson.py:85-93)."""
    n = mesh.n_nodes
    mask = np.zeros(n, dtype=bool)
    vals = np.zeros(n)
    for name, value in (bc or {}).items():
        ids = np.arange(n) if name == "all" else np.asarray(mesh.boundaries[name], dtype=int)
        if ids.size == 0:
            continue
        vals[ids] = eval_scalar(value, np.asarray(mesh.nodes)[ids])
        mask[ids] = True
    return mask, vals


def dirichlet_vec3(mesh, bc):
    """(mask (3N,), values (3N,)) of a vector bc dict, None = free
    (elastic3d.py:204-218)."""
    n = mesh.n_nodes
    mask = np.zeros(3 * n, dtype=bool)
    vals = np.zeros(3 * n)
    for name, comps in (bc or {}).items():
        ids = np.arange(n) if name == "all" else np.asarray(mesh.boundaries[name], dtype=int)
        if ids.size == 0:
            continue
        pts = np.asarray(mesh.nodes)[ids]
        for c, v in enumerate(comps):
            if v is None:
                continue
            vals[3 * ids + c] = eval_scalar(v, pts)
            mask[3 * ids + c] = True
    return mask, vals
