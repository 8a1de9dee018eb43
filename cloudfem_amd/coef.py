"""Host evaluation of the reference's float-or-callable coefficients.

The reference calls user callables f(x, y, z) once per centroid, quadrature
point or node, with numpy float64 scalars (tet_poisson.py:48-56,
tet10.py:152-153, elastic_tet10.py:176-186).  Python callables cannot run on
the GPU, so they are evaluated here into arrays that are then uploaded.

A callable is first tried once on whole coordinate arrays, which is only a
shortcut: a vectorised result is kept only if it has one value per point and
agrees with per-point calls at sampled points (first, last and a spread of
others).  A callable that returns a single value for array inputs is called
point by point, since it may reduce over its inputs (``np.linalg.norm([x, y,
z])`` does).  Any other outcome (an exception, a wrong shape, a mismatch at a
sampled point) falls back to one call per point, which is exactly what the
reference does.
"""

import numpy as np

# Sampled points compared with per-point calls before a vectorised result is
# trusted; relative tolerance of that comparison (array math may round the
# last bit differently from scalar math).
_N_CHECK = 16
_CHECK_RTOL = 1e-13


def _per_point(v, pts):
    return np.array([float(v(x, y, z)) for x, y, z in pts], dtype=float)


def _sample_ids(n):
    if n <= _N_CHECK:
        return np.arange(n)
    rng = np.random.default_rng(n)
    mid = rng.choice(np.arange(1, n - 1), size=_N_CHECK - 2, replace=False)
    return np.unique(np.concatenate([[0, n - 1], mid]))


class LazyPoints:
    """Evaluation points built on first use: a constant coefficient needs
    only their count (384,000 centroids take ~30 ms of host time)."""

    def __init__(self, n, build):
        self.n, self._build, self._pts = int(n), build, None

    def get(self):
        if self._pts is None:
            self._pts = np.asarray(self._build(), dtype=float)
        return self._pts


def eval_scalar(v, pts):
    """Values of a float / callable(x, y, z) at points (P, 3) (or
    LazyPoints) -> (P,)."""
    if isinstance(pts, LazyPoints):
        if not callable(v):
            return np.full(pts.n, float(v))
        pts = pts.get()
    pts = np.asarray(pts, dtype=float)
    n = pts.shape[0]
    if not callable(v):
        return np.full(n, float(v))
    if n == 0:
        return np.zeros(0)
    try:
        out = np.asarray(v(pts[:, 0], pts[:, 1], pts[:, 2]), dtype=float)
    except Exception:  # noqa: BLE001 - any failure means "not vectorisable"
        out = None
    if out is None or out.shape != (n,):
        return _per_point(v, pts)
    ids = _sample_ids(n)
    ref = _per_point(v, pts[ids])
    tol = _CHECK_RTOL * np.maximum(np.abs(ref), np.abs(out[ids])) + 1e-300
    if not np.all(np.abs(out[ids] - ref) <= tol):
        return _per_point(v, pts)
    return out


def eval_components(vals, pts):
    """3-component spec (each float / callable / None) at points (P, 3) (or
    LazyPoints) -> (P, 3) (elastic3d.py:88-102)."""
    if not isinstance(pts, LazyPoints):
        pts = np.asarray(pts, dtype=float)
    out = np.zeros((pts.n if isinstance(pts, LazyPoints) else pts.shape[0], 3))
    for c, v in enumerate(vals):
        if v is not None:
            out[:, c] = eval_scalar(v, pts)
    return out


def centroids(mesh):
    """Corner centroids (M, 3) as p[:, i].mean() (tet_poisson.py:54)."""
    return np.asarray(mesh.nodes)[np.asarray(mesh.elements)[:, :4]].mean(axis=1)


def centroid_points(mesh):
    """The centroids as LazyPoints."""
    return LazyPoints(np.asarray(mesh.elements).shape[0], lambda: centroids(mesh))


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
