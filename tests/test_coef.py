"""Host evaluation of user coefficients: every value must equal a per-point
call, as the reference makes one call per centroid / quadrature point / node
(tet_poisson.py:48-56, 85-93; elastic_tet10.py:176-186; elastic3d.py:204-218).
Reducing callables (which return one value for array inputs) and callables
whose array form differs from their scalar form must not be broadcast."""

import numpy as np

from cloudfem_amd import coef
from cloudfem_amd.mesh import box_mesh, box_mesh_tet10

# one value for array inputs, a different value per point
NORM = lambda x, y, z: 1.0 + np.linalg.norm([x - 0.5, y - 0.5, z - 0.5])  # noqa: E731
# shape (n,) for array inputs, but not the per-point value
DEMEAN = lambda x, y, z: 2.0 + np.asarray(x) - np.mean(x)  # noqa: E731
# vectorises correctly
SMOOTH = lambda x, y, z: 1.0 + x + 2.0 * y * z  # noqa: E731


def _per_point(fn, pts):
    return np.array([float(fn(*p)) for p in pts])


def test_reducing_callable_is_called_per_point():
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.5, 0.5, 0.5]])
    got = coef.eval_scalar(NORM, pts)
    np.testing.assert_array_equal(got, _per_point(NORM, pts))
    assert got[2] == 1.0 and got[0] > 1.8


def test_shape_preserving_but_wrong_vectorisation_falls_back():
    pts = np.random.default_rng(0).random((1000, 3))
    got = coef.eval_scalar(DEMEAN, pts)
    np.testing.assert_array_equal(got, _per_point(DEMEAN, pts))
    np.testing.assert_allclose(got, 2.0, rtol=1e-15)


def test_vectorised_callable_matches_per_point():
    pts = np.random.default_rng(1).random((5000, 3))
    got = coef.eval_scalar(SMOOTH, pts)
    np.testing.assert_allclose(got, _per_point(SMOOTH, pts), rtol=1e-15, atol=0)


def test_constant_callable_and_float():
    pts = np.random.default_rng(2).random((7, 3))
    np.testing.assert_array_equal(coef.eval_scalar(lambda x, y, z: 3.5, pts), np.full(7, 3.5))
    np.testing.assert_array_equal(coef.eval_scalar(2.0, pts), np.full(7, 2.0))
    assert coef.eval_scalar(NORM, np.zeros((0, 3))).shape == (0,)


def test_scalar_only_callable():
    def scalar_only(x, y, z):
        if np.ndim(x):
            raise TypeError("scalars only")
        return x * y + z
    pts = np.random.default_rng(3).random((50, 3))
    np.testing.assert_array_equal(coef.eval_scalar(scalar_only, pts), _per_point(scalar_only, pts))


def test_components_traction_and_body():
    pts = np.random.default_rng(4).random((200, 3))
    out = coef.eval_components((NORM, None, DEMEAN), pts)
    np.testing.assert_array_equal(out[:, 0], _per_point(NORM, pts))
    np.testing.assert_array_equal(out[:, 1], 0.0)
    np.testing.assert_array_equal(out[:, 2], _per_point(DEMEAN, pts))


def test_dirichlet_scalar_reducing_bc():
    m = box_mesh(3, 3, 3)
    mask, vals = coef.dirichlet_scalar(m, {"xmin": NORM, "xmax": 0.0})
    ids = np.asarray(m.boundaries["xmin"])
    ids = ids[~np.isin(ids, m.boundaries["xmax"])]
    np.testing.assert_array_equal(vals[ids], _per_point(NORM, m.nodes[ids]))
    assert mask[ids].all()


def test_dirichlet_vec3_reducing_bc():
    m = box_mesh_tet10(2, 2, 2)
    mask, vals = coef.dirichlet_vec3(m, {"zmax": (None, NORM, 0.0)})
    ids = np.asarray(m.boundaries["zmax"])
    np.testing.assert_array_equal(vals[3 * ids + 1], _per_point(NORM, m.nodes[ids]))
    assert not mask[3 * ids].any() and mask[3 * ids + 2].all()
    assert len(set(vals[3 * ids + 1])) > 1


def test_lazy_points_built_only_for_callables():
    """A constant needs only the point count: LazyPoints builds its points
    for callables only, and the values equal those at the built points."""
    import numpy as np
    from cloudfem_amd.coef import LazyPoints, eval_components, eval_scalar
    built = []
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0], [0.5, 0.5, 0.5]])

    def build():
        built.append(1)
        return pts
    lp = LazyPoints(3, build)
    assert np.array_equal(eval_scalar(2.5, lp), np.full(3, 2.5)) and not built
    assert np.array_equal(eval_components((1.0, None, 2.0), lp), eval_components((1.0, None, 2.0), pts))
    assert not built
    f = lambda x, y, z: x + 2 * y - z  # noqa: E731
    assert np.array_equal(eval_scalar(f, lp), eval_scalar(f, pts)) and len(built) == 1
    assert np.array_equal(eval_components((f, 0.0, f), lp), eval_components((f, 0.0, f), pts))
    assert len(built) == 1
