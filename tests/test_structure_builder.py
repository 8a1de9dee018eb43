"""The vectorised pattern builders the large-mesh GPU parity tests compare
against, pinned here by brute force and by scipy's coo->csr merge of the
oracle's element triplets on small meshes."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import elastic as oel
from oracle import meshes as om
from tests._helpers import dof_pattern3, node_pattern, sampled_row_parity


def _brute_pattern(elements, n):
    rows = [set() for _ in range(n)]
    for el in np.asarray(elements):
        for a in el:
            for b in el:
                rows[int(a)].add(int(b))
    indptr = np.cumsum([0] + [len(r) for r in rows])
    indices = np.concatenate([sorted(r) for r in rows])
    return indptr, indices


@pytest.mark.parametrize("builder,dims", [("box_mesh", (3, 2, 2)), ("box_mesh_tet10", (2, 2, 3)),
                                          ("box_mesh_tet10", (4, 1, 1))])
def test_node_pattern_brute_force(builder, dims):
    m = getattr(om, builder)(*dims)
    ptr, idx = node_pattern(m.elements, m.n_nodes)
    bptr, bidx = _brute_pattern(m.elements, m.n_nodes)
    assert ptr.dtype == idx.dtype == np.int32
    assert np.array_equal(ptr, bptr) and np.array_equal(idx, bidx)


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 1)])
def test_dof_pattern_matches_oracle_elasticity(dims):
    """The DOF pattern of the oracle's Tet10 elasticity matrix (scipy coo->csr
    of the 30x30 element triplets) equals the expanded node pattern."""
    m = om.box_mesh_tet10(*dims)
    A0, _ = oel.assemble_elasticity_tet10(m, 210e9, 0.3)
    ptr, idx = dof_pattern3(*node_pattern(m.elements, m.n_nodes))
    assert np.array_equal(ptr, A0.indptr) and np.array_equal(idx, A0.indices)
    # and against a direct coo->csr of the expanded node pairs
    e = np.asarray(m.elements)
    d = (3 * e[:, :, None] + np.arange(3)).reshape(len(e), -1)
    r = np.repeat(d, 30, axis=1).ravel()
    c = np.tile(d, (1, 30)).ravel()
    B = sp.coo_matrix((np.ones(r.size), (r, c)), shape=A0.shape).tocsr()
    assert np.array_equal(ptr, B.indptr) and np.array_equal(idx, B.indices)


def test_sampled_rows_catch_a_wrong_value():
    """The sampled-row check passes on the oracle's own matrix and fails when
    one value of a sampled row is perturbed."""
    m = om.box_mesh_tet10(3, 2, 2)
    build = lambda sub: oel.assemble_elasticity_tet10(sub, 210e9, 0.3)[0]  # noqa: E731
    A0 = build(m)
    ids = np.array([0, 17, m.n_nodes - 1])
    assert sampled_row_parity(A0, m, ids, build) > 0
    A1 = A0.copy()
    A1.data[A1.indptr[3 * 17 + 1] + 2] *= 1.0 + 1e-9
    with pytest.raises(AssertionError):
        sampled_row_parity(A1, m, ids, build)
