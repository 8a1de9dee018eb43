"""Shared comparison helpers for the parity tests."""

import numpy as np
import scipy.sparse as sp

FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def assert_csr_parity(A, A0, rtol=1e-12):
    """Structure bit-exact (type, index dtype, indptr, indices) and values
    within rtol of each row's largest reference entry."""
    assert isinstance(A, sp.csr_matrix) and not isinstance(A, sp.csr_array)
    assert A.shape == A0.shape
    assert A.indices.dtype == A0.indices.dtype == np.int32
    assert A.indptr.dtype == A0.indptr.dtype == np.int32
    assert np.array_equal(A.indptr, A0.indptr), "indptr differs"
    assert np.array_equal(A.indices, A0.indices), "indices differ"
    diff = np.abs(A.data - A0.data)
    rows = np.repeat(np.arange(A0.shape[0]), np.diff(A0.indptr))
    scale = np.zeros(A0.shape[0])
    np.maximum.at(scale, rows, np.abs(A0.data))
    scale[scale == 0] = 1.0
    rel = diff / scale[rows] if diff.size else diff
    assert rel.size == 0 or rel.max() <= rtol, f"max row-relative value error {rel.max():.3e}"


def assert_vec_close(v, v0, rtol=1e-12):
    v, v0 = np.asarray(v), np.asarray(v0)
    assert v.shape == v0.shape
    scale = max(np.abs(v0).max(), 1e-300)
    err = np.abs(v - v0).max() / scale
    assert err <= rtol, f"relative error {err:.3e}"


def rel_err(u, u0):
    return float(np.linalg.norm(np.asarray(u) - np.asarray(u0)) / max(np.linalg.norm(u0), 1e-300))


def fp64_floor(Aff, x_f, b_f):
    """eps ||A_ff| |x_f|| / ||b_f||: the smallest relative residual FP64
    arithmetic resolves for this system (rounding of A_ff x_f)."""
    return float(np.finfo(np.float64).eps * np.linalg.norm(abs(Aff) @ np.abs(x_f))
                 / np.linalg.norm(b_f))


def assert_residual_bar(st, tol, floor):
    """The true FP64 residual a solver reports meets tol, or the solver says
    it stopped at the FP64 floor and the residual is within 3x that floor."""
    r = st["rel_residual"]
    if r <= tol:
        return
    assert st.get("floor_limited"), f"residual {r:.3e} > tol {tol:.1e} without floor_limited"
    assert r <= 3.0 * floor, f"residual {r:.3e} > 3 x FP64 floor {floor:.3e}"


def node_pattern(elements, n_nodes):
    """Vectorised CSR pattern of the node graph of the elements: row i lists
    every node that shares an element with node i, sorted and unique, as
    scipy's coo_matrix((data, (rows, cols))).tocsr() stores the element
    triplets (duplicates merged, explicit zeros kept).  Returns int32
    (indptr, indices)."""
    e = np.asarray(elements, dtype=np.int64)
    k = e.shape[1]
    rows = np.repeat(e, k, axis=1).ravel()
    cols = np.tile(e, (1, k)).ravel()
    key = np.unique(rows * n_nodes + cols)
    r, c = np.divmod(key, n_nodes)
    indptr = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n_nodes), out=indptr[1:])
    return indptr.astype(np.int32), c.astype(np.int32)


def dof_pattern3(indptr, indices):
    """DOF-level pattern of a 3-DOF-per-node matrix whose node blocks are
    dense (DOF 3i+c): row 3i+c lists 3j+d for every j in node row i and
    d = 0, 1, 2.  Returns int32 (indptr, indices)."""
    indptr = np.asarray(indptr, dtype=np.int64)
    lens = np.diff(indptr)
    rowlen = np.repeat(3 * lens, 3)
    dptr = np.zeros(rowlen.size + 1, dtype=np.int64)
    np.cumsum(rowlen, out=dptr[1:])
    expanded = (3 * np.asarray(indices, dtype=np.int64)[:, None] + np.arange(3)).ravel()
    drow = np.repeat(np.arange(rowlen.size, dtype=np.int64), rowlen)
    pos = np.arange(dptr[-1], dtype=np.int64) - dptr[drow]
    src = 3 * indptr[drow // 3] + pos
    del drow, pos
    return dptr.astype(np.int32), expanded[src].astype(np.int32)


def sampled_row_parity(A, mesh, node_ids, element_matrix, rtol=1e-12, ndof=3):
    """Rows of `A` for the DOFs of `node_ids` against the oracle assembled
    over only the elements that touch those nodes (those rows are complete
    there).  element_matrix(submesh) -> oracle CSR of the same shape."""
    node_ids = np.asarray(node_ids)
    el = np.asarray(mesh.elements)
    touch = np.isin(el, node_ids).any(axis=1)
    sub = type(mesh).__new__(type(mesh))
    sub.__dict__.update(mesh.__dict__)
    sub.elements = el[touch]
    A0 = element_matrix(sub)
    rows = (ndof * node_ids[:, None] + np.arange(ndof)).ravel()
    for r in rows:
        a = A.indices[A.indptr[r]:A.indptr[r + 1]], A.data[A.indptr[r]:A.indptr[r + 1]]
        b = A0.indices[A0.indptr[r]:A0.indptr[r + 1]], A0.data[A0.indptr[r]:A0.indptr[r + 1]]
        assert np.array_equal(a[0], b[0]), f"row {r}: columns differ"
        scale = max(np.abs(b[1]).max(), 1e-300)
        err = np.abs(a[1] - b[1]).max() / scale
        assert err <= rtol, f"row {r}: row-relative value error {err:.3e}"
    return int(touch.sum())
