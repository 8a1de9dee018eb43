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
