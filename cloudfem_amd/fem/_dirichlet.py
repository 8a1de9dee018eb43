"""Scalar Dirichlet elimination on the GPU (tet_poisson.py:72-110 semantics)."""

import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native as N
from ..device import ptr, require_gpu, stream_handle, to_dev, workspace


def apply(A, rhs, mask, fixed_vals):
    """(csr_matrix, rhs) with identity fixed rows and every fixed column
    stored in every row (explicit zeros), exactly as the reference's
    lil/csr assignments leave them."""
    dev = require_gpu()
    lib = N.load()
    A = sp.csr_matrix(A)
    if not A.has_canonical_format:
        A = A.copy()
        A.sum_duplicates()
    n = A.shape[0]
    rhs = np.asarray(rhs, dtype=float)
    flist = np.nonzero(mask)[0].astype(np.int32)
    d_ptr = to_dev(A.indptr.astype(np.int32), torch.int32, dev)
    d_idx = to_dev(A.indices.astype(np.int32), torch.int32, dev)
    d_val = to_dev(A.data.astype(np.float64), torch.float64, dev)
    d_rhs = to_dev(rhs, torch.float64, dev)
    d_fix = to_dev(mask.astype(np.uint8), torch.uint8, dev)
    d_fval = to_dev(np.asarray(fixed_vals, dtype=np.float64), torch.float64, dev)
    d_flist = to_dev(flist if flist.size else np.zeros(1, np.int32), torch.int32, dev)
    nb = ctypes.c_size_t()
    N.check(lib.cf_dirichlet_workspace_bytes(n, ctypes.byref(nb)), "cf_dirichlet_workspace_bytes")
    ws = workspace(nb.value, dev)
    ptr_out = torch.empty(n + 1, dtype=torch.int32, device=dev)
    rhs_out = torch.empty(n, dtype=torch.float64, device=dev)
    nnz_dev = torch.zeros(1, dtype=torch.int64, device=dev)
    s = stream_handle()
    N.check(lib.cf_dirichlet_count(n, ptr(d_ptr), ptr(d_idx), ptr(d_val), ptr(d_rhs),
                                   ptr(d_fix), ptr(d_fval), int(flist.size), ptr(ws),
                                   nb.value, ptr(ptr_out), ptr(rhs_out), ptr(nnz_dev), s),
            "cf_dirichlet_count")
    nnz = int(nnz_dev.item())
    idx_out = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    val_out = torch.empty(max(nnz, 1), dtype=torch.float64, device=dev)
    N.check(lib.cf_dirichlet_fill(n, ptr(d_ptr), ptr(d_idx), ptr(d_val), ptr(d_fix),
                                  ptr(d_flist), int(flist.size), ptr(ptr_out), ptr(idx_out),
                                  ptr(val_out), s), "cf_dirichlet_fill")
    out = sp.csr_matrix((val_out[:nnz].cpu().numpy(), idx_out[:nnz].cpu().numpy(),
                         ptr_out.cpu().numpy()), shape=A.shape)
    return out, rhs_out.cpu().numpy()
