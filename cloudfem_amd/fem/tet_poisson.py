"""P1 scalar assembly, Dirichlet conditions and steady solve on the GPU.

Drop-in for cloudfem/fem/tet_poisson.py: same signatures, same returned
types (scipy csr_matrix with int32 indices, FP64 data, explicit zeros kept;
NumPy vectors).  The arithmetic runs in libcloudfem_hip.
"""

import warnings

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native as N
from ..coef import centroid_points, dirichlet_scalar, eval_scalar
from ..device import DeviceMesh, scalar_csr, to_dev
from ..solver import DeviceMatrix, pcg
from . import _dirichlet

_GREF = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def _tet_gradients(p):
    """P1 gradients (4, 3) and volume of one tetrahedron (tet_poisson.py:17-32).

    Single-element helper kept for API compatibility; assembly computes the
    same quantities on the GPU (cf_tet_geometry).
    """
    p = np.asarray(p, dtype=float)
    jac = np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1)
    vol = abs(np.linalg.det(jac)) / 6.0
    return np.linalg.solve(jac.T, _GREF.T).T, vol


def _assemble_device(mesh, k=1.0, f=0.0, dm=None):
    """(DeviceMesh, stiffness values on the node pattern, rhs) on the GPU."""
    dm = dm or DeviceMesh(mesh)
    cen = centroid_points(mesh)
    kv = to_dev(eval_scalar(k, cen), torch.float64, dm.dev)
    fv = to_dev(eval_scalar(f, cen), torch.float64, dm.dev)
    vals = dm.matrix_values(N.OP_P1_STIFF, coef=kv)
    rhs = torch.zeros(dm.n_nodes, dtype=torch.float64, device=dm.dev)
    dm.add_rhs(N.RHS_P1_SOURCE, fv, rhs)
    return dm, vals, rhs


def assemble(mesh, k=1.0, f=0.0):
    """Stiffness K (csr_matrix (N, N)) and load vector (N,) (tet_poisson.py:35-69).

    K = sum_e k(x_c) V_e G_e G_e^T, rhs_i += f(x_c) V_e / 4, coefficients at
    element centroids.
    """
    dm, vals, rhs = _assemble_device(mesh, k, f)
    return scalar_csr(dm.pattern(), vals), rhs.cpu().numpy()


def apply_dirichlet(A, rhs, mesh, bc):
    """Identity rows / zero columns for Dirichlet nodes (tet_poisson.py:72-110).

    Returns (csr_matrix, rhs) with the reference's structure: every row
    stores every fixed column (explicit zeros), fixed rows are e_i.
    """
    mask, vals = dirichlet_scalar(mesh, bc)
    return _dirichlet.apply(A, rhs, mask, vals)


def _solve_scalar(dm, vals, rhs, mesh, bc):
    n = dm.n_nodes
    if not bc:
        # the reference's spsolve on the singular pure-Neumann stiffness
        warnings.warn("Matrix is exactly singular", sp.linalg.MatrixRankWarning, stacklevel=3)
        return np.full(n, np.nan)
    mask, fv = dirichlet_scalar(mesh, bc)
    x = to_dev(np.where(mask, fv, 0.0), torch.float64, dm.dev)
    fixed = to_dev(mask.astype(np.uint8), torch.uint8, dm.dev)
    A = DeviceMatrix(N.FMT_CSR, dm.pattern(), vals)
    pcg(A, rhs, x, fixed)
    return x.cpu().numpy()


def solve_tet_poisson(mesh, k=1.0, f=0.0, bc=None):
    """Solve -div(k grad u) = f with Dirichlet bc (tet_poisson.py:113-122)."""
    dm, vals, rhs = _assemble_device(mesh, k, f)
    return _solve_scalar(dm, vals, rhs, mesh, bc)
