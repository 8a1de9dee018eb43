"""P1 vector mass matrices on the GPU.  Drop-in for the mass half of
cloudfem/physics/modal3d.py (the eigen-solver is outside this path)."""

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native as N
from ..device import DeviceMesh, to_dev, vector_csr


def assemble_mass_3d(mesh, rho, lumped=False):
    """(3N, 3N) mass: consistent rho V (1+d_ab)/20 on each component block,
    or lumped rho V / 4 on the diagonal (modal3d.py:40-75)."""
    dm = DeviceMesh(mesh)
    if dm.k != 4:
        raise ValueError("assemble_mass_3d needs a 4-node mesh")
    if lumped:
        diag = torch.zeros(3 * dm.n_nodes, dtype=torch.float64, device=dm.dev)
        rv = to_dev(np.full(3 * dm.n_elem, float(rho)), torch.float64, dm.dev)
        dm.add_rhs(N.RHS_P1_BODY, rv, diag)
        return sp.diags(diag.cpu().numpy(), format="csr")
    coef = to_dev(np.full(dm.n_elem, float(rho)), torch.float64, dm.dev)
    vals = dm.matrix_values(N.OP_P1_VMASS, coef=coef)
    return vector_csr(dm.pattern(), vals, dense3=False)
