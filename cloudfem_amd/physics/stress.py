"""Element strain/stress post-processing on the GPU (cf_element_stress)."""

import numpy as np
import torch

from .. import _native as N
from ..coef import centroids, eval_scalar
from ..device import DeviceMesh, ptr, stream_handle, to_dev
from .elastic3d import lame

# Tet10 shape functions at the centroid: corners -1/8, midpoints 1/4
# (elastic_tet10.py:255-259 with L = (1/4, 1/4, 1/4, 1/4)).
_NC10 = np.array([-0.125] * 4 + [0.25] * 6)


def as_vector(u, n):
    """(3N,) interleaved displacement from a vector or a (ux, uy, uz) tuple."""
    if isinstance(u, tuple):
        uv = np.empty(3 * n)
        uv[0::3], uv[1::3], uv[2::3] = u
        return uv
    return np.asarray(u, dtype=float).ravel()


def element_stress(mesh, u, E, nu, T=None, alpha=0.0, T_ref=0.0, quadratic=False, dm=None):
    """(sig, eps), each (M, 6), at element centroids."""
    dm = dm or DeviceMesh(mesh)
    if quadratic and dm.k != 10:
        raise ValueError("element_stress_tet10 needs a 10-node mesh")
    lam, mu = lame(E, nu)
    uv = to_dev(as_vector(u, dm.n_nodes), torch.float64, dm.dev)
    s_th = None
    if T is not None and alpha != 0.0:
        if callable(T):
            Tc = eval_scalar(T, centroids(mesh))
        else:
            Tn = np.asarray(T, dtype=float)[np.asarray(mesh.elements)]
            Tc = Tn @ _NC10 if dm.k == 10 else Tn.mean(axis=1)
        s_th = to_dev(alpha * (Tc - T_ref), torch.float64, dm.dev)
    m = dm.n_elem
    sig = torch.empty((max(m, 1), 6), dtype=torch.float64, device=dm.dev)
    eps = torch.empty((max(m, 1), 6), dtype=torch.float64, device=dm.dev)
    N.check(N.load().cf_element_stress(ptr(dm.conn), m, dm.k, ptr(dm.geometry()), ptr(uv),
                                       ptr(s_th), lam, mu, ptr(sig), ptr(eps),
                                       stream_handle()), "cf_element_stress")
    return sig[:m].cpu().numpy(), eps[:m].cpu().numpy()
