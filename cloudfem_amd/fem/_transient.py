"""Theta-method time stepping on the GPU (shared by P1 and Tet10 heat).

Replaces the LU-once loop of tet_heat.py:234-274 / tet10.py:504-548:
    (C + th h K) T+ = (C - (1-th) h K) T + h F,
with the Dirichlet elimination done by solving the free-DOF system (the
same system the reference's apply_dirichlet + per-step rhs - A[:, F] v
produces).  Each step is one SpMV for the right-hand side and one PCG
solve that starts from the extrapolation 2 T_n - T_{n-1} of the last two
states; the history stays in HBM until the end.
"""

import numpy as np
import torch

from .. import _native as N
from ..device import ptr, stream_handle, to_dev
from ..solver import DeviceMatrix, pcg

THETA = {"be": 1.0, "implicit_euler": 1.0, "cn": 0.5, "crank_nicolson": 0.5}


def check_method(method):
    if method not in THETA:
        raise ValueError(f"unknown method {method!r}; use one of {sorted(THETA)}")
    return THETA[method]


def step_count(t_end, dt):
    steps = int(np.ceil(t_end / dt - 1e-12))
    if steps < 1:
        raise ValueError("t_end must be positive")
    return steps, t_end / steps


def axpby(a, x, b, y):
    """y = a x + b y on the device."""
    N.check(N.load().cf_axpby(y.numel(), float(a), ptr(x), float(b), ptr(y), stream_handle()),
            "cf_axpby")
    return y


def theta_run(dm, K_vals, C_vals, F, theta, t_end, steps, h, T0, mask, fvals,
              step_load=None):
    """Run the theta recursion; returns (times (S+1,), T_hist (S+1, N)) on the host.

    step_load(s) -> device tensor of the extra load of step s (time-dependent
    Neumann), or None.
    """
    n = dm.n_nodes
    pat = dm.pattern()
    A_vals = C_vals.clone()
    axpby(theta * h, K_vals, 1.0, A_vals)
    R_vals = C_vals.clone()
    axpby(-(1.0 - theta) * h, K_vals, 1.0, R_vals)
    A = DeviceMatrix(N.FMT_CSR, pat, A_vals)
    R = DeviceMatrix(N.FMT_CSR, pat, R_vals)
    if np.isscalar(T0):
        T = torch.full((n,), float(T0), dtype=torch.float64, device=dm.dev)
    else:
        T = to_dev(np.asarray(T0, dtype=float).copy(), torch.float64, dm.dev)
    fixed = None
    if mask is not None and mask.any():
        fixed = to_dev(mask.astype(np.uint8), torch.uint8, dm.dev)
        fv = to_dev(fvals, torch.float64, dm.dev)
        N.check(N.load().cf_copy_masked(n, ptr(fixed), ptr(fv), ptr(T), stream_handle()),
                "cf_copy_masked")
    hist = torch.empty((steps + 1, n), dtype=torch.float64, device=dm.dev)
    hist[0].copy_(T)
    for s in range(steps):
        b = R.matvec(T)
        axpby(h, F, 1.0, b)
        extra = step_load(s) if step_load is not None else None
        if extra is not None:
            axpby(h, extra, 1.0, b)
        if s > 0:
            # start from the extrapolation 2 T_n - T_{n-1} (prescribed
            # values are kept: they are the same in both states)
            axpby(-1.0, hist[s - 1], 2.0, T)
        pcg(A, b, T, fixed)
        hist[s + 1].copy_(T)
    times = np.cumsum(np.concatenate([[0.0], np.full(steps, h)]))
    times[-1] = t_end
    return times, hist.cpu().numpy()
