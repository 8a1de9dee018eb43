"""Adjoint dJ/dk (hand-written kernel behind torch.autograd) against the
oracle's SciPy FP64 adjoint and central finite differences."""

import numpy as np
import pytest

from oracle import adjoint as oadj
from oracle import meshes as om

pytestmark = pytest.mark.gpu


def _problem(n, n_sens):
    m0 = om.box_mesh(n, n, n)
    k = np.random.default_rng(0).lognormal(0.0, 0.5, m0.n_elements)
    cand = np.setdiff1d(np.arange(m0.n_nodes), m0.boundaries["xmin"])
    sens = np.random.default_rng(1).choice(cand, n_sens, replace=False)
    target = np.random.default_rng(2).uniform(0.0, 1.0, n_sens)
    return m0, k, sens, target


def _gpu_grad(n, k, sens, target, f):
    import torch
    from cloudfem_amd.adjoint import steady_diffusion
    from cloudfem_amd.mesh import box_mesh
    m = box_mesh(n, n, n)
    kap = torch.tensor(k, dtype=torch.float64, device="cuda", requires_grad=True)
    T = steady_diffusion(m, kap, f=f, bc={"xmin": 0.0}, tol=1e-13)
    s = torch.as_tensor(sens, device="cuda")
    J = 0.5 * ((T[s] - torch.as_tensor(target, device="cuda")) ** 2).sum()
    J.backward()
    return float(J), kap.grad.cpu().numpy(), T.detach().cpu().numpy()


def test_adjoint_matches_scipy_adjoint(gpu):
    m0, k, sens, target = _problem(10, 100)
    J0, dk0 = oadj.gradient(m0, k, 1e4, {"xmin": 0.0}, sens, target)
    J, dk, _ = _gpu_grad(10, k, sens, target, 1e4)
    assert abs(J - J0) <= 1e-10 * abs(J0)
    assert np.abs(dk - dk0).max() <= 1e-8 * np.abs(dk0).max()


def test_adjoint_matches_central_differences(gpu):
    import torch
    from cloudfem_amd.adjoint import steady_diffusion
    from cloudfem_amd.mesh import box_mesh
    n = 8
    m0, k, sens, target = _problem(n, 60)
    J, dk, _ = _gpu_grad(n, k, sens, target, 1e4)
    m = box_mesh(n, n, n)

    def Jof(kv):
        kap = torch.tensor(kv, dtype=torch.float64, device="cuda")
        T = steady_diffusion(m, kap, f=1e4, bc={"xmin": 0.0}, tol=1e-14).cpu().numpy()
        return 0.5 * float(((T[sens] - target) ** 2).sum())

    for e in np.random.default_rng(3).choice(m.n_elements, 20, replace=False):
        h = 1e-6 * k[e]
        kp, km = k.copy(), k.copy()
        kp[e] += h
        km[e] -= h
        fd = (Jof(kp) - Jof(km)) / (2 * h)
        assert abs(fd - dk[e]) <= 1e-4 * max(abs(dk[e]), 1e-3 * np.abs(dk).max()), (e, fd, dk[e])


def test_adjoint_40cubed_workload(gpu):
    """40^3 Kuhn box (384,000 tets, 68,921 nodes), k ~ lognormal(0, 0.5) per
    element, f = 1e4, T = 0 on xmin, 500 sensors: dJ/dk against the SciPy FP64
    adjoint everywhere and against central differences (step 1e-6 k_e) on 20
    random elements with a resolvable sensitivity."""
    import torch
    from cloudfem_amd.adjoint import steady_diffusion
    from cloudfem_amd.device import DeviceMesh
    from cloudfem_amd.mesh import box_mesh
    n = 40
    m0, k, sens, target = _problem(n, 500)
    J0, dk0 = oadj.gradient_vec(m0, k, 1e4, {"xmin": 0.0}, sens, target)
    J, dk, _ = _gpu_grad(n, k, sens, target, 1e4)
    assert abs(J - J0) <= 1e-9 * abs(J0)
    assert np.abs(dk - dk0).max() <= 1e-7 * np.abs(dk0).max()
    m = box_mesh(n, n, n)
    dm = DeviceMesh(m)

    def T_sens(kv):
        kap = torch.tensor(kv, dtype=torch.float64, device="cuda")
        T = steady_diffusion(m, kap, f=1e4, bc={"xmin": 0.0}, tol=1e-14, dm=dm).cpu().numpy()
        return T[sens]

    # J is ~3e9 here, so J(k+h) - J(k-h) at h = 1e-6 k_e is formed from the
    # sensor temperatures, 0.5 sum (a - b)(a + b - 2 target), not from two J
    # values; and only elements whose dJ/dk is resolvable at that step
    # (|dJ/dk_e| >= 1e-2 max) are drawn.  The 1e-3 bar is the FD noise of a
    # 1e-6 step on solves converged to 1e-14 at this size; the adjoint itself
    # is held to 1e-7 against the SciPy adjoint above.
    big = np.nonzero(np.abs(dk) >= 1e-2 * np.abs(dk).max())[0]
    for e in np.random.default_rng(3).choice(big, 20, replace=False):
        h = 1e-6 * k[e]
        kp, km = k.copy(), k.copy()
        kp[e] += h
        km[e] -= h
        a, b = T_sens(kp), T_sens(km)
        fd = 0.5 * float((a - b) @ (a + b - 2.0 * target)) / (2 * h)
        assert abs(fd - dk[e]) <= 1e-3 * abs(dk[e]), (e, fd, dk[e])
