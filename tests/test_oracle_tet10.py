"""Pins the Tet10 and elasticity oracles.

Known answers: the 1/420 Tet10 mass table, finite-difference derivatives of
the shape functions, an independent collapsed-Gauss (Duffy) quadrature of
the element stiffness, and the reference's analytic validations
(tests/test_tet10.py:75-466, tests/test_elastic3d.py, tests/test_modal3d.py).
"""

import os

import numpy as np
import pytest

from oracle import elastic as oel
from oracle import meshes as om
from oracle import p1 as op1
from oracle import tet10 as ot10
from tests._helpers import FACES

E_S, NU_S = 200e9, 0.3


def collapsed_gauss(n):
    """Points (P, 3) and weights (sum 1/6) on the unit tet, exact for
    polynomials of degree <= 2n - 3 (Duffy map of an n^3 Gauss rule)."""
    g, w = np.polynomial.legendre.leggauss(n)
    g, w = 0.5 * (g + 1.0), 0.5 * w
    pts, wts = [], []
    for u, wu in zip(g, w):
        for v, wv in zip(g, w):
            for t, wt in zip(g, w):
                pts.append([u, v * (1 - u), t * (1 - u) * (1 - v)])
                wts.append(wu * wv * wt * (1 - u) ** 2 * (1 - v))
    return np.array(pts), np.array(wts)


def _sphere10():
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "sphere_mesh_0p5_0p22.npz"))
    m4 = om.OracleTetMesh(d["nodes"], d["elements"], {"boundary": d["boundary"]})
    return m4, om.tet4_to_tet10(m4)


def _bmask(m):
    mask = np.zeros(m.n_nodes, bool)
    for s in m.boundaries.values():
        mask[s] = True
    return mask


def test_shape_functions_and_quadrature():
    Lnodes = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1],
                       [.5, .5, 0, 0], [0, .5, .5, 0], [.5, 0, .5, 0],
                       [.5, 0, 0, .5], [0, .5, 0, .5], [0, 0, .5, .5]])
    Nn, _ = ot10.shape_functions(Lnodes)
    assert np.abs(Nn - np.eye(10)).max() < 1e-14
    Lq, wq = ot10.quadrature()
    Nq, dN = ot10.shape_functions(Lq)
    assert np.abs(Nq.sum(axis=1) - 1).max() < 1e-14 and abs(wq.sum() - 1) < 1e-15
    # dN/dL against central differences
    h = 1e-6
    for i in range(4):
        dL = np.zeros(4)
        dL[i] = h
        fd = (ot10.shape_functions(Lq + dL)[0] - ot10.shape_functions(Lq - dL)[0]) / (2 * h)
        assert np.abs(fd - dN[..., i]).max() < 1e-8
    # Keast rule exact for degree <= 3 monomials (6 a!/(|a|+3)!)
    from math import factorial
    for a in np.ndindex(4, 4, 4, 4):
        if sum(a) <= 3:
            num = np.prod([factorial(x) for x in a])
            exact = 6.0 * num / factorial(sum(a) + 3)
            got = np.sum(wq * np.prod(Lq ** np.array(a), axis=1))
            assert abs(got - exact) < 1e-15


def test_mass_table_known_answer():
    T = np.array([
        [6, 1, 1, 1, -4, -6, -4, -4, -6, -6], [1, 6, 1, 1, -4, -4, -6, -6, -4, -6],
        [1, 1, 6, 1, -6, -4, -4, -6, -6, -4], [1, 1, 1, 6, -6, -6, -6, -4, -4, -4],
        [-4, -4, -6, -6, 32, 16, 16, 16, 16, 8], [-6, -4, -4, -6, 16, 32, 16, 8, 16, 16],
        [-4, -6, -4, -6, 16, 16, 32, 16, 8, 16], [-4, -6, -6, -4, 16, 8, 16, 32, 16, 16],
        [-6, -4, -6, -4, 16, 16, 8, 16, 32, 16], [-6, -6, -4, -4, 8, 16, 16, 16, 16, 32],
    ]) / 420.0
    assert np.abs(ot10.MASS_REF - T).max() < 1e-16
    corners = np.array([[0., 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    mids = np.array([0.5 * (corners[i] + corners[j]) for i, j in om.TET10_EDGES])
    m = om.OracleTet10Mesh(np.vstack([corners, mids]), np.arange(10)[None, :], n_corners=4)
    C = ot10.assemble_mass_tet10(m, rho_cp=2.0)
    V = 2.0 / 6.0
    assert np.abs(C @ np.ones(10) - V * np.array([-1 / 20] * 4 + [1 / 5] * 6)).max() < 1e-15
    assert np.linalg.eigvalsh(C.toarray())[0] > 0


def test_element_stiffness_vs_independent_rule():
    p = np.array([[0.1, 0, 0], [1.3, 0.1, 0.2], [0.2, 1.1, 0.3], [0.1, 0.2, 0.9]])
    mids = np.array([0.5 * (p[i] + p[j]) for i, j in om.TET10_EDGES])
    m = om.OracleTet10Mesh(np.vstack([p, mids]), np.arange(10)[None, :], n_corners=4)
    jac = np.column_stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]])
    pts, wts = collapsed_gauss(4)
    L = np.column_stack([1 - pts.sum(axis=1), pts])
    _, dL = ot10.shape_functions(L)
    dxi = np.stack([dL[..., 1] - dL[..., 0], dL[..., 2] - dL[..., 0],
                    dL[..., 3] - dL[..., 0]], axis=-1)
    G = dxi @ np.linalg.inv(jac)
    det = abs(np.linalg.det(jac))
    # scalar stiffness
    K_ref = det * np.einsum("q,qai,qbi->ab", wts, G, G)
    K, _ = ot10.assemble_tet10(m)
    assert np.abs(K.toarray() - K_ref).max() < 1e-12 * np.abs(K_ref).max()
    # elasticity
    D = oel.constitutive_3d(E_S, NU_S)
    B = np.stack([oel.strain_displacement(G[q]) for q in range(len(wts))])
    Ke_ref = det * np.einsum("q,qai,ab,qbj->ij", wts, B, D, B)
    Ke, _ = oel.assemble_elasticity_tet10(m, E_S, NU_S)
    assert np.abs(Ke.toarray() - Ke_ref).max() < 1e-12 * np.abs(Ke_ref).max()
    # rigid-body modes (3 translations, 3 rotations) in the null space
    X = m.nodes
    modes = [np.tile(np.eye(3)[c], 10) for c in range(3)]
    for a, b in ((0, 1), (1, 2), (0, 2)):
        v = np.zeros((10, 3))
        v[:, a], v[:, b] = -X[:, b], X[:, a]
        modes.append(v.ravel())
    for v in modes:
        assert np.abs(Ke @ v).max() < 1e-12 * np.abs(Ke_ref).max()


def test_heat_patch_and_orders():
    ul = lambda x, y, z: 1.0 + 2.0 * x - y + 0.5 * z  # noqa: E731
    uq = lambda x, y, z: x ** 2 + 2.0 * y ** 2 - z ** 2 + x * y  # noqa: E731
    m = om.box_mesh_tet10(2, 2, 2)
    u = ot10.solve_tet10_poisson(m, f=0.0, bc={s: ul for s in FACES})
    assert np.abs(u - ul(*m.nodes.T)).max() < 1e-11
    m4, m10 = _sphere10()
    u10 = ot10.solve_tet10_poisson(m10, f=-4.0, bc={"boundary": uq})
    u4 = op1.solve_tet_poisson(m4, f=-4.0, bc={"boundary": uq})
    assert np.abs(u10 - uq(*m10.nodes.T))[~_bmask(m10)].max() < 1e-11
    assert np.abs(u4 - uq(*m4.nodes.T))[~_bmask(m4)].max() > 1e-3
    # L2 / H1 orders on the manufactured sine (tests/test_tet10.py:180-218)
    Lq, wq = ot10.quadrature()
    ue = lambda X: np.prod(np.sin(np.pi * X), axis=1)  # noqa: E731

    def ge(X):
        s, c = np.sin(np.pi * X), np.cos(np.pi * X)
        return np.pi * np.stack([c[:, 0] * s[:, 1] * s[:, 2], s[:, 0] * c[:, 1] * s[:, 2],
                                 s[:, 0] * s[:, 1] * c[:, 2]], axis=1)

    errs = []
    for n in (2, 4, 6, 8):
        m = om.box_mesh_tet10(n, n, n)
        u = ot10.solve_tet10_poisson(m, 1.0, lambda x, y, z: 3 * np.pi ** 2 * np.sin(np.pi * x)
                                     * np.sin(np.pi * y) * np.sin(np.pi * z),
                                     {s: 0.0 for s in FACES})
        L2 = H1 = 0.0
        for el in m.elements:
            G, N, xq, vol = ot10.element_operators(m, el)
            L2 += vol * np.sum(wq * (N @ u[el] - ue(xq)) ** 2)
            H1 += vol * np.sum(wq * np.sum((np.einsum("a,qai->qi", u[el], G) - ge(xq)) ** 2, 1))
        errs.append((np.sqrt(L2), np.sqrt(H1)))
    hs = np.log([1.0 / n for n in (2, 4, 6, 8)])
    o_l2 = np.polyfit(hs, np.log([e[0] for e in errs]), 1)[0]
    o_h1 = np.polyfit(hs, np.log([e[1] for e in errs]), 1)[0]
    assert 2.6 < o_l2 < 3.4 and 1.7 < o_h1 < 2.3, (o_l2, o_h1)


def test_cantilever_and_patches():
    L, B, H, P = 0.5, 0.05, 0.05, 1000.0
    m = om.box_mesh_tet10(4, 1, 1, xlim=(0, L), ylim=(0, B), zlim=(0, H))
    _, rhs = oel.assemble_elasticity_tet10(m, E_S, NU_S, traction={"xmax": (0, 0, -P / (B * H))})
    assert abs(rhs[2::3].sum() + P) < 1e-9 * P
    m = om.box_mesh_tet10(50, 2, 4, xlim=(0, L), ylim=(0, B), zlim=(0, H))
    _, _, uz = oel.solve_elasticity_tet10(m, E_S, NU_S, bc={"xmin": (0.0, 0.0, 0.0)},
                                          traction={"xmax": (0.0, 0.0, -P / (B * H))})
    tip = abs(float(uz[m.boundaries["xmax"]].mean()))
    d_eb = P * L ** 3 / (3 * E_S * B * H ** 3 / 12.0)
    assert abs(tip - d_eb) / d_eb < 0.005
    # constant strain patch
    A = np.array([[1e-3, 2e-4, 0.0], [0.0, -5e-4, 3e-4], [1e-4, 0.0, 8e-4]])
    c = np.array([1e-3, -2e-3, 5e-4])
    bc = {s: tuple((lambda x, y, z, r=r: A[r, 0] * x + A[r, 1] * y + A[r, 2] * z + c[r])
                   for r in range(3)) for s in FACES}
    m = om.box_mesh_tet10(2, 2, 2)
    u = np.stack(oel.solve_elasticity_tet10(m, E_S, NU_S, bc=bc), axis=1)
    assert np.abs(u - (m.nodes @ A.T + c)).max() < 1e-12
    # quadratic displacement on the sphere with the compensating body force
    lam = E_S * NU_S / ((1 + NU_S) * (1 - 2 * NU_S))
    mu = E_S / (2 * (1 + NU_S))
    s = 1e-3
    m4, m10 = _sphere10()
    u10 = np.stack(oel.solve_elasticity_tet10(
        m10, E_S, NU_S, bc={"boundary": (lambda x, y, z: s * x * x, 0.0, 0.0)},
        body=(-2.0 * (lam + 2 * mu) * s, 0.0, 0.0)), axis=1)
    ex = np.zeros_like(u10)
    ex[:, 0] = s * m10.nodes[:, 0] ** 2
    assert np.abs(u10 - ex)[~_bmask(m10)].max() < 1e-14
    # fully constrained uniform heating: sigma = -E alpha dT / (1 - 2 nu)
    alpha, dT = 12e-6, 100.0
    m = om.box_mesh_tet10(2, 2, 2)
    u = oel.solve_elasticity_tet10(m, E_S, NU_S, bc={f: (0.0, 0.0, 0.0) for f in FACES},
                                   T=lambda x, y, z: dT, alpha=alpha)
    sig, _ = oel.element_stress_tet10(m, u, E_S, NU_S, T=lambda x, y, z: dT, alpha=alpha)
    sex = -E_S * alpha * dT / (1 - 2 * NU_S)
    assert np.abs(sig[:, :3] - sex).max() / abs(sex) < 1e-9
    assert np.abs(sig[:, 3:]).max() / abs(sex) < 1e-9


def test_tet10_transient():
    m = om.box_mesh_tet10(4, 4, 4)
    C = ot10.assemble_mass_tet10(m)
    rng = np.random.default_rng(7)
    T0 = rng.random(m.n_nodes)
    _, TH = ot10.transient_heat_tet10(m, bc=None, t_end=0.05, dt=0.01, T0=T0)
    heat = TH @ (C @ np.ones(m.n_nodes))
    assert np.abs(heat - heat[0]).max() / abs(heat[0]) < 1e-12
    cen = int(np.argmin(np.sum((m.nodes - 0.5) ** 2, axis=1)))
    _, TH = ot10.transient_heat_tet10(m, bc={s: 0.0 for s in FACES}, t_end=0.1, dt=0.005,
                                      T0=1.0)
    n = np.arange(200)
    slab = np.sum(4.0 / (np.pi * (2 * n + 1)) * (-1.0) ** n
                  * np.exp(-((2 * n + 1) * np.pi) ** 2 * 0.1))
    assert abs(TH[-1, cen] - slab ** 3) < 5e-3


def test_p1_elasticity_known_answers():
    # uniaxial tension with symmetry rollers: the exact linear field
    t = 1e6
    m = om.box_mesh(4, 2, 2, xlim=(0, 2.0))
    bc = {"xmin": (0.0, None, None), "ymin": (None, 0.0, None), "zmin": (None, None, 0.0)}
    ux, uy, uz = oel.solve_elasticity_3d(m, E_S, NU_S, bc=bc, traction={"xmax": (t, 0, 0)})
    x, y, z = m.nodes.T
    assert np.abs(ux - t * x / E_S).max() < 1e-12 * t / E_S * 2
    assert np.abs(uy + NU_S * t * y / E_S).max() < 1e-12 * t / E_S * 2
    # constrained heating
    alpha, dT = 12e-6, 100.0
    m = om.box_mesh(2, 2, 2)
    u = oel.solve_elasticity_3d(m, E_S, NU_S, bc={f: (0.0, 0.0, 0.0) for f in FACES},
                                T=np.full(m.n_nodes, dT), alpha=alpha)
    sig, _ = oel.element_stress_strain_3d(m, u, E_S, NU_S, T=np.full(m.n_nodes, dT),
                                          alpha=alpha)
    assert np.abs(sig[:, :3] + E_S * alpha * dT / (1 - 2 * NU_S)).max() < 1e-6
    # vector mass vs the 4-point Gauss rule, lumped diagonal, totals
    p = np.array([[0.0, 0, 0], [1.3, 0.1, 0.2], [0.2, 1.1, 0.3], [0.1, 0.2, 0.9]])
    one = om.OracleTetMesh(p, np.array([[0, 1, 2, 3]]))
    M = oel.assemble_mass_3d(one, 7850.0).toarray()
    V = one.element_volumes()[0]
    a, b = 0.5854101966249685, 0.1381966011250105
    lamb = np.full((4, 4), b) + np.eye(4) * (a - b)
    ref = 7850.0 * (lamb.T @ lamb) * (V / 4.0)
    for c in range(3):
        assert np.abs(M[c::3, c::3] - ref).max() < 1e-12 * 7850 * V
    assert abs(M[0, 1]) + abs(M[0, 2]) + abs(M[1, 2]) == 0.0
    Ml = oel.assemble_mass_3d(one, 7850.0, lumped=True).toarray()
    assert np.abs(Ml - np.eye(12) * 7850.0 * V / 4.0).max() < 1e-12 * 7850 * V
