"""Pins the P1 oracle (assembly, Dirichlet, mass, face terms, transient).

Known answers: closed-form element matrices, an independent 4-point Gauss
rule (tests/test_modal3d.py:47-78), the reference's analytic validations
(tests/test_tet.py:24-70; tests/test_transient3d.py:83-272) and SciPy's
canonical CSR semantics checked by brute force.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import meshes as om
from oracle import p1 as op1
from tests._helpers import FACES

GAUSS4_A, GAUSS4_B = 0.5854101966249685, 0.1381966011250105
BARY4 = np.full((4, 4), GAUSS4_B) + np.eye(4) * (GAUSS4_A - GAUSS4_B)


def _one_tet(p):
    return om.OracleTetMesh(np.asarray(p, float), np.array([[0, 1, 2, 3]]))


def test_element_stiffness_closed_form():
    # reference tet: grads are the rows of g_ref, V = 1/6
    K, _ = op1.assemble(_one_tet([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]))
    g = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    assert np.abs(K.toarray() - g @ g.T / 6.0).max() < 1e-15
    # a skewed tet: K 1 = 0, symmetric PSD of rank 3, G p = I for linear fields
    p = np.array([[0.1, 0, 0], [1.3, 0.1, 0.2], [0.2, 1.1, 0.3], [0.1, 0.2, 0.9]])
    K = op1.assemble(_one_tet(p), k=2.5)[0].toarray()
    assert np.abs(K @ np.ones(4)).max() < 1e-14 and np.abs(K - K.T).max() < 1e-15
    ev = np.linalg.eigvalsh(K)
    assert abs(ev[0]) < 1e-13 and ev[1] > 0
    G, vol = op1.tet_gradients(p)
    assert np.abs(G.T @ p - np.eye(3)).max() < 1e-13


def test_pattern_is_sorted_node_pairs():
    m = om.box_mesh(3, 2, 2)
    K, _ = op1.assemble(m)
    pairs = {(int(a), int(b)) for el in m.elements for a in el for b in el}
    rows = np.repeat(np.arange(m.n_nodes), np.diff(K.indptr))
    got = set(zip(rows.tolist(), K.indices.tolist()))
    assert got == pairs and K.nnz == len(pairs)
    assert K.has_canonical_format and K.indices.dtype == np.int32
    assert isinstance(K, sp.csr_matrix)


def test_affine_exact_and_rates():
    ex = lambda x, y, z: x + 2.0 * y + 3.0 * z - 1.0  # noqa: E731
    m = om.box_mesh(6, 6, 6)
    T = op1.solve_tet_poisson(m, 1.0, 0.0, {n: ex for n in FACES})
    assert np.sqrt(np.mean((T - ex(*m.nodes.T)) ** 2)) < 1e-10
    ue = lambda x, y, z: np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)  # noqa: E731
    errs = []
    for n in (4, 8, 16):
        m = om.box_mesh(n, n, n)
        T = op1.solve_tet_poisson(m, 1.0, lambda x, y, z: 3 * np.pi ** 2 * ue(x, y, z),
                                  {f: 0.0 for f in FACES})
        errs.append(np.sqrt(np.mean((T - ue(*m.nodes.T)) ** 2)))
    rates = np.log(np.array(errs[:-1]) / np.array(errs[1:])) / np.log(2.0)
    assert np.all((rates > 1.7) & (rates < 2.3)), rates


def test_apply_dirichlet_semantics():
    m = om.box_mesh(3, 3, 2)
    A, rhs = op1.assemble(m, 1.0, 1.0)
    bc = {"xmin": 2.0, "zmax": lambda x, y, z: x}
    B, r = op1.apply_dirichlet(A, rhs, m, bc)
    fixed = op1.dirichlet_values(m, bc)
    F = np.array(sorted(fixed))
    Bd = B.toarray()
    # every row stores every fixed column; fixed rows are identity rows
    rows = np.repeat(np.arange(m.n_nodes), np.diff(B.indptr))
    stored = set(zip(rows.tolist(), B.indices.tolist()))
    assert all((i, j) in stored for i in range(m.n_nodes) for j in F)
    for i in F:
        e = np.zeros(m.n_nodes)
        e[i] = 1.0
        assert np.array_equal(Bd[i], e) and r[i] == fixed[i]
    free = np.setdiff1d(np.arange(m.n_nodes), F)
    assert np.all(Bd[np.ix_(free, F)] == 0.0)
    # same solution as the reduced system
    v = np.array([fixed[i] for i in F])
    Ad = A.toarray()
    u = np.zeros(m.n_nodes)
    u[F] = v
    u[free] = np.linalg.solve(Ad[np.ix_(free, free)], rhs[free] - Ad[np.ix_(free, F)] @ v)
    assert np.abs(spla.spsolve(B.tocsc(), r) - u).max() < 1e-12


def test_mass_vs_gauss_rule():
    p = np.array([[0.0, 0, 0], [1.3, 0.1, 0.2], [0.2, 1.1, 0.3], [0.1, 0.2, 0.9]])
    m = _one_tet(p)
    C = op1.assemble_mass(m, rho_cp=3.0).toarray()
    V = m.element_volumes()[0]
    ref = 3.0 * (BARY4.T @ BARY4) * (V / 4.0)
    assert np.abs(C - ref).max() < 1e-15 * 3.0 * V * 10
    mb = om.box_mesh(4, 4, 4, xlim=(0, 2.0), zlim=(0, 0.5))
    Cb = op1.assemble_mass(mb, rho_cp=3.0)
    one = np.ones(mb.n_nodes)
    assert abs(one @ (Cb @ one) - 3.0) < 1e-12
    assert np.linalg.eigvalsh(Cb.toarray())[0] > 0


def test_boundary_terms_totals():
    m = om.box_mesh(3, 3, 3)
    K, F = op1.assemble_boundary_terms(m, neumann={s: 2.0 for s in FACES})
    assert K.nnz == 0 and abs(F.sum() - 12.0) < 1e-12
    K, F = op1.assemble_boundary_terms(m, robin={"xmax": (20.0, 300.0)})
    assert abs(K.sum() - 20.0) < 1e-12 and abs(F.sum() - 6000.0) < 1e-9
    # Robin face mass is the exact P1 face mass h A (1 + d)/12
    f = op1.boundary_faces(m, "xmax")[0]
    cen, area = m.nodes[f].mean(axis=0), 0.5 / 9.0
    sub = K.toarray()[np.ix_(f, f)]
    assert sub[0, 0] >= 20.0 * area * 2 / 12 - 1e-12


def test_transient_known_answers():
    # adiabatic: 1^T C T is conserved exactly
    m = om.box_mesh(4, 4, 4)
    x, y, z = m.nodes.T
    T0 = 50 + 100 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
    C = op1.assemble_mass(m)
    _, H = op1.transient_heat(m, t_end=0.1, dt=0.02, T0=T0, method="cn")
    heat = H @ (C @ np.ones(m.n_nodes))
    assert np.abs(heat - heat[0]).max() / heat[0] < 1e-10
    # steady limit: BE with large steps lands on the steady solver
    m = om.box_mesh(4, 4, 4)
    bc = {n: 0.0 for n in FACES}
    T_ss = op1.solve_tet_poisson(m, 1.0, 8.0, bc)
    _, H = op1.transient_heat(m, f=8.0, bc=bc, t_end=12.0, dt=2.0, method="be")
    assert np.abs(H[-1] - T_ss).max() / np.abs(T_ss).max() < 1e-9
    # uniform Neumann flux q on all faces of the unit cube: E(t) = E0 + 6 q t
    _, H = op1.transient_heat(m, neumann={s: 3.0 for s in FACES}, t_end=1.0, dt=0.25,
                              T0=10.0, method="be")
    C = op1.assemble_mass(m)
    assert abs(np.ones(m.n_nodes) @ (C @ H[-1]) - (10.0 + 18.0)) < 1e-9
    with pytest.raises(ValueError):
        op1.transient_heat(m, method="rk4")
    with pytest.raises(ValueError):
        op1.transient_heat(m, t_end=0.0)


def test_adjoint_oracle_vs_central_differences():
    from oracle import adjoint as oadj
    m = om.box_mesh(4, 4, 4)
    rng = np.random.default_rng(0)
    k = rng.lognormal(0.0, 0.5, m.n_elements)
    bc = {"xmin": 0.0}
    cand = np.setdiff1d(np.arange(m.n_nodes), m.boundaries["xmin"])
    sens = np.random.default_rng(1).choice(cand, 20, replace=False)
    target = np.random.default_rng(2).uniform(0, 1, 20)
    f = 10.0
    J, dk = oadj.gradient(m, k, f, bc, sens, target)
    for e in np.random.default_rng(3).choice(m.n_elements, 6, replace=False):
        h = 1e-6 * k[e]
        kp, km = k.copy(), k.copy()
        kp[e] += h
        km[e] -= h
        Jp = oadj.objective(oadj.forward(m, kp, f, bc)[0], sens, target)
        Jm = oadj.objective(oadj.forward(m, km, f, bc)[0], sens, target)
        fd = (Jp - Jm) / (2 * h)
        assert abs(fd - dk[e]) <= 1e-5 * max(abs(dk[e]), 1e-3 * np.abs(dk).max())


def test_vectorised_adjoint_matches_loop_adjoint():
    """The element-vectorised adjoint used at 40^3 equals the per-element
    restatement (same formulas, different summation order)."""
    from oracle import adjoint as oadj
    m = om.box_mesh(5, 4, 3)
    k = np.random.default_rng(0).lognormal(0.0, 0.5, m.n_elements)
    cand = np.setdiff1d(np.arange(m.n_nodes), m.boundaries["xmin"])
    sens = np.random.default_rng(1).choice(cand, 15, replace=False)
    target = np.random.default_rng(2).uniform(0, 1, 15)
    J0, dk0 = oadj.gradient(m, k, 1e4, {"xmin": 0.0}, sens, target)
    J1, dk1 = oadj.gradient_vec(m, k, 1e4, {"xmin": 0.0}, sens, target)
    assert abs(J1 - J0) <= 1e-12 * abs(J0)
    assert np.abs(dk1 - dk0).max() <= 1e-11 * np.abs(dk0).max()
