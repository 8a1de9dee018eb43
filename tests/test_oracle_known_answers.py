"""Known-answer pins for the oracle functions that no closed form elsewhere
fixes: the Tet10 Neumann / Robin face terms and the theta-method driver that
uses them (re-hosting the checks of the reference's tests/test_tet10_bc.py),
the theta-weighted time-dependent Neumann load, von Mises stress
(tests/test_elastic3d.py:96, 120) and the P1 body load.  Each check is one a
plausible mistake (a dropped term, a wrong shape-function index, a wrong
theta weight, a transposed Voigt order) fails."""

import numpy as np
import pytest

from oracle import elastic as oel
from oracle import meshes as om
from oracle import p1 as op1
from oracle import tet10 as ot10

FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")

# Closed-form P2 triangle mass matrix / area, local order [c0, c1, c2, m01,
# m12, m20]: corner diagonal 6, corner-corner -1, corner and the midpoint of
# the opposite edge -4, midpoint diagonal 32, midpoint-midpoint 16, all / 180.
M6 = np.array([[6, -1, -1, 0, -4, 0],
               [-1, 6, -1, 0, 0, -4],
               [-1, -1, 6, -4, 0, 0],
               [0, 0, -4, 32, 16, 16],
               [-4, 0, 0, 16, 32, 16],
               [0, -4, 0, 16, 16, 32]], dtype=float) / 180.0


def _total_heat(mesh, TH):
    C = ot10.assemble_mass_tet10(mesh, rho_cp=1.0)
    return TH @ (C @ np.ones(mesh.n_nodes))


def test_tet10_face_totals():
    """test_tet10_bc.py:59-75: the face terms integrate q, h T_inf and h
    over the unit face exactly."""
    m = om.box_mesh_tet10(4, 4, 4)
    for s in FACES:
        assert len(ot10.boundary_faces10(m, s)) == 32
    _, F = ot10.assemble_boundary_terms_tet10(m, neumann={"xmax": 3.0})
    assert abs(F.sum() - 3.0) < 1e-12
    K, F = ot10.assemble_boundary_terms_tet10(m, robin={"xmax": (2.0, 1.0)})
    assert abs(F.sum() - 2.0) < 1e-12
    assert abs(K.sum() - 2.0) < 1e-12
    # linear fluxes: int_{z=1} (1 + x) dA = 1.5, int_{y=0} (2 + 3 x z) dA = 2.75
    _, F = ot10.assemble_boundary_terms_tet10(m, neumann={"zmax": lambda x, y, z: 1.0 + x})
    assert abs(F.sum() - 1.5) < 1e-12
    _, F = ot10.assemble_boundary_terms_tet10(
        m, neumann={"ymin": lambda x, y, z: 2.0 + 3.0 * x * z})
    assert abs(F.sum() - 2.75) < 1e-12


def test_tet10_face_element_closed_forms():
    """Per face: constant flux puts 0 on the corners and q A / 3 on each
    midpoint; the Robin face matrix is h A M6 (the exact P2 triangle mass)."""
    m = om.box_mesh_tet10(2, 1, 1)
    faces = ot10.boundary_faces10(m, "xmax")
    n = m.n_nodes
    for face in faces:
        pc = m.nodes[face[:3]]
        area = 0.5 * np.linalg.norm(np.cross(pc[1] - pc[0], pc[2] - pc[0]))
        one = om.OracleTet10Mesh(m.nodes, m.elements, dict(m.boundaries), m.n_corners)
        one.boundaries = {"f": face}
        # a boundary set holding only this face's nodes selects only it
        assert len(ot10.boundary_faces10(one, "f")) == 1
        _, F = ot10.assemble_boundary_terms_tet10(one, neumann={"f": 2.0})
        np.testing.assert_allclose(F[face], 2.0 * area * np.array([0, 0, 0, 1, 1, 1]) / 3.0,
                                   rtol=0, atol=1e-15)
        assert np.abs(np.delete(F, face)).max() == 0.0
        K, F = ot10.assemble_boundary_terms_tet10(one, robin={"f": (5.0, 0.0)})
        Kf = K.toarray()[np.ix_(face, face)]
        np.testing.assert_allclose(Kf, 5.0 * area * M6, rtol=0, atol=1e-15 * 5.0 * area)
        assert abs(K.sum() - 5.0 * area) < 1e-14
        assert F.shape == (n,) and np.abs(F).max() == 0.0


def test_tet10_neumann_energy_balance():
    """test_tet10_bc.py:106-134: inflow q A t raises the consistent-mass
    total heat exactly; q(t) = 2 t gives t_end^2 under the theta weighting
    (Crank-Nicolson integrates the linear flux exactly); q = 1 + x on z = 1
    gives 1.5 t; equal in- and outflow conserve the heat."""
    m = om.box_mesh_tet10(4, 4, 4)
    _, TH = ot10.transient_heat_tet10(m, k=1.0, rho_cp=1.0, neumann={"xmax": 2.5},
                                      t_end=0.4, dt=0.02, T0=0.0, method="cn")
    assert abs(_total_heat(m, TH)[-1] - 1.0) < 1e-12
    _, TH = ot10.transient_heat_tet10(m, k=1.0, rho_cp=1.0,
                                      neumann={"xmax": lambda x, y, z, t: 2.0 * t},
                                      t_end=0.4, dt=0.02, T0=0.0, method="cn")
    heat = _total_heat(m, TH)
    assert abs(heat[-1] - 0.16) / 0.16 < 1e-12
    # the theta weighting per step: heat after step s is sum of h (q_s + q_s+1)/2
    np.testing.assert_allclose(heat, np.linspace(0.0, 0.4, 21) ** 2, rtol=0, atol=1e-13)
    # backward Euler weights the end of the step: sum_s h * 2 (s + 1) h
    _, TH = ot10.transient_heat_tet10(m, k=1.0, rho_cp=1.0,
                                      neumann={"xmax": lambda x, y, z, t: 2.0 * t},
                                      t_end=0.4, dt=0.02, T0=0.0, method="be")
    h = 0.02
    assert abs(_total_heat(m, TH)[-1] - sum(2.0 * (s + 1) * h * h for s in range(20))) < 1e-13
    _, TH = ot10.transient_heat_tet10(m, k=1.0, rho_cp=1.0,
                                      neumann={"zmax": lambda x, y, z: 1.0 + x},
                                      t_end=0.4, dt=0.02, T0=0.0, method="be")
    assert abs(_total_heat(m, TH)[-1] - 0.6) / 0.6 < 1e-12
    _, TH = ot10.transient_heat_tet10(m, k=1.0, rho_cp=1.0,
                                      neumann={"xmin": 3.0, "xmax": -3.0},
                                      t_end=0.5, dt=0.01, T0=0.0, method="cn")
    heat = _total_heat(m, TH)
    assert np.abs(heat - heat[0]).max() < 1e-12
    assert TH[-1].max() - TH[-1].min() > 1.0


def test_tet10_robin_steady_state():
    """test_tet10_bc.py:137-153: slab with T = 0 at x = 0 and convection
    (h = 2, T_inf = 1) at x = 1 reaches T = (2/3) x; a cube convecting to
    T_inf = 1 on every face reaches T = 1."""
    ms = om.box_mesh_tet10(8, 2, 2)
    _, TR = ot10.transient_heat_tet10(ms, k=1.0, rho_cp=1.0, bc={"xmin": 0.0},
                                      robin={"xmax": (2.0, 1.0)}, t_end=5.0, dt=0.25,
                                      T0=0.0, method="be")
    assert np.abs(TR[-1] - (2.0 / 3.0) * ms.nodes[:, 0]).max() < 1e-6
    m = om.box_mesh_tet10(4, 4, 4)
    _, TU = ot10.transient_heat_tet10(m, k=1.0, rho_cp=1.0,
                                      robin={s: (1.5, 1.0) for s in FACES},
                                      t_end=8.0, dt=0.2, T0=0.0, method="be")
    assert np.abs(TU[-1] - 1.0).max() < 1e-12


def test_tet10_combined_matches_p1():
    """test_tet10_bc.py:177-190: Dirichlet + Neumann + Robin in one run; the
    Dirichlet values are exact and the corner field agrees with P1."""
    kw = dict(k=1.0, rho_cp=1.0, bc={"xmin": 0.3}, neumann={"xmax": 1.2},
              robin={"zmax": (0.8, 0.1)}, t_end=0.3, dt=0.01, T0=0.0, method="cn")
    mc = om.box_mesh_tet10(6, 3, 3)
    _, TA = ot10.transient_heat_tet10(mc, **kw)
    _, TB = op1.transient_heat(om.box_mesh(6, 3, 3), **kw)
    assert np.abs(TA[-1][mc.boundaries["xmin"]] - 0.3).max() == 0.0
    assert np.abs(TA[-1, :TB.shape[1]] - TB[-1]).max() < 5e-2


def test_von_mises_closed_forms():
    """test_elastic3d.py:96, 120: uniaxial stress p gives p, a hydrostatic
    state gives 0; pure shear tau gives sqrt(3) tau; in general
    sqrt(3/2 s':s') of the deviator of the full tensor."""
    p = 1e8
    assert oel.von_mises_3d(np.array([p, 0, 0, 0, 0, 0])) == pytest.approx(p, rel=1e-15)
    assert oel.von_mises_3d(np.array([0, 0, p, 0, 0, 0])) == pytest.approx(p, rel=1e-15)
    assert oel.von_mises_3d(np.array([-3e8, -3e8, -3e8, 0, 0, 0])) == 0.0
    for c in (3, 4, 5):
        s = np.zeros(6)
        s[c] = p
        assert oel.von_mises_3d(s) == pytest.approx(np.sqrt(3.0) * p, rel=1e-15)
    rng = np.random.default_rng(5)
    sig = rng.normal(size=(200, 6)) * 1e7
    full = np.zeros((200, 3, 3))
    for c, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]):
        full[:, i, j] = full[:, j, i] = sig[:, c]
    dev = full - np.trace(full, axis1=1, axis2=2)[:, None, None] / 3.0 * np.eye(3)
    vm = np.sqrt(1.5 * np.einsum("kij,kij->k", dev, dev))
    np.testing.assert_allclose(oel.von_mises_3d(sig), vm, rtol=1e-13)


def test_p1_body_load_totals():
    """P1 body load b(x_c) V / 4 per node (elastic3d.py:157-162): totals
    equal int b dV for linear b (the centroid rule is exact), component by
    component, and a constant b spreads V/4 per corner."""
    m = om.box_mesh(4, 3, 2, xlim=(0.0, 2.0))
    body = (lambda x, y, z: 1.0 + x, 0.0, lambda x, y, z: -2.0 * y + z)
    _, rhs = oel.assemble_elasticity_3d(m, 200e9, 0.3, body=body)
    # int over [0,2]x[0,1]x[0,1]: (1 + x) -> 4, (-2 y + z) -> -1 + 1 = -1
    assert abs(rhs[0::3].sum() - 4.0) < 1e-13
    assert abs(rhs[1::3].sum()) == 0.0
    assert abs(rhs[2::3].sum() + 1.0) < 1e-13
    one = om.OracleTetMesh(m.nodes, m.elements[:1], m.boundaries)
    _, r1 = oel.assemble_elasticity_3d(one, 200e9, 0.3, body=(0.0, 0.0, 8.0))
    v = one.element_volumes()[0]
    np.testing.assert_allclose(r1[2::3][m.elements[0]], 2.0 * v, rtol=1e-15)
    assert np.count_nonzero(r1) == 4
