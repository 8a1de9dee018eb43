"""GPU tests of the standalone drop-ins against the oracle: Tet10 steady
Poisson (tet10.py:168-178), apply_dirichlet_vec3 on a caller's matrix
(elastic3d.py:192-233), the P1 elasticity solve (elastic3d.py:236-249), and
coefficients given as callables that reduce over array inputs (the reference
calls them once per point, tet_poisson.py:48-56, elastic_tet10.py:176-186)."""

import numpy as np
import pytest

from oracle import elastic as oel
from oracle import meshes as om
from oracle import p1 as op1
from oracle import tet10 as ot10
from tests._helpers import FACES, assert_csr_parity, assert_vec_close, rel_err

pytestmark = pytest.mark.gpu

KVAR = lambda x, y, z: 1.0 + x + 2.0 * y * z  # noqa: E731
SRC = lambda x, y, z: 3.0 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)  # noqa: E731
# one value for array inputs, another per point
NORM = lambda x, y, z: 1.0 + np.linalg.norm([x - 0.5, y - 0.5, z - 0.5])  # noqa: E731


def test_solve_tet10_poisson_matches_oracle(gpu):
    from cloudfem_amd.fem import tet10
    from cloudfem_amd.mesh import box_mesh_tet10
    bc = {"xmin": 0.0, "xmax": lambda x, y, z: y + z, "zmax": 0.5}
    m, m0 = box_mesh_tet10(4, 3, 3), om.box_mesh_tet10(4, 3, 3)
    u = tet10.solve_tet10_poisson(m, k=KVAR, f=SRC, bc=bc)
    u0 = ot10.solve_tet10_poisson(m0, k=KVAR, f=SRC, bc=bc)
    assert rel_err(u, u0) < 1e-10


def test_apply_dirichlet_vec3_on_a_caller_matrix(gpu):
    """The caller's own (oracle-assembled) Tet10 elasticity matrix, uploaded
    as scalar CSR and solved on the free DOFs with prescribed nonzero and
    free (None) components."""
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic3d
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    m0 = om.box_mesh_tet10(12, 2, 2, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction={"xmax": (0.0, 0.0, -1e6)},
                                             body=(0.0, 0.0, -7850.0 * 9.81))
    bc = {"xmin": (0.0, 0.0, 0.0), "xmax": (1e-5, None, None)}
    u = elastic3d.apply_dirichlet_vec3(A0, rhs0, box_mesh_tet10(12, 2, 2, **lims), bc)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
    assert rel_err(u, u0) < 1e-10
    assert np.array_equal(u[3 * m0.boundaries["xmax"]], np.full(len(m0.boundaries["xmax"]), 1e-5))


def test_solve_elasticity_3d_matches_oracle(gpu):
    from cloudfem_amd.mesh import box_mesh
    from cloudfem_amd.physics import elastic3d
    lims = dict(xlim=(0.0, 1.0), ylim=(0.0, 0.2), zlim=(0.0, 0.2))
    args = dict(bc={"xmin": (0.0, 0.0, 0.0)}, traction={"xmax": (0.0, lambda x, y, z: -1e5 * z, -1e6)},
                body=(0.0, 0.0, -1e4), T=lambda x, y, z: 20.0 + 30.0 * x, alpha=1.2e-5, T_ref=20.0)
    u = np.stack(elastic3d.solve_elasticity_3d(box_mesh(10, 3, 3, **lims), 200e9, 0.3, **args))
    u0 = np.stack(oel.solve_elasticity_3d(om.box_mesh(10, 3, 3, **lims), 200e9, 0.3, **args))
    assert rel_err(u.ravel(), u0.ravel()) < 1e-10


def test_reducing_callables_reach_the_kernels_per_point(gpu):
    """k, f, Dirichlet values and traction given by a callable that returns
    one value for array inputs: the drop-ins must give the oracle's
    per-point results."""
    from cloudfem_amd.fem import tet_poisson
    from cloudfem_amd.mesh import box_mesh, box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    m, m0 = box_mesh(4, 4, 3), om.box_mesh(4, 4, 3)
    A, rhs = tet_poisson.assemble(m, k=NORM, f=NORM)
    A0, rhs0 = op1.assemble(m0, k=NORM, f=NORM)
    assert_csr_parity(A, A0)
    assert_vec_close(rhs, rhs0)
    bc = {f: NORM for f in FACES}
    u = tet_poisson.solve_tet_poisson(m, k=NORM, f=NORM, bc=bc)
    u0 = op1.solve_tet_poisson(m0, k=NORM, f=NORM, bc=bc)
    assert rel_err(u, u0) < 1e-10
    m, m0 = box_mesh_tet10(3, 2, 2), om.box_mesh_tet10(3, 2, 2)
    trac = {"xmax": (0.0, NORM, -1e6), "zmax": (NORM, 0.0, 0.0)}
    A, rhs = elastic_tet10.assemble_elasticity_tet10(m, 210e9, 0.3, traction=trac,
                                                     body=(NORM, 0.0, 0.0))
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac,
                                             body=(NORM, 0.0, 0.0))
    assert_csr_parity(A, A0)
    assert_vec_close(rhs, rhs0)
