"""GPU assembly parity against the CPU oracle: CSR structure bit-exact,
values within 1e-12 row-relative, load vectors within 1e-12."""

import numpy as np
import pytest

from oracle import elastic as oel
from oracle import meshes as om
from oracle import p1 as op1
from oracle import tet10 as ot10
from tests._helpers import assert_csr_parity, assert_vec_close

pytestmark = pytest.mark.gpu

SRC = lambda x, y, z: 3.0 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)  # noqa: E731
KVAR = lambda x, y, z: 1.0 + x + 2.0 * y * z  # noqa: E731


def _pair(builder, *dims, **kw):
    from cloudfem_amd import mesh as cm
    return getattr(cm, builder)(*dims, **kw), getattr(om, builder)(*dims, **kw)


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 4), (20, 20, 20)])
def test_p1_stiffness_and_rhs(gpu, dims):
    from cloudfem_amd.fem import tet_poisson
    m, m0 = _pair("box_mesh", *dims)
    A, rhs = tet_poisson.assemble(m, k=KVAR, f=SRC)
    A0, rhs0 = op1.assemble(m0, k=KVAR, f=SRC)
    assert_csr_parity(A, A0)
    assert_vec_close(rhs, rhs0)


def test_p1_box20_pattern_size(gpu):
    from cloudfem_amd.fem import tet_poisson
    from cloudfem_amd.mesh import box_mesh
    A, _ = tet_poisson.assemble(box_mesh(20, 20, 20), k=1.0, f=SRC)
    assert A.shape == (9261, 9261) and A.nnz == 128581
    # K 1 = 0 (constants in the null space)
    assert np.abs(A @ np.ones(9261)).max() < 1e-12 * np.abs(A.data).max()


@pytest.mark.parametrize("dims", [(2, 2, 2), (4, 3, 2)])
def test_p1_mass(gpu, dims):
    from cloudfem_amd.fem import tet_heat
    m, m0 = _pair("box_mesh", *dims, xlim=(0, 2.0), zlim=(0, 0.5))
    assert_csr_parity(tet_heat.assemble_mass_tet(m, rho_cp=KVAR),
                      op1.assemble_mass(m0, rho_cp=KVAR))


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 2)])
def test_tet10_stiffness_mass_rhs(gpu, dims):
    from cloudfem_amd.fem import tet10
    m, m0 = _pair("box_mesh_tet10", *dims)
    A, rhs = tet10.assemble_tet10(m, k=KVAR, f=SRC)
    A0, rhs0 = ot10.assemble_tet10(m0, k=KVAR, f=SRC)
    assert_csr_parity(A, A0)
    assert_vec_close(rhs, rhs0)
    assert_csr_parity(tet10.assemble_mass_tet10(m, rho_cp=KVAR),
                      ot10.assemble_mass_tet10(m0, rho_cp=KVAR))


@pytest.mark.parametrize("dims", [(1, 1, 1), (4, 2, 2)])
def test_tet10_elasticity(gpu, dims):
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    m, m0 = _pair("box_mesh_tet10", *dims, **lims)
    body = (lambda x, y, z: 1e3 * x, 0.0, None)
    trac = {"xmax": (0.0, 0.0, -1e6), "ymin": (lambda x, y, z: 1e5 * x, None, 0.0)}
    T = lambda x, y, z: 20.0 + 100.0 * x  # noqa: E731
    A, rhs = elastic_tet10.assemble_elasticity_tet10(m, 210e9, 0.3, body=body, traction=trac,
                                                     T=T, alpha=12e-6, T_ref=20.0)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, body=body, traction=trac, T=T,
                                             alpha=12e-6, T_ref=20.0)
    assert_csr_parity(A, A0)
    assert_vec_close(rhs, rhs0)


def test_tet10_elasticity_80x10x10_structure(gpu):
    """80x10x10 Tet10 box: the structure counts of the full benchmark mesh."""
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    m = box_mesh_tet10(80, 10, 10, xlim=(0, 0.4), ylim=(0, 0.05), zlim=(0, 0.05))
    A, rhs = elastic_tet10.assemble_elasticity_tet10(m, 210e9, 0.3,
                                                     traction={"xmax": (0.0, 0.0, -1e6)})
    assert A.shape == (213003, 213003) and A.nnz == 17271009
    assert abs(rhs[2::3].sum() + 1e6 * 0.05 * 0.05) < 1e-9 * 2500.0
    # rigid translations and rotations are in the null space
    x, y, z = m.nodes.T
    modes = []
    for c in range(3):
        v = np.zeros(A.shape[0])
        v[c::3] = 1.0
        modes.append(v)
    rot = np.zeros(A.shape[0])
    rot[0::3], rot[1::3] = -y, x
    modes.append(rot)
    amax = np.abs(A.data).max()
    for v in modes:
        assert np.abs(A @ v).max() < 1e-9 * amax * max(1.0, np.abs(v).max())
    # symmetric
    D = (A - A.T).tocsr()
    assert (np.abs(D.data).max() if D.nnz else 0.0) <= 1e-14 * amax


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 1, 2)])
def test_p1_elasticity_and_vector_mass(gpu, dims):
    from cloudfem_amd.physics import elastic3d, modal3d
    m, m0 = _pair("box_mesh", *dims)
    body = (0.0, lambda x, y, z: 5.0 * y, 1.0)
    trac = {"zmax": (1e3, 0.0, lambda x, y, z: -2e3 * x)}
    A, rhs = elastic3d.assemble_elasticity_3d(m, 200e9, 0.3, body=body, traction=trac,
                                              T=np.linspace(0, 50, m.n_nodes), alpha=1e-5)
    A0, rhs0 = oel.assemble_elasticity_3d(m0, 200e9, 0.3, body=body, traction=trac,
                                          T=np.linspace(0, 50, m0.n_nodes), alpha=1e-5)
    assert_csr_parity(A, A0)
    assert_vec_close(rhs, rhs0)
    assert_csr_parity(modal3d.assemble_mass_3d(m, 7850.0), oel.assemble_mass_3d(m0, 7850.0))
    Ml = modal3d.assemble_mass_3d(m, 7850.0, lumped=True)
    Ml0 = oel.assemble_mass_3d(m0, 7850.0, lumped=True)
    assert_csr_parity(Ml, Ml0)


def test_boundary_terms(gpu):
    from cloudfem_amd.fem import tet10, tet_heat
    m, m0 = _pair("box_mesh", 3, 3, 2)
    neu = {"xmin": 5.0, "zmax": lambda x, y, z: 1.0 + x}
    rob = {"xmax": (20.0, 300.0), "ymin": (lambda x, y, z: 10.0 + y, 250.0)}
    K, F = tet_heat.assemble_boundary_terms(m, neumann=neu, robin=rob)
    K0, F0 = op1.assemble_boundary_terms(m0, neumann=neu, robin=rob)
    assert_csr_parity(K, K0)
    assert_vec_close(F, F0)
    m, m0 = _pair("box_mesh_tet10", 2, 3, 2)
    K, F = tet10.assemble_boundary_terms_tet10(m, neumann=neu, robin=rob)
    K0, F0 = ot10.assemble_boundary_terms_tet10(m0, neumann=neu, robin=rob)
    assert K.shape == K0.shape
    assert np.abs((K - K0)).max() <= 1e-12 * np.abs(K0).max()
    assert_vec_close(F, F0)


def test_assembly_is_deterministic(gpu):
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    m = box_mesh_tet10(10, 4, 4)
    A1, r1 = elastic_tet10.assemble_elasticity_tet10(m, 1.0, 0.25, traction={"xmax": (0, 0, 1)})
    A2, r2 = elastic_tet10.assemble_elasticity_tet10(m, 1.0, 0.25, traction={"xmax": (0, 0, 1)})
    assert np.array_equal(A1.data, A2.data) and np.array_equal(r1, r2)
