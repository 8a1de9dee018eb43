"""GPU theta-method transient heat against the oracle's LU-once recursion
and the reference's analytic checks (tests/test_transient3d.py,
tests/test_tet10.py:344-378)."""

import numpy as np
import pytest

from oracle import meshes as om
from oracle import p1 as op1
from oracle import tet10 as ot10
from tests._helpers import FACES

pytestmark = pytest.mark.gpu


def test_p1_transient_matches_oracle(gpu):
    from cloudfem_amd.fem import tet_heat
    from cloudfem_amd.mesh import box_mesh
    kw = dict(k=lambda x, y, z: 1.0 + x, rho_cp=2.0, f=3.0, bc={"xmin": 1.0},
              neumann={"ymax": 0.5}, robin={"xmax": (4.0, 2.0), "zmin": (lambda x, y, z: 1 + y, 0.0)},
              t_end=0.3, dt=0.07, T0=0.25, method="cn")
    t, H = tet_heat.transient_heat_tet(box_mesh(5, 4, 3), **kw)
    t0, H0 = op1.transient_heat(om.box_mesh(5, 4, 3), **kw)
    assert np.array_equal(t, t0)
    assert np.abs(H - H0).max() / np.abs(H0).max() < 1e-10


def test_p1_transient_adiabatic_and_steady(gpu):
    from cloudfem_amd.fem import tet_heat, tet_poisson
    from cloudfem_amd.mesh import box_mesh
    m = box_mesh(4, 4, 4)
    x, y, z = m.nodes.T
    T0 = 50 + 100 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
    C = tet_heat.assemble_mass_tet(m)
    for method in ("be", "cn"):
        _, H = tet_heat.transient_heat_tet(m, t_end=0.1, dt=0.02, T0=T0, method=method)
        heat = H @ (C @ np.ones(m.n_nodes))
        assert np.abs(heat - heat[0]).max() / heat[0] < 1e-10
    bc = {n: 0.0 for n in FACES}
    T_ss = tet_poisson.solve_tet_poisson(m, 1.0, 8.0, bc)
    _, H = tet_heat.transient_heat_tet(m, f=8.0, bc=bc, t_end=12.0, dt=2.0, method="be")
    assert np.abs(H[-1] - T_ss).max() / np.abs(T_ss).max() < 1e-9
    with pytest.raises(ValueError):
        tet_heat.transient_heat_tet(m, method="rk4")
    with pytest.raises(ValueError):
        tet_heat.transient_heat_tet(m, t_end=0.0)


def test_tet10_transient_matches_oracle(gpu):
    from cloudfem_amd.fem import tet10
    from cloudfem_amd.mesh import box_mesh_tet10
    kw = dict(k=2.0, rho_cp=lambda x, y, z: 1.0 + z, f=1.0, bc={"zmin": 0.0},
              neumann={"xmax": lambda x, y, z, t: 1.0 + t}, robin={"ymin": (3.0, 1.5)},
              t_end=0.2, dt=0.05, T0=0.5, method="be")
    t, H = tet10.transient_heat_tet10(box_mesh_tet10(3, 2, 2), **kw)
    t0, H0 = ot10.transient_heat_tet10(om.box_mesh_tet10(3, 2, 2), **kw)
    assert np.array_equal(t, t0)
    assert np.abs(H - H0).max() / np.abs(H0).max() < 1e-10


def test_tet10_cooling_vs_series(gpu):
    from cloudfem_amd.fem import tet10
    from cloudfem_amd.mesh import box_mesh_tet10
    m = box_mesh_tet10(4, 4, 4)
    cen = int(np.argmin(np.sum((m.nodes - 0.5) ** 2, axis=1)))
    _, TH = tet10.transient_heat_tet10(m, bc={s: 0.0 for s in FACES}, t_end=0.1, dt=0.005, T0=1.0)
    n = np.arange(200)
    slab = np.sum(4.0 / (np.pi * (2 * n + 1)) * (-1.0) ** n
                  * np.exp(-((2 * n + 1) * np.pi) ** 2 * 0.1))
    assert abs(TH[-1, cen] - slab ** 3) < 5e-3
