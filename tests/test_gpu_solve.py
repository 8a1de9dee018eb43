"""GPU solves against SciPy SuperLU on the oracle's systems and against the
reference's analytic validations (tests/test_tet.py, tests/test_tet10.py)."""

import numpy as np
import pytest

from oracle import elastic as oel
from oracle import meshes as om
from oracle import p1 as op1
from tests._helpers import (FACES, assert_csr_parity, assert_residual_bar, assert_vec_close,
                            fp64_floor, rel_err)

pytestmark = pytest.mark.gpu


def test_p1_affine_exact(gpu):
    from cloudfem_amd.fem.tet_poisson import solve_tet_poisson
    from cloudfem_amd.mesh import box_mesh
    ex = lambda x, y, z: x + 2.0 * y + 3.0 * z - 1.0  # noqa: E731
    m = box_mesh(6, 6, 6)
    T = solve_tet_poisson(m, k=1.0, f=0.0, bc={n: ex for n in FACES})
    assert np.sqrt(np.mean((T - ex(*m.nodes.T)) ** 2)) < 1e-10


def test_p1_manufactured_rates(gpu):
    from cloudfem_amd.fem.tet_poisson import solve_tet_poisson
    from cloudfem_amd.mesh import box_mesh
    ue = lambda x, y, z: np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)  # noqa: E731
    errs = []
    for n in (4, 8, 16):
        m = box_mesh(n, n, n)
        T = solve_tet_poisson(m, 1.0, lambda x, y, z: 3 * np.pi ** 2 * ue(x, y, z),
                              {f: 0.0 for f in FACES})
        errs.append(np.sqrt(np.mean((T - ue(*m.nodes.T)) ** 2)))
    rates = [np.log(errs[i] / errs[i + 1]) / np.log(2.0) for i in range(2)]
    assert all(1.7 < r < 2.3 for r in rates), rates


def test_p1_box20_poisson_matches_spsolve(gpu):
    """20^3 Kuhn box, f = 3 pi^2 sin sin sin, T = 0 on all faces."""
    from cloudfem_amd.fem.tet_poisson import solve_tet_poisson
    from cloudfem_amd.mesh import box_mesh
    src = lambda x, y, z: 3 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)  # noqa: E731
    bc = {f: 0.0 for f in FACES}
    u = solve_tet_poisson(box_mesh(20, 20, 20), 1.0, src, bc)
    u0 = op1.solve_tet_poisson(om.box_mesh(20, 20, 20), 1.0, src, bc)
    assert rel_err(u, u0) < 1e-10


def test_apply_dirichlet_structure(gpu):
    from cloudfem_amd.fem import tet_poisson
    from cloudfem_amd.mesh import box_mesh
    bc = {"xmin": 1.0, "ymax": lambda x, y, z: x + z}
    m, m0 = box_mesh(4, 3, 3), om.box_mesh(4, 3, 3)
    A, rhs = tet_poisson.assemble(m, 1.0, 2.0)
    A0, rhs0 = op1.assemble(m0, 1.0, 2.0)
    B, r = tet_poisson.apply_dirichlet(A, rhs, m, bc)
    B0, r0 = op1.apply_dirichlet(A0, rhs0, m0, bc)
    assert_csr_parity(B, B0)
    assert_vec_close(r, r0)


@pytest.mark.parametrize("precision,precond", [("fp64", "jacobi"), ("fp64", "block3")])
def test_tet10_elasticity_matches_spsolve(gpu, precision, precond):
    from cloudfem_amd import solver_options
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(16, 2, 2, **lims)
    with solver_options(precision=precision, precond=precond):
        u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, traction=trac),
                     axis=1).ravel()
    m0 = om.box_mesh_tet10(16, 2, 2, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
    assert rel_err(u, u0) < 1e-10


@pytest.mark.parametrize("precision", ["mixed", "mixed16"])
@pytest.mark.parametrize("dims", [(8, 2, 2), (20, 3, 3)])
def test_mixed_precision_meets_fp64_residual(gpu, precision, dims):
    """Mixed precision (FP64 flexible CG, FP32 Chebyshev preconditioner on a
    FP32- or FP16-stored copy of the scaled matrix) reaches the FP64 residual
    bar and agrees with SuperLU to 1e-10.  tol 1e-12 lies below the FP64 floor
    eps ||A_ff| |x_f|| / ||b_f|| of these cantilevers (7.7e-11 at 8x2x2,
    1.3e-10 at 20x3x3): the solver restarts from the true residual until it
    no longer falls, and must then report floor_limited with the true
    residual within 3x the floor."""
    tol = 1e-12
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(*dims, **lims)
    with solver.solver_options(precision=precision, tol=tol):
        u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, traction=trac),
                     axis=1).ravel()
    st = solver.last_stats.value
    m0 = om.box_mesh_tet10(*dims, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
    err = rel_err(u, u0)
    Aff, bf, free = oel.reduced_system_vec3(A0, rhs0, m0, bc)[:3]
    floor = fp64_floor(Aff, u0[free], bf)
    print(f"{precision} {dims}: {st['iterations']} FP64 iterations, {st['residual_updates']} "
          f"low-precision products, {st['restarts']} restarts, rel. residual "
          f"{st['rel_residual']:.2e} (FP64 floor {floor:.2e}), error vs SuperLU {err:.2e}")
    assert_residual_bar(st, tol, floor)
    assert st["residual_updates"] >= 1
    assert err < 1e-10


@pytest.mark.parametrize("dims", [(8, 2, 2), (20, 3, 3)])
def test_matfree_mixed_meets_fp64_residual(gpu, dims):
    """FP64 flexible CG on the matrix-free operator with the FP32 matrix-free
    Chebyshev preconditioner: FP64 residual bar (tol 1e-12, or floor_limited
    within 3x the FP64 floor), error against SuperLU below 1e-10."""
    tol = 1e-12
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(*dims, **lims)
    with solver.solver_options(operator="matfree", precision="mixed", tol=tol):
        u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, traction=trac),
                     axis=1).ravel()
    st = solver.last_stats.value
    assert st["operator"] == "matfree" and st["precision"] == "mixed"
    m0 = om.box_mesh_tet10(*dims, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
    err = rel_err(u, u0)
    Aff, bf, free = oel.reduced_system_vec3(A0, rhs0, m0, bc)[:3]
    floor = fp64_floor(Aff, u0[free], bf)
    print(f"matfree mixed {dims}: {st['iterations']} FP64 iterations, "
          f"{st['residual_updates']} FP32 products, {st['restarts']} restarts, rel. residual "
          f"{st['rel_residual']:.2e} (floor {floor:.2e}), error vs SuperLU {err:.2e}")
    assert_residual_bar(st, tol, floor)
    assert st["residual_updates"] >= 1
    assert err < 1e-10


@pytest.mark.parametrize("dims", [(8, 2, 2), (20, 3, 3), (40, 5, 5)])
def test_matfree_twolevel_meets_fp64_residual(gpu, dims):
    """Two-level mixed solver (rigid-body coarse correction around the FP32
    matrix-free Chebyshev polynomial): FP64 residual bar (tol 1e-12, or
    floor_limited within 3x the FP64 floor) and the solution against SuperLU
    below 1e-10.  8x2x2 has one aggregate, 40x5x5 four (groups of 1,536
    elements)."""
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    tol = 1e-12
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(*dims, **lims)
    with solver.solver_options(operator="matfree", precision="mixed", precond="twolevel",
                               tol=tol):
        u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, traction=trac),
                     axis=1).ravel()
    st = solver.last_stats.value
    assert st["operator"] == "matfree" and st["precond"] == "two-level"
    m0 = om.box_mesh_tet10(*dims, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
    Aff, bf, free = oel.reduced_system_vec3(A0, rhs0, m0, bc)[:3]
    floor = fp64_floor(Aff, u0[free], bf)
    err = rel_err(u, u0)
    print(f"two-level {dims}: {st['iterations']} FP64 iterations, {st['residual_updates']} "
          f"FP32 products, {st['restarts']} restarts, rel. residual {st['rel_residual']:.2e} "
          f"(floor {floor:.2e}), error vs SuperLU {err:.2e}")
    assert_residual_bar(st, tol, floor)
    assert err < 1e-10


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("dims", [(8, 2, 2), (40, 5, 5)])
def test_twolevel_assembled_outer_matches_spsolve(gpu, precision, dims):
    """Two-level solver whose outer FP64 products use the assembled NB3
    matrix (the system the reference solves), with the Chebyshev polynomial
    and coarse correction matrix-free in FP32 ("mixed") or FP64 ("fp64", the
    precision twin): residual bar and 1e-10 against SuperLU."""
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(*dims, **lims)
    with solver.solver_options(operator="assembled", precision=precision, precond="twolevel",
                               tol=1e-12):
        u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, traction=trac),
                     axis=1).ravel()
    st = solver.last_stats.value
    assert st["operator"] == "assembled" and st["precond"] == "two-level"
    assert st["precision"] == precision
    m0 = om.box_mesh_tet10(*dims, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
    Aff, bf, free = oel.reduced_system_vec3(A0, rhs0, m0, bc)[:3]
    floor = fp64_floor(Aff, u0[free], bf)
    err = rel_err(u, u0)
    print(f"two-level assembled {precision} {dims}: {st['iterations']} iterations, "
          f"{st['residual_updates']} inner products, {st['restarts']} restarts, rel. residual "
          f"{st['rel_residual']:.2e} (floor {floor:.2e}), error vs SuperLU {err:.2e}")
    assert_residual_bar(st, 1e-12, floor)
    assert err < 1e-10


@pytest.mark.parametrize("n_free", [1, 2])
def test_twolevel_aggregate_with_few_free_nodes(gpu, n_free):
    """Every node clamped but one or two: the one aggregate's six rigid
    modes then span only 3 or 5 independent directions on its free DOFs,
    although every coarse diagonal entry is positive.  The dependent modes
    must be dropped before the coarse factorisation, and the solve must
    match SuperLU."""
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    m, m0 = box_mesh_tet10(4, 2, 2), om.box_mesh_tet10(4, 2, 2)
    keep = [m.n_corners + 7, m.n_corners + 40][:n_free]      # interior midpoints
    fixed_ids = np.setdiff1d(np.arange(m.n_nodes), keep)
    for mesh in (m, m0):
        mesh.boundaries = dict(mesh.boundaries, clamp=fixed_ids)
    bc = {"clamp": (0.0, 0.0, 0.0)}
    body = (1e3, -2e3, -7850.0 * 9.81)
    for op in ("matfree", "assembled"):
        with solver.solver_options(operator=op, precision="mixed", precond="twolevel",
                                   tol=1e-12):
            u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, body=body),
                         axis=1).ravel()
        st = solver.last_stats.value
        A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, body=body)
        u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
        err = rel_err(u, u0)
        print(f"{n_free} free node(s), {op}: {st['iterations']} iterations, error {err:.2e}")
        assert err < 1e-10


def test_twolevel_jittered_mesh_body_load(gpu):
    """Two-level solver on an irregular Tet10 mesh (interior and boundary
    nodes of a 36x6x6 Kuhn box jittered by up to 0.2 h, 7,776 elements, six
    aggregates) under a body load, clamped at x = 0, against SuperLU."""
    from cloudfem_amd import solver
    from cloudfem_amd.physics import elastic_tet10
    m4 = om.box_mesh(36, 6, 6, xlim=(0.0, 1.0), ylim=(0.0, 0.2), zlim=(0.0, 0.2))
    rng = np.random.default_rng(7)
    h = np.array([1.0 / 36, 0.2 / 6, 0.2 / 6])
    nodes = m4.nodes + rng.uniform(-0.2, 0.2, size=m4.nodes.shape) * h
    nodes[m4.boundaries["xmin"], 0] = 0.0   # keep the clamped face planar
    m = om.tet4_to_tet10(om.OracleTetMesh(nodes, m4.elements, m4.boundaries))
    body = (0.0, 0.0, -7850.0 * 9.81)
    bc = {"xmin": (0.0, 0.0, 0.0)}
    with solver.solver_options(operator="matfree", precision="mixed", precond="twolevel",
                               tol=1e-12):
        u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, body=body),
                     axis=1).ravel()
    st = solver.last_stats.value
    A0, rhs0 = oel.assemble_elasticity_tet10(m, 210e9, 0.3, body=body)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m, bc)
    err = rel_err(u, u0)
    Aff, bf, free = oel.reduced_system_vec3(A0, rhs0, m, bc)[:3]
    floor = fp64_floor(Aff, u0[free], bf)
    print(f"two-level jittered: {st['iterations']} FP64 iterations, rel. residual "
          f"{st['rel_residual']:.2e} (floor {floor:.2e}), error vs SuperLU {err:.2e}")
    assert_residual_bar(st, 1e-12, floor)
    assert err < 1e-10


def test_twolevel_unstructured_sphere(gpu):
    """Default Tet10 solver (two-level mixed, assembled outside) on the
    unstructured sphere fixture (Delaunay + smoothing, irregular
    connectivity), clamped on the surface nodes below z = -0.25, under
    gravity, against SuperLU."""
    import os
    from cloudfem_amd import solver
    from cloudfem_amd.physics import elastic_tet10
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "sphere_mesh_0p5_0p22.npz"))
    bnd = z["boundary"]
    cap = bnd[z["nodes"][bnd, 2] < -0.25]
    m = om.tet4_to_tet10(om.OracleTetMesh(z["nodes"], z["elements"],
                                          {"boundary": bnd, "cap": cap}))
    body = (0.0, 0.0, -7850.0 * 9.81)
    bc = {"cap": (0.0, 0.0, 0.0)}
    with solver.solver_options(tol=1e-12):
        u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, body=body),
                     axis=1).ravel()
    st = solver.last_stats.value
    assert st["precond"] == "two-level" and st["precision"] == "mixed"
    A0, rhs0 = oel.assemble_elasticity_tet10(m, 210e9, 0.3, body=body)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m, bc)
    err = rel_err(u, u0)
    Aff, bf, free = oel.reduced_system_vec3(A0, rhs0, m, bc)[:3]
    floor = fp64_floor(Aff, u0[free], bf)
    print(f"sphere: {len(cap)} clamped corner nodes, {st['iterations']} outer iterations, "
          f"rel. residual {st['rel_residual']:.2e} (floor {floor:.2e}), error {err:.2e}")
    assert_residual_bar(st, 1e-12, floor)
    assert err < 1e-10


@pytest.mark.parametrize("path", ["fp64", "mixed_2l"])
def test_tol_below_fp64_floor_stops_at_floor(gpu, path):
    """tol = 1e-15 lies far below the FP64 floor eps || |A_ff| |x_f| || /
    ||b_f|| of a 40x5x5 cantilever: the solver measures the floor when its
    true residual misses tol and stops once it is within 2x that floor
    (floor_limited; no restart is spent below it), and the solution still
    meets 1e-10 against SuperLU."""
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(40, 5, 5, **lims)
    opts = (dict(operator="assembled", precision="fp64", precond="jacobi") if path == "fp64"
            else dict(operator="assembled", precision="mixed", precond="twolevel"))
    with solver.solver_options(tol=1e-15, **opts):
        u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, traction=trac),
                     axis=1).ravel()
    st = solver.last_stats.value
    m0 = om.box_mesh_tet10(40, 5, 5, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
    Aff, bf, free = oel.reduced_system_vec3(A0, rhs0, m0, bc)[:3]
    floor = fp64_floor(Aff, u0[free], bf)
    err = rel_err(u, u0)
    print(f"{path} tol 1e-15: {st['iterations']} iterations, {st['restarts']} restarts, "
          f"rel. residual {st['rel_residual']:.2e} (floor {floor:.2e}), floor_limited "
          f"{st['floor_limited']}, error vs SuperLU {err:.2e}")
    assert st["floor_limited"] == 1
    assert st["rel_residual"] <= 2.0 * floor * 1.05
    # a restart is taken only while the true residual is above 2x the floor
    # (FP64 Jacobi: its recursion drifts over 2,300 iterations, one restart
    # brings it to 8.8e-11); none is spent below it
    assert st["restarts"] <= 1
    assert err < 1e-10


@pytest.mark.parametrize("dims", [(40, 5, 5), (80, 6, 6)])
def test_twolevel_coarse_factor_paths(gpu, monkeypatch, dims):
    """The three factorisations of the coarse matrix (banded Cholesky with
    the register or the LDS-ring inverse, rocSOLVER's dense potrf/potri)
    give the same solve: each meets the FP64 residual bar, and the solutions
    agree with SuperLU.  80x6x6 has 23,040 elements, 15 aggregates."""
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(*dims, **lims)
    m0 = om.box_mesh_tet10(*dims, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
    Aff, bf, free = oel.reduced_system_vec3(A0, rhs0, m0, bc)[:3]
    floor = fp64_floor(Aff, u0[free], bf)
    its = {}
    for path in ("auto", "ring", "dense"):
        monkeypatch.setenv("CF_COARSE_FACTOR", path)
        with solver.solver_options(operator="matfree", precision="mixed", precond="twolevel",
                                   tol=1e-12):
            u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc,
                                                              traction=trac),
                         axis=1).ravel()
        st = solver.last_stats.value
        its[path] = st["iterations"]
        err = rel_err(u, u0)
        print(f"{dims} coarse factor {path}: {st['iterations']} iterations, rel. residual "
              f"{st['rel_residual']:.2e} (floor {floor:.2e}), error vs SuperLU {err:.2e}")
        assert_residual_bar(st, 1e-12, floor)
        assert err < 1e-10
    assert max(its.values()) - min(its.values()) <= 0.1 * min(its.values()) + 2


def test_twolevel_scattered_node_numbering(gpu):
    """A 160x12x12 cantilever (138,240 elements, 90 aggregates) with its node
    ids randomly permuted: the element-order aggregates then all touch each
    other and need more than 64 colours, so the solver falls back to slabs
    along the longest axis.  The solution, mapped back, must agree with the
    solve on the original numbering, and both meet the FP64 residual bar."""
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.mesh.tet10_mesh import Tet10Mesh
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(160, 12, 12, **lims)
    perm = np.random.default_rng(3).permutation(m.n_nodes)   # old id -> new id
    nodes = np.empty_like(m.nodes)
    nodes[perm] = m.nodes
    mp = Tet10Mesh(nodes, perm[m.elements], {k: perm[v] for k, v in m.boundaries.items()},
                   n_corners=m.n_corners)
    sol = {}
    for name, mesh in (("original", m), ("permuted", mp)):
        with solver.solver_options(operator="matfree", precision="mixed", precond="twolevel",
                                   tol=1e-12):
            u = np.stack(elastic_tet10.solve_elasticity_tet10(mesh, 210e9, 0.3, bc=bc,
                                                              traction=trac), axis=1)
        st = solver.last_stats.value
        print(f"{name}: {st['iterations']} outer iterations, {st['restarts']} restarts, "
              f"rel. residual {st['rel_residual']:.2e}, floor_limited {st['floor_limited']}")
        assert st["rel_residual"] <= 1e-12 or st["floor_limited"]
        assert st["rel_residual"] <= 1e-9
        sol[name] = u if name == "original" else u[perm]
    err = rel_err(sol["permuted"].ravel(), sol["original"].ravel())
    print(f"permuted vs original numbering: {err:.2e}")
    assert err < 1e-10


def test_twolevel_fewer_iterations_than_chebyshev(gpu):
    """The coarse correction must pay: on the 40x5x5 cantilever the two-level
    solver needs fewer outer iterations than the Chebyshev-only solver."""
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(40, 5, 5, **lims)
    its = {}
    for pc in ("jacobi", "twolevel"):
        with solver.solver_options(operator="matfree", precision="mixed", precond=pc, tol=1e-10):
            elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, traction=trac)
        its[pc] = solver.last_stats.value["iterations"]
    print("outer iterations", its)
    assert its["twolevel"] < its["jacobi"]


@pytest.mark.parametrize("precision", ["mixed", "mixed16"])
def test_p1_mixed_precision_matches_spsolve(gpu, precision):
    """Scalar CSR systems take the FP32 inner CG with FP64 residual
    replacement (both mixed options): 14^3 Kuhn box Poisson vs SuperLU."""
    from cloudfem_amd import solver
    from cloudfem_amd.fem.tet_poisson import solve_tet_poisson
    from cloudfem_amd.mesh import box_mesh
    src = lambda x, y, z: 3 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)  # noqa: E731
    bc = {f: 0.0 for f in FACES}
    with solver.solver_options(precision=precision, tol=1e-12):
        u = solve_tet_poisson(box_mesh(14, 14, 14), 1.0, src, bc)
    st = solver.last_stats.value
    u0 = op1.solve_tet_poisson(om.box_mesh(14, 14, 14), 1.0, src, bc)
    assert st["rel_residual"] <= 1e-12
    assert rel_err(u, u0) < 1e-10


def test_tet10_cantilever_vs_beam_theory(gpu):
    """BIG-M6 geometry, Tet10 (50, 2, 4): tip within 0.5 % of Euler-Bernoulli
    (tests/test_tet10.py:225-256, 424-438)."""
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    E, nu, L, B, H, P = 200e9, 0.3, 0.5, 0.05, 0.05, 1000.0
    m = box_mesh_tet10(50, 2, 4, xlim=(0, L), ylim=(0, B), zlim=(0, H))
    _, _, uz = elastic_tet10.solve_elasticity_tet10(m, E, nu, bc={"xmin": (0.0, 0.0, 0.0)},
                                                    traction={"xmax": (0.0, 0.0, -P / (B * H))})
    tip = abs(float(uz[m.boundaries["xmax"]].mean()))
    d_eb = P * L ** 3 / (3 * E * B * H ** 3 / 12.0)
    assert abs(tip - d_eb) / d_eb < 0.005


def test_tet10_constant_strain_patch(gpu):
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    A = np.array([[1e-3, 2e-4, 0.0], [0.0, -5e-4, 3e-4], [1e-4, 0.0, 8e-4]])
    c = np.array([1e-3, -2e-3, 5e-4])
    bc = {s: tuple((lambda x, y, z, r=r: A[r, 0] * x + A[r, 1] * y + A[r, 2] * z + c[r])
                   for r in range(3)) for s in FACES}
    m = box_mesh_tet10(2, 2, 2)
    u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 200e9, 0.3, bc=bc), axis=1)
    assert np.abs(u - (m.nodes @ A.T + c)).max() < 1e-12
    sig, _ = elastic_tet10.element_stress_tet10(m, tuple(u.T), 200e9, 0.3)
    eps_v = np.array([A[0, 0], A[1, 1], A[2, 2], A[1, 2] + A[2, 1], A[0, 2] + A[2, 0],
                      A[0, 1] + A[1, 0]])
    sig_ex = oel.constitutive_3d(200e9, 0.3) @ eps_v
    assert np.abs(sig - sig_ex).max() / np.abs(sig_ex).max() < 1e-9


def test_default_tet10_elasticity_solver_is_two_level_mixed(gpu):
    """Without solver options, solve_elasticity_tet10 uses the two-level
    solver with FP32 inner products on the assembled matrix, and matches
    SuperLU (the drop-in's default path)."""
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac, bc = {"xmax": (0.0, 0.0, -1e6)}, {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(24, 3, 3, **lims)
    u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, traction=trac),
                 axis=1).ravel()
    st = solver.last_stats.value
    assert (st["precond"], st["precision"], st["operator"]) == ("two-level", "mixed", "assembled")
    m0 = om.box_mesh_tet10(24, 3, 3, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac)
    assert rel_err(u, oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)) < 1e-10


def test_solver_failures_raise(gpu):
    from cloudfem_amd import solver_options
    from cloudfem_amd._native import SolverBreakdown, SolverNotConverged
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    m = box_mesh_tet10(8, 2, 2)
    args = dict(bc={"xmin": (0.0, 0.0, 0.0)}, traction={"xmax": (0.0, 0.0, -1.0)})
    with solver_options(maxit=3):
        with pytest.raises(SolverNotConverged):
            elastic_tet10.solve_elasticity_tet10(m, 1.0, 0.3, **args)
    with pytest.raises(SolverBreakdown):
        elastic_tet10.solve_elasticity_tet10(m, 1.0, 0.3, bc=args["bc"],
                                             traction={"xmax": (0.0, 0.0, np.nan)})


def test_matfree_operator_matches_assembled(gpu):
    import torch
    from cloudfem_amd import solver_options
    from cloudfem_amd._native import FMT_NB3
    from cloudfem_amd.device import DeviceMesh
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    from cloudfem_amd.physics.elastic3d import lame
    from cloudfem_amd.solver import DeviceMatrix, matfree_t10_apply
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    m = box_mesh_tet10(12, 3, 2, **lims)
    lam, mu = lame(210e9, 0.3)
    dm = DeviceMesh(m)
    vals = dm.matrix_values(5, lam=lam, mu=mu)
    A = DeviceMatrix(FMT_NB3, dm.pattern(), vals)
    g = torch.Generator(device="cpu").manual_seed(3)
    x = torch.randn(3 * m.n_nodes, dtype=torch.float64, generator=g).to("cuda")
    y_asm = A.matvec(x)
    for chunked in (False, True):
        y_mf = matfree_t10_apply(dm, lam, mu, x, chunked=chunked)
        err = float((y_mf - y_asm).abs().max() / y_asm.abs().max())
        assert err < 1e-13, (chunked, err)
    fixed = torch.zeros(3 * m.n_nodes, dtype=torch.uint8, device="cuda")
    fixed[::7] = 1
    y_fx = matfree_t10_apply(dm, lam, mu, x, fixed=fixed, chunked=True)
    assert float(y_fx[fixed.bool()].abs().max()) == 0.0
    assert torch.equal(y_fx[~fixed.bool()], y_mf[~fixed.bool()])
    plan, counts = dm.matfree_plan()
    assert counts[0] > 0 and 0 < counts[1] <= counts[0] and 0 < counts[2] <= m.n_nodes
    bc, trac = {"xmin": (0.0, 0.0, 0.0)}, {"xmax": (0.0, 0.0, -1e6)}
    with solver_options(precision="fp64", precond="jacobi"):
        u_a = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc, traction=trac))
    for pc in ("jacobi", "block3"):
        with solver_options(operator="matfree", precision="fp64", precond=pc):
            u_m = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc,
                                                                traction=trac))
        assert rel_err(u_m.ravel(), u_a.ravel()) < 1e-10


def test_inner_precision_operator_matches_assembled(gpu):
    """The solvers' inner product S K S x (matrix-free, run plan, tile-packed
    inputs) against the FP64 assembled matrix: the FP64 twin to rounding,
    the FP32 one to FP32 rounding, on a jittered-free 12x3x2 box with a
    random S."""
    import torch
    from cloudfem_amd._native import FMT_NB3
    from cloudfem_amd.device import DeviceMesh
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics.elastic3d import lame
    from cloudfem_amd.solver import DeviceMatrix, matfree_t10_lp_apply
    m = box_mesh_tet10(12, 3, 2, xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    lam, mu = lame(210e9, 0.3)
    dm = DeviceMesh(m)
    A = DeviceMatrix(FMT_NB3, dm.pattern(), dm.matrix_values(5, lam=lam, mu=mu))
    g = torch.Generator(device="cpu").manual_seed(5)
    x = torch.randn(3 * m.n_nodes, dtype=torch.float64, generator=g).to("cuda")
    sc = (torch.rand(3 * m.n_nodes, dtype=torch.float64, generator=g) + 0.5).to("cuda") * 1e-5
    y_ref = sc * A.matvec(sc * x)
    y64, _ = matfree_t10_lp_apply(dm, lam, mu, x, precision="fp64", scale=sc)
    y32, _ = matfree_t10_lp_apply(dm, lam, mu, x.float(), precision="mixed", scale=sc)
    e64 = float((y64 - y_ref).abs().max() / y_ref.abs().max())
    e32 = float((y32.double() - y_ref).abs().max() / y_ref.abs().max())
    print(f"inner product vs assembled: FP64 {e64:.2e}, FP32 {e32:.2e}")
    assert e64 < 1e-13
    assert e32 < 1e-5


@pytest.mark.parametrize("dims", [(12, 3, 2), (5, 5, 5), (24, 5, 5)])
def test_inner_product_mfma_matches_valu(gpu, monkeypatch, dims):
    """The matrix-core element kernel (mf_t10_elem_mfma, CF_INNER_KERNEL=mfma)
    against the VALU kernel and the FP64 assembled matrix: S K S x with a
    random S (the power iteration's form, S gathered in the kernel) on boxes
    whose last 256-element tile is partial (432, 750 and 3,600 elements)."""
    import torch
    from cloudfem_amd._native import FMT_NB3
    from cloudfem_amd.device import DeviceMesh
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics.elastic3d import lame
    from cloudfem_amd.solver import DeviceMatrix, matfree_t10_lp_apply
    m = box_mesh_tet10(*dims, xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    lam, mu = lame(210e9, 0.3)
    dm = DeviceMesh(m)
    A = DeviceMatrix(FMT_NB3, dm.pattern(), dm.matrix_values(5, lam=lam, mu=mu))
    g = torch.Generator(device="cpu").manual_seed(sum(dims))
    x = torch.randn(3 * m.n_nodes, dtype=torch.float64, generator=g).to("cuda")
    sc = (torch.rand(3 * m.n_nodes, dtype=torch.float64, generator=g) + 0.5).to("cuda") * 1e-5
    y_ref = sc * A.matvec(sc * x)
    ys = {}
    for kern in ("valu", "mfma"):
        monkeypatch.setenv("CF_INNER_KERNEL", kern)
        ys[kern], _ = matfree_t10_lp_apply(dm, lam, mu, x.float(), precision="mixed", scale=sc)
    scale = float(y_ref.abs().max())
    e_ref = float((ys["mfma"].double() - y_ref).abs().max()) / scale
    e_valu = float((ys["mfma"] - ys["valu"]).abs().max()) / scale
    print(f"{dims} MFMA inner product: vs assembled FP64 {e_ref:.2e}, vs VALU {e_valu:.2e}")
    # FP32 rounding of x, S, geometry and the sums (the VALU kernel: ~1.6e-7)
    assert e_ref < 1e-6
    assert e_valu < 1e-6


@pytest.mark.parametrize("dims", [(8, 2, 2), (40, 5, 5)])
def test_twolevel_mfma_inner_matches_spsolve(gpu, monkeypatch, dims):
    """The two-level mixed solver with its FP32 Chebyshev products on the
    matrix cores: residual bar and 1e-10 against SuperLU, and within a few
    outer iterations of the VALU kernel's solve."""
    from cloudfem_amd import solver
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics import elastic_tet10
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    trac = {"xmax": (0.0, 0.0, -1e6)}
    bc = {"xmin": (0.0, 0.0, 0.0)}
    m = box_mesh_tet10(*dims, **lims)
    m0 = om.box_mesh_tet10(*dims, **lims)
    A0, rhs0 = oel.assemble_elasticity_tet10(m0, 210e9, 0.3, traction=trac)
    u0 = oel.apply_dirichlet_vec3(A0, rhs0, m0, bc)
    Aff, bf, free = oel.reduced_system_vec3(A0, rhs0, m0, bc)[:3]
    floor = fp64_floor(Aff, u0[free], bf)
    its = {}
    for kern in ("valu", "mfma"):
        monkeypatch.setenv("CF_INNER_KERNEL", kern)
        with solver.solver_options(operator="assembled", precision="mixed", precond="twolevel",
                                   tol=1e-12):
            u = np.stack(elastic_tet10.solve_elasticity_tet10(m, 210e9, 0.3, bc=bc,
                                                              traction=trac),
                         axis=1).ravel()
        st = solver.last_stats.value
        its[kern] = st["iterations"]
        err = rel_err(u, u0)
        print(f"{dims} inner kernel {kern}: {st['iterations']} iterations, rel. residual "
              f"{st['rel_residual']:.2e} (floor {floor:.2e}), error vs SuperLU {err:.2e}")
        assert_residual_bar(st, 1e-12, floor)
        assert err < 1e-10
    # FP32 rounding differs between the kernels, so the iteration counts may
    # differ a little (measured: 189 / 193 at 8x2x2, 178 / 159 at 40x5x5)
    assert its["mfma"] <= 1.25 * its["valu"] + 5


@pytest.mark.parametrize("dims", [(24, 3, 3), (5, 5, 5)])
def test_chunked_matfree_solve_matches_two_pass(gpu, dims):
    """The LDS-chunked operator (Morton chunks, shared-node partials) and the
    two-pass operator give the same CG solution; several chunks per mesh."""
    import torch
    from cloudfem_amd.device import DeviceMesh
    from cloudfem_amd.mesh import box_mesh_tet10
    from cloudfem_amd.physics.elastic3d import lame
    from cloudfem_amd.solver import pcg_matfree_t10
    from cloudfem_amd.coef import dirichlet_vec3
    m = box_mesh_tet10(*dims, xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    lam, mu = lame(210e9, 0.3)
    dm = DeviceMesh(m)
    mask, _ = dirichlet_vec3(m, {"xmin": (0.0, 0.0, 0.0)})
    fixed = torch.as_tensor(mask.astype(np.uint8), device="cuda")
    g = torch.Generator(device="cpu").manual_seed(sum(dims))
    b = torch.randn(3 * m.n_nodes, dtype=torch.float64, generator=g).to("cuda")
    b[fixed.bool()] = 0.0
    xs = []
    for chunked in (False, True):
        x = torch.zeros_like(b)
        st = pcg_matfree_t10(dm, lam, mu, b, x, fixed, tol=1e-12, chunked=chunked)
        # true residual: tol, or the FP64 floor of these cantilevers (~1e-10)
        assert st["rel_residual"] <= 1e-12 or (st["floor_limited"] and st["rel_residual"] < 1e-9)
        xs.append(x)
    assert float((xs[0] - xs[1]).abs().max() / xs[0].abs().max()) < 1e-10
