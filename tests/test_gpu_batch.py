"""Batched theta-method heat (many jittered P1 box meshes, per-element k,
Robin on all faces) against the oracle's per-mesh LU recursion
(tet_heat.py:197-274) and against the single-mesh GPU path."""

import numpy as np
import pytest

from oracle import meshes as om  # noqa: F401  (oracle package import check)
from oracle import p1 as op1
from oracle.adjoint import centroid_lookup

pytestmark = pytest.mark.gpu

FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
RHO_CP = 3.6e6
ROBIN = {f: (20.0, 20.0) for f in FACES}


def _oracle(mesh, k_e, t_end, dt):
    cen = np.asarray(mesh.nodes)[np.asarray(mesh.elements)].mean(axis=1)
    return op1.transient_heat(mesh, k=centroid_lookup(cen, k_e), rho_cp=RHO_CP, robin=ROBIN,
                              t_end=t_end, dt=dt, T0=100.0, method="cn")


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_oracle(gpu, precision):
    from cloudfem_amd.fem.batch_heat import jittered_box_batch, last_batch_stats
    from cloudfem_amd.fem.batch_heat import transient_heat_batch
    meshes, ks = jittered_box_batch(5, lo=2, hi=5, seed=1)
    out = transient_heat_batch(meshes, k=ks, rho_cp=RHO_CP, robin=ROBIN, t_end=1.0, dt=0.1,
                               T0=100.0, method="cn", precision=precision)
    st = dict(last_batch_stats)
    assert st["failed"] == 0 and st["n_steps"] == 10
    for m, k_e, (times, hist) in zip(meshes, ks, out):
        t0, h0 = _oracle(m, k_e, 1.0, 0.1)
        assert np.allclose(times, t0, rtol=0, atol=1e-15)
        assert hist.shape == h0.shape
        assert _rel(hist, h0) < 1e-9, (precision, _rel(hist, h0))


def test_batch_matches_single_mesh_path(gpu):
    """Same mesh through transient_heat_tet (one system, PCG per step)."""
    from cloudfem_amd.fem.batch_heat import jittered_box_batch, transient_heat_batch
    from cloudfem_amd.fem.tet_heat import transient_heat_tet
    meshes, ks = jittered_box_batch(2, lo=6, hi=9, seed=3)
    out = transient_heat_batch(meshes, k=ks, rho_cp=RHO_CP, robin=ROBIN, t_end=0.5, dt=0.1,
                               T0=100.0)
    for m, k_e, (_, hist) in zip(meshes, ks, out):
        cen = np.asarray(m.nodes)[np.asarray(m.elements)].mean(axis=1)
        _, h1 = transient_heat_tet(m, k=centroid_lookup(cen, k_e), rho_cp=RHO_CP, robin=ROBIN,
                                   t_end=0.5, dt=0.1, T0=100.0, method="cn")
        assert _rel(hist, h1) < 1e-9


def test_batch_64_jittered_meshes(gpu):
    """The 64-mesh workload (sides U{10..40}, 50 CN steps of 0.1 s): every
    system converges each step, mixed agrees with FP64, heat content falls
    monotonically, temperatures stay within 1 K of [T_inf, T0] (a consistent
    capacity matrix may over/undershoot slightly), and the two smallest
    meshes match the oracle."""
    from cloudfem_amd.fem.batch_heat import jittered_box_batch, last_batch_stats
    from cloudfem_amd.fem.batch_heat import transient_heat_batch
    meshes, ks = jittered_box_batch()
    kw = dict(k=ks, rho_cp=RHO_CP, robin=ROBIN, t_end=5.0, dt=0.1, T0=100.0, method="cn")
    out64 = transient_heat_batch(meshes, precision="fp64", **kw)
    st64 = dict(last_batch_stats)
    outmx = transient_heat_batch(meshes, precision="mixed", **kw)
    stmx = dict(last_batch_stats)
    print(f"fp64: {st64['iterations']} CG iterations, worst step {st64['worst_step']}, "
          f"{st64['graph_replays']} graph launches; mixed: {stmx['iterations']} FP32 "
          f"iterations, {stmx['sweeps']} sweeps; tiles {st64['autotune']}")
    assert st64["failed"] == 0 and stmx["failed"] == 0
    worst = max(_rel(a[1], b[1]) for a, b in zip(outmx, out64))
    assert worst < 1e-9, worst
    for (_, hist) in out64:
        assert hist.shape[0] == 51
        assert hist.min() > 19.0 and hist.max() < 101.0
        mean = hist.mean(axis=1)
        assert np.all(np.diff(mean) < 0)
    order = np.argsort([m.n_nodes for m in meshes])[:2]
    for i in order:
        _, h0 = _oracle(meshes[i], ks[i], 5.0, 0.1)
        assert _rel(out64[i][1], h0) < 1e-9


def test_batch_per_mesh_constants(gpu):
    """One float per mesh for k and rho_cp (expanded per element on the
    device) and a nonzero constant source, against the oracle."""
    from cloudfem_amd.fem.batch_heat import jittered_box_batch, transient_heat_batch
    meshes, _ = jittered_box_batch(3, lo=2, hi=4, seed=7)
    kv, rcv = [1.0, 2.5, 0.7], [RHO_CP, 0.5 * RHO_CP, 2.0 * RHO_CP]
    out = transient_heat_batch(meshes, k=kv, rho_cp=rcv, f=5e3, robin=ROBIN, t_end=0.5,
                               dt=0.1, T0=100.0, method="cn")
    for m, kk, rc, (_, hist) in zip(meshes, kv, rcv, out):
        _, h0 = op1.transient_heat(m, k=kk, rho_cp=rc, f=5e3, robin=ROBIN, t_end=0.5, dt=0.1,
                                   T0=100.0, method="cn")
        assert _rel(hist, h0) < 1e-9
