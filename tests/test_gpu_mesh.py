"""Mesh processing on the GPU against the host restatements (which the CPU
suite pins to brute force and the reference's rules) and, on a mesh large
enough that the drop-ins route it to the GPU, against the oracle directly:
Tet4 -> Tet10 edge numbering, midpoint coordinates, boundary sets and
exterior face lists must be identical (bit-exact integers, exact
midpoints)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def _meshes():
    from cloudfem_amd.fem.batch_heat import jittered_box_batch
    from cloudfem_amd.mesh import box_mesh
    out = [box_mesh(3, 2, 4), box_mesh(7, 5, 6, xlim=(0, 2), ylim=(-1, 1), zlim=(0, 0.5))]
    out += jittered_box_batch(2, lo=4, hi=9, seed=5)[0]
    return out


def test_tet10_edges_match_host(gpu):
    from cloudfem_amd.mesh import tet4_to_tet10
    from cloudfem_amd.mesh.device_ops import tet4_to_tet10_device
    for m in _meshes():
        h = tet4_to_tet10(m)
        d = tet4_to_tet10_device(m)
        assert np.array_equal(d.elements, h.elements)
        assert np.array_equal(d.nodes, h.nodes)
        assert d.n_corners == h.n_corners
        assert set(d.boundaries) == set(h.boundaries)
        for name in h.boundaries:
            assert np.array_equal(d.boundaries[name], h.boundaries[name])


def test_boundary_faces_match_host(gpu):
    from cloudfem_amd.mesh import box_mesh_tet10, tet4_to_tet10
    from cloudfem_amd.mesh.device_ops import boundary_faces10_device, boundary_faces_device
    from cloudfem_amd.mesh.faces import boundary_faces, boundary_faces10
    for m in _meshes():
        for name in FACES:
            assert np.array_equal(boundary_faces_device(m, name), boundary_faces(m, name))
        m10 = tet4_to_tet10(m)
        for name in FACES:
            assert np.array_equal(boundary_faces10_device(m10, name), boundary_faces10(m10, name))
    m10 = box_mesh_tet10(6, 3, 2, xlim=(0, 0.4), ylim=(0, 0.05), zlim=(0, 0.05))
    for name in FACES:
        assert np.array_equal(boundary_faces10_device(m10, name), boundary_faces10(m10, name))


def test_multi_name_face_scan_matches_host(gpu):
    """One device scan for several boundary names (the batched solver's
    union mesh) gives each name's host face list, in order."""
    from cloudfem_amd.fem.batch_heat import _union
    from cloudfem_amd.mesh.device_ops import boundary_faces_multi_device
    from cloudfem_amd.mesh.faces import boundary_faces
    for m in _meshes():
        got = boundary_faces_multi_device(m, FACES)
        for name in FACES:
            assert np.array_equal(got[name], boundary_faces(m, name)), name
    ms = _meshes()
    um, offs = _union(ms)
    for name in FACES:
        um.boundaries[name] = np.concatenate(
            [np.asarray(m.boundaries[name]) + offs[i] for i, m in enumerate(ms)])
    got = boundary_faces_multi_device(um, FACES)
    for name in FACES:
        want = np.concatenate([boundary_faces(m, name) + offs[i] for i, m in enumerate(ms)])
        assert np.array_equal(got[name], want), name


def test_sphere_fixture(gpu):
    """The unstructured sphere fixture (Delaunay + smoothing, irregular
    connectivity): Tet10 numbering and exterior faces as on the host."""
    import os
    from cloudfem_amd.mesh import TetMesh, tet4_to_tet10
    from cloudfem_amd.mesh.device_ops import (boundary_faces10_device, boundary_faces_device,
                                              tet4_to_tet10_device)
    from cloudfem_amd.mesh.faces import boundary_faces, boundary_faces10
    f = os.path.join(os.path.dirname(__file__), "golden", "sphere_mesh_0p5_0p22.npz")
    z = np.load(f)
    m = TetMesh(z["nodes"], z["elements"], {"boundary": z["boundary"]})
    assert np.array_equal(boundary_faces_device(m, "boundary"), boundary_faces(m, "boundary"))
    h, d = tet4_to_tet10(m), tet4_to_tet10_device(m)
    assert np.array_equal(d.elements, h.elements) and np.array_equal(d.nodes, h.nodes)
    assert np.array_equal(d.boundaries["boundary"], h.boundaries["boundary"])
    assert np.array_equal(boundary_faces10_device(h, "boundary"), boundary_faces10(h, "boundary"))


def test_large_box_drop_ins_run_on_gpu_and_match_oracle(gpu, monkeypatch):
    """box_mesh_tet10 / tet4_to_tet10 / boundary_faces10 of a 96,000-element
    box take the device path (tet10_mesh.py:77-135, tet10.py:329-352 and
    elastic_tet10.py:68-91 restated on the GPU) and equal the oracle."""
    from cloudfem_amd.mesh import box_mesh_tet10, device_ops
    from cloudfem_amd.mesh import faces as faces_mod
    from cloudfem_amd.mesh import tet10_mesh
    from oracle import meshes as om
    from oracle import tet10 as ot10
    calls = {"edges": 0, "faces": 0}
    real_edges, real_faces = device_ops.tet4_to_tet10_device, device_ops.boundary_faces10_device

    def edges(m):
        calls["edges"] += 1
        return real_edges(m)

    def faces(m, name):
        calls["faces"] += 1
        return real_faces(m, name)
    monkeypatch.setattr(device_ops, "tet4_to_tet10_device", edges)
    monkeypatch.setattr(device_ops, "boundary_faces10_device", faces)
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.1), zlim=(0.0, 0.1))
    m = box_mesh_tet10(40, 20, 20, **lims)
    assert m.n_elements >= tet10_mesh.DEVICE_MIN_ELEMENTS and calls["edges"] == 1
    m0 = om.box_mesh_tet10(40, 20, 20, **lims)
    assert np.array_equal(m.elements, m0.elements)
    assert np.array_equal(m.nodes, m0.nodes)
    assert m.n_corners == m0.n_corners
    for name in FACES:
        assert np.array_equal(m.boundaries[name], m0.boundaries[name])
        f = faces_mod.boundary_faces10(m, name)
        assert np.array_equal(f, ot10.boundary_faces10(m0, name))
    assert calls["faces"] == len(FACES)
