"""Mesh processing on the GPU against the host restatements (which the CPU
suite pins to brute force and the reference's rules): Tet4 -> Tet10 edge
numbering, midpoint coordinates, boundary sets and exterior face lists must
be identical (bit-exact integers, exact midpoints)."""

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
