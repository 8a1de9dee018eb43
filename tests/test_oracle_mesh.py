"""Pins the oracle's mesh numbering and the product's host mesh generators.

Known answers: node index formula and Kuhn split (tet_mesh.py:61-82),
lexicographic midpoint ranks (tet10_mesh.py:92-97), the reference's own
checks (tests/test_tet.py:16-21; tests/test_tet10.py:107-117, 398-399;
tests/test_tet10_bc.py:59-75).
"""

import itertools

import numpy as np
import pytest

from cloudfem_amd import mesh as cm
from cloudfem_amd.mesh import faces as cf_faces
from oracle import meshes as om
from oracle import p1 as op1
from oracle import tet10 as ot10
from tests._helpers import FACES

DIMS = [(1, 1, 1), (2, 3, 4), (5, 2, 3)]


@pytest.mark.parametrize("dims", DIMS)
def test_box_mesh_brute_force(dims):
    nx, ny, nz = dims
    m = om.box_mesh(*dims, xlim=(0, 2.0), ylim=(0, 1.0), zlim=(0, 0.5))
    # node (i, j, k) sits at index (i (ny+1) + j)(nz+1) + k
    for i, j, k in itertools.product(range(nx + 1), range(ny + 1), range(nz + 1)):
        idx = (i * (ny + 1) + j) * (nz + 1) + k
        assert np.allclose(m.nodes[idx], [2.0 * i / nx, 1.0 * j / ny, 0.5 * k / nz])
    assert m.n_elements == 6 * nx * ny * nz
    vols = m.element_volumes()
    assert np.all(vols > 0) and abs(vols.sum() - 1.0) < 1e-12
    # every tet of cell (i, j, k) (cells k-outer, i-inner) holds the cell's
    # (0,0,0) and (1,1,1) corners
    e = 0
    for k, j, i in itertools.product(range(nz), range(ny), range(nx)):
        c0 = (i * (ny + 1) + j) * (nz + 1) + k
        c7 = ((i + 1) * (ny + 1) + j + 1) * (nz + 1) + k + 1
        for t in range(6):
            assert m.elements[e + t][0] == c0 and c7 in m.elements[e + t]
        assert len({tuple(sorted(t)) for t in m.elements[e:e + 6]}) == 6
        e += 6


@pytest.mark.parametrize("dims", DIMS)
def test_tet10_midpoint_numbering(dims):
    m4 = om.box_mesh(*dims)
    m10 = om.tet4_to_tet10(m4)
    n0 = m4.n_nodes
    assert np.array_equal(m10.nodes[:n0], m4.nodes)
    assert np.array_equal(m10.elements[:, :4], m4.elements)
    edges = sorted({tuple(sorted((int(el[a]), int(el[b]))))
                    for el in m4.elements for a, b in om.TET10_EDGES})
    rank = {e: r for r, e in enumerate(edges)}
    for el, el10 in zip(m4.elements, m10.elements):
        for s, (a, b) in enumerate(om.TET10_EDGES):
            key = tuple(sorted((int(el[a]), int(el[b]))))
            assert el10[4 + s] == n0 + rank[key]
            assert np.array_equal(m10.nodes[el10[4 + s]],
                                  0.5 * (m4.nodes[key[0]] + m4.nodes[key[1]]))
    assert abs(m10.element_volumes().sum() - 1.0) < 1e-12
    for s in FACES:
        assert set(m4.boundaries[s]) <= set(m10.boundaries[s])


def test_tet10_box_face_counts():
    mb = om.box_mesh_tet10(2, 2, 2)
    assert all(len(mb.boundaries[s]) == 25 for s in FACES)


@pytest.mark.parametrize("dims", DIMS)
def test_product_meshes_identical(dims):
    kw = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(-1.0, 1.0))
    for name in ("box_mesh", "box_mesh_tet10"):
        a, b = getattr(cm, name)(*dims, **kw), getattr(om, name)(*dims, **kw)
        assert np.array_equal(a.nodes, b.nodes) and np.array_equal(a.elements, b.elements)
        assert a.boundaries.keys() == b.boundaries.keys()
        for s in b.boundaries:
            assert np.array_equal(a.boundaries[s], b.boundaries[s])
        assert np.allclose(a.element_volumes(), b.element_volumes(), rtol=1e-15, atol=0)


def _sphere():
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "sphere_mesh_0p5_0p22.npz"))
    return d["nodes"], d["elements"], {"boundary": d["boundary"]}


def test_sphere_fixture_tet10_conversion():
    nodes, els, bnd = _sphere()
    a = cm.tet4_to_tet10(cm.TetMesh(nodes, els, bnd))
    b = om.tet4_to_tet10(om.OracleTetMesh(nodes, els, bnd))
    assert np.array_equal(a.nodes, b.nodes) and np.array_equal(a.elements, b.elements)
    assert np.array_equal(a.boundaries["boundary"], b.boundaries["boundary"])


@pytest.mark.parametrize("dims", DIMS)
def test_boundary_faces_identical(dims):
    m4, o4 = cm.box_mesh(*dims), om.box_mesh(*dims)
    m10, o10 = cm.box_mesh_tet10(*dims), om.box_mesh_tet10(*dims)
    for s in FACES:
        assert np.array_equal(cf_faces.boundary_faces(m4, s), op1.boundary_faces(o4, s))
        assert np.array_equal(cf_faces.boundary_faces10(m10, s), ot10.boundary_faces10(o10, s))
    with pytest.raises(KeyError):
        cf_faces.boundary_faces(m4, "nope")


def test_tet10_face_areas():
    """32 six-node faces per side of a 4^3 cube, areas summing to 1
    (tests/test_tet10_bc.py:59-67)."""
    m = om.box_mesh_tet10(4, 4, 4)
    for s in FACES:
        f = ot10.boundary_faces10(m, s)
        assert f.shape == (32, 6)
        p = m.nodes[f[:, :3]]
        area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
        assert abs(area.sum() - 1.0) < 1e-12
