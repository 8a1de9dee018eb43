"""Exterior boundary faces of a named node set (host-side mesh processing).

Same rule and order as the reference's dictionary scans
(tet_heat.py:97-120, tet10.py:329-352, elastic_tet10.py:68-91): a face is
exterior iff its sorted corner triple occurs in exactly one element; it is
kept iff all three corners are in the set; faces are listed in order of
first occurrence (element order, then local face order) with the node
order of that occurrence.  Done with one sort of 64-bit face keys.
"""

import numpy as np

# From this many elements up the scan runs on the GPU when one is visible
# (cf_boundary_faces: one device sort of the face keys); same faces, same order.
DEVICE_MIN_ELEMENTS = 50_000


def _use_device(mesh):
    if np.asarray(mesh.elements).shape[0] < DEVICE_MIN_ELEMENTS:
        return False
    import torch
    return torch.cuda.is_available()

# P1 local faces (a,b,c), (a,b,d), (a,c,d), (b,c,d) (tet_heat.py:111)
P1_FACES = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))
# Tet10 faces as corners + edge midpoints (tet10.py:287-292)
T10_FACES = (((0, 1, 2), (4, 5, 6)), ((0, 3, 1), (7, 8, 4)),
             ((0, 2, 3), (6, 9, 7)), ((1, 3, 2), (8, 9, 5)))


def _exterior(tris, n_nodes):
    """Mask of faces (rows of `tris`, in occurrence order) whose sorted
    triple occurs exactly once."""
    s = np.sort(tris, axis=1).astype(np.int64)
    n = np.int64(max(n_nodes, 1))
    keys = (s[:, 0] * n + s[:, 1]) * n + s[:, 2]
    _, inv, counts = np.unique(keys, return_inverse=True, return_counts=True)
    return counts[inv.ravel()] == 1


def _exterior_faces(mesh, local_faces, width):
    """All exterior faces (occurrence order) of the mesh, memoised on the mesh
    object for repeated calls with different boundary names."""
    els = np.asarray(mesh.elements)
    key = (width, id(mesh.elements), els.shape, mesh.n_nodes)
    cache = getattr(mesh, "_cf_exterior", None)
    if cache is not None and cache[0] == key:
        return cache[1]
    full = np.stack([els[:, list(f)] for f in local_faces], axis=1).reshape(-1, width)
    ext = full[_exterior(full[:, :3], mesh.n_nodes)]
    try:
        mesh._cf_exterior = (key, ext)
    except AttributeError:
        pass
    return ext


def boundary_faces(mesh, name):
    """(F, 3) P1 faces of boundary `name` (KeyError if unknown)."""
    ids = np.asarray(mesh.boundaries[name], dtype=int)
    if _use_device(mesh):
        from .device_ops import boundary_faces_device
        return boundary_faces_device(mesh, name)
    tris = _exterior_faces(mesh, P1_FACES, 3)
    member = np.zeros(mesh.n_nodes, dtype=bool)
    member[ids] = True
    return tris[member[tris].all(axis=1)].astype(int).reshape(-1, 3)


def boundary_faces10(mesh, name):
    """(F, 6) Tet10 faces [c0, c1, c2, m01, m12, m20] of boundary `name`."""
    ids = np.asarray(mesh.boundaries[name], dtype=int)
    if _use_device(mesh):
        from .device_ops import boundary_faces10_device
        return boundary_faces10_device(mesh, name)
    full = _exterior_faces(mesh, [c + m for c, m in T10_FACES], 6)
    member = np.zeros(mesh.n_nodes, dtype=bool)
    member[ids] = True
    return full[member[full[:, :3]].all(axis=1)].astype(int).reshape(-1, 6)
