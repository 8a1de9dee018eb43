"""Mesh processing on the GPU (same numbering and order as the host versions).

tet4_to_tet10_device mirrors tet4_to_tet10 (tet10_mesh.py:77-112) and
boundary_faces_device / boundary_faces10_device mirror the exterior-face scans
of tet_heat.py:97-120 / tet10.py:329-352, with the sorts and scans on the
device (cf_tet10_edges, cf_edge_midpoints, cf_boundary_faces).  The results
are returned as host arrays, identical to the host versions.
"""

import ctypes

import numpy as np
import torch

from .. import _native as N
from ..device import ptr, require_gpu, stream_handle, to_dev, workspace
from .tet10_mesh import Tet10Mesh


def _ws(n_elem, dev):
    nb = ctypes.c_size_t()
    N.check(N.load().cf_mesh_workspace_bytes(int(n_elem), ctypes.byref(nb)),
            "cf_mesh_workspace_bytes")
    return workspace(nb.value, dev), nb.value


def tet4_to_tet10_device(mesh):
    """Tet10Mesh with edge midpoints numbered n0 + lexicographic edge rank."""
    dev = require_gpu()
    els = np.asarray(mesh.elements)
    m, n0 = els.shape[0], mesh.n_nodes
    conn4 = to_dev(els[:, :4].astype(np.int32), torch.int32, dev)
    conn10 = torch.empty((m, 10), dtype=torch.int32, device=dev)
    ea = torch.empty(6 * m, dtype=torch.int32, device=dev)
    eb = torch.empty(6 * m, dtype=torch.int32, device=dev)
    ne = torch.zeros(1, dtype=torch.int64, device=dev)
    ws, nb = _ws(m, dev)
    lib = N.load()
    N.check(lib.cf_tet10_edges(m, ptr(conn4), 4, n0, ptr(conn10), ptr(ea), ptr(eb), ptr(ne),
                               ptr(ws), nb, stream_handle()), "cf_tet10_edges")
    n_edges = int(ne.item())
    xyz = torch.empty((3, n0 + n_edges), dtype=torch.float64, device=dev)
    xyz[:, :n0] = to_dev(np.asarray(mesh.nodes, dtype=float).T.copy(), torch.float64, dev)
    N.check(lib.cf_edge_midpoints(n_edges, n0, ptr(ea), ptr(eb), ptr(xyz[0]), ptr(xyz[1]),
                                  ptr(xyz[2]), stream_handle()), "cf_edge_midpoints")
    a, b = ea[:n_edges].long(), eb[:n_edges].long()
    boundaries = {}
    for name, ids in mesh.boundaries.items():
        ids = np.asarray(ids, dtype=np.int64)
        member = torch.zeros(n0, dtype=torch.bool, device=dev)
        member[to_dev(ids, torch.int64, dev)] = True
        extra = (n0 + torch.nonzero(member[a] & member[b]).ravel()).cpu().numpy()
        boundaries[name] = np.sort(np.concatenate([ids, extra])).astype(int)
    return Tet10Mesh(xyz.T.cpu().numpy(), conn10.cpu().numpy().astype(int), boundaries,
                     n_corners=n0)


def _faces(mesh, name, width, conn=None):
    dev = require_gpu()
    if conn is None:
        conn = to_dev(np.asarray(mesh.elements).astype(np.int32), torch.int32, dev)
    m, k = conn.shape
    member = None
    if name is not None:
        member = torch.zeros(mesh.n_nodes, dtype=torch.uint8, device=dev)
        member[to_dev(np.asarray(mesh.boundaries[name], dtype=np.int64), torch.int64, dev)] = 1
    faces = torch.empty((4 * m, width), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    ws, nb = _ws(m, dev)
    N.check(N.load().cf_boundary_faces(m, ptr(conn), k, mesh.n_nodes, ptr(member), width,
                                       ptr(faces), ptr(cnt), ptr(ws), nb, stream_handle()),
            "cf_boundary_faces")
    return faces[:int(cnt.item())].cpu().numpy().astype(int)


def boundary_faces_multi_device(mesh, names, conn=None):
    """{name: (F, 3) exterior P1 faces of boundary `name`} for several names
    from one device scan: a name's faces are the exterior faces whose three
    nodes all lie on it, in the scan's order (the subsequence a scan with
    that name's node set returns)."""
    dev = require_gpu()
    if conn is None:
        conn = to_dev(np.asarray(mesh.elements).astype(np.int32), torch.int32, dev)
    m, k = conn.shape
    faces = torch.empty((4 * m, 3), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    ws, nb = _ws(m, dev)
    N.check(N.load().cf_boundary_faces(m, ptr(conn), k, mesh.n_nodes, None, 3, ptr(faces),
                                       ptr(cnt), ptr(ws), nb, stream_handle()),
            "cf_boundary_faces")
    ext = faces[:int(cnt.item())].long()
    out = {}
    for name in names:
        member = torch.zeros(mesh.n_nodes, dtype=torch.bool, device=dev)
        member[to_dev(np.asarray(mesh.boundaries[name], dtype=np.int64), torch.int64, dev)] = True
        keep = member[ext[:, 0]] & member[ext[:, 1]] & member[ext[:, 2]]
        out[name] = ext[keep].cpu().numpy()
    return out


def boundary_faces_device(mesh, name, conn=None):
    """(F, 3) exterior P1 faces of boundary `name` (KeyError if unknown);
    conn: the connectivity already on the device (int32), if at hand."""
    return _faces(mesh, name, 3, conn)


def boundary_faces10_device(mesh, name):
    """(F, 6) exterior Tet10 faces [c0, c1, c2, m01, m12, m20] of `name`."""
    return _faces(mesh, name, 6)
