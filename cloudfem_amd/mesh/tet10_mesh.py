"""Ten-node tetrahedral meshes (Tet4 -> Tet10 midpoint enrichment).

Same numbering as cloudfem/mesh/tet10_mesh.py:33-135: corners keep their
ids; the midpoint of the sorted corner pair (a, b) gets id
n_corners + (rank of (a, b) among the unique edges in lexicographic
order).  The lexicographic rank is computed from the 64-bit key a*N + b,
which orders pairs exactly like np.unique(axis=0).
"""

import numpy as np

from .tet_mesh import TetMesh, box_mesh

# Meshes from this many elements up are enriched on the GPU when one is
# visible (cf_tet10_edges / cf_edge_midpoints: radix sort of the edge keys on
# the device); smaller ones on the host, where the transfers would dominate.
# Both give the same numbering, coordinates and boundary sets.
DEVICE_MIN_ELEMENTS = 50_000


def _use_device(n_elements):
    if n_elements < DEVICE_MIN_ELEMENTS:
        return False
    import torch
    return torch.cuda.is_available()

# Local edges of midpoint slots 4..9 (tet10_mesh.py:33)
EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))


class Tet10Mesh:
    """nodes (N, 3), elements (M, 10) [4 corners + 6 edge midpoints]."""

    def __init__(self, nodes, elements, boundaries=None, n_corners=None):
        self.nodes = np.asarray(nodes, dtype=float)
        self.elements = np.asarray(elements, dtype=int)
        self.boundaries = dict(boundaries) if boundaries else {}
        if n_corners is None:
            n_corners = len(np.unique(self.elements[:, :4]))
        self.n_corners = int(n_corners)
        self._volumes = None

    @property
    def n_nodes(self):
        return self.nodes.shape[0]

    @property
    def n_elements(self):
        return self.elements.shape[0]

    def element_volumes(self):
        """Corner-Jacobian volumes abs(det) / 6 (tet10_mesh.py:66-74)."""
        if self._volumes is None:
            p = self.nodes[self.elements[:, :4]]
            jac = np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=2)
            self._volumes = np.abs(np.linalg.det(jac)) / 6.0
        return self._volumes


def _unique_edges(els, n0):
    a = np.stack([els[:, i] for i, _ in EDGES], axis=1).astype(np.int64)
    b = np.stack([els[:, j] for _, j in EDGES], axis=1).astype(np.int64)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    keys = (lo * n0 + hi).ravel()
    uniq, inv = np.unique(keys, return_inverse=True)
    return np.column_stack([uniq // n0, uniq % n0]), inv.reshape(-1, 6)


def tet4_to_tet10(mesh):
    """Enrich a P1 mesh with edge-midpoint nodes (tet10_mesh.py:77-112)."""
    els = np.asarray(mesh.elements, dtype=int)
    if _use_device(els.shape[0]):
        from .device_ops import tet4_to_tet10_device
        return tet4_to_tet10_device(mesh)
    n0 = mesh.n_nodes
    uniq, inv = _unique_edges(els, max(n0, 1))
    mids = 0.5 * (mesh.nodes[uniq[:, 0]] + mesh.nodes[uniq[:, 1]])
    elements = np.column_stack([els, n0 + inv])
    nodes = np.vstack([mesh.nodes, mids])
    boundaries = {}
    for name, ids in mesh.boundaries.items():
        ids = np.asarray(ids, dtype=int)
        member = np.zeros(n0, dtype=bool)
        member[ids] = True
        keep = member[uniq[:, 0]] & member[uniq[:, 1]]
        extra = n0 + np.nonzero(keep)[0]
        boundaries[name] = np.sort(np.concatenate([ids, extra])).astype(int)
    return Tet10Mesh(nodes, elements, boundaries, n_corners=n0)


def box_mesh_tet10(nx, ny, nz, xlim=(0.0, 1.0), ylim=(0.0, 1.0), zlim=(0.0, 1.0)):
    """Kuhn Tet10 box; face sets selected by coordinate with
    tol = 1e-9 * max(1, ptp(nodes)) (tet10_mesh.py:115-135)."""
    mesh = tet4_to_tet10(box_mesh(nx, ny, nz, xlim=xlim, ylim=ylim, zlim=zlim))
    tol = 1e-9 * max(1.0, float(np.ptp(mesh.nodes)))
    lims = {"xmin": (0, xlim[0]), "xmax": (0, xlim[1]), "ymin": (1, ylim[0]),
            "ymax": (1, ylim[1]), "zmin": (2, zlim[0]), "zmax": (2, zlim[1])}
    mesh.boundaries = {
        name: np.nonzero(np.abs(mesh.nodes[:, ax] - v) <= tol)[0].astype(int)
        for name, (ax, v) in lims.items()}
    return mesh
