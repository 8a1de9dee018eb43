"""Oracle mesh containers and structured box meshes (CPU, test-only).

Restates cloudfem/mesh/tet_mesh.py:11-102 and
cloudfem/mesh/tet10_mesh.py:33-135.
"""

import numpy as np

# Local Tet10 edge order of the midpoint slots 4..9
# (cloudfem/mesh/tet10_mesh.py:33).
TET10_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))

FACE_NAMES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def _volumes(nodes, corners):
    p = nodes[corners]
    jac = np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]],
                   axis=2)
    return np.abs(np.linalg.det(jac)) / 6.0


class OracleTetMesh:
    """P1 mesh container (cloudfem/mesh/tet_mesh.py:11-44)."""

    def __init__(self, nodes, elements, boundaries=None):
        self.nodes = np.asarray(nodes, dtype=float)
        self.elements = np.asarray(elements, dtype=int)
        self.boundaries = dict(boundaries) if boundaries else {}

    @property
    def n_nodes(self):
        return self.nodes.shape[0]

    @property
    def n_elements(self):
        return self.elements.shape[0]

    def element_volumes(self):
        return _volumes(self.nodes, self.elements[:, :4])


class OracleTet10Mesh(OracleTetMesh):
    """Tet10 container (cloudfem/mesh/tet10_mesh.py:36-74)."""

    def __init__(self, nodes, elements, boundaries=None, n_corners=None):
        super().__init__(nodes, elements, boundaries)
        if n_corners is None:
            n_corners = len(np.unique(self.elements[:, :4]))
        self.n_corners = int(n_corners)


def box_mesh(nx, ny, nz, xlim=(0.0, 1.0), ylim=(0.0, 1.0), zlim=(0.0, 1.0)):
    """Kuhn box mesh (cloudfem/mesh/tet_mesh.py:47-102).

    Node (i, j, k) has index (i*(ny+1) + j)*(nz+1) + k (:61-63); cells are
    visited k-outer, j, i-inner (:66-69); each cell gives the 6 tets of
    :75-82 around its (0,0,0)-(1,1,1) diagonal.
    """
    xs = np.linspace(xlim[0], xlim[1], nx + 1)
    ys = np.linspace(ylim[0], ylim[1], ny + 1)
    zs = np.linspace(zlim[0], zlim[1], nz + 1)
    gx, gy, gz = np.meshgrid(xs, ys, zs, indexing="ij")
    nodes = np.column_stack([gx.ravel(), gy.ravel(), gz.ravel()])

    def nid(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    tets = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = [nid(i + a, j + b, k + d)
                     for d in (0, 1) for b in (0, 1) for a in (0, 1)]
                tets.append((c[0], c[1], c[3], c[7]))
                tets.append((c[0], c[1], c[7], c[5]))
                tets.append((c[0], c[5], c[7], c[4]))
                tets.append((c[0], c[3], c[2], c[7]))
                tets.append((c[0], c[2], c[6], c[7]))
                tets.append((c[0], c[6], c[4], c[7]))
    elements = np.array(tets, dtype=int).reshape(-1, 4)

    bnd = {
        "xmin": [nid(0, j, k) for k in range(nz + 1) for j in range(ny + 1)],
        "xmax": [nid(nx, j, k) for k in range(nz + 1) for j in range(ny + 1)],
        "ymin": [nid(i, 0, k) for k in range(nz + 1) for i in range(nx + 1)],
        "ymax": [nid(i, ny, k) for k in range(nz + 1) for i in range(nx + 1)],
        "zmin": [nid(i, j, 0) for j in range(ny + 1) for i in range(nx + 1)],
        "zmax": [nid(i, j, nz) for j in range(ny + 1) for i in range(nx + 1)],
    }
    bnd = {k: np.array(v, dtype=int) for k, v in bnd.items()}
    return OracleTetMesh(nodes, elements, bnd)


def tet4_to_tet10(mesh):
    """Midpoint enrichment (cloudfem/mesh/tet10_mesh.py:77-112).

    Midpoint id = n_corners + lexicographic rank of the sorted (min, max)
    corner pair among all unique edges (:92-97); a midpoint joins a named
    boundary iff both parents do (:104-110).
    """
    els = np.asarray(mesh.elements, dtype=int)
    n0 = mesh.n_nodes
    pairs = np.stack([els[:, [a, b]] for a, b in TET10_EDGES], axis=1)
    pairs = np.sort(pairs.reshape(-1, 2), axis=1)
    uniq, inv = np.unique(pairs, axis=0, return_inverse=True)
    inv = np.asarray(inv).reshape(-1)
    mids = 0.5 * (mesh.nodes[uniq[:, 0]] + mesh.nodes[uniq[:, 1]])
    elements = np.concatenate([els, n0 + inv.reshape(-1, 6)], axis=1)
    nodes = np.concatenate([mesh.nodes, mids], axis=0)
    boundaries = {}
    for name, ids in mesh.boundaries.items():
        members = set(int(i) for i in ids)
        keep = [e for e, (a, b) in enumerate(uniq)
                if int(a) in members and int(b) in members]
        extra = n0 + np.array(keep, dtype=int)
        boundaries[name] = np.sort(np.concatenate(
            [np.asarray(ids, dtype=int), extra])).astype(int)
    return OracleTet10Mesh(nodes, elements, boundaries, n_corners=n0)


def box_mesh_tet10(nx, ny, nz, xlim=(0.0, 1.0), ylim=(0.0, 1.0),
                   zlim=(0.0, 1.0)):
    """Tet10 box (cloudfem/mesh/tet10_mesh.py:115-135): face sets are
    re-selected by coordinate with tol = 1e-9 * max(1, ptp(nodes))."""
    m = tet4_to_tet10(box_mesh(nx, ny, nz, xlim, ylim, zlim))
    tol = 1e-9 * max(1.0, float(np.ptp(m.nodes)))
    planes = {"xmin": (0, xlim[0]), "xmax": (0, xlim[1]),
              "ymin": (1, ylim[0]), "ymax": (1, ylim[1]),
              "zmin": (2, zlim[0]), "zmax": (2, zlim[1])}
    m.boundaries = {
        name: np.nonzero(np.abs(m.nodes[:, ax] - v) <= tol)[0].astype(int)
        for name, (ax, v) in planes.items()}
    return m
