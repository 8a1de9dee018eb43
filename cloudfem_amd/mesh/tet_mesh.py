"""P1 tetrahedral mesh container and structured Kuhn box mesh.

Same public surface and numbering as cloudfem/mesh/tet_mesh.py:11-102
(node (i, j, k) -> (i*(ny+1) + j)*(nz+1) + k; cells k-outer, i-inner; six
tets per cell around the (0,0,0)-(1,1,1) diagonal), built with array
operations instead of Python loops.  The numbering fixes the CSR pattern,
so it is reproduced exactly.
"""

import numpy as np

FACE_NAMES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")

# Kuhn split of a cell, local corners c[d*4 + b*2 + a] = (i+a, j+b, k+d)
# (tet_mesh.py:72-82)
_KUHN = np.array([[0, 1, 3, 7], [0, 1, 7, 5], [0, 5, 7, 4],
                  [0, 3, 2, 7], [0, 2, 6, 7], [0, 6, 4, 7]])


class TetMesh:
    """nodes (N, 3) float, elements (M, 4) int, boundaries name -> node ids."""

    def __init__(self, nodes, elements, boundaries=None):
        self.nodes = np.asarray(nodes, dtype=float)
        self.elements = np.asarray(elements, dtype=int)
        self.boundaries = dict(boundaries) if boundaries else {}
        self._volumes = None

    @classmethod
    def wrap(cls, nodes, elements, boundaries=None):
        """A mesh over existing arrays, without dtype conversion or copies
        (e.g. the int32 connectivity of a union of meshes)."""
        self = cls.__new__(cls)
        self.nodes = nodes
        self.elements = elements
        self.boundaries = dict(boundaries) if boundaries else {}
        self._volumes = None
        return self

    @property
    def n_nodes(self):
        return self.nodes.shape[0]

    @property
    def n_elements(self):
        return self.elements.shape[0]

    def element_volumes(self):
        """|det J| / 6 per element, (M,) (tet_mesh.py:36-44)."""
        if self._volumes is None:
            p = self.nodes[self.elements[:, :4]]
            jac = np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=2)
            self._volumes = np.abs(np.linalg.det(jac)) / 6.0
        return self._volumes


def box_mesh(nx, ny, nz, xlim=(0.0, 1.0), ylim=(0.0, 1.0), zlim=(0.0, 1.0)):
    """Uniform Kuhn tetrahedral mesh of a box with faces xmin..zmax."""
    x = np.linspace(xlim[0], xlim[1], nx + 1)
    y = np.linspace(ylim[0], ylim[1], ny + 1)
    z = np.linspace(zlim[0], zlim[1], nz + 1)
    xx, yy, zz = np.meshgrid(x, y, z, indexing="ij")
    nodes = np.column_stack([xx.ravel(), yy.ravel(), zz.ravel()])
    sj, si = nz + 1, (ny + 1) * (nz + 1)  # strides of j and i

    # cell origins in k-outer, j, i-inner order
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    origin = (i * si + j * sj + k).ravel()
    corner_off = np.array([a * si + b * sj + d
                           for d in (0, 1) for b in (0, 1) for a in (0, 1)])
    corners = origin[:, None] + corner_off[None, :]          # (cells, 8)
    elements = corners[:, _KUHN].reshape(-1, 4).astype(int)  # (6 cells, 4)

    def plane(fix_axis, val):
        # node ids of a face, ordered as the reference's comprehensions
        # (outer loop over the later axis, inner over the earlier one)
        if fix_axis == 0:
            kk, jj = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), indexing="ij")
            return (val * si + jj * sj + kk).ravel()
        if fix_axis == 1:
            kk, ii = np.meshgrid(np.arange(nz + 1), np.arange(nx + 1), indexing="ij")
            return (ii * si + val * sj + kk).ravel()
        jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
        return (ii * si + jj * sj + val).ravel()

    boundaries = {"xmin": plane(0, 0), "xmax": plane(0, nx), "ymin": plane(1, 0),
                  "ymax": plane(1, ny), "zmin": plane(2, 0), "zmax": plane(2, nz)}
    boundaries = {k_: v.astype(int) for k_, v in boundaries.items()}
    return TetMesh(nodes, elements, boundaries)
