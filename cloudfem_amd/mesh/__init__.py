"""Host-side mesh containers and generators (same numbering as CloudFEM)."""

from .tet_mesh import TetMesh, box_mesh, FACE_NAMES
from .tet10_mesh import Tet10Mesh, tet4_to_tet10, box_mesh_tet10, EDGES

__all__ = ["TetMesh", "box_mesh", "FACE_NAMES", "Tet10Mesh", "tet4_to_tet10",
           "box_mesh_tet10", "EDGES"]
