"""cloudfem_amd: MI355X (gfx950) tetrahedral P1/Tet10 assembly and sparse SPD
solve behind CloudFEM's function signatures.

Modules mirror the reference layout:
  cloudfem_amd.mesh               TetMesh, box_mesh, Tet10Mesh, tet4_to_tet10, box_mesh_tet10
  cloudfem_amd.fem.tet_poisson    assemble, apply_dirichlet, solve_tet_poisson
  cloudfem_amd.fem.tet_heat       assemble_mass_tet, assemble_boundary_terms, transient_heat_tet
  cloudfem_amd.fem.tet10          Tet10 scalar assembly, mass, face terms, transient
  cloudfem_amd.physics.elastic_tet10  assemble/solve_elasticity_tet10, element_stress_tet10
  cloudfem_amd.physics.elastic3d  P1 elasticity, apply_dirichlet_vec3, stresses
  cloudfem_amd.physics.modal3d    assemble_mass_3d
Compute runs in libcloudfem_hip.so (HIP, gfx950) through a C ABI
(include/cloudfem_hip.h); there is no CPU fallback.
"""

__version__ = "0.1.0"

from .solver import solver_options  # noqa: E402,F401
