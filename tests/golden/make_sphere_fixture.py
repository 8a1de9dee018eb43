"""Write tests/golden/sphere_mesh_0p5_0p22.npz from the reference's own mesh
generator (cloudfem/mesh/unstructured3d.py:277-300, Delaunay + smoothing).

The Tet10 patch tests of the reference (tests/test_tet10.py:143-152,
311-322) run on sphere_mesh(0.5, 0.22); nothing but the generator itself
fixes that mesh, so its output is stored here as a fixture (nodes, elements,
the 'boundary' node set).  Run once in a container that has the reference
at /root/reference:  python tests/golden/make_sphere_fixture.py
"""

import os
import sys

import numpy as np

REF = os.environ.get("CLOUDFEM_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere_mesh_0p5_0p22.npz")


def main():
    sys.path.insert(0, REF)
    from cloudfem.mesh.unstructured3d import sphere_mesh
    m = sphere_mesh(0.5, 0.22)
    np.savez_compressed(OUT, nodes=m.nodes, elements=m.elements.astype(np.int64),
                        boundary=np.asarray(m.boundaries["boundary"], dtype=np.int64))
    print(f"wrote {OUT}: {m.nodes.shape[0]} nodes, {m.elements.shape[0]} tets")


if __name__ == "__main__":
    main()
