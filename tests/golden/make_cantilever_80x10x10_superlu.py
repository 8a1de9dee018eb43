"""Writes cantilever_80x10x10_superlu.npz: the SciPy SuperLU solution of the
80x10x10 Tet10 cantilever (0.4 x 0.05 x 0.05 m, E = 210 GPa, nu = 0.3,
clamped at x = 0, 1 MPa traction in -z on x = 0.4) on the oracle's reduced
system (the spsolve of elastic3d.py:220-230).  SuperLU needs minutes and
several GB here, too long for a GPU test, so the solution is stored.

Run from the repository root:  python tests/golden/make_cantilever_80x10x10_superlu.py
"""

import os
import sys
import time

import numpy as np
import scipy.sparse.linalg as spla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import elastic as oel  # noqa: E402
from oracle import meshes as om  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cantilever_80x10x10_superlu.npz")


def main():
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    m = om.box_mesh_tet10(80, 10, 10, **lims)
    A, rhs = oel.assemble_elasticity_tet10(m, 210e9, 0.3, traction={"xmax": (0.0, 0.0, -1e6)})
    Aff, bf, free = oel.reduced_system_vec3(A, rhs, m, {"xmin": (0.0, 0.0, 0.0)})[:3]
    t = time.time()
    xf = spla.spsolve(Aff.tocsc(), bf, permc_spec="MMD_AT_PLUS_A")
    res = np.linalg.norm(bf - Aff @ xf) / np.linalg.norm(bf)
    print(f"SuperLU {time.time() - t:.1f} s, relative residual {res:.3e}")
    u = np.zeros(A.shape[0])
    u[free] = xf
    np.savez_compressed(OUT, u=u, residual=res)


if __name__ == "__main__":
    main()
