"""Writes cantilever_80x10x10_superlu.npz for the 80x10x10 Tet10 cantilever
(0.4 x 0.05 x 0.05 m, E = 210 GPa, nu = 0.3, clamped at x = 0, 1 MPa
traction in -z on x = 0.4), on the oracle's reduced system (the spsolve of
elastic3d.py:220-230):

  u_exact     the solution of that FP64 system to ~1e-13 relative residual:
              SuperLU's solution refined with residuals accumulated in
              extended precision (x86 long double) and corrections from the
              same LU factors;
  d_spsolve   spsolve's own solution minus u_exact (float32: the difference
              is ~1e-10 of u, so float32 keeps it to ~1e-17 of u).

SuperLU needs minutes and several GB here, too long for a GPU test, so the
vectors are stored.  Run from the repository root:
    python tests/golden/make_cantilever_80x10x10_superlu.py
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


def residual_long(A, x, b):
    """b - A x with products and sums in long double."""
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    prod = A.data.astype(np.longdouble) * x[A.indices]
    s = np.zeros(A.shape[0], dtype=np.longdouble)
    np.add.at(s, rows, prod)
    return b.astype(np.longdouble) - s


def main():
    lims = dict(xlim=(0.0, 0.4), ylim=(0.0, 0.05), zlim=(0.0, 0.05))
    m = om.box_mesh_tet10(80, 10, 10, **lims)
    A, rhs = oel.assemble_elasticity_tet10(m, 210e9, 0.3, traction={"xmax": (0.0, 0.0, -1e6)})
    Aff, bf, free = oel.reduced_system_vec3(A, rhs, m, {"xmin": (0.0, 0.0, 0.0)})[:3]
    Aff = Aff.tocsr()
    t = time.time()
    x_sp = spla.spsolve(Aff.tocsc(), bf, permc_spec="MMD_AT_PLUS_A")
    print(f"spsolve {time.time() - t:.1f} s, relative residual "
          f"{np.linalg.norm(bf - Aff @ x_sp) / np.linalg.norm(bf):.3e}", flush=True)
    t = time.time()
    lu = spla.splu(Aff.tocsc(), permc_spec="MMD_AT_PLUS_A")
    x = lu.solve(bf).astype(np.longdouble)
    for it in range(4):
        r = residual_long(Aff, x, bf)
        print(f"refinement {it}: relative residual "
              f"{float(np.linalg.norm(r.astype(float)) / np.linalg.norm(bf)):.3e}", flush=True)
        x = x + lu.solve(r.astype(np.float64)).astype(np.longdouble)
    x_ex = x.astype(np.float64)
    print(f"splu + refinement {time.time() - t:.1f} s; spsolve vs refined "
          f"{np.linalg.norm(x_sp - x_ex) / np.linalg.norm(x_ex):.3e}")
    u = np.zeros(A.shape[0])
    u[free] = x_ex
    d = np.zeros(A.shape[0], dtype=np.float32)
    d[free] = (x_sp - x_ex).astype(np.float32)
    np.savez_compressed(OUT, u_exact=u, d_spsolve=d)


if __name__ == "__main__":
    main()
