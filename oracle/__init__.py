"""CPU oracle for the tetrahedral assembly + sparse SPD solve path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package
(`cloudfem_amd`) may import, call or link anything in this directory.
Only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of
`bench.py` use it, and only as the checker (or the timed CPU baseline),
never as the thing that is measured or shipped.

What it is: a NumPy/SciPy restatement, written for this project, of the
algorithms of CloudFEM's hot path.  Every function cites the reference
file:line it follows (paths relative to the CloudFEM repository root).
Assembly keeps the reference's per-element structure and its
`scipy.sparse.coo_matrix(...).tocsr()` merge, so the CSR structure it
returns is the one the reference returns; solves use the same SciPy
SuperLU routines (`spsolve`, `factorized`).

How it is pinned (tests/test_oracle_*.py):
  * mesh numbering  -- brute-force enumeration on tiny boxes, counts,
                       volumes tiling the box, lexicographic edge ranks;
  * CSR structure   -- brute-force set of element node pairs;
  * element values  -- closed forms (P1 V*G G^T, V(1+d_ij)/20, the Tet10
                       1/420 mass table), an independent 4-point degree-2
                       rule and a degree-5 collapsed Gauss rule for the
                       Tet10 stiffness, invariants (K 1 = 0, rigid-body
                       null space, symmetry);
  * solutions       -- the reference's analytic validations: affine
                       exactness, O(h^2) P1 rates, Tet10 orders, patch
                       tests, cantilever vs Euler-Bernoulli, slab series,
                       adiabatic conservation, energy balances;
  * sphere fixture  -- tests/golden/sphere_mesh_0p5_0p22.npz, written by
                       tests/golden/make_sphere_fixture.py from the
                       reference's own mesh generator (the only function
                       on this path that no closed form pins).
Functions with no such check are marked "parity unpinned" in their
docstring.
"""
