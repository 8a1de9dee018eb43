"""Host-side checks of the native library (no GPU needed).

* libcloudfem_hip.so loads and exports every function include/cloudfem_hip.h
  declares; argument validation returns CF_ERR_ARG with a message.
* The compile-time reference tables of cf_tables.hpp (compiled here with
  g++ into a tiny dumper) equal the oracle's tables: Keast weights, shape
  values and gradients, the reference stiffness tensor R_ab, the exact Tet10
  mass, centroid gradients and the two triangle rules.
"""

import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from cloudfem_amd import _native as N
from oracle import tet10 as ot10
from oracle.elastic import TRI_L3

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cloudfem_hip.h")
CSRC = os.path.join(ROOT, "cloudfem_amd", "csrc")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(cf_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = N.load()
    declared = header_functions()
    assert len(declared) >= 20
    missing = [f for f in declared if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(N.EXPORTED) == declared
    assert lib.cf_version() >= 100


def test_argument_errors_without_gpu():
    lib = N.load()
    rc = lib.cf_pattern_sort(None, 10, 4, 5, None, 0, None, None, None)
    assert rc == N.CF_ERR_ARG and "cf_pattern_sort" in N.last_error()
    rc = lib.cf_pcg_solve(None, None, None, None, 0, 0, 1e-10, 10, None, 0, None, None)
    assert rc == N.CF_ERR_ARG
    with pytest.raises(N.CFError):
        N.check(rc, "cf_pcg_solve")
    A = N.CfMatrix(7, 10, 5, 1, 1, 1)
    nb = ctypes.c_size_t()
    assert lib.cf_pcg_workspace_bytes(ctypes.byref(A), 0, 0, ctypes.byref(nb)) == N.CF_ERR_ARG


DUMPER = r"""
#include <cstdio>
#include "cf_tables.hpp"
int main() {
  constexpr cf::Tet10Tables t = cf::build_tet10_tables();
  constexpr cf::FaceTables f = cf::build_face_tables();
  auto out = [](const double* p, int n) { for (int i = 0; i < n; ++i) std::printf("%.17g\n", p[i]); };
  out(t.w, 5); out(&t.N[0][0], 50); out(&t.D[0][0][0], 150); out(&t.R[0][0][0][0], 900);
  out(&t.M[0][0], 100); out(&t.Dc[0][0], 30);
  out(f.w6, 6); out(&f.N6[0][0], 36); out(f.w3, 3); out(&f.N3[0][0], 18);
  return 0;
}
"""


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("tables")
    src = d / "dump.cpp"
    src.write_text(DUMPER)
    exe = d / "dump"
    subprocess.run([gxx, "-std=c++17", "-O0", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    vals = np.array([float(v) for v in subprocess.run([str(exe)], check=True,
                                                      capture_output=True, text=True).stdout.split()])
    sizes = [5, 50, 150, 900, 100, 30, 6, 36, 3, 18]
    parts = np.split(vals, np.cumsum(sizes)[:-1])
    return dict(zip(["w", "N", "D", "R", "M", "Dc", "w6", "N6", "w3", "N3"], parts))


def test_tet10_tables_match_oracle(tables):
    Lq, wq = ot10.quadrature()
    Nq, dxi = ot10.ref_gradients()
    assert np.array_equal(tables["w"], wq)
    assert np.abs(tables["N"] - Nq.ravel()).max() < 1e-16
    assert np.abs(tables["D"] - dxi.ravel()).max() < 1e-15
    R = np.einsum("q,qam,qbn->abmn", wq, dxi, dxi)
    assert np.abs(tables["R"] - R.ravel()).max() < 1e-14
    assert np.abs(tables["M"] - ot10.MASS_REF.ravel()).max() < 1e-17
    _, dL = ot10.shape_functions(np.array([[0.25] * 4]))
    dc = np.stack([dL[0, :, 1] - dL[0, :, 0], dL[0, :, 2] - dL[0, :, 0],
                   dL[0, :, 3] - dL[0, :, 0]], axis=-1)
    assert np.array_equal(tables["Dc"], dc.ravel())


def test_face_tables_match_oracle(tables):
    assert np.array_equal(tables["w6"], ot10.TRI_W6)
    assert np.abs(tables["N6"] - ot10.FACE_N6.ravel()).max() < 1e-16
    assert np.allclose(tables["w3"], 1.0 / 3.0, rtol=0, atol=0)
    assert np.abs(tables["N3"] - ot10.face_shape6(TRI_L3).ravel()).max() < 1e-16
