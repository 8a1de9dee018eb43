"""cf_pcg_solve_noalloc (static scratch array, no allocation) against
cf_pcg_solve (caller workspace) on the same device systems: the 20^3 Kuhn
box Poisson problem (f = 3 pi^2 sin sin sin, T = 0 on all faces) and then a
smaller 8x6x5 box with a variable conductivity, in that order, so the second
call runs on a scratch array that still holds the first call's data.  Both
entry points run the same kernels in the same order, so the solutions must
agree bit for bit."""

import ctypes

import numpy as np
import pytest

from tests._helpers import FACES

pytestmark = pytest.mark.gpu

SRC = lambda x, y, z: 3 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)  # noqa: E731


def _system(dims, k):
    import torch
    from cloudfem_amd import _native as N
    from cloudfem_amd.coef import dirichlet_scalar
    from cloudfem_amd.device import to_dev
    from cloudfem_amd.fem.tet_poisson import _assemble_device
    from cloudfem_amd.mesh import box_mesh
    from cloudfem_amd.solver import DeviceMatrix
    m = box_mesh(*dims)
    dm, vals, rhs = _assemble_device(m, k, SRC)
    mask, fv = dirichlet_scalar(m, {f: 0.0 for f in FACES})
    x0 = to_dev(np.where(mask, fv, 0.0), torch.float64, dm.dev)
    fixed = to_dev(mask.astype(np.uint8), torch.uint8, dm.dev)
    return DeviceMatrix(N.FMT_CSR, dm.pattern(), vals), rhs, x0, fixed


@pytest.mark.parametrize("dims,k", [((20, 20, 20), 1.0),
                                    ((8, 6, 5), lambda x, y, z: 1.0 + x + 2.0 * y * z)])
def test_noalloc_matches_workspace_solve(gpu, dims, k):
    import torch
    from cloudfem_amd.solver import pcg, pcg_noalloc
    A, rhs, x0, fixed = _system(dims, k)
    x_ws = x0.clone()
    st_ws = pcg(A, rhs, x_ws, fixed, precision="fp64", precond="jacobi", tol=1e-12)
    x_na = x0.clone()
    st_na = pcg_noalloc(A, rhs, x_na, fixed, tol=1e-12)
    torch.cuda.synchronize()
    print(f"{dims}: {st_ws['iterations']} / {st_na['iterations']} iterations, residual "
          f"{st_ws['rel_residual']:.2e} / {st_na['rel_residual']:.2e}")
    assert st_na["iterations"] == st_ws["iterations"]
    assert torch.equal(x_na, x_ws)
    assert st_na["rel_residual"] == st_ws["rel_residual"] <= 1e-12


def test_noalloc_rejects_oversized_and_nb3(gpu):
    """Limits are checked on the host before any device work."""
    import torch
    from cloudfem_amd import _native as N
    from cloudfem_amd.device import ptr
    lib = N.load()
    dummy = torch.zeros(4, dtype=torch.int32, device="cuda")
    vals = torch.zeros(4, dtype=torch.float64, device="cuda")
    st = N.CfPcgStats()
    big = N.CfMatrix(N.FMT_CSR, 300000, 1, dummy.data_ptr(), dummy.data_ptr(), vals.data_ptr())
    rc = lib.cf_pcg_solve_noalloc(ctypes.byref(big), ptr(vals), ptr(vals), None, 1e-10, 10,
                                  ctypes.byref(st), None)
    assert rc == 1 and b"exceed" in lib.cf_last_error()
    nb3 = N.CfMatrix(N.FMT_NB3, 1, 1, dummy.data_ptr(), dummy.data_ptr(), vals.data_ptr())
    rc = lib.cf_pcg_solve_noalloc(ctypes.byref(nb3), ptr(vals), ptr(vals), None, 1e-10, 10,
                                  ctypes.byref(st), None)
    assert rc == 1 and b"CSR" in lib.cf_last_error()
