"""Batched theta-method transient heat on many P1 meshes at once.

The reference runs `transient_heat_tet` (tet_heat.py:197-274) one mesh at a
time: assemble K, F, C, factorize C + th h K once, then one triangular solve
per step.  `transient_heat_batch` runs the same recursion with the same
per-mesh semantics for a list of meshes: the meshes become one
block-diagonal system on the GPU (one pattern build, one assembly launch per
operator for all meshes), and cf_batch_theta_run steps every mesh together
with per-mesh CG scalars, one replayed hipGraph per time step and a launch
shape chosen by the autotuner (cloudfem_amd.autotune).

Per mesh i, like transient_heat_tet: k, rho_cp, f evaluated at element
centroids (a float, a callable(x, y, z), or here also an (M_i,) array of
per-element values, the data the reference feeds through its nearest-centroid
lookup, induction3d.py:111-123); Neumann q A/3 and Robin h A (1+d)/12,
h T_inf A/3 on the named boundaries (tet_heat.py:136-172); h = t_end /
ceil(t_end / dt); theta from `method`.  Dirichlet conditions are not part of
the batched path (the batched configuration has none): use
transient_heat_tet for them.
"""

import ctypes
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from .. import _native as N
from ..autotune import batch_spmv_config
from ..coef import centroids, eval_scalar
from ..device import DeviceMesh, ptr, require_gpu, stream_handle, to_dev, workspace
from ..mesh.device_ops import boundary_faces_multi_device
from ..mesh.faces import boundary_faces
from ..mesh.tet_mesh import TetMesh, box_mesh
from . import _transient

__all__ = ["transient_heat_batch", "jittered_box_batch", "last_batch_stats"]

last_batch_stats = {}


def _per_mesh(v, n_mesh, what):
    if isinstance(v, (list, tuple)):
        if len(v) != n_mesh:
            raise ValueError(f"{what}: {len(v)} entries for {n_mesh} meshes")
        return list(v)
    return [v] * n_mesh


def _element_values(v, mesh):
    """Per-element values of a float / callable (at centroids) / (M,) array."""
    if isinstance(v, np.ndarray) or (isinstance(v, (list, tuple)) and not callable(v)):
        a = np.asarray(v, dtype=float).ravel()
        if a.shape != (mesh.n_elements,):
            raise ValueError(f"per-element array of shape {a.shape} for "
                             f"{mesh.n_elements} elements")
        return a
    if not callable(v):
        return np.full(mesh.n_elements, float(v))
    return eval_scalar(v, centroids(mesh))


def _union(meshes, pinned=False):
    """One mesh of all meshes (node ids offset per mesh), int32 connectivity
    written in place (the device layer's type), the meshes' slices filled by
    a few host threads (NumPy releases the GIL in these copies).  pinned:
    the coordinates (as x, y, z rows) and the connectivity live in
    page-locked host tensors, returned as well, so that they upload by DMA
    without a staging copy."""
    offs = np.concatenate([[0], np.cumsum([m.n_nodes for m in meshes])]).astype(np.int64)
    if offs[-1] >= 2 ** 31:
        raise ValueError(f"transient_heat_batch: {offs[-1]} nodes exceed int32 node ids")
    eoffs = np.concatenate([[0], np.cumsum([m.n_elements for m in meshes])]).astype(np.int64)
    n, ne = int(offs[-1]), int(eoffs[-1])
    if pinned:
        xyz_t = torch.empty((3, n), dtype=torch.float64, pin_memory=True)
        conn_t = torch.empty((ne, 4), dtype=torch.int32, pin_memory=True)
        nodes, elems = xyz_t.numpy().T, conn_t.numpy()
    else:
        xyz_t = conn_t = None
        nodes = np.empty((n, 3), dtype=float)
        elems = np.empty((ne, 4), dtype=np.int32)

    def fill(i):
        m = meshes[i]
        nodes[offs[i]:offs[i + 1]] = np.asarray(m.nodes, dtype=float)
        np.add(np.asarray(m.elements)[:, :4], offs[i], out=elems[eoffs[i]:eoffs[i + 1]],
               casting="unsafe")

    workers = min(8, os.cpu_count() or 1, len(meshes))
    if workers > 1:
        with ThreadPoolExecutor(workers) as pool:
            list(pool.map(fill, range(len(meshes))))
    else:
        for i in range(len(meshes)):
            fill(i)
    um = TetMesh.wrap(nodes, elems)
    return (um, offs, xyz_t, conn_t) if pinned else (um, offs)


def _device_union(meshes, dev):
    """(union TetMesh, node offsets, DeviceMesh) with the union's arrays
    uploaded from page-locked memory (the meshes' ids were range-checked
    when the meshes were built; the union's are checked on the device)."""
    um, offs, xyz_t, conn_t = _union(meshes, pinned=True)
    xyz = xyz_t.to(dev, non_blocking=True)
    conn = conn_t.to(dev, non_blocking=True)
    if conn.numel() and bool(((conn < 0) | (conn >= um.n_nodes)).any()):
        raise IndexError("element node index out of range")
    return um, offs, DeviceMesh.from_device(xyz[0], xyz[1], xyz[2], conn)


def _needs_points(v):
    """Whether a coefficient spec (float, callable, or a tuple of them) is
    evaluated at points; constants only need the face count."""
    if isinstance(v, (list, tuple)):
        return any(callable(x) for x in v)
    return callable(v)


def _union_faces(um, dm, meshes, offs, names):
    """{name: (F, 3) union faces} from one device scan of the union's
    exterior faces (disjoint meshes: the union's exterior faces are the
    meshes' exterior faces, in mesh order then occurrence order)."""
    for name in names:
        um.boundaries[name] = np.concatenate(
            [np.asarray(m.boundaries[name], dtype=np.int64) + offs[i]
             for i, m in enumerate(meshes)])
    return boundary_faces_multi_device(um, names, conn=dm.conn)


def _faces_and_coef(meshes, offs, specs, pick, union=None):
    """Concatenated union faces and per-face values over all meshes."""
    if union is not None and specs and all(sp is specs[0] for sp in specs):
        faces, vals = [], []
        for name, v in (specs[0] or {}).items():
            fc = union[name]
            if fc.shape[0]:
                if _needs_points(v):
                    key = ("_centroids", name)
                    if key not in union:   # once per face set, shared by the callers
                        union[key] = union["_nodes"][fc].mean(axis=1)
                    cen = union[key]
                else:
                    cen = np.broadcast_to(0.0, (fc.shape[0], 3))
                faces.append(fc)
                vals.append(pick(v, cen))
        if not faces:
            return None, None
        return np.concatenate(faces), np.concatenate(vals)
    faces, vals = [], []
    for i, (m, spec) in enumerate(zip(meshes, specs)):
        for name, v in (spec or {}).items():
            fc = boundary_faces(m, name)
            if not fc.shape[0]:
                continue
            cen = np.asarray(m.nodes)[fc[:, :3]].mean(axis=1)
            faces.append(fc + offs[i])
            vals.append(pick(v, cen))
    if not faces:
        return None, None
    return np.concatenate(faces), np.concatenate(vals)


def _is_const(v):
    return isinstance(v, (int, float, np.integer, np.floating)) and not isinstance(v, bool)


def _element_coef(vals, meshes, dev):
    """Per-element values over the union on the device.  When every mesh's
    value is a constant, one value per mesh crosses the bus and the device
    expands it (64 jittered boxes: 6.8 M elements, 55 MB per coefficient
    not built on the host nor uploaded)."""
    if all(_is_const(v) for v in vals):
        per = torch.tensor([float(v) for v in vals], dtype=torch.float64, device=dev)
        counts = torch.tensor([m.n_elements for m in meshes], dtype=torch.int64, device=dev)
        return torch.repeat_interleave(per, counts,
                                       output_size=sum(m.n_elements for m in meshes))
    return to_dev(np.concatenate([_element_values(v, m) for v, m in zip(vals, meshes)]),
                  torch.float64, dev)


def transient_heat_batch(meshes, k=1.0, rho_cp=1.0, f=0.0, neumann=None, robin=None,
                         t_end=1.0, dt=0.1, T0=0.0, method="cn", precision="fp64", tol=1e-12,
                         maxit=10000, history=True, bc=None):
    """[(times (S+1,), T_hist (S+1, N_i) or T_final (N_i,)) for each mesh].

    k, rho_cp, f, T0, neumann, robin: one value for every mesh, or a list
    with one entry per mesh.  precision: "fp64" or "mixed" (FP32 CG inside
    FP64 iterative refinement; same residual bar).  tol: relative residual
    of every system's step solve (default 1e-12, so histories agree with the
    reference's LU recursion to ~1e-10).
    """
    if bc:
        raise NotImplementedError("transient_heat_batch: Dirichlet conditions are not "
                                  "supported; use transient_heat_tet")
    meshes = list(meshes)
    if not meshes:
        return []
    t_start = time.perf_counter()
    nm = len(meshes)
    theta = _transient.check_method(method)
    steps, h = _transient.step_count(t_end, dt)
    prec = {"fp64": N.PREC_FP64, "mixed": N.PREC_MIXED}[precision]
    ks, rcs, fs = (_per_mesh(v, nm, w) for v, w in ((k, "k"), (rho_cp, "rho_cp"), (f, "f")))
    dev = require_gpu()
    um, offs, dm = _device_union(meshes, dev)
    K = dm.matrix_values(N.OP_P1_STIFF, coef=_element_coef(ks, meshes, dev))
    C = dm.matrix_values(N.OP_P1_MASS, coef=_element_coef(rcs, meshes, dev))
    F = torch.zeros(dm.n_nodes, dtype=torch.float64, device=dev)
    if not all(_is_const(v) and float(v) == 0.0 for v in fs):
        f_e = _element_coef(fs, meshes, dev)
        if bool(torch.any(f_e != 0)):
            dm.add_rhs(N.RHS_P1_SOURCE, f_e, F)
    nspec = _per_mesh(neumann, nm, "neumann")
    rspec = _per_mesh(robin, nm, "robin")
    # same spec for every mesh: one device face scan per name over the union
    names = set()
    for spec in (nspec[0], rspec[0]):
        names |= set(spec or {})
    union = None
    if not isinstance(neumann, (list, tuple)) and not isinstance(robin, (list, tuple)) and names:
        union = _union_faces(um, dm, meshes, offs, sorted(names))
        union["_nodes"] = um.nodes
    faces, q = _faces_and_coef(meshes, offs, nspec, lambda v, cen: eval_scalar(v, cen), union)
    if faces is not None:
        dm.add_face_rhs(N.FACE_P1_FLUX, faces, q, F)
    faces, hv = _faces_and_coef(meshes, offs, rspec, lambda v, cen: eval_scalar(v[0], cen),
                                union)
    if faces is not None:
        _, ht = _faces_and_coef(meshes, offs, rspec,
                                lambda v, cen: eval_scalar(v[0], cen) * eval_scalar(v[1], cen),
                                union)
        dm.add_face_rhs(N.FACE_P1_FLUX, faces, ht, F)
        dm.add_subpattern(K, *dm.face_matrix(N.FACE_P1_ROBIN_K, faces, hv))
    A = C.clone()
    _transient.axpby(theta * h, K, 1.0, A)
    R = C.clone()
    _transient.axpby(-(1.0 - theta) * h, K, 1.0, R)
    T0s = _per_mesh(T0, nm, "T0")
    T = torch.empty(dm.n_nodes, dtype=torch.float64, device=dev)
    for i, (m, t0) in enumerate(zip(meshes, T0s)):
        seg = T[offs[i]:offs[i + 1]]
        if np.isscalar(t0):
            seg.fill_(float(t0))
        else:
            seg.copy_(to_dev(np.asarray(t0, dtype=float).ravel(), torch.float64, dev))
    pat = dm.pattern()
    sys_rows = offs.astype(np.int32)
    lanes, rows, tune = batch_spmv_config(sys_rows, pat.indptr, pat.indices, A)
    hist = (torch.empty((steps + 1, dm.n_nodes), dtype=torch.float64, device=dev)
            if history else None)
    lib = N.load()
    nb = ctypes.c_size_t()
    N.check(lib.cf_batch_workspace_bytes(nm, sys_rows.ctypes.data, pat.nnz, prec,
                                         ctypes.byref(nb)), "cf_batch_workspace_bytes")
    ws = workspace(nb.value, dev)
    st = N.CfBatchStats()
    torch.cuda.synchronize(dev)
    t_run = time.perf_counter()
    rc = lib.cf_batch_theta_run(nm, sys_rows.ctypes.data, ptr(pat.indptr), ptr(pat.indices),
                                pat.nnz, ptr(A), ptr(R), ptr(F), float(h), ptr(T), ptr(hist),
                                steps, prec, float(tol), int(maxit), lanes, rows, ptr(ws),
                                nb.value, ctypes.byref(st), stream_handle())
    torch.cuda.synchronize(dev)
    t_done = time.perf_counter()
    last_batch_stats.clear()
    last_batch_stats.update({name: getattr(st, name) for name, _ in N.CfBatchStats._fields_})
    last_batch_stats.update({"autotune": tune, "setup_s": t_run - t_start,
                             "run_s": t_done - t_run, "rows": int(dm.n_nodes),
                             "nnz": int(pat.nnz)})
    N.check(rc, "cf_batch_theta_run")
    times = np.cumsum(np.concatenate([[0.0], np.full(steps, h)]))
    times[-1] = t_end
    out_dev = hist if history else T
    host = out_dev.cpu().numpy()
    res = []
    for i in range(nm):
        a, b = offs[i], offs[i + 1]
        res.append((times, host[:, a:b].copy() if history else host[a:b].copy()))
    return res


def jittered_box_batch(n_meshes=64, lo=10, hi=40, size=0.1, jitter=0.2, seed=0):
    """Batch of box meshes [0, size]^3 with (nx, ny, nz) ~ U{lo..hi} per axis,
    interior nodes moved by U(-jitter h, +jitter h) per axis (boundary nodes
    fixed, so the faces stay planar), and per-element conductivities
    k ~ lognormal(0, 0.5).  Returns (meshes, k arrays)."""
    rng = np.random.default_rng(seed)
    dims = rng.integers(lo, hi + 1, size=(n_meshes, 3))
    meshes, ks = [], []
    for nx, ny, nz in dims:
        m = box_mesh(int(nx), int(ny), int(nz), xlim=(0.0, size), ylim=(0.0, size),
                     zlim=(0.0, size))
        on_bnd = np.zeros(m.n_nodes, dtype=bool)
        for ids in m.boundaries.values():
            on_bnd[np.asarray(ids)] = True
        hstep = size / np.array([nx, ny, nz], dtype=float)
        d = rng.uniform(-jitter, jitter, size=(m.n_nodes, 3)) * hstep
        nodes = np.asarray(m.nodes, dtype=float).copy()
        nodes[~on_bnd] += d[~on_bnd]
        meshes.append(TetMesh(nodes, m.elements, m.boundaries))
        ks.append(rng.lognormal(0.0, 0.5, size=m.n_elements))
    return meshes, ks
