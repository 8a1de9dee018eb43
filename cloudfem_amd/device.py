"""Device-resident mesh data and the assembly pipeline on one GPU.

PyTorch provides device memory and the current HIP stream; every byte of
arithmetic runs in libcloudfem_hip through the C ABI.  A DeviceMesh holds
(in HBM): node coordinates as FP64 structure-of-arrays, int32
connectivity, the per-element geometry record (inverse Jacobian + volume),
the node-pair CSR pattern with its contributor lists, and the node incidence
lists used by the deterministic right-hand-side gathers.
"""

import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _native as N


def require_gpu():
    if not torch.cuda.is_available():
        raise N.CFError("cloudfem_amd needs a ROCm GPU (torch.cuda.is_available() "
                        "is False); there is no CPU fallback")
    N.load()
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def to_dev(a, dtype, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype)


def workspace(nbytes, dev):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)


class NodePattern:
    """Node-pair CSR pattern + contributor lists on the device."""

    def __init__(self, indptr, indices, seg, contrib, n_rows, nnz, k):
        self.indptr, self.indices, self.seg, self.contrib = indptr, indices, seg, contrib
        self.n_rows, self.nnz, self.k = n_rows, nnz, k


def build_pattern(conn, n_nodes, dev):
    """Pattern of the node graph of `conn` (device int32 (m, k))."""
    lib = N.load()
    m, k = int(conn.shape[0]), int(conn.shape[1])
    nb = ctypes.c_size_t()
    N.check(lib.cf_pattern_workspace_bytes(m, k, n_nodes, ctypes.byref(nb)),
            "cf_pattern_workspace_bytes")
    ws = workspace(nb.value, dev)
    contrib = torch.empty(max(m * k * k, 1), dtype=torch.int32, device=dev)
    nnz_dev = torch.zeros(1, dtype=torch.int64, device=dev)
    s = stream_handle()
    N.check(lib.cf_pattern_sort(ptr(conn), m, k, n_nodes, ptr(ws), nb.value, ptr(contrib),
                                ptr(nnz_dev), s), "cf_pattern_sort")
    nnz = int(nnz_dev.item())
    indptr = torch.empty(n_nodes + 1, dtype=torch.int32, device=dev)
    indices = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    seg = torch.empty(nnz + 1, dtype=torch.int32, device=dev)
    N.check(lib.cf_pattern_finalize(m, k, n_nodes, nnz, ptr(ws), ptr(indptr), ptr(indices),
                                    ptr(seg), s), "cf_pattern_finalize")
    return NodePattern(indptr, indices[:nnz], seg, contrib, n_nodes, nnz, k)


def build_incidence(items, n_nodes, dev):
    """(offs (n_nodes+1), perm) such that node i owns items perm[offs[i]:offs[i+1]]
    (flat item ids of `items`, ascending)."""
    lib = N.load()
    n = int(items.numel())
    nb = ctypes.c_size_t()
    N.check(lib.cf_incidence_workspace_bytes(n, n_nodes, ctypes.byref(nb)),
            "cf_incidence_workspace_bytes")
    ws = workspace(nb.value, dev)
    perm = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    offs = torch.empty(n_nodes + 1, dtype=torch.int32, device=dev)
    N.check(lib.cf_incidence_build(ptr(items), n, n_nodes, ptr(ws), nb.value, ptr(perm),
                                   ptr(offs), stream_handle()), "cf_incidence_build")
    return offs, perm


class DeviceMesh:
    """A TetMesh/Tet10Mesh (duck-typed: .nodes (N,3), .elements (M,k)) in HBM."""

    def __init__(self, mesh, dev=None):
        self.dev = dev or require_gpu()
        nodes = np.asarray(mesh.nodes, dtype=np.float64)
        els = np.asarray(mesh.elements)
        self.n_nodes = int(nodes.shape[0])
        self.n_elem = int(els.shape[0])
        self.k = int(els.shape[1])
        if self.k not in (4, 10):
            raise ValueError(f"elements must have 4 or 10 nodes, got {self.k}")
        if self.n_nodes >= 2 ** 31 or (els.dtype.kind != "i" and els.dtype.kind != "u"):
            raise ValueError("element node ids must be integers below 2**31")
        xyz = to_dev(nodes.T.copy(), torch.float64, self.dev)
        self.x, self.y, self.z = xyz[0], xyz[1], xyz[2]
        if self.n_elem and els.dtype.itemsize > 4 and (els.min() < 0 or
                                                       els.max() >= self.n_nodes):
            raise IndexError("element node index out of range")
        self.conn = to_dev(np.ascontiguousarray(els, dtype=np.int32), torch.int32, self.dev)
        # int32 ids (no host pass above): checked on the device before any kernel uses them
        if self.n_elem and els.dtype.itemsize <= 4 and bool(
                ((self.conn < 0) | (self.conn >= self.n_nodes)).any()):
            raise IndexError("element node index out of range")
        self._geo = None
        self._pattern = None
        self._incidence = None

    @classmethod
    def from_device(cls, x, y, z, conn):
        """Wrap coordinates/connectivity already resident in HBM."""
        self = cls.__new__(cls)
        self.dev = x.device
        self.n_nodes = int(x.numel())
        self.n_elem, self.k = int(conn.shape[0]), int(conn.shape[1])
        self.x, self.y, self.z, self.conn = x, y, z, conn
        self._geo = self._pattern = self._incidence = self._mf = None
        return self

    def add_face_rhs_device(self, op, faces, coef, rhs, incidence=None):
        """rhs += face loads for device-resident faces (F, nf) and coefficients."""
        lib = N.load()
        offs, perm = incidence or build_incidence(faces.reshape(-1), self.n_nodes, self.dev)
        N.check(lib.cf_face_rhs(op, ptr(self.x), ptr(self.y), ptr(self.z), ptr(faces),
                                faces.shape[0], faces.shape[1], ptr(offs), ptr(perm),
                                self.n_nodes, ptr(coef), ptr(rhs), stream_handle()),
                "cf_face_rhs")

    def corner_conn(self):
        return self.conn if self.k == 4 else self.conn[:, :4].contiguous()

    def geometry(self):
        """(M, 10) FP64: inverse corner Jacobian (row-major) and volume."""
        if self._geo is None:
            lib = N.load()
            geo = torch.empty((max(self.n_elem, 1), 10), dtype=torch.float64, device=self.dev)
            bad = torch.zeros(1, dtype=torch.int32, device=self.dev)
            N.check(lib.cf_tet_geometry(ptr(self.x), ptr(self.y), ptr(self.z), ptr(self.conn),
                                        self.n_elem, self.k, ptr(geo), ptr(bad),
                                        stream_handle()), "cf_tet_geometry")
            nbad = int(bad.item())
            if nbad:
                # the reference's np.linalg.solve/inv raise on a singular Jacobian
                raise np.linalg.LinAlgError(f"Singular matrix: {nbad} degenerate element(s)")
            self._geo = geo
        return self._geo

    def pattern(self):
        if self._pattern is None:
            self._pattern = build_pattern(self.conn, self.n_nodes, self.dev)
        return self._pattern

    def incidence(self):
        if self._incidence is None:
            self._incidence = build_incidence(self.conn.reshape(-1), self.n_nodes, self.dev)
        return self._incidence

    def matfree_arrays(self):
        """(conn, geo, offs, perm) in node-local element order for the
        matrix-free operator: elements stably sorted by their first corner
        (one radix sort), so a tile of consecutive elements touches few nodes
        and its per-node partial sums are long.  The operator is a sum over
        elements, so the order changes only rounding."""
        if getattr(self, "_mf", None) is None:
            _, order = build_incidence(self.conn[:, 0].contiguous(), self.n_nodes, self.dev)
            order = order[:self.n_elem].long()
            # the sort returns element ids; check before they index anything
            # (one read-back for both bounds)
            lo, hi = (torch.stack(torch.aminmax(order)).tolist() if order.numel() else (0, -1))
            if order.numel() != self.n_elem or lo < 0 or hi >= self.n_elem:
                raise N.CFError(f"element order out of range [{lo}, {hi}] for {self.n_elem}")
            conn = self.conn.index_select(0, order).contiguous()
            geo = self.geometry().index_select(0, order).contiguous()
            offs, perm = build_incidence(conn.reshape(-1), self.n_nodes, self.dev)
            self._mf = (conn, geo, offs, perm)
        return self._mf

    def matfree_plan(self):
        """(plan buffer, counts (3,) int64 host array) of the chunked
        matrix-free Tet10 operator (cf_mf_t10_plan_build): Morton-ordered
        chunks of 256 elements, node sums in LDS."""
        if getattr(self, "_mfplan", None) is None:
            lib = N.load()
            offs, _ = self.incidence()
            bbox = torch.stack([self.x.min(), self.y.min(), self.z.min(),
                                self.x.max(), self.y.max(), self.z.max()]).cpu().numpy()
            nb = ctypes.c_size_t()
            N.check(lib.cf_mf_t10_plan_bytes(self.n_elem, ctypes.byref(nb)), "cf_mf_t10_plan_bytes")
            plan = workspace(nb.value, self.dev)
            counts = np.zeros(3, dtype=np.int64)
            N.check(lib.cf_mf_t10_plan_build(self.n_nodes, self.n_elem, ptr(self.conn),
                                             ptr(self.geometry()), ptr(self.x), ptr(self.y),
                                             ptr(self.z), ptr(offs),
                                             bbox.astype(np.float64).ctypes.data, ptr(plan),
                                             nb.value, counts.ctypes.data, stream_handle()),
                    "cf_mf_t10_plan_build")
            self._mfplan = (plan, counts)
        return self._mfplan

    # ----------------------------------------------------------- assembly

    def matrix_values(self, op, coef=None, lam=0.0, mu=0.0):
        """Values of an element operator on the node pattern (see
        cf_assemble_matrix): scalar -> (nnz,), *_ELAST -> NB3 (9 nnz,),
        P1_VMASS -> (3 nnz,)."""
        lib = N.load()
        pat = self.pattern()
        geo = self.geometry()
        width = {N.OP_T10_ELAST: 9, N.OP_P1_ELAST: 9, N.OP_P1_VMASS: 3}.get(op, 1)
        vals = torch.empty(max(width * pat.nnz, 1), dtype=torch.float64, device=self.dev)
        N.check(lib.cf_assemble_matrix(op, ptr(pat.indptr), ptr(pat.indices), ptr(pat.seg),
                                       ptr(pat.contrib), ptr(self.conn), self.n_nodes,
                                       pat.nnz, self.k,
                                       ptr(geo), ptr(coef), float(lam), float(mu), ptr(vals),
                                       stream_handle()), "cf_assemble_matrix")
        return vals[:width * pat.nnz]

    def add_rhs(self, op, coef, rhs, lam=0.0, mu=0.0):
        lib = N.load()
        offs, perm = self.incidence()
        N.check(lib.cf_assemble_rhs(op, ptr(offs), ptr(perm), self.n_nodes, self.k,
                                    ptr(self.geometry()), ptr(coef), float(lam), float(mu),
                                    ptr(rhs), stream_handle()), "cf_assemble_rhs")

    def add_face_rhs(self, op, faces_np, coef_np, rhs):
        """rhs += face loads of `faces_np` (F, nf) with per-face coefficients."""
        faces_np = np.asarray(faces_np)
        if faces_np.shape[0] == 0:
            return
        lib = N.load()
        faces = to_dev(faces_np.astype(np.int32), torch.int32, self.dev)
        coef = to_dev(np.asarray(coef_np, dtype=np.float64).ravel(), torch.float64, self.dev)
        offs, perm = build_incidence(faces.reshape(-1), self.n_nodes, self.dev)
        N.check(lib.cf_face_rhs(op, ptr(self.x), ptr(self.y), ptr(self.z), ptr(faces),
                                faces.shape[0], faces.shape[1], ptr(offs), ptr(perm),
                                self.n_nodes, ptr(coef), ptr(rhs), stream_handle()),
                "cf_face_rhs")

    def face_matrix(self, op, faces_np, coef_np):
        """(NodePattern, values) of a Robin face matrix over `faces_np`."""
        lib = N.load()
        faces = to_dev(np.asarray(faces_np).astype(np.int32), torch.int32, self.dev)
        coef = to_dev(np.asarray(coef_np, dtype=np.float64).ravel(), torch.float64, self.dev)
        pat = build_pattern(faces, self.n_nodes, self.dev)
        vals = torch.empty(max(pat.nnz, 1), dtype=torch.float64, device=self.dev)
        N.check(lib.cf_face_matrix(op, ptr(self.x), ptr(self.y), ptr(self.z), ptr(faces),
                                   faces.shape[1], ptr(pat.indptr), ptr(pat.indices),
                                   ptr(pat.seg), ptr(pat.contrib), self.n_nodes, pat.nnz,
                                   ptr(coef), ptr(vals), stream_handle()), "cf_face_matrix")
        return pat, vals[:pat.nnz]

    def add_subpattern(self, vals, sub_pat, sub_vals, alpha=1.0):
        """vals (on this mesh's scalar pattern) += alpha * sub (a sub-pattern)."""
        lib = N.load()
        pat = self.pattern()
        missing = torch.zeros(1, dtype=torch.int32, device=self.dev)
        N.check(lib.cf_csr_add_subpattern(self.n_nodes, ptr(pat.indptr), ptr(pat.indices),
                                          ptr(vals), ptr(sub_pat.indptr), ptr(sub_pat.indices),
                                          ptr(sub_vals), float(alpha), ptr(missing),
                                          stream_handle()), "cf_csr_add_subpattern")
        if int(missing.item()):
            raise N.CFError("face pattern is not contained in the element pattern")


# ---------------------------------------------------------------- export

def scalar_csr(pat, vals):
    """scipy csr_matrix (N, N) from a node pattern and its values."""
    n = pat.n_rows
    A = sp.csr_matrix((vals.cpu().numpy(), pat.indices.cpu().numpy(),
                       pat.indptr.cpu().numpy()), shape=(n, n))
    return A


def vector_csr(pat, vals, dense3=True):
    """scipy csr_matrix (3N, 3N) of an NB3 (dense3) or component-diagonal
    vector matrix; the DOF-level pattern is expanded on the device."""
    lib = N.load()
    n = pat.n_rows
    dev = vals.device
    width = 9 if dense3 else 3
    dptr = torch.empty(3 * n + 1, dtype=torch.int32, device=dev)
    didx = torch.empty(max(width * pat.nnz, 1), dtype=torch.int32, device=dev)
    N.check(lib.cf_expand_pattern3(n, ptr(pat.indptr), ptr(pat.indices), 1 if dense3 else 0,
                                   ptr(dptr), ptr(didx), stream_handle()), "cf_expand_pattern3")
    return sp.csr_matrix((vals.cpu().numpy(), didx[:width * pat.nnz].cpu().numpy(),
                          dptr.cpu().numpy()), shape=(3 * n, 3 * n))


def upload_csr(A, dev):
    """DeviceMatrix (scalar CSR) of a host scipy matrix (canonicalised)."""
    from .solver import DeviceMatrix
    A = sp.csr_matrix(A)
    if not A.has_canonical_format:
        A = A.copy()
        A.sum_duplicates()
    n = A.shape[0]
    pat = NodePattern(to_dev(A.indptr.astype(np.int32), torch.int32, dev),
                      to_dev(A.indices.astype(np.int32), torch.int32, dev), None, None, n,
                      int(A.nnz), 1)
    return DeviceMatrix(N.FMT_CSR, pat, to_dev(A.data.astype(np.float64), torch.float64, dev))
