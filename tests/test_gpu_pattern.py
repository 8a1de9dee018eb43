"""The node-pair pattern builder (cf_pattern_sort / cf_pattern_finalize)
against a NumPy restatement of what scipy's coo->csr does with the element
triplets (tet_poisson.py:68, elastic_tet10.py:208-210): CSR indptr/indices,
and the contributor lists the value kernels sum, i.e. the pair ids
e k^2 + a k + b stably sorted by (row, col), with the first contributor of
every stored entry.  Synthetic connectivities give rows of every length the
per-row register sort handles (up to 64, 128, 256, 512 and 1,024 pairs) and
rows beyond it, which switch the whole pattern to the global radix sort; all
must give identical arrays."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _expected(conn, n_nodes):
    conn = np.asarray(conn, dtype=np.int64)
    m, k = conn.shape
    rows = np.repeat(conn, k, axis=1).ravel()
    cols = np.tile(conn, (1, k)).ravel()
    key = rows * n_nodes + cols
    order = np.argsort(key, kind="stable")
    skey = key[order]
    head = np.ones(skey.size, dtype=bool)
    head[1:] = skey[1:] != skey[:-1]
    uk = skey[head]
    r, c = np.divmod(uk, n_nodes)
    indptr = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n_nodes), out=indptr[1:])
    seg = np.append(np.nonzero(head)[0], skey.size)
    return (indptr.astype(np.int32), c.astype(np.int32), order.astype(np.uint32),
            seg.astype(np.int32))


def _hub_connectivity(valences, k, n_nodes, seed):
    """Elements of k distinct nodes; hub node h lies in valences[h] of them,
    the other slots are filled from the non-hub nodes at random."""
    rng = np.random.default_rng(seed)
    hubs = len(valences)
    els = []
    for h, v in enumerate(valences):
        for _ in range(v):
            others = rng.choice(np.arange(hubs, n_nodes), k - 1, replace=False)
            row = np.insert(others, rng.integers(0, k), h)
            els.append(row)
    # plain elements over the non-hub nodes
    for _ in range(60):
        els.append(rng.choice(np.arange(hubs, n_nodes), k, replace=False))
    els = np.array(els, dtype=np.int32)
    return els[rng.permutation(len(els))]


def _check(conn, n_nodes):
    import torch
    from cloudfem_amd.device import build_pattern
    dev = torch.device("cuda", 0)
    pat = build_pattern(torch.from_numpy(conn).to(dev), n_nodes, dev)
    ptr, idx, contrib, seg = _expected(conn, n_nodes)
    assert pat.nnz == idx.size
    assert np.array_equal(pat.indptr.cpu().numpy(), ptr)
    assert np.array_equal(pat.indices.cpu().numpy(), idx)
    assert np.array_equal(pat.contrib.cpu().numpy().view(np.uint32), contrib)
    assert np.array_equal(pat.seg.cpu().numpy(), seg)


@pytest.mark.parametrize("k", [4, 10])
def test_rows_of_every_register_sort_width(gpu, k):
    # pairs per hub row: v k; with k = 10: 1000, 500, 250, 120, 60 and the
    # plain nodes' short rows
    valences = [100, 50, 25, 12, 6] if k == 10 else [250, 120, 60, 30, 15]
    conn = _hub_connectivity(valences, k, 400, seed=3)
    assert (np.bincount(conn.ravel()) * k).max() <= 1024
    _check(conn, 400)


def test_rows_beyond_the_register_sort_use_the_global_sort(gpu):
    conn = _hub_connectivity([110, 40, 7], 10, 500, seed=4)
    assert (np.bincount(conn.ravel()) * 10).max() > 1024
    _check(conn, 500)


def test_box_meshes_and_unused_nodes(gpu):
    from cloudfem_amd.mesh import box_mesh, box_mesh_tet10
    for m in (box_mesh(3, 2, 4), box_mesh_tet10(4, 3, 2)):
        _check(np.asarray(m.elements, dtype=np.int32), m.n_nodes)
    # nodes that no element uses (empty rows) at the start, middle and end
    m = box_mesh_tet10(2, 2, 2)
    els = np.asarray(m.elements, dtype=np.int32)
    remap = np.arange(m.n_nodes) * 2 + 1
    _check(remap[els].astype(np.int32), 2 * m.n_nodes + 3)
