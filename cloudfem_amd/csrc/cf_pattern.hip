// Node-pair CSR pattern and node incidence lists, built on the GPU.
//
// The reference gets its CSR pattern from scipy's coo_matrix(...).tocsr()
// over the per-element triplets (tet_poisson.py:68, tet10.py:162-164,
// elastic_tet10.py:208-210): sorted unique (row, col) pairs per row, duplicates
// merged, explicit zeros kept.  Here the same set is produced by one stable
// LSD radix sort of the 64-bit keys (row << sh | col) of all element node
// pairs; the sorted pair ids double as the contributor lists that the value
// kernels sum in a fixed (ascending element) order, so no atomics are needed
// and assembly is bitwise reproducible.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "cf_common.hpp"

namespace {

constexpr int kBlock = 256;

int bits_for(int64_t n) {
  int b = 1;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

__global__ void pair_keys(const int32_t* __restrict__ conn, int64_t n_pairs,
                          int k, int sh, uint64_t* __restrict__ keys,
                          uint32_t* __restrict__ ids) {
  const int kk = k * k;
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n_pairs;
       t += int64_t(gridDim.x) * blockDim.x) {
    int64_t e = t / kk;
    int ab = int(t - e * kk);
    int a = ab / k, b = ab - (ab / k) * k;
    uint64_t row = uint32_t(conn[e * k + a]);
    uint64_t col = uint32_t(conn[e * k + b]);
    keys[t] = (row << sh) | col;
    ids[t] = uint32_t(t);
  }
}

__global__ void run_heads(const uint64_t* __restrict__ keys, int64_t n,
                          int32_t* __restrict__ head) {
  for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < n;
       j += int64_t(gridDim.x) * blockDim.x)
    head[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1 : 0;
}

__global__ void store_count(const int32_t* __restrict__ incl, int64_t n,
                            int64_t* __restrict__ nnz) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *nnz = n > 0 ? int64_t(incl[n - 1]) : 0;
}

// One thread per sorted pair; run heads write the column and the contributor
// segment start of their stored entry.
__global__ void pattern_fill(const uint64_t* __restrict__ keys,
                             const int32_t* __restrict__ incl, int64_t n,
                             int sh, int64_t nnz, int32_t* __restrict__ indices,
                             int32_t* __restrict__ seg) {
  const uint64_t mask = (uint64_t(1) << sh) - 1;
  for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < n;
       j += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t key = keys[j];
    if (j > 0 && keys[j - 1] == key) continue;
    const int32_t s = incl[j] - 1;
    indices[s] = int32_t(key & mask);
    seg[s] = int32_t(j);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) seg[nnz] = int32_t(n);
}

// indptr[r] = slot of the first stored entry with row >= r: a lower bound
// over the sorted keys (that entry is always a run head), nnz past the end.
// One binary search per row, so empty and trailing rows cost nothing extra.
__global__ void pattern_rowptr(const uint64_t* __restrict__ keys,
                               const int32_t* __restrict__ incl, int64_t n, int sh,
                               int64_t n_rows, int64_t nnz, int32_t* __restrict__ indptr) {
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r <= n_rows;
       r += int64_t(gridDim.x) * blockDim.x) {
    int64_t lo = 0, hi = n;
    const uint64_t first = uint64_t(r) << sh;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (keys[mid] < first)
        lo = mid + 1;
      else
        hi = mid;
    }
    indptr[r] = lo < n ? incl[lo] - 1 : int32_t(nnz);
  }
}

__global__ void item_keys(const int32_t* __restrict__ targets, int64_t n,
                          uint32_t* __restrict__ keys,
                          uint32_t* __restrict__ ids) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x) {
    keys[t] = uint32_t(targets[t]);
    ids[t] = uint32_t(t);
  }
}

// offs[q] = first position whose sorted key is >= q (binary search per node).
__global__ void incidence_offsets(const uint32_t* __restrict__ keys, int64_t n,
                                  int64_t n_nodes, int32_t* __restrict__ offs) {
  for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q <= n_nodes;
       q += int64_t(gridDim.x) * blockDim.x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (int64_t(keys[mid]) < q)
        lo = mid + 1;
      else
        hi = mid;
    }
    offs[q] = int32_t(lo);
  }
}

// Layout of the pattern workspace; identical in every phase.
struct PatternWs {
  uint64_t* keys_in;
  uint64_t* keys_out;
  uint32_t* ids_in;
  int32_t* head;
  int32_t* incl;
  void* tmp;
  size_t tmp_bytes;
  size_t total;
};

int pattern_layout(void* base, int64_t n, int end_bit, PatternWs* w) {
  size_t sort_bytes = 0, scan_bytes = 0;
  CF_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint64_t*)nullptr,
                                   (uint64_t*)nullptr, (uint32_t*)nullptr,
                                   (uint32_t*)nullptr, size_t(n), 0, end_bit));
  CF_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, (int32_t*)nullptr,
                                 (int32_t*)nullptr, size_t(n),
                                 rocprim::plus<int32_t>()));
  cf::Carve c(base);
  w->keys_in = c.take<uint64_t>(n);
  w->keys_out = c.take<uint64_t>(n);
  w->ids_in = c.take<uint32_t>(n);
  w->head = c.take<int32_t>(n);
  w->incl = c.take<int32_t>(n);
  w->tmp_bytes = sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
  w->tmp = c.take<char>(w->tmp_bytes);
  w->total = c.used;
  return CF_OK;
}

}  // namespace

extern "C" int cf_pattern_workspace_bytes(int64_t n_elem, int32_t k,
                                          int64_t n_nodes, size_t* bytes) {
  CF_REQUIRE(bytes && n_elem >= 0 && k > 0 && n_nodes > 0,
             "cf_pattern_workspace_bytes: bad arguments");
  int64_t n = n_elem * k * k;
  CF_REQUIRE(n < (int64_t(1) << 31), "pattern too large for int32 ids: %lld pairs",
             (long long)n);
  PatternWs w;
  int rc = pattern_layout(nullptr, n, 2 * bits_for(n_nodes), &w);
  if (rc) return rc;
  *bytes = w.total;
  return CF_OK;
}

extern "C" int cf_pattern_sort(const int32_t* conn, int64_t n_elem, int32_t k,
                               int64_t n_nodes, void* ws, size_t ws_bytes,
                               uint32_t* contrib, int64_t* nnz_dev,
                               void* stream) {
  CF_REQUIRE(conn || n_elem == 0, "cf_pattern_sort: null connectivity");
  CF_REQUIRE(ws && contrib && nnz_dev && k > 0 && n_nodes > 0,
             "cf_pattern_sort: bad arguments");
  const int64_t n = n_elem * k * k;
  const int sh = bits_for(n_nodes);
  CF_REQUIRE(2 * sh <= 64, "cf_pattern_sort: too many nodes");
  PatternWs w;
  int rc = pattern_layout(ws, n, 2 * sh, &w);
  if (rc) return rc;
  if (ws_bytes < w.total) {
    cf::set_error("cf_pattern_sort: workspace %zu < %zu bytes", ws_bytes, w.total);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  if (n == 0) {
    store_count<<<1, 1, 0, s>>>(w.incl, 0, nnz_dev);
    CF_LAUNCHED();
    return CF_OK;
  }
  int grid = cf::stride_grid(n, kBlock);
  pair_keys<<<grid, kBlock, 0, s>>>(conn, n, k, sh, w.keys_in, w.ids_in);
  CF_LAUNCHED();
  size_t tb = w.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs(w.tmp, tb, w.keys_in, w.keys_out, w.ids_in,
                                   contrib, size_t(n), 0, 2 * sh, s));
  run_heads<<<grid, kBlock, 0, s>>>(w.keys_out, n, w.head);
  CF_LAUNCHED();
  tb = w.tmp_bytes;
  CF_HIP(rocprim::inclusive_scan(w.tmp, tb, w.head, w.incl, size_t(n),
                                 rocprim::plus<int32_t>(), s));
  store_count<<<1, 1, 0, s>>>(w.incl, n, nnz_dev);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_pattern_finalize(int64_t n_elem, int32_t k, int64_t n_nodes,
                                   int64_t nnz, const void* ws,
                                   int32_t* indptr, int32_t* indices,
                                   int32_t* seg, void* stream) {
  CF_REQUIRE(ws && indptr && seg && (indices || nnz == 0) && k > 0 && n_nodes > 0,
             "cf_pattern_finalize: bad arguments");
  const int64_t n = n_elem * k * k;
  const int sh = bits_for(n_nodes);
  PatternWs w;
  int rc = pattern_layout(const_cast<void*>(ws), n, 2 * sh, &w);
  if (rc) return rc;
  hipStream_t s = cf::as_stream(stream);
  int grid = cf::stride_grid(n > 0 ? n : 1, kBlock);
  pattern_fill<<<grid, kBlock, 0, s>>>(w.keys_out, w.incl, n, sh, nnz, indices, seg);
  CF_LAUNCHED();
  pattern_rowptr<<<cf::stride_grid(n_nodes + 1, kBlock), kBlock, 0, s>>>(w.keys_out, w.incl, n,
                                                                          sh, n_nodes, nnz,
                                                                          indptr);
  CF_LAUNCHED();
  return CF_OK;
}

// ---------------------------------------------------------------- incidence

namespace {
struct IncidenceWs {
  uint32_t* keys_in;
  uint32_t* keys_out;
  uint32_t* ids_in;
  void* tmp;
  size_t tmp_bytes;
  size_t total;
};

int incidence_layout(void* base, int64_t n, int end_bit, IncidenceWs* w) {
  size_t sort_bytes = 0;
  CF_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint32_t*)nullptr,
                                   (uint32_t*)nullptr, (uint32_t*)nullptr,
                                   (uint32_t*)nullptr, size_t(n), 0, end_bit));
  cf::Carve c(base);
  w->keys_in = c.take<uint32_t>(n);
  w->keys_out = c.take<uint32_t>(n);
  w->ids_in = c.take<uint32_t>(n);
  w->tmp_bytes = sort_bytes;
  w->tmp = c.take<char>(sort_bytes);
  w->total = c.used;
  return CF_OK;
}
}  // namespace

extern "C" int cf_incidence_workspace_bytes(int64_t n_items, int64_t n_nodes,
                                            size_t* bytes) {
  CF_REQUIRE(bytes && n_items >= 0 && n_nodes > 0 && n_items < (int64_t(1) << 31),
             "cf_incidence_workspace_bytes: bad arguments");
  IncidenceWs w;
  int rc = incidence_layout(nullptr, n_items, bits_for(n_nodes + 1), &w);
  if (rc) return rc;
  *bytes = w.total;
  return CF_OK;
}

extern "C" int cf_incidence_build(const int32_t* targets, int64_t n_items,
                                  int64_t n_nodes, void* ws, size_t ws_bytes,
                                  uint32_t* perm, int32_t* offs, void* stream) {
  CF_REQUIRE(ws && offs && n_nodes > 0 && (n_items == 0 || (targets && perm)),
             "cf_incidence_build: bad arguments");
  IncidenceWs w;
  const int end_bit = bits_for(n_nodes + 1);
  int rc = incidence_layout(ws, n_items, end_bit, &w);
  if (rc) return rc;
  if (ws_bytes < w.total) {
    cf::set_error("cf_incidence_build: workspace %zu < %zu bytes", ws_bytes, w.total);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  int grid = cf::stride_grid(n_items > 0 ? n_items : 1, kBlock);
  if (n_items > 0) {
    item_keys<<<grid, kBlock, 0, s>>>(targets, n_items, w.keys_in, w.ids_in);
    CF_LAUNCHED();
    size_t tb = w.tmp_bytes;
    CF_HIP(rocprim::radix_sort_pairs(w.tmp, tb, w.keys_in, w.keys_out, w.ids_in,
                                     perm, size_t(n_items), 0, end_bit, s));
  }
  incidence_offsets<<<cf::stride_grid(n_nodes + 1, kBlock), kBlock, 0, s>>>(w.keys_out, n_items,
                                                                           n_nodes, offs);
  CF_LAUNCHED();
  return CF_OK;
}
