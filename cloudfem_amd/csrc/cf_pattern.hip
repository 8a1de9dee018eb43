// Node-pair CSR pattern and node incidence lists, built on the GPU.
//
// The reference gets its CSR pattern from scipy's coo_matrix(...).tocsr()
// over the per-element triplets (tet_poisson.py:68, tet10.py:162-164,
// elastic_tet10.py:208-210): sorted unique (row, col) pairs per row, duplicates
// merged, explicit zeros kept.  Here the pairs of one row are the pairs of the
// elements that hold its node: the node incidence (one radix sort of the
// element items by node) lists them, and one wave per row sorts the row's
// (column, pair id) keys in registers and counts the distinct columns.  The
// sorted pair ids double as the contributor lists that the value kernels sum
// in a fixed (ascending element) order, so no atomics are needed and assembly
// is bitwise reproducible.  Rows too long for the register sort (more than
// kRowCap pairs: a node in over 100 Tet10 elements) switch the whole pattern
// to one stable LSD radix sort of the 64-bit keys (row << sh | col) of all
// element node pairs, which gives the same arrays.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "cf_common.hpp"

// Device radix sorts run rocPRIM's onesweep algorithm at every size above a
// single block: its default falls back to a merge sort up to 1 M items (about
// 20 short launches per sort, against one per 8-bit digit here).
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;

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
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(nullptr, sort_bytes, (uint64_t*)nullptr,
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

// ---- row-wise pattern -----------------------------------------------------
//
// Row i holds the k pairs (i, conn[e][b]) of each of the d_i elements e that
// hold node i (items e k + a of the node incidence, ascending, so ascending
// in e).  Their pair ids p = k (e k + a) + b sit at contrib[k offs[i] ...],
// which is where the global sort by (row, col) puts row i too, and sorting
// the keys (col << 32 | p) orders them by (col, element) exactly as the
// stable global sort does (p grows with e for a fixed row).
constexpr int kRowCap = 1024;  // pairs per row sorted in registers (16 per lane)

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  const int lo = __shfl_xor(int(uint32_t(v)), m, 64);
  const int hi = __shfl_xor(int(uint32_t(v >> 32)), m, 64);
  return (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo);
}

__device__ __forceinline__ uint64_t shfl_up1_u64(uint64_t v) {
  const int lo = __shfl_up(int(uint32_t(v)), 1, 64);
  const int hi = __shfl_up(int(uint32_t(v >> 32)), 1, 64);
  return (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo);
}

// Ascending bitonic sort of the wave's 64 E keys, key g = lane E + e in
// v[e] of its lane: compare-exchanges at distance < E inside the lane,
// larger ones across lanes (lane ^ distance / E).
template <int E>
__device__ __forceinline__ void wave_bitonic(uint64_t (&v)[E]) {
  const int lane = threadIdx.x & 63;
  constexpr int P = 64 * E;
#pragma unroll
  for (int size = 2; size <= P; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride < E) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int pe = e ^ stride;
          if (pe > e) {
            const bool asc = ((lane * E + e) & size) == 0;
            const uint64_t a = v[e], b = v[pe];
            const bool sw = asc ? (a > b) : (a < b);
            v[e] = sw ? b : a;
            v[pe] = sw ? a : b;
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int g = lane * E + e;
          const bool asc = (g & size) == 0;
          const bool lower = (g & stride) == 0;
          const uint64_t o = shfl_xor_u64(v[e], stride / E);
          const bool keep_min = lower == asc;
          v[e] = keep_min ? (o < v[e] ? o : v[e]) : (o > v[e] ? o : v[e]);
        }
      }
    }
  }
}

// One wave: the c <= 64 E pairs of one row, sorted into contrib[cbase ...];
// the row's distinct columns and the positions of their first contributors
// go to hcol / hpos[cbase + u], u = 0 .. count - 1.  Returns the count (in
// every lane).
template <int E>
__device__ int row_sort(int c, int k, int64_t item0, const int32_t* __restrict__ conn,
                        const uint32_t* __restrict__ perm, uint32_t* __restrict__ contrib,
                        int32_t* __restrict__ hcol, int32_t* __restrict__ hpos, int64_t cbase) {
  const int lane = threadIdx.x & 63;
  uint64_t v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int g = lane * E + e;
    v[e] = ~uint64_t(0);
    if (g < c) {
      const int j = g / k, b = g - (g / k) * k;
      const uint32_t item = perm[item0 + j];
      const uint32_t p = item * uint32_t(k) + uint32_t(b);
      const int64_t row0 = int64_t(item) - int64_t(item % uint32_t(k));
      v[e] = (uint64_t(uint32_t(conn[row0 + b])) << 32) | p;
    }
  }
  wave_bitonic<E>(v);
  const uint64_t before = shfl_up1_u64(v[E - 1]);
  bool head[E];
  int mine = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int g = lane * E + e;
    const uint64_t prev = e > 0 ? v[e > 0 ? e - 1 : 0] : before;
    head[e] = g < c && (g == 0 || (prev >> 32) != (v[e] >> 32));
    mine += head[e] ? 1 : 0;
    if (g < c) contrib[cbase + g] = uint32_t(v[e]);
  }
  // exclusive prefix of the head counts over the lanes
  int incl = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  int u = incl - mine;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (head[e]) {
      hcol[cbase + u] = int32_t(v[e] >> 32);
      hpos[cbase + u] = int32_t(cbase + lane * E + e);
      ++u;
    }
  }
  return __shfl(incl, 63, 64);
}

// One wave per row (4 rows per block): contributor lists and row lengths
// (rowlen[n_nodes] = 0 for the scan); rows beyond kRowCap pairs set
// flags[0].  LARGE = false sorts the rows of at most 256 pairs (4 per lane:
// every row of a Tet10 box mesh) and sets flags[2] for longer ones, which
// the LARGE = true launch then sorts (8 or 16 per lane); kept apart, the
// common kernel needs 35 instead of 78 VGPRs (Tet10 160x20x20: 446 against
// 481 us).
template <bool LARGE>
__global__ void __launch_bounds__(kBlock)
row_pattern(int64_t n_nodes, int k, const int32_t* __restrict__ conn,
            const int32_t* __restrict__ offs, const uint32_t* __restrict__ perm,
            uint32_t* __restrict__ contrib, int32_t* __restrict__ hcol,
            int32_t* __restrict__ hpos, int32_t* __restrict__ rowlen,
            int32_t* __restrict__ flags) {
  const int lane = threadIdx.x & 63;
  const int64_t i = blockIdx.x * int64_t(kBlock / 64) + threadIdx.x / 64;
  if (i > n_nodes) return;
  int cnt = 0;
  if (i < n_nodes) {
    const int64_t o0 = offs[i];
    const int64_t c = (offs[i + 1] - o0) * k;
    const int64_t cbase = o0 * k;
    if constexpr (LARGE) {
      if (c <= 256 || c > kRowCap) return;
      if (c > 512) {
        cnt = row_sort<16>(int(c), k, o0, conn, perm, contrib, hcol, hpos, cbase);
      } else {
        cnt = row_sort<8>(int(c), k, o0, conn, perm, contrib, hcol, hpos, cbase);
      }
    } else {
      if (c > kRowCap) {
        if (lane == 0) flags[0] = 1;
      } else if (c > 256) {
        if (lane == 0) flags[2] = 1;
        return;   // the LARGE launch writes this row
      } else if (c > 128) {
        cnt = row_sort<4>(int(c), k, o0, conn, perm, contrib, hcol, hpos, cbase);
      } else if (c > 64) {
        cnt = row_sort<2>(int(c), k, o0, conn, perm, contrib, hcol, hpos, cbase);
      } else if (c > 0) {
        cnt = row_sort<1>(int(c), k, o0, conn, perm, contrib, hcol, hpos, cbase);
      }
    }
  } else if (LARGE) {
    return;
  }
  if (lane == 0) rowlen[i] = cnt;
}

// One wave per row: its distinct columns and first-contributor positions
// moved from the per-row staging (at k offs[i]) to the CSR slots.
__global__ void __launch_bounds__(kBlock)
row_compact(int64_t n_nodes, int k, const int32_t* __restrict__ offs,
            const int32_t* __restrict__ hcol, const int32_t* __restrict__ hpos,
            const int32_t* __restrict__ rowptr, int32_t* __restrict__ indices,
            int32_t* __restrict__ seg) {
  const int lane = threadIdx.x & 63;
  const int64_t i = blockIdx.x * int64_t(kBlock / 64) + threadIdx.x / 64;
  if (i >= n_nodes) return;
  const int64_t cbase = int64_t(offs[i]) * k;
  const int64_t r0 = rowptr[i], len = int64_t(rowptr[i + 1]) - r0;
  for (int64_t u = lane; u < len; u += 64) {
    indices[r0 + u] = hcol[cbase + u];
    seg[r0 + u] = hpos[cbase + u];
  }
}

__global__ void set_last(int32_t* __restrict__ at, int32_t v) { *at = v; }

__global__ void store_total(const int32_t* __restrict__ rowptr, int64_t n,
                            int64_t* __restrict__ nnz) {
  *nnz = int64_t(rowptr[n]);
}

// Row-wise path's arrays, placed after the global sort's (both phases).
struct RowWs {
  uint32_t* keys_in;
  uint32_t* keys_out;
  uint32_t* ids_in;
  uint32_t* perm;
  int32_t* offs;
  int32_t* rowlen;
  int32_t* rowptr;
  int32_t* flags;  // [0] a row beyond kRowCap, [1] 1: the global sort ran,
                   // [2] a row of 257 .. kRowCap pairs
  void* tmp;
  size_t tmp_bytes;
};

int row_layout(cf::Carve& c, int64_t n_items, int64_t n_nodes, RowWs* w) {
  size_t sort_bytes = 0, scan_bytes = 0;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(nullptr, sort_bytes, (uint32_t*)nullptr,
                                   (uint32_t*)nullptr, (uint32_t*)nullptr,
                                   (uint32_t*)nullptr, size_t(n_items), 0,
                                   bits_for(n_nodes + 1)));
  CF_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, 0,
                                 size_t(n_nodes + 1), rocprim::plus<int32_t>()));
  w->keys_in = c.take<uint32_t>(n_items);
  w->keys_out = c.take<uint32_t>(n_items);
  w->ids_in = c.take<uint32_t>(n_items);
  w->perm = c.take<uint32_t>(n_items);
  w->offs = c.take<int32_t>(n_nodes + 1);
  w->rowlen = c.take<int32_t>(n_nodes + 1);
  w->rowptr = c.take<int32_t>(n_nodes + 1);
  w->flags = c.take<int32_t>(3);
  w->tmp_bytes = sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
  w->tmp = c.take<char>(w->tmp_bytes);
  return CF_OK;
}

// Both layouts in one workspace: the global sort's first, then the row path's.
int full_layout(void* base, int64_t n_elem, int k, int64_t n_nodes, PatternWs* g, RowWs* r) {
  const int64_t n = n_elem * k * k;
  int rc = pattern_layout(base, n, 2 * bits_for(n_nodes), g);
  if (rc) return rc;
  cf::Carve c(base ? static_cast<char*>(base) + g->total : nullptr);
  if ((rc = row_layout(c, n_elem * k, n_nodes, r))) return rc;
  g->total += c.used;
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
  RowWs r;
  int rc = full_layout(nullptr, n_elem, k, n_nodes, &w, &r);
  if (rc) return rc;
  *bytes = w.total;
  return CF_OK;
}

namespace {

// The global radix sort of all pairs (rows beyond the register sort).
int global_sort(const int32_t* conn, int64_t n, int k, int sh, const PatternWs& w,
                uint32_t* contrib, int64_t* nnz_dev, hipStream_t s) {
  int grid = cf::stride_grid(n, kBlock);
  pair_keys<<<grid, kBlock, 0, s>>>(conn, n, k, sh, w.keys_in, w.ids_in);
  CF_LAUNCHED();
  size_t tb = w.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(w.tmp, tb, w.keys_in, w.keys_out, w.ids_in,
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

}  // namespace

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
  RowWs r;
  int rc = full_layout(ws, n_elem, k, n_nodes, &w, &r);
  if (rc) return rc;
  if (ws_bytes < w.total) {
    cf::set_error("cf_pattern_sort: workspace %zu < %zu bytes", ws_bytes, w.total);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  CF_HIP(hipMemsetAsync(r.flags, 0, 3 * sizeof(int32_t), s));
  if (n == 0) {
    store_count<<<1, 1, 0, s>>>(w.incl, 0, nnz_dev);
    CF_LAUNCHED();
    return CF_OK;
  }
  // node incidence of the element items, then one wave per row
  const int64_t items = n_elem * k;
  const int igrid = cf::stride_grid(items, kBlock);
  item_keys<<<igrid, kBlock, 0, s>>>(conn, items, r.keys_in, r.ids_in);
  CF_LAUNCHED();
  size_t tb = r.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(r.tmp, tb, r.keys_in, r.keys_out, r.ids_in, r.perm,
                                   size_t(items), 0, bits_for(n_nodes + 1), s));
  incidence_offsets<<<cf::stride_grid(n_nodes + 1, kBlock), kBlock, 0, s>>>(r.keys_out, items,
                                                                           n_nodes, r.offs);
  CF_LAUNCHED();
  const int64_t rblocks = (n_nodes + 1 + kBlock / 64 - 1) / (kBlock / 64);
  // the global sort's key arrays are free on this path: they stage the
  // distinct columns and positions of each row
  int32_t* hcol = reinterpret_cast<int32_t*>(w.keys_in);
  int32_t* hpos = reinterpret_cast<int32_t*>(w.keys_out);
  row_pattern<false><<<int(rblocks), kBlock, 0, s>>>(n_nodes, k, conn, r.offs, r.perm, contrib,
                                                       hcol, hpos, r.rowlen, r.flags);
  CF_LAUNCHED();
  int32_t fl[3] = {0, 0, 0};
  CF_HIP(hipMemcpyAsync(fl, r.flags, sizeof(fl), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  const int32_t over = fl[0];
  if (!over && fl[2]) {
    row_pattern<true><<<int(rblocks), kBlock, 0, s>>>(n_nodes, k, conn, r.offs, r.perm, contrib,
                                                        hcol, hpos, r.rowlen, r.flags);
    CF_LAUNCHED();
  }
  if (over) {
    set_last<<<1, 1, 0, s>>>(r.flags + 1, 1);
    CF_LAUNCHED();
    return global_sort(conn, n, k, sh, w, contrib, nnz_dev, s);
  }
  tb = r.tmp_bytes;
  CF_HIP(rocprim::exclusive_scan(r.tmp, tb, r.rowlen, r.rowptr, 0, size_t(n_nodes + 1),
                                 rocprim::plus<int32_t>(), s));
  store_total<<<1, 1, 0, s>>>(r.rowptr, n_nodes, nnz_dev);
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
  RowWs r;
  int rc = full_layout(const_cast<void*>(ws), n_elem, k, n_nodes, &w, &r);
  if (rc) return rc;
  hipStream_t s = cf::as_stream(stream);
  int32_t global = 0;
  if (n > 0) {
    CF_HIP(hipMemcpyAsync(&global, r.flags + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CF_HIP(hipStreamSynchronize(s));
  }
  if (n == 0 || global) {
    int grid = cf::stride_grid(n > 0 ? n : 1, kBlock);
    pattern_fill<<<grid, kBlock, 0, s>>>(w.keys_out, w.incl, n, sh, nnz, indices, seg);
    CF_LAUNCHED();
    pattern_rowptr<<<cf::stride_grid(n_nodes + 1, kBlock), kBlock, 0, s>>>(w.keys_out, w.incl,
                                                                            n, sh, n_nodes, nnz,
                                                                            indptr);
    CF_LAUNCHED();
    return CF_OK;
  }
  CF_HIP(hipMemcpyAsync(indptr, r.rowptr, sizeof(int32_t) * (n_nodes + 1),
                        hipMemcpyDeviceToDevice, s));
  const int64_t rblocks = (n_nodes + kBlock / 64 - 1) / (kBlock / 64);
  row_compact<<<int(rblocks), kBlock, 0, s>>>(n_nodes, k, r.offs,
                                                reinterpret_cast<const int32_t*>(w.keys_in),
                                                reinterpret_cast<const int32_t*>(w.keys_out),
                                                r.rowptr, indices, seg);
  CF_LAUNCHED();
  set_last<<<1, 1, 0, s>>>(seg + nnz, int32_t(n));
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
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(nullptr, sort_bytes, (uint32_t*)nullptr,
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
    CF_HIP(rocprim::radix_sort_pairs<SortConfig>(w.tmp, tb, w.keys_in, w.keys_out, w.ids_in,
                                     perm, size_t(n_items), 0, end_bit, s));
  }
  incidence_offsets<<<cf::stride_grid(n_nodes + 1, kBlock), kBlock, 0, s>>>(w.keys_out, n_items,
                                                                           n_nodes, offs);
  CF_LAUNCHED();
  return CF_OK;
}
