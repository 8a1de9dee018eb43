// Mesh processing on the GPU: Tet4 -> Tet10 edge-midpoint numbering and
// exterior boundary faces of a node set.
//
// tet10_mesh.py:77-112 numbers the midpoint of the sorted corner pair (a, b)
// n_corners + (rank of (a, b) among the unique pairs, lexicographic), the
// order np.unique(axis=0) produces; here one stable radix sort of the 64-bit
// keys a * n + b gives the same order.  tet_heat.py:97-120 (and the Tet10
// and elasticity copies) keep a face when its sorted corner triple occurs in
// exactly one element and all three corners are in the named set, listed in
// first-occurrence order (element order, then local face order) with that
// occurrence's node order; here a radix sort of 63-bit triple keys finds the
// singletons and a scan over the face slots keeps occurrence order.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// local edges of the midpoint slots 4..9 (tet10_mesh.py:33)
__constant__ int kEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
// P1 faces (tet_heat.py:111) and Tet10 faces: corners + midpoints (tet10.py:287-292)
__constant__ int kFace4[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
__constant__ int kFace10[4][6] = {
    {0, 1, 2, 4, 5, 6}, {0, 3, 1, 7, 8, 4}, {0, 2, 3, 6, 9, 7}, {1, 3, 2, 8, 9, 5}};

int bits_for(uint64_t n) {
  int b = 1;
  while (b < 64 && (uint64_t(1) << b) < n) ++b;
  return b;
}

__global__ void edge_keys(int64_t m, const int32_t* __restrict__ conn4, int32_t stride,
                          int64_t n0, uint64_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < 6 * m;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = t / 6;
    const int l = int(t - 6 * e);
    const int64_t a = conn4[e * stride + kEdge[l][0]], b = conn4[e * stride + kEdge[l][1]];
    const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
    keys[t] = uint64_t(lo) * uint64_t(n0) + uint64_t(hi);
    ids[t] = uint32_t(t);
  }
}

__global__ void run_heads(int64_t n, const uint64_t* __restrict__ keys,
                          int32_t* __restrict__ heads) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// conn10[e][4 + l] = n0 + rank; unique edge endpoints; count
__global__ void edge_fill(int64_t m6, int64_t n0, const uint64_t* __restrict__ keys,
                          const uint32_t* __restrict__ ids, const int32_t* __restrict__ scan,
                          int32_t* __restrict__ conn10, int32_t* __restrict__ ea,
                          int32_t* __restrict__ eb, int64_t* __restrict__ n_edges) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < m6;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t rank = scan[i] - 1;
    const int64_t t = ids[i];
    const int64_t e = t / 6, l = t - 6 * e;
    conn10[e * 10 + 4 + l] = int32_t(n0 + rank);
    if (i == 0 || keys[i] != keys[i - 1]) {
      ea[rank] = int32_t(keys[i] / uint64_t(n0));
      eb[rank] = int32_t(keys[i] % uint64_t(n0));
    }
    if (i == m6 - 1) *n_edges = rank + 1;
  }
}

__global__ void copy_corners(int64_t m, const int32_t* __restrict__ conn4, int32_t stride,
                             int32_t* __restrict__ conn10) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < 4 * m;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = t / 4, a = t - 4 * e;
    conn10[e * 10 + a] = conn4[e * stride + a];
  }
}

// x[n0 + i] = (x[a_i] + x[b_i]) / 2, same for y, z
__global__ void edge_midpoints(int64_t n_edges, int64_t n0, const int32_t* __restrict__ ea,
                               const int32_t* __restrict__ eb, double* __restrict__ x,
                               double* __restrict__ y, double* __restrict__ z) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_edges;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t a = ea[i], b = eb[i];
    x[n0 + i] = 0.5 * (x[a] + x[b]);
    y[n0 + i] = 0.5 * (y[a] + y[b]);
    z[n0 + i] = 0.5 * (z[a] + z[b]);
  }
}

__global__ void face_keys(int64_t m, const int32_t* __restrict__ conn, int32_t k, int64_t n,
                          uint64_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < 4 * m;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = t / 4;
    const int f = int(t - 4 * e);
    uint64_t v[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = uint64_t(uint32_t(conn[e * k + kFace4[f][j]]));
    // sort the triple
    if (v[0] > v[1]) { const uint64_t s = v[0]; v[0] = v[1]; v[1] = s; }
    if (v[1] > v[2]) { const uint64_t s = v[1]; v[1] = v[2]; v[2] = s; }
    if (v[0] > v[1]) { const uint64_t s = v[0]; v[0] = v[1]; v[1] = s; }
    keys[t] = (v[0] * uint64_t(n) + v[1]) * uint64_t(n) + v[2];
    ids[t] = uint32_t(t);
  }
}

// keep[slot] = the slot's triple occurs once and its corners are members
__global__ void face_keep(int64_t nf, const uint64_t* __restrict__ keys,
                          const uint32_t* __restrict__ ids, const int32_t* __restrict__ conn,
                          int32_t k, const uint8_t* __restrict__ member,
                          int32_t* __restrict__ keep) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nf;
       i += int64_t(gridDim.x) * blockDim.x) {
    const bool single = (i == 0 || keys[i - 1] != keys[i]) && (i == nf - 1 || keys[i + 1] != keys[i]);
    const int64_t t = ids[i];
    bool in = single;
    if (in && member) {
      const int64_t e = t / 4;
      const int f = int(t - 4 * e);
#pragma unroll
      for (int j = 0; j < 3; ++j) in = in && member[conn[e * k + kFace4[f][j]]] != 0;
    }
    keep[t] = in ? 1 : 0;
  }
}

// faces[pos] = nodes of kept slot t in occurrence order (pos = scan[t] - 1)
__global__ void face_fill(int64_t nf, const int32_t* __restrict__ keep,
                          const int32_t* __restrict__ scan, const int32_t* __restrict__ conn,
                          int32_t k, int32_t width, int32_t* __restrict__ faces,
                          int64_t* __restrict__ count) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < nf;
       t += int64_t(gridDim.x) * blockDim.x) {
    if (t == nf - 1) *count = scan[t];
    if (!keep[t]) continue;
    const int64_t pos = scan[t] - 1;
    const int64_t e = t / 4;
    const int f = int(t - 4 * e);
    for (int j = 0; j < width; ++j)
      faces[pos * width + j] = conn[e * k + (width == 3 ? kFace4[f][j] : kFace10[f][j])];
  }
}

struct MeshWs {
  uint64_t *keys_a, *keys_b;
  uint32_t *ids_a, *ids_b;
  int32_t *flags, *scan;
  void* tmp;
  size_t tmp_bytes;
  size_t total;
};

MeshWs mesh_layout(int64_t items, void* base) {
  cf::Carve c(base);
  MeshWs W{};
  W.keys_a = c.take<uint64_t>(items);
  W.keys_b = c.take<uint64_t>(items);
  W.ids_a = c.take<uint32_t>(items);
  W.ids_b = c.take<uint32_t>(items);
  W.flags = c.take<int32_t>(items);
  W.scan = c.take<int32_t>(items);
  size_t a = 0, b = 0;
  rocprim::radix_sort_pairs<SortConfig>(nullptr, a, (uint64_t*)nullptr, (uint64_t*)nullptr,
                            (uint32_t*)nullptr, (uint32_t*)nullptr, size_t(items), 0, 64);
  rocprim::inclusive_scan(nullptr, b, (int32_t*)nullptr, (int32_t*)nullptr, size_t(items),
                          rocprim::plus<int32_t>());
  W.tmp_bytes = std::max(a, b);
  W.tmp = c.take<char>(W.tmp_bytes);
  W.total = c.used;
  return W;
}

}  // namespace

extern "C" int cf_mesh_workspace_bytes(int64_t n_elem, size_t* bytes) {
  CF_REQUIRE(n_elem > 0 && bytes, "cf_mesh_workspace_bytes: bad arguments");
  CF_REQUIRE(6 * n_elem < (int64_t(1) << 32), "cf_mesh_workspace_bytes: too many elements");
  *bytes = mesh_layout(6 * n_elem, nullptr).total;
  return CF_OK;
}

extern "C" int cf_tet10_edges(int64_t n_elem, const int32_t* conn4, int32_t stride,
                              int64_t n_corners, int32_t* conn10, int32_t* edge_a,
                              int32_t* edge_b, int64_t* n_edges, void* ws, size_t ws_bytes,
                              void* stream) {
  CF_REQUIRE(n_elem > 0 && n_corners > 0 && conn4 && conn10 && edge_a && edge_b && n_edges && ws,
             "cf_tet10_edges: bad arguments");
  CF_REQUIRE(stride >= 4, "cf_tet10_edges: stride must be >= 4");
  CF_REQUIRE(6 * n_elem < (int64_t(1) << 32), "cf_tet10_edges: too many elements");
  const int64_t m6 = 6 * n_elem;
  MeshWs W = mesh_layout(m6, ws);
  if (ws_bytes < W.total) {
    cf::set_error("cf_tet10_edges: workspace %zu < %zu bytes", ws_bytes, W.total);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  const int g = cf::stride_grid(m6, kBlock);
  edge_keys<<<g, kBlock, 0, s>>>(n_elem, conn4, stride, n_corners, W.keys_a, W.ids_a);
  CF_LAUNCHED();
  const uint64_t max_key = uint64_t(n_corners) * uint64_t(n_corners);
  size_t tb = W.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(W.tmp, tb, W.keys_a, W.keys_b, W.ids_a, W.ids_b, size_t(m6), 0,
                                   bits_for(max_key), s));
  run_heads<<<g, kBlock, 0, s>>>(m6, W.keys_b, W.flags);
  CF_LAUNCHED();
  tb = W.tmp_bytes;
  CF_HIP(rocprim::inclusive_scan(W.tmp, tb, W.flags, W.scan, size_t(m6), rocprim::plus<int32_t>(),
                                 s));
  copy_corners<<<cf::stride_grid(4 * n_elem, kBlock), kBlock, 0, s>>>(n_elem, conn4, stride,
                                                                       conn10);
  CF_LAUNCHED();
  edge_fill<<<g, kBlock, 0, s>>>(m6, n_corners, W.keys_b, W.ids_b, W.scan, conn10, edge_a,
                                 edge_b, n_edges);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_edge_midpoints(int64_t n_edges, int64_t n_corners, const int32_t* edge_a,
                                 const int32_t* edge_b, double* x, double* y, double* z,
                                 void* stream) {
  CF_REQUIRE(n_edges >= 0 && n_corners > 0 && (n_edges == 0 || (edge_a && edge_b && x && y && z)),
             "cf_edge_midpoints: bad arguments");
  if (n_edges == 0) return CF_OK;
  edge_midpoints<<<cf::stride_grid(n_edges, kBlock), kBlock, 0, cf::as_stream(stream)>>>(
      n_edges, n_corners, edge_a, edge_b, x, y, z);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_boundary_faces(int64_t n_elem, const int32_t* conn, int32_t k, int64_t n_nodes,
                                 const uint8_t* member, int32_t width, int32_t* faces,
                                 int64_t* count, void* ws, size_t ws_bytes, void* stream) {
  CF_REQUIRE(n_elem > 0 && conn && faces && count && ws, "cf_boundary_faces: bad arguments");
  CF_REQUIRE((k == 4 && width == 3) || (k == 10 && (width == 3 || width == 6)),
             "cf_boundary_faces: (k, width) must be (4, 3), (10, 3) or (10, 6)");
  CF_REQUIRE(n_nodes > 0 && n_nodes < (int64_t(1) << 21),
             "cf_boundary_faces: node count must be below 2^21 (63-bit face keys)");
  const int64_t nf = 4 * n_elem;
  MeshWs W = mesh_layout(6 * n_elem, ws);
  if (ws_bytes < W.total) {
    cf::set_error("cf_boundary_faces: workspace %zu < %zu bytes", ws_bytes, W.total);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  const int g = cf::stride_grid(nf, kBlock);
  face_keys<<<g, kBlock, 0, s>>>(n_elem, conn, k, n_nodes, W.keys_a, W.ids_a);
  CF_LAUNCHED();
  size_t tb = W.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(W.tmp, tb, W.keys_a, W.keys_b, W.ids_a, W.ids_b, size_t(nf), 0,
                                   bits_for(uint64_t(n_nodes) * uint64_t(n_nodes) *
                                            uint64_t(n_nodes)),
                                   s));
  face_keep<<<g, kBlock, 0, s>>>(nf, W.keys_b, W.ids_b, conn, k, member, W.flags);
  CF_LAUNCHED();
  tb = W.tmp_bytes;
  CF_HIP(rocprim::inclusive_scan(W.tmp, tb, W.flags, W.scan, size_t(nf), rocprim::plus<int32_t>(),
                                 s));
  face_fill<<<g, kBlock, 0, s>>>(nf, W.flags, W.scan, conn, k, width, faces, count);
  CF_LAUNCHED();
  return CF_OK;
}
