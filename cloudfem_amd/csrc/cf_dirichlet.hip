// Scalar Dirichlet elimination with the reference's CSR structure, and a
// vector axpby used by the time-stepping drivers.
//
// tet_poisson.apply_dirichlet (tet_poisson.py:72-110) turns fixed rows into
// identity rows, subtracts A[:, F] v from the right-hand side, and then
// assigns A[:, F] = 0, which makes scipy store an explicit zero for every
// (row, fixed column) pair.  The output pattern of row r is therefore
// (cols(r) \ F) U F, merged in ascending order; fixed rows hold F with a 1 on
// the diagonal.  Phase 1 counts rows and builds the new right-hand side;
// phase 2 writes the merged rows, one 64-lane wave per row.
#include <cstring>
#include <rocprim/device/device_scan.hpp>

#include "cf_common.hpp"

namespace {

constexpr int kBlock = 256;

__global__ void dirichlet_count(int64_t n, const int32_t* __restrict__ ptr,
                                const int32_t* __restrict__ idx,
                                const double* __restrict__ val,
                                const double* __restrict__ rhs,
                                const uint8_t* __restrict__ fixed,
                                const double* __restrict__ fval, int64_t n_fixed,
                                int32_t* __restrict__ counts,
                                double* __restrict__ rhs_out) {
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r < n;
       r += int64_t(gridDim.x) * blockDim.x) {
    if (fixed[r]) {
      counts[r] = int32_t(n_fixed);
      rhs_out[r] = fval[r];
      continue;
    }
    int32_t keep = 0;
    double acc = 0.0;
    for (int32_t t = ptr[r]; t < ptr[r + 1]; ++t) {
      const int32_t c = idx[t];
      if (fixed[c])
        acc += val[t] * fval[c];
      else
        ++keep;
    }
    counts[r] = keep + int32_t(n_fixed);
    rhs_out[r] = rhs[r] - acc;
  }
}

__global__ void shift_scan(int64_t n, const int32_t* __restrict__ incl,
                           int32_t* __restrict__ ptr_out, int64_t* __restrict__ nnz) {
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r <= n;
       r += int64_t(gridDim.x) * blockDim.x)
    ptr_out[r] = r == 0 ? 0 : incl[r - 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) *nnz = n > 0 ? int64_t(incl[n - 1]) : 0;
}

// number of entries of sorted list a[0..na) that are < key
__device__ __forceinline__ int32_t count_less(const int32_t* __restrict__ a, int32_t na,
                                              int32_t key) {
  int32_t lo = 0, hi = na;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void dirichlet_fill(int64_t n, const int32_t* __restrict__ ptr,
                               const int32_t* __restrict__ idx,
                               const double* __restrict__ val,
                               const uint8_t* __restrict__ fixed,
                               const int32_t* __restrict__ flist, int32_t n_fixed,
                               const int32_t* __restrict__ ptr_out,
                               int32_t* __restrict__ idx_out,
                               double* __restrict__ val_out) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * int64_t(blockDim.x) + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  for (int64_t r = wave; r < n; r += nwaves) {
    const int32_t ob = ptr_out[r];
    if (fixed[r]) {
      for (int32_t t = lane; t < n_fixed; t += 64) {
        idx_out[ob + t] = flist[t];
        val_out[ob + t] = flist[t] == r ? 1.0 : 0.0;
      }
      continue;
    }
    const int32_t b = ptr[r], e = ptr[r + 1];
    // kept (non-fixed) columns of the row: position among kept entries plus
    // the number of fixed columns below it
    for (int32_t t = b + lane; t < e; t += 64) {
      const int32_t c = idx[t];
      if (fixed[c]) continue;
      int32_t before = 0;
      for (int32_t u = b; u < t; ++u) before += fixed[idx[u]] ? 0 : 1;
      const int32_t pos = ob + before + count_less(flist, n_fixed, c);
      idx_out[pos] = c;
      val_out[pos] = val[t];
    }
    // fixed columns: their rank in F plus the kept columns below them
    for (int32_t t = lane; t < n_fixed; t += 64) {
      const int32_t c = flist[t];
      int32_t below = 0;
      for (int32_t u = b; u < e && idx[u] < c; ++u) below += fixed[idx[u]] ? 0 : 1;
      idx_out[ob + t + below] = c;
      val_out[ob + t + below] = 0.0;
    }
  }
}

__global__ void axpby(int64_t n, double a, const double* __restrict__ x, double b,
                      double* __restrict__ y) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

__global__ void fill_where(int64_t n, const uint8_t* __restrict__ mask,
                           const double* __restrict__ src, double* __restrict__ dst) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    if (mask[i]) dst[i] = src[i];
}

}  // namespace

extern "C" int cf_dirichlet_workspace_bytes(int64_t n, size_t* bytes) {
  CF_REQUIRE(bytes && n > 0, "cf_dirichlet_workspace_bytes: bad arguments");
  size_t scan = 0;
  CF_HIP(rocprim::inclusive_scan(nullptr, scan, (int32_t*)nullptr, (int32_t*)nullptr, size_t(n),
                                 rocprim::plus<int32_t>()));
  *bytes = cf::align256(scan) + 2 * cf::align256(sizeof(int32_t) * n);
  return CF_OK;
}

extern "C" int cf_dirichlet_count(int64_t n, const int32_t* ptr, const int32_t* idx,
                                  const double* val, const double* rhs, const uint8_t* fixed,
                                  const double* fixed_val, int64_t n_fixed, void* ws,
                                  size_t ws_bytes, int32_t* ptr_out, double* rhs_out,
                                  int64_t* nnz_dev, void* stream) {
  CF_REQUIRE(n > 0 && ptr && rhs && fixed && fixed_val && ws && ptr_out && rhs_out && nnz_dev,
             "cf_dirichlet_count: bad arguments");
  size_t need = 0;
  int rc = cf_dirichlet_workspace_bytes(n, &need);
  if (rc) return rc;
  if (ws_bytes < need) {
    cf::set_error("cf_dirichlet_count: workspace %zu < %zu bytes", ws_bytes, need);
    return CF_ERR_WORKSPACE;
  }
  cf::Carve c(ws);
  int32_t* counts = c.take<int32_t>(n);
  int32_t* incl = c.take<int32_t>(n);
  size_t tb = need - c.used;
  void* tmp = c.take<char>(tb);
  hipStream_t s = cf::as_stream(stream);
  const int grid = cf::stride_grid(n + 1, kBlock);
  dirichlet_count<<<grid, kBlock, 0, s>>>(n, ptr, idx, val, rhs, fixed, fixed_val, n_fixed,
                                          counts, rhs_out);
  CF_LAUNCHED();
  CF_HIP(rocprim::inclusive_scan(tmp, tb, counts, incl, size_t(n), rocprim::plus<int32_t>(), s));
  shift_scan<<<grid, kBlock, 0, s>>>(n, incl, ptr_out, nnz_dev);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_dirichlet_fill(int64_t n, const int32_t* ptr, const int32_t* idx,
                                 const double* val, const uint8_t* fixed,
                                 const int32_t* fixed_list, int64_t n_fixed,
                                 const int32_t* ptr_out, int32_t* idx_out, double* val_out,
                                 void* stream) {
  CF_REQUIRE(n > 0 && ptr && fixed && ptr_out && idx_out && val_out &&
                 (n_fixed == 0 || fixed_list),
             "cf_dirichlet_fill: bad arguments");
  hipStream_t s = cf::as_stream(stream);
  dirichlet_fill<<<cf::stride_grid(n * 64, kBlock), kBlock, 0, s>>>(
      n, ptr, idx, val, fixed, fixed_list, int32_t(n_fixed), ptr_out, idx_out, val_out);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_axpby(int64_t n, double a, const double* x, double b, double* y,
                        void* stream) {
  CF_REQUIRE(n >= 0 && x && y, "cf_axpby: bad arguments");
  if (n == 0) return CF_OK;
  hipStream_t s = cf::as_stream(stream);
  axpby<<<cf::stride_grid(n, kBlock), kBlock, 0, s>>>(n, a, x, b, y);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_copy_masked(int64_t n, const uint8_t* mask, const double* src, double* dst,
                              void* stream) {
  CF_REQUIRE(n >= 0 && mask && src && dst, "cf_copy_masked: bad arguments");
  if (n == 0) return CF_OK;
  hipStream_t s = cf::as_stream(stream);
  fill_where<<<cf::stride_grid(n, kBlock), kBlock, 0, s>>>(n, mask, src, dst);
  CF_LAUNCHED();
  return CF_OK;
}
