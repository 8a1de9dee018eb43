// Sparse matrix-vector product and preconditioned CG on the free-DOF system.
//
// Replaces scipy's SuperLU spsolve behind apply_dirichlet_vec3
// (elastic3d.py:220-232) and solve_tet_poisson (tet_poisson.py:121).  The
// free system A_ff x_f = b_f - A_fF x_F is never formed: fixed DOFs are
// masked (their rows of A p are zeroed and their entries of r, z, p stay 0),
// which is the same arithmetic as CG on A[free][:, free].
//
// Kernels per CG iteration (fixed grids, so the per-block partial sums --
// and therefore every dot product -- are bitwise reproducible; each kernel
// that needs a reduction folds the previous kernel's block partials itself,
// in the same order in every block; no floating-point atomics):
//   1. spmv      Ap = A p (masked), p.Ap partials
//   2. update    alpha; x += alpha p, r -= alpha Ap, z = M^-1 r, r.r, r.z
//   3. direction convergence test, beta; p = z + beta p
// A device-side halt flag makes the remaining kernels of a launched batch
// exit at once after convergence, so the host polls only once per batch.
//
// Mixed precision (CF_PREC_MIXED): the inner iteration runs entirely in FP32
// on the Jacobi-scaled matrix S A S (S = diag(A)^-1/2, FP32 copy of the
// values, 4 bytes per stored entry instead of 8) with FP64 accumulation of
// dot products.  The solution is accumulated in FP64 and the residual is
// periodically replaced by the FP64 residual b - A x computed with the FP64
// matrix ("reliable updates", Clark et al., Comput. Phys. Commun. 181 (2010)
// 1517), so the stopping test is applied to the true FP64 residual.
#include <cmath>
#include <cstdlib>
#include <limits>
#include <cstring>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocsolver/rocsolver.h>
#include <type_traits>

#include "cf_common.hpp"
#include "cf_tables.hpp"

// Device radix sorts run rocPRIM's onesweep algorithm at every size above a
// single block: its default falls back to a merge sort up to 1 M items (about
// 20 short launches per sort, against one per 8-bit digit here).
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;  // MI355X: 8 accelerator complex dies, 32 CUs each

enum Status : int32_t { kRunning = 0, kConverged = 1, kMaxit = 2, kBreakdown = 3, kRefresh = 4 };

struct PcgState {
  double rz;      // r.z of the current residual
  double rr;      // r.r (recursive residual; scaled residual in mixed mode)
  double pAp;
  double alpha;
  double beta;
  double stop2;   // stop when rr <= stop2
  double refresh2;// mixed: request an FP64 residual replacement when rr <= refresh2
  double pad;
  int32_t iter;
  int32_t maxit;
  int32_t halt;   // Status
  uint32_t ticket;
  double rzs[2];  // r.z by iteration parity (the fused CG kernels read one, write the other)
  double floor2;  // || |A_ff| |x_f| ||^2 at the last floor check
};

// three consecutive values of a node (one 12- or 24-byte load)
template <class T>
struct Vec3 {
  T v[3];
};

// two values staged as one aligned 8- or 16-byte LDS access
template <class T>
struct alignas(2 * sizeof(T)) Vec2 {
  T v[2];
};

// ------------------------------------------------------------- reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int L>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int L>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of NV values over W waves; result valid in every thread.
template <int NV, int W = kWaves>
__device__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][W];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][wave] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) s += sh[k][w];
    v[k] = s;
  }
}

// Per-block partial sums for a grid-wide reduction: thread 0 of each block
// stores its block's sums to partials[k * gridDim.x + blockIdx.x] with plain
// stores.  A 1-block `finalize` kernel launched next folds them in a fixed
// order (deterministic).  No in-kernel cross-block handshake: on gfx950 an
// agent-scope fence writes back and invalidates the XCD's L2, which, issued
// once per block, ruined the bandwidth of every kernel that reduced.
template <int NV>
__device__ void block_partials(double (&v)[NV], double* __restrict__ partials) {
  block_sum<NV>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
}

// Contiguous row range of a block.  Blocks are NOT regrouped per XCD: the
// hardware deals workgroups round-robin to the 8 XCDs, and keeping that order
// makes the blocks resident at any moment stream one contiguous region of the
// matrix.  Regrouping so that each XCD walks its own contiguous eighth (to
// keep its part of x in its L2) measured 1.6x slower on the Tet10 SpMV
// (341 vs 212 us at 539,601 node rows): HBM streaming locality matters more
// here than L2 reuse of x, which the 256 MB last-level cache serves anyway.
__device__ __forceinline__ void block_rows(int64_t n_rows, int64_t& r0, int64_t& r1) {
  const int64_t G = gridDim.x;
  const int64_t chunk = (n_rows + G - 1) / G;
  r0 = blockIdx.x * chunk;
  r1 = r0 + chunk < n_rows ? r0 + chunk : n_rows;
}

// Tiled launches: grid = tile_grid(tiles) blocks, block b takes tile b, so
// every block gets at most one tile (no tail round on a few blocks).
__device__ __forceinline__ int64_t tile_index(int64_t b) { return b; }

// Workgroups are dispatched round-robin over the kXcd dies, each with its own
// L2.  xcd_tile(b, nb) maps launch block b of nb to a tile so that each die
// walks one contiguous range of tiles (a bijection on [0, nb)).  Used where
// the work per tile is uniform (elements): node neighbourhoods then stay in
// one die's L2 (FP32 element forces 24.7 vs 26.0 us at 384,000 elements).
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
  const int64_t q = nb / kXcd, r = nb - q * kXcd;
  const int64_t d = b % kXcd, i = b / kXcd;
  return d * q + (d < r ? d : r) + i;
}

// ------------------------------------------------------------- SpMV

template <class T>
__device__ __forceinline__ T ld(const T* p) { return __builtin_nontemporal_load(p); }

// NB3 SpMV, one node row per 64-lane wave.  The row's column indices (up to
// 128 blocks) are loaded once, coalesced, and handed to the lanes that need
// them with cross-lane permutes, so the x gathers do not wait on a
// per-value index load; each lane then issues the value and x loads of up
// to 4 slots (256 values per pass) before the first FMA.  Rows longer than
// 128 blocks fall back to direct index loads.
template <class VT, class XT, class AT>
__global__ void __launch_bounds__(kBlock)
spmv_nb3_wave(int64_t n_nodes, const int32_t* __restrict__ ptr,
              const int32_t* __restrict__ idx, const VT* __restrict__ val,
              const XT* __restrict__ x, XT* __restrict__ y,
              const uint8_t* __restrict__ fixed, PcgState* state,
              double* __restrict__ partials) {
  if (state && state->halt) return;
  constexpr int kW = 64;
  const int lane = threadIdx.x & (kW - 1);
  const int wave = threadIdx.x / kW;
  constexpr int kWaves = kBlock / kW;
  int64_t r0, r1;
  block_rows(n_nodes, r0, r1);
  double dot = 0.0;
  for (int64_t i = r0 + wave; i < r1; i += kWaves) {
    const int32_t base = ptr[i];
    const int32_t len = ptr[i + 1] - base;
    const int32_t l3 = 3 * len, nv = 9 * len;
    const VT* v = val + int64_t(9) * base;
    const int32_t* col = idx + base;
    const int c_lo = lane < len ? col[lane] : 0;
    const int c_hi = lane + kW < len ? col[lane + kW] : 0;
    const bool direct = len > 2 * kW;
    AT a0 = 0, a1 = 0, a2 = 0;
    for (int32_t f0 = 0; f0 < nv; f0 += 4 * kW) {
      VT vv[4];
      XT xv[4];
      int cc[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int32_t f = f0 + u * kW + lane;
        const bool ok = f < nv;
        const int c = (f >= l3) + (f >= 2 * l3);
        const int32_t q = f - c * l3;
        const int32_t kq = q / 3;
        const int cp = q - 3 * kq;
        // the source lane (kq mod 64) holds col[kq] in c_lo (kq < 64) or
        // c_hi (64 <= kq < 128): fetch both, select on the requesting lane
        const int from_lo = __shfl(c_lo, kq & (kW - 1), kW);
        const int from_hi = __shfl(c_hi, kq & (kW - 1), kW);
        int cj = kq < kW ? from_lo : from_hi;
        if (direct && ok) cj = col[kq];
        cc[u] = ok ? c : 3;
        vv[u] = ok ? ld(v + f) : VT(0);
        xv[u] = ok ? x[3 * int64_t(cj) + cp] : XT(0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const AT prod = AT(vv[u]) * AT(xv[u]);
        a0 += cc[u] == 0 ? prod : AT(0);
        a1 += cc[u] == 1 ? prod : AT(0);
        a2 += cc[u] == 2 ? prod : AT(0);
      }
    }
    a0 = group_sum<kW>(a0);
    a1 = group_sum<kW>(a1);
    a2 = group_sum<kW>(a2);
    if (lane == 0) {
      AT o[3] = {a0, a1, a2};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int64_t d = 3 * i + c;
        if (fixed && fixed[d]) o[c] = AT(0);
        y[d] = XT(o[c]);
        if (state) dot += double(x[d]) * double(XT(o[c]));
      }
    }
  }
  if (!state) return;
  double v1[1] = {dot};
  block_partials<1>(v1, partials);
}

template <class VT, class XT, class AT, int L>
__global__ void __launch_bounds__(kBlock)
spmv_csr(int64_t n_rows, const int32_t* __restrict__ ptr,
         const int32_t* __restrict__ idx, const VT* __restrict__ val,
         const XT* __restrict__ x, XT* __restrict__ y,
         const uint8_t* __restrict__ fixed, PcgState* state,
         double* __restrict__ partials) {
  if (state && state->halt) return;
  const int lane = threadIdx.x % L;
  const int group = threadIdx.x / L;
  constexpr int kGroups = kBlock / L;
  int64_t r0, r1;
  block_rows(n_rows, r0, r1);
  double dot = 0.0;
  for (int64_t i = r0 + group; i < r1; i += kGroups) {
    const int32_t b = ptr[i], e = ptr[i + 1];
    AT a = 0;
    for (int32_t t = b + lane; t < e; t += L) a += AT(ld(val + t)) * AT(x[idx[t]]);
    a = group_sum<L>(a);
    if (lane == 0) {
      if (fixed && fixed[i]) a = AT(0);
      y[i] = XT(a);
      if (state) dot += double(x[i]) * double(XT(a));
    }
  }
  if (!state) return;
  double v1[1] = {dot};
  block_partials<1>(v1, partials);
}

// ------------------------------------------------------------- preconditioners

// Jacobi: dinv[i] = 1 / A_ii (0 on fixed DOFs).  Block3: per node the inverse
// of the 3x3 diagonal block restricted to the free components.
__global__ void diag_csr(int64_t n, const int32_t* __restrict__ ptr,
                         const int32_t* __restrict__ idx, const double* __restrict__ val,
                         const uint8_t* __restrict__ fixed, double* __restrict__ dinv,
                         int32_t* __restrict__ bad) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double d = 0.0;
    for (int32_t t = ptr[i]; t < ptr[i + 1]; ++t)
      if (idx[t] == i) d = val[t];
    if (fixed && fixed[i]) {
      dinv[i] = 0.0;
    } else if (d > 0.0) {
      dinv[i] = 1.0 / d;
    } else {
      dinv[i] = 0.0;
      atomicAdd(bad, 1);
    }
  }
}

__device__ __forceinline__ bool inv3(const double a[3][3], double o[3][3]) {
  const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  const double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
  const double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
  if (!(det > 0.0)) return false;
  const double r = 1.0 / det;
  o[0][0] = c00 * r;
  o[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * r;
  o[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * r;
  o[1][0] = c01 * r;
  o[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * r;
  o[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * r;
  o[2][0] = c02 * r;
  o[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * r;
  o[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * r;
  return true;
}

__global__ void diag_nb3(int64_t n_nodes, const int32_t* __restrict__ ptr,
                         const int32_t* __restrict__ idx, const double* __restrict__ val,
                         const uint8_t* __restrict__ fixed, int block3,
                         double* __restrict__ dinv, int32_t* __restrict__ bad) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t base = ptr[i], len = ptr[i + 1] - base;
    int32_t kd = -1;
    for (int32_t t = 0; t < len; ++t)
      if (idx[base + t] == i) kd = t;
    double blk[3][3] = {};
    if (kd >= 0)
      for (int c = 0; c < 3; ++c)
        for (int cp = 0; cp < 3; ++cp)
          blk[c][cp] = val[int64_t(9) * base + c * 3 * len + 3 * kd + cp];
    bool fx[3];
    for (int c = 0; c < 3; ++c) fx[c] = fixed && fixed[3 * i + c];
    if (!block3) {
      for (int c = 0; c < 3; ++c) {
        const double d = blk[c][c];
        if (fx[c]) {
          dinv[3 * i + c] = 0.0;
        } else if (d > 0.0) {
          dinv[3 * i + c] = 1.0 / d;
        } else {
          dinv[3 * i + c] = 0.0;
          atomicAdd(bad, 1);
        }
      }
      continue;
    }
    // Restrict to free components: fixed rows/columns -> identity, then zero
    // them in the inverse.
    for (int c = 0; c < 3; ++c)
      if (fx[c])
        for (int t = 0; t < 3; ++t) blk[c][t] = blk[t][c] = (c == t) ? 1.0 : 0.0;
    double o[3][3];
    if (!inv3(blk, o)) {
      atomicAdd(bad, 1);
      for (int t = 0; t < 9; ++t) dinv[9 * i + t] = 0.0;
      continue;
    }
    for (int c = 0; c < 3; ++c)
      for (int cp = 0; cp < 3; ++cp) dinv[9 * i + 3 * c + cp] = (fx[c] || fx[cp]) ? 0.0 : o[c][cp];
  }
}

template <int PC>
__device__ __forceinline__ void precond(const double* __restrict__ dinv,
                                        const double* __restrict__ r, int64_t node,
                                        double z[3]) {
  if constexpr (PC == CF_PC_BLOCK3) {
    const double* m = dinv + 9 * node;
    const double r0 = r[3 * node], r1 = r[3 * node + 1], r2 = r[3 * node + 2];
    z[0] = m[0] * r0 + m[1] * r1 + m[2] * r2;
    z[1] = m[3] * r0 + m[4] * r1 + m[5] * r2;
    z[2] = m[6] * r0 + m[7] * r1 + m[8] * r2;
  } else {
    z[0] = dinv[node] * r[node];
  }
}

// ------------------------------------------------------------- FP64 CG

// Fused CG scalars.  Each block of the kernel after a reduction folds that
// reduction's block partials itself (same order in every block, so every
// block computes the same value) instead of waiting for a one-block
// `finalize` launch; block 0 alone writes the scalars back.  r.z lives in
// rzs[slot] / rzs[slot ^ 1] so no block reads a value another block of the
// same kernel is overwriting.
template <int NV>
__device__ __forceinline__ void fold_partials(const double* __restrict__ parts, int n,
                                              double (&v)[NV]) {
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += parts[k * n + b];
  block_sum<NV>(v);
}

// alpha = rzs[slot] / p.Ap (p.Ap from the product's n_pa partials); then as
// cg_update, writing (r.r, r.z) partials to pb.
template <int PC>
__global__ void __launch_bounds__(kBlock)
cg_update_fused(int64_t n_units, double* __restrict__ x, double* __restrict__ r,
                const double* __restrict__ p, const double* __restrict__ Ap,
                const double* __restrict__ dinv, PcgState* state,
                const double* __restrict__ pa, int n_pa, double* __restrict__ pb, int slot) {
  if (state->halt) return;
  double f[1];
  fold_partials<1>(pa, n_pa, f);
  const double pAp = f[0];
  const bool ok = pAp > 0.0 && isfinite(pAp);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    state->pAp = pAp;
    if (ok)
      state->alpha = state->rzs[slot] / pAp;
    else
      state->halt = kBreakdown;
  }
  if (!ok) return;
  constexpr int W = (PC == CF_PC_BLOCK3) ? 3 : 1;
  const double alpha = state->rzs[slot] / pAp;
  double rr = 0.0, rz = 0.0;
  for (int64_t u = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; u < n_units;
       u += int64_t(gridDim.x) * blockDim.x) {
#pragma unroll
    for (int c = 0; c < W; ++c) {
      const int64_t d = W * u + c;
      x[d] += alpha * p[d];
      const double rn = r[d] - alpha * Ap[d];
      r[d] = rn;
      rr += rn * rn;
    }
    double z[3];
    precond<PC>(dinv, r, u, z);
#pragma unroll
    for (int c = 0; c < W; ++c) rz += r[W * u + c] * z[c];
  }
  double v[2] = {rr, rz};
  block_partials<2>(v, pb);
}

// (r.r, r.z) from cg_update_fused's partials -> stop test and beta; block 0
// records iter / rr / rzs[slot ^ 1] / halt; p = z + beta p.
template <int PC>
__global__ void __launch_bounds__(kBlock)
cg_direction_fused(int64_t n_units, const double* __restrict__ r, double* __restrict__ p,
                   const double* __restrict__ dinv, PcgState* state,
                   const double* __restrict__ pb, int n_pb, int slot) {
  if (state->halt) return;
  double v[2];
  fold_partials<2>(pb, n_pb, v);
  const double rr = v[0], rz = v[1];
  const bool finite = isfinite(rr) && isfinite(rz);
  const bool done = !finite || rr <= state->stop2;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int32_t it = state->iter + 1;
    state->iter = it;
    state->rr = rr;
    state->rzs[slot ^ 1] = rz;
    state->rz = rz;
    if (!finite)
      state->halt = kBreakdown;
    else if (rr <= state->stop2)
      state->halt = kConverged;
    else if (it >= state->maxit)
      state->halt = kMaxit;
  }
  if (done) return;
  constexpr int W = (PC == CF_PC_BLOCK3) ? 3 : 1;
  const double beta = rz / state->rzs[slot];
  for (int64_t u = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; u < n_units;
       u += int64_t(gridDim.x) * blockDim.x) {
    double z[3];
    precond<PC>(dinv, r, u, z);
#pragma unroll
    for (int c = 0; c < W; ++c) p[W * u + c] = z[c] + beta * p[W * u + c];
  }
}

// ------------------------------------------------------------- setup helpers

// out = (fixed ? keep_fixed : keep_free) ? x : 0
__global__ void split_vector(int64_t n, const double* __restrict__ x,
                             const uint8_t* __restrict__ fixed, int want_fixed,
                             double* __restrict__ out) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const bool f = fixed && fixed[i];
    out[i] = (f == bool(want_fixed)) ? x[i] : 0.0;
  }
}

// r = (b - Ax) on free DOFs, 0 on fixed ones; reduce r.r into state->pad
// (mode 0: also store as the reference norm) or into rr (mode 1).
__global__ void __launch_bounds__(kBlock)
residual(int64_t n, const double* __restrict__ b, const double* __restrict__ Ax,
         const uint8_t* __restrict__ fixed, double* __restrict__ r, PcgState* state,
         double* __restrict__ partials, double* __restrict__ out_norm2) {
  double rr = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double v = (fixed && fixed[i]) ? 0.0 : b[i] - Ax[i];
    r[i] = v;
    rr += v * v;
  }
  double v1[1] = {rr};
  block_partials<1>(v1, partials);
  (void)state;
  (void)out_norm2;
}

// p = z = M^-1 r, state->rz = r.z, state->rr = r.r
template <int PC>
__global__ void __launch_bounds__(kBlock)
cg_start(int64_t n_units, const double* __restrict__ r, double* __restrict__ p,
         const double* __restrict__ dinv, PcgState* state, double* __restrict__ partials) {
  constexpr int W = (PC == CF_PC_BLOCK3) ? 3 : 1;
  double rr = 0.0, rz = 0.0;
  for (int64_t u = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; u < n_units;
       u += int64_t(gridDim.x) * blockDim.x) {
    double z[3];
    precond<PC>(dinv, r, u, z);
#pragma unroll
    for (int c = 0; c < W; ++c) {
      p[W * u + c] = z[c];
      rz += r[W * u + c] * z[c];
      rr += r[W * u + c] * r[W * u + c];
    }
  }
  double v[2] = {rr, rz};
  block_partials<2>(v, partials);
  (void)state;
}

// ------------------------------------------------------------- mixed precision

// S_i = 1/sqrt(A_ii) on free DOFs, 0 on fixed (from the Jacobi dinv).
__global__ void make_scale(int64_t n, const double* __restrict__ dinv, double* __restrict__ s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    s[i] = dinv[i] > 0.0 ? sqrt(dinv[i]) : 0.0;
}

__global__ void scale_csr_to_f32(int64_t n, const int32_t* __restrict__ ptr,
                                 const int32_t* __restrict__ idx,
                                 const double* __restrict__ val, const double* __restrict__ s,
                                 float* __restrict__ out) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    for (int32_t t = ptr[i]; t < ptr[i + 1]; ++t) out[t] = float(s[i] * val[t] * s[idx[t]]);
}

// rh = fp32(S r), ph = rh (if start), rr = rh.rh; used at start and at every
// residual replacement.  `state->refresh2` is set from the new rr.
__global__ void __launch_bounds__(kBlock)
mx_load_residual(int64_t n, const double* __restrict__ r, const double* __restrict__ s,
                 float* __restrict__ rh, float* __restrict__ ph, int start,
                 PcgState* state, double* __restrict__ partials, double refresh_factor2) {
  double rr = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float v = float(s[i] * r[i]);
    rh[i] = v;
    if (start) ph[i] = v;
    rr += double(v) * double(v);
  }
  double v1[1] = {rr};
  block_partials<1>(v1, partials);
  (void)state;
  (void)refresh_factor2;
}

// y += alpha p; r -= alpha Ap; reduce r.r (FP32 vectors, FP64 sums).
__global__ void __launch_bounds__(kBlock)
mx_update(int64_t n, float* __restrict__ y, float* __restrict__ rh,
          const float* __restrict__ ph, const float* __restrict__ Aph, PcgState* state,
          double* __restrict__ partials) {
  if (state->halt) return;
  const float alpha = float(state->alpha);
  double rr = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    y[i] += alpha * ph[i];
    const float rn = rh[i] - alpha * Aph[i];
    rh[i] = rn;
    rr += double(rn) * double(rn);
  }
  double v1[1] = {rr};
  block_partials<1>(v1, partials);
}

__global__ void mx_direction(int64_t n, const float* __restrict__ rh, float* __restrict__ ph,
                             const PcgState* state) {
  if (state->halt) return;
  const float beta = float(state->beta);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    ph[i] = rh[i] + beta * ph[i];
}

// After a residual replacement the direction continues with the new beta.
__global__ void mx_direction_forced(int64_t n, const float* __restrict__ rh,
                                    float* __restrict__ ph, const PcgState* state) {
  const float beta = float(state->beta);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    ph[i] = rh[i] + beta * ph[i];
}

// x += S y; y = 0
__global__ void mx_accumulate(int64_t n, double* __restrict__ x, float* __restrict__ y,
                              const double* __restrict__ s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    x[i] += s[i] * double(y[i]);
    y[i] = 0.0f;
  }
}

// ----------------------------------------- mixed precision: FP64 flexible CG
// with a Chebyshev polynomial of the Jacobi-scaled matrix S A S (FP32
// vectors, FP32 or FP16 stored values) as the preconditioner:
//   z = S p(S A S) S r,  p = degree-(K-1) Chebyshev iteration on [b/ratio, b].
// The outer recurrence, its residual and the solution stay FP64.

// x += alpha p; r -= alpha Ap (skipped when !apply_step); r.r partials; then
// the first Chebyshev step, which needs no product: w = fp32(S r),
// d = w / theta, z = d.  SQ (three-term form, mf_gather_cheb32): the
// Chebyshev vectors are kept pre-multiplied by S, w = fp32(S^2 r) and the
// first iterate d = w / theta; z is not written (the steps build it from d),
// so a matrix-free product reads its input without scaling it and the final
// z = S zh is zh itself.
template <bool SQ, class T>
__global__ void __launch_bounds__(kBlock)
fcg_update(int64_t n, double* __restrict__ x, double* __restrict__ r,
           const double* __restrict__ p, const double* __restrict__ Ap,
           const double* __restrict__ sc, T* __restrict__ w, T* __restrict__ d,
           T* __restrict__ z, T inv_theta, int apply_step, PcgState* state,
           double* __restrict__ partials) {
  if (apply_step && state->halt) return;
  const double alpha = apply_step ? state->alpha : 0.0;
  double rr = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double rn = r[i];
    if (apply_step) {
      x[i] += alpha * p[i];
      rn -= alpha * Ap[i];
      r[i] = rn;
    }
    rr += rn * rn;
    const T wv = T(SQ ? sc[i] * sc[i] * rn : sc[i] * rn);
    w[i] = wv;
    const T dv = wv * inv_theta;
    d[i] = dv;
    if (!SQ) z[i] = dv;
  }
  double v1[1] = {rr};
  block_partials<1>(v1, partials);
}

// z = S zh (FP64); partials of r.z and r.z_old
__global__ void __launch_bounds__(kBlock)
fcg_zdots(int64_t n, const double* __restrict__ r, const double* __restrict__ sc,
          const float* __restrict__ zh, double* __restrict__ z, const double* __restrict__ zold,
          const PcgState* state, double* __restrict__ partials) {
  if (state->halt) return;
  double rz = 0.0, rzo = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double zv = sc[i] * double(zh[i]);
    z[i] = zv;
    rz += r[i] * zv;
    rzo += r[i] * zold[i];
  }
  double v[2] = {rz, rzo};
  block_partials<2>(v, partials);
}

// p = z + beta p
__global__ void __launch_bounds__(kBlock)
fcg_direction(int64_t n, const double* __restrict__ z, double* __restrict__ p,
              const PcgState* state) {
  if (state->halt) return;
  const double beta = state->beta;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// Fused forms of the three kernels above: each block folds the reduction
// it needs from the previous kernel's block partials (same order in every
// block) and block 0 records the scalars, so an outer iteration needs no
// one-block finalize launches.  rzs[slot] holds r.z of the iteration.

// alpha = rzs[slot] / p.Ap (n_pa product partials); x, r update; w, d, z
// start of the Chebyshev iteration; r.r partials into pb.
template <bool SQ, class T>
__global__ void __launch_bounds__(kBlock)
fcg_update_fused(int64_t n, double* __restrict__ x, double* __restrict__ r,
                 const double* __restrict__ p, const double* __restrict__ Ap,
                 const double* __restrict__ sc, T* __restrict__ w, T* __restrict__ d,
                 T* __restrict__ z, T inv_theta, PcgState* state,
                 const double* __restrict__ pa, int n_pa, double* __restrict__ pb, int slot) {
  if (state->halt) return;
  double f[1];
  fold_partials<1>(pa, n_pa, f);
  const double pAp = f[0];
  const bool ok = pAp > 0.0 && isfinite(pAp);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    state->pAp = pAp;
    if (ok)
      state->alpha = state->rzs[slot] / pAp;
    else
      state->halt = kBreakdown;
  }
  if (!ok) return;
  const double alpha = state->rzs[slot] / pAp;
  double rr = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    x[i] += alpha * p[i];
    const double rn = r[i] - alpha * Ap[i];
    r[i] = rn;
    rr += rn * rn;
    const T wv = T(SQ ? sc[i] * sc[i] * rn : sc[i] * rn);
    w[i] = wv;
    const T dv = wv * inv_theta;
    d[i] = dv;
    if (!SQ) z[i] = dv;
  }
  double v1[1] = {rr};
  block_partials<1>(v1, pb);
}

// r.r from fcg_update_fused (when `count`: stop test, block 0 records iter,
// rr, halt); z = S zh with (r.z, r.z_old) partials into pc.
template <bool SQ, class T>
__global__ void __launch_bounds__(kBlock)
fcg_zdots_fused(int64_t n, const double* __restrict__ r, const double* __restrict__ sc,
                const T* __restrict__ zh, double* __restrict__ z,
                const double* __restrict__ zold, PcgState* state, const double* __restrict__ pb,
                int n_pb, int count, double* __restrict__ pc) {
  if (state->halt) return;
  if (count) {
    double f[1];
    fold_partials<1>(pb, n_pb, f);
    const double rr = f[0];
    const bool fin = isfinite(rr);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const int32_t it = state->iter + 1;
      state->iter = it;
      state->rr = rr;
      if (!fin)
        state->halt = kBreakdown;
      else if (rr <= state->stop2)
        state->halt = kConverged;
      else if (it >= state->maxit)
        state->halt = kMaxit;
    }
    if (!fin || rr <= state->stop2) return;
  }
  double rz = 0.0, rzo = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double zv = SQ ? double(zh[i]) : sc[i] * double(zh[i]);
    z[i] = zv;
    rz += r[i] * zv;
    rzo += r[i] * zold[i];
  }
  double v[2] = {rz, rzo};
  block_partials<2>(v, pc);
}

// (r.z, r.z_old) -> Polak-Ribiere beta with rzs[slot]; block 0 records
// rzs[slot ^ 1]; p = z + beta p.
__global__ void __launch_bounds__(kBlock)
fcg_direction_fused(int64_t n, const double* __restrict__ z, double* __restrict__ p,
                    PcgState* state, const double* __restrict__ pc, int n_pc, int slot) {
  if (state->halt) return;
  double v[2];
  fold_partials<2>(pc, n_pc, v);
  const bool ok = v[0] > 0.0 && isfinite(v[0]) && isfinite(v[1]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (!ok) {
      state->halt = kBreakdown;
    } else {
      state->beta = (v[0] - v[1]) / state->rzs[slot];
      state->rzs[slot ^ 1] = v[0];
      state->rz = v[0];
    }
  }
  if (!ok) return;
  const double beta = (v[0] - v[1]) / state->rzs[slot];
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// Two-level form of the outer update: alpha from the product partials (when
// APPLY), x += alpha p, r -= alpha Ap, r.r partials into pb, and an FP32 copy
// of r for the preconditioner.
template <bool APPLY, class T>
__global__ void __launch_bounds__(kBlock)
fcg_update2l(int64_t n, double* __restrict__ x, double* __restrict__ r,
             const double* __restrict__ p, const double* __restrict__ Ap,
             T* __restrict__ r32, PcgState* state, const double* __restrict__ pa, int n_pa,
             double* __restrict__ pb, int slot) {
  if (state->halt) return;
  double alpha = 0.0;
  if (APPLY) {
    double f[1];
    fold_partials<1>(pa, n_pa, f);
    const double pAp = f[0];
    const bool ok = pAp > 0.0 && isfinite(pAp);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      state->pAp = pAp;
      if (ok)
        state->alpha = state->rzs[slot] / pAp;
      else
        state->halt = kBreakdown;
    }
    if (!ok) return;
    alpha = state->rzs[slot] / pAp;
  }
  double rr = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double rn = r[i];
    if (APPLY) {
      x[i] += alpha * p[i];
      rn -= alpha * Ap[i];
      r[i] = rn;
    }
    rr += rn * rn;
    r32[i] = T(rn);
  }
  double v1[1] = {rr};
  block_partials<1>(v1, pb);
}

template <class T>
__global__ void vec_add(int64_t n, T* __restrict__ u, const T* __restrict__ v,
                        const PcgState* state) {
  if (state && state->halt) return;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    u[i] += v[i];
}

// Power iteration on S A S for the top of the spectrum: v0 = S 1 spread by a
// fixed hash, then v = y / ||y|| with y = S A S v.
template <class T>
__global__ void power_start(int64_t n, const double* __restrict__ sc, T* __restrict__ v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t hsh = uint32_t(i) * 2654435761u;
    v[i] = sc[i] > 0.0 ? T(1.0f + float(hsh >> 22) * (1.0f / 1024.0f)) : T(0);
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
norm2_partials(int64_t n, const T* __restrict__ y, double* __restrict__ partials) {
  double a = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    a += double(y[i]) * double(y[i]);
  double v1[1] = {a};
  block_partials<1>(v1, partials);
}

template <class T>
__global__ void power_scale(int64_t n, const T* __restrict__ y, T* __restrict__ v,
                            const PcgState* state) {
  const T inv = state->pad > 0.0 ? T(1.0 / sqrt(state->pad)) : T(0);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    v[i] = y[i] * inv;
}

// Gershgorin bound of S A S: per DOF row sum_j |S_i A_ij S_j|, block maxima
// in partials (folded by finalize kFinMax).
__global__ void __launch_bounds__(kBlock)
gershgorin_nb3(int64_t n_nodes, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
               const double* __restrict__ val, const double* __restrict__ sc,
               double* __restrict__ partials) {
  double mx = 0.0;
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < 3 * n_nodes;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = t / 3;
    const int c = int(t - 3 * i);
    const int32_t base = ptr[i], len = ptr[i + 1] - base;
    const double si = sc[t];
    double acc = 0.0;
    for (int32_t k = 0; k < len; ++k) {
      const int64_t j3 = 3 * int64_t(idx[base + k]);
      const double* v = val + int64_t(9) * base + c * 3 * len + 3 * k;
      acc += fabs(v[0] * sc[j3]) + fabs(v[1] * sc[j3 + 1]) + fabs(v[2] * sc[j3 + 2]);
    }
    mx = fmax(mx, si * acc);
  }
  __shared__ double red[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < kBlock / 64; ++k) m = fmax(m, red[k]);
    partials[blockIdx.x] = m;
  }
}

// ------------------------------------------------------------- matrix-free Tet10

// Reference tables for the matrix-free operator (same values as the
// assembly kernels' copy; each translation unit owns its constant memory).
__constant__ cf::Tet10Tables kT10m = cf::build_tet10_tables();

// Element forces f = K_e x_e of isotropic Tet10 elasticity, K_e = V sum_q w_q
// B_q^T D B_q, for the affine elements the reference builds (midpoints on
// straight edges, tet10_mesh.py:77-110).  The integrand is then a polynomial
// of degree 2, which the symmetric 4-point rule (barycentric (a, b, b, b) and
// permutations, a = (5 + 3 sqrt 5) / 20, b = (5 - sqrt 5) / 20, weights 1/4)
// integrates exactly, as does the reference's 5-point Keast rule
// (tet10.py:84-100, elastic_tet10.py:164-170): the same matrix up to
// rounding.  With N_i = L_i (2 L_i - 1), N_ij = 4 L_i L_j the barycentric
// derivatives of u at point q (L_q = a, the others b) are
//   g_q = (4a - 1) x_q + 4b S_q,   g_i = T_i + 4(a - b) x_iq   (i != q),
//   S_i = sum of the 3 midpoints on edges at i, T_i = (4b - 1) x_i + 4b S_i,
// and the reference gradient is U[c][k] = g_{k+1}[c] - g_0[c].  With
// H = U Jinv, sigma = lam tr(H) I + mu (H + H^T) and P = (V/4) sigma Jinv^T,
// Q_{k+1} = P[.][k], Q_0 = -sum_k P[.][k], the forces are the transpose:
//   f_i  = (4b - 1) R_i + 4(a - b) Q_i(q = i),
//   f_ij = 4b (R_i + R_j) + 4(a - b) (Q_i(q = j) + Q_j(q = i)),
// R_i = sum_q Q_i(q).  About 580 instead of 960 multiply-adds per element
// (the 5-point barycentric form), the FP32 element kernel being bound by
// them (26 us at 384,000 elements with its gathers and stores removed).
constexpr double kT4a = 0.58541019662496845446;  // (5 + 3 sqrt 5) / 20
constexpr double kT4b = 0.13819660112501051518;  // (5 - sqrt 5) / 20
constexpr int kTE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
// edge of the corner pair (i, j), i != j
constexpr int kEdgeOf[4][4] = {{-1, 0, 2, 3}, {0, -1, 1, 4}, {2, 1, -1, 5}, {3, 4, 5, -1}};

template <class T>
__device__ __forceinline__ void t10_forces_4pt(const T (&xa)[10][3], const T (&J)[3][3], T V,
                                               T lam, T mu, T (&f)[10][3]) {
  const T cb = T(4.0 * kT4b - 1.0), c4b = T(4.0 * kT4b), cab = T(4.0 * (kT4a - kT4b));
  const T lq = T(0.25) * V * lam, mq = T(0.25) * V * mu;
  T Tg[4][3];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T sum = T(0);
      bool first = true;
#pragma unroll
      for (int e = 0; e < 6; ++e)
        if (kTE[e][0] == i || kTE[e][1] == i) {
          sum = first ? xa[4 + e][c] : sum + xa[4 + e][c];
          first = false;
        }
      Tg[i][c] = cb * xa[i][c] + c4b * sum;
    }
  T R[4][3], Ecr[6][3];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    T g[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        g[i][c] = Tg[i][c] + cab * (i == q ? xa[i][c] : xa[4 + kEdgeOf[i][q]][c]);
    T H[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T u0 = g[1][c] - g[0][c], u1 = g[2][c] - g[0][c], u2 = g[3][c] - g[0][c];
#pragma unroll
      for (int i = 0; i < 3; ++i) H[c][i] = u0 * J[0][i] + u1 * J[1][i] + u2 * J[2][i];
    }
    // (V/4) sigma, symmetric
    const T ltr = lq * (H[0][0] + H[1][1] + H[2][2]);
    T sg[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sg[c][c] = T(2) * mq * H[c][c] + ltr;
#pragma unroll
      for (int i = c + 1; i < 3; ++i) sg[c][i] = sg[i][c] = mq * (H[c][i] + H[i][c]);
    }
    T Q[4][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T P[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) P[k] = sg[c][0] * J[k][0] + sg[c][1] * J[k][1] + sg[c][2] * J[k][2];
      Q[1][c] = P[0];
      Q[2][c] = P[1];
      Q[3][c] = P[2];
      Q[0][c] = -(P[0] + P[1] + P[2]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        R[i][c] = q == 0 ? Q[i][c] : R[i][c] + Q[i][c];
        if (i == q) {
          f[i][c] = cab * Q[i][c];
        } else {
          // edge (i, q) collects Q_i at q here and Q_q at i when q' = i
          const int e = kEdgeOf[i][q];
          const bool first = (i < q ? i : q) == q;  // the smaller corner's point comes first
          Ecr[e][c] = first ? Q[i][c] : Ecr[e][c] + Q[i][c];
        }
      }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) f[i][c] = cb * R[i][c] + f[i][c];
#pragma unroll
  for (int e = 0; e < 6; ++e)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      f[4 + e][c] = c4b * (R[kTE[e][0]][c] + R[kTE[e][1]][c]) + cab * Ecr[e][c];
}

// Element forces of one element: gathers its 10 nodal x from the FP64
// vector and applies t10_forces_4pt; f receives the 10 nodal forces.
__device__ __forceinline__ void t10_forces(const int32_t* __restrict__ cn,
                                           const double* __restrict__ g, double lam, double mu,
                                           const double* __restrict__ x, double (&f)[10][3]) {
  double J[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) J[r][c] = g[3 * r + c];
  const double V = g[9];
  double xa[10][3];
#pragma unroll
  for (int a = 0; a < 10; ++a) {
    const int64_t nd = cn[a];
#pragma unroll
    for (int c = 0; c < 3; ++c) xa[a][c] = x[3 * nd + c];
  }
  t10_forces_4pt<double>(xa, J, V, lam, mu, f);
}

// Run sums of the matrix-free operator (FP64 and FP32).  Scattered stores of
// the element forces into node-sorted slots cost one L2 write request each
// (FP32: 3.0 M per product at 384,000 elements, the L2 busy for the whole
// kernel: 25.6 us).
// Instead the block of a tile of kBlock consecutive elements stages the
// tile's forces in LDS sorted by (node, element) -- item (e, a) at its rank
// k -- so that the items of one node form a contiguous run; it sums each
// run in element order and writes one partial per (tile, node) pair,
// coalesced (0.78 M partials instead of 3.84 M forces at 384,000 elements).
// The node gather then adds its node's partials in tile order (1.4 per node
// on Tet10 box meshes).  Deterministic: fixed orders throughout.
//   rank[e * 10 + a]  k of item (e, a) in its tile
//   toff[t]           first run of tile t (runs numbered tile by tile, by
//                     node inside a tile); tstart[q] = first k of run q
//   nq[j]             runs of node n at j in [nro[n], nro[n+1]), tile order
// Elements per tile (threads per element block).  Smaller tiles need fewer
// waves per barrier but give more runs: FP32 product at 384,000 elements
// 25.3 us with 256, 27.0 with 128, 29.4 with 64.
constexpr int kTile = 256;
constexpr int kTileItems = kTile * 10;
// the two-level aggregates count elements in units of 256, whatever the tile
constexpr int kAggUnitItems = 2560;

// Element-kernel grid: one block per tile, at most 8 blocks of 256 threads
// per CU (more tiles: blocks take several, xcd_tile order).
int elem_grid(int64_t m) {
  const int64_t n_tiles = (m + kTile - 1) / kTile;
  const int64_t cap = int64_t(cf::cu_count()) * 8 * (kBlock / kTile);
  return int(n_tiles < 1 ? 1 : (n_tiles < cap ? n_tiles : cap));
}

__global__ void iota_u32(int64_t n, uint32_t* __restrict__ v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    v[i] = uint32_t(i);
}

// keys[p] = tile of the item in node-sorted slot p (perm[p] = e * 10 + a)
__global__ void mf_store_keys(int64_t n_items, const uint32_t* __restrict__ perm,
                              uint32_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p < n_items;
       p += int64_t(gridDim.x) * blockDim.x) {
    keys[p] = perm[p] / uint32_t(kTileItems);
    ids[p] = uint32_t(p);
  }
}

// tslot: the slots sorted by tile (stable, so by slot inside a tile); the
// rank of an item is its position in its tile's list (every tile but the
// last holds exactly kTileItems items)
__global__ void mf_store_rank(int64_t n_items, const uint32_t* __restrict__ perm,
                              const uint32_t* __restrict__ tslot, uint16_t* __restrict__ rank) {
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < n_items;
       k += int64_t(gridDim.x) * blockDim.x)
    rank[perm[tslot[k]]] = uint16_t(k % kTileItems);
}

// head[p] = 1 where a run starts: slot p is the first of its node
// (mf_node_heads) or its item lies in another tile than the item of slot
// p - 1 (mf_tile_heads); head[n_items] = 0.
__global__ void mf_tile_heads(int64_t n_items, const uint32_t* __restrict__ perm,
                              uint32_t* __restrict__ head) {
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p <= n_items;
       p += int64_t(gridDim.x) * blockDim.x) {
    uint32_t h = 0;
    if (p < n_items)
      h = p == 0 || perm[p] / uint32_t(kTileItems) != perm[p - 1] / uint32_t(kTileItems);
    head[p] = h;
  }
}

__global__ void mf_node_heads(int64_t n_nodes, const int32_t* __restrict__ offs,
                              uint32_t* __restrict__ head) {
  for (int64_t n = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; n < n_nodes;
       n += int64_t(gridDim.x) * blockDim.x)
    if (offs[n] < offs[n + 1]) head[offs[n]] = 1;
}

// run r starting at slot p (ridx[p] = r): rk = position of the slot's item in
// the tile-major layout, rs = p; nro[n] = ridx[offs[n]]
__global__ void mf_run_fill(int64_t n_items, const uint32_t* __restrict__ perm,
                            const uint16_t* __restrict__ rank, const uint32_t* __restrict__ head,
                            const uint32_t* __restrict__ ridx, uint32_t* __restrict__ rk,
                            uint32_t* __restrict__ rs) {
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p < n_items;
       p += int64_t(gridDim.x) * blockDim.x) {
    if (!head[p]) continue;
    const uint32_t it = perm[p];
    rk[ridx[p]] = (it / uint32_t(kTileItems)) * uint32_t(kTileItems) + rank[it];
    rs[ridx[p]] = uint32_t(p);
  }
}

__global__ void mf_run_len(int64_t n_items, const uint32_t* __restrict__ ridx,
                           const uint32_t* __restrict__ rs, uint16_t* __restrict__ rl,
                           int64_t n_nodes, const int32_t* __restrict__ offs,
                           int32_t* __restrict__ nro) {
  const int64_t n_runs = ridx[n_items];
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r < n_runs;
       r += int64_t(gridDim.x) * blockDim.x)
    rl[r] = uint16_t((r + 1 < n_runs ? int64_t(rs[r + 1]) : n_items) - int64_t(rs[r]));
  for (int64_t n = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; n <= n_nodes;
       n += int64_t(gridDim.x) * blockDim.x)
    nro[n] = int32_t(ridx[offs[n]]);
}

// run4[i] = the node's run ids in tile order, -1 padded, when it has at most
// four; *over counts the nodes with more (the plan then keeps nro / nq only)
__global__ void mf_run4_fill(int64_t n_nodes, const int32_t* __restrict__ nro,
                             const uint32_t* __restrict__ nq, int4* __restrict__ run4,
                             int32_t* __restrict__ over) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t b = nro[i], n = nro[i + 1] - b;
    if (n > 4) atomicAdd(over, 1);
    int q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = u < n ? int32_t(nq[b + u]) : -1;
    run4[i] = make_int4(q[0], q[1], q[2], q[3]);
  }
}

// tile order of the runs: rq[q] = node-order run id of the q-th smallest
// start rk; nq[r] = q, tstart[q] = start inside the tile
__global__ void mf_run_tile_order(int64_t n_runs, const uint32_t* __restrict__ rk_sorted,
                                  const uint32_t* __restrict__ rq, uint32_t* __restrict__ nq,
                                  uint16_t* __restrict__ tstart) {
  for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < n_runs;
       q += int64_t(gridDim.x) * blockDim.x) {
    nq[rq[q]] = uint32_t(q);
    tstart[q] = uint16_t(rk_sorted[q] % uint32_t(kTileItems));
  }
}

// toff[t] = first q with rk_sorted[q] >= t * kTileItems (t = 0 .. n_tiles)
__global__ void mf_tile_run_off(int64_t n_tiles, int64_t n_runs,
                                const uint32_t* __restrict__ rk_sorted, int32_t* __restrict__ toff) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t <= n_tiles;
       t += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t target = uint64_t(t) * kTileItems;
    int64_t lo = 0, hi = n_runs;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (uint64_t(rk_sorted[mid]) >= target) hi = mid; else lo = mid + 1;
    }
    toff[t] = int32_t(lo);
  }
}

// FP64 element forces -> run partials part[c * n_runs + q] (component
// planes), the tile's forces staged in LDS one component at a time (20 KB per
// block, so the 3 waves per SIMD its registers allow stay resident).
__global__ void __launch_bounds__(kTile)
mf_t10_elem(int64_t m, const int32_t* __restrict__ conn, const double* __restrict__ geo,
            double lam, double mu, const double* __restrict__ x,
            const uint16_t* __restrict__ rank, const int32_t* __restrict__ toff,
            const uint16_t* __restrict__ tstart, int64_t n_runs, double* __restrict__ part,
            const PcgState* state) {
  if (state && state->halt) return;
  __shared__ double fl[kTileItems];
  const int64_t n_tiles = (m + kTile - 1) / kTile;
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < n_tiles; t += gridDim.x) {
    const int64_t e = t * kTile + threadIdx.x;
    double f[10][3];
    int rk[10];
    if (e < m) {
      const double* g = geo + e * 10;
      double J[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[r][c] = g[3 * r + c];
      const double V = g[9];
      double xa[10][3];
#pragma unroll
      for (int a = 0; a < 10; ++a) {
        const int64_t nd = conn[e * 10 + a];
#pragma unroll
        for (int c = 0; c < 3; ++c) xa[a][c] = x[3 * nd + c];
      }
      t10_forces_4pt<double>(xa, J, V, lam, mu, f);
#pragma unroll
      for (int a = 0; a < 10; ++a) rk[a] = rank[e * 10 + a];
    }
    const int64_t k0 = t * kTileItems;
    const int nt = int((k0 + kTileItems < 10 * m ? k0 + kTileItems : 10 * m) - k0);
    const int32_t q0 = toff[t], q1 = toff[t + 1];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (e < m) {
#pragma unroll
        for (int a = 0; a < 10; ++a) fl[rk[a]] = f[a][c];
      }
      __syncthreads();
      for (int32_t q = q0 + threadIdx.x; q < q1; q += kTile) {
        const int kb = tstart[q], ke = q + 1 < q1 ? int(tstart[q + 1]) : nt;
        double sum = 0.0;
        for (int k = kb; k < ke; ++k) sum += fl[k];
        part[c * n_runs + q] = sum;
      }
      __syncthreads();
    }
  }
}

// y_i = sum of node i's run partials in tile order, fixed rows zeroed; with
// `state`: p.Ap partials.
__global__ void __launch_bounds__(kBlock)
mf_gather_runs(int64_t n_nodes, const int32_t* __restrict__ nro, const uint32_t* __restrict__ nq,
               int64_t n_runs, const double* __restrict__ part,
               const uint8_t* __restrict__ fixed, const double* __restrict__ x,
               double* __restrict__ y, PcgState* state, double* __restrict__ partials) {
  if (state && state->halt) return;
  double dot = 0.0;
  // two nodes per thread, their partial chains interleaved (as in
  // mf_gather_cheb32)
  const int64_t i0 = tile_index(blockIdx.x) * (2 * kBlock) + threadIdx.x;
  double acc[2][3] = {};
  int32_t jb[2] = {0, 0}, je[2] = {0, 0};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = i0 + h * kBlock;
    if (i >= n_nodes) continue;
    jb[h] = nro[i];
    je[h] = nro[i + 1];
  }
  for (int32_t j = 0;; ++j) {
    const bool m0 = jb[0] + j < je[0], m1 = jb[1] + j < je[1];
    if (!m0 && !m1) break;
    if (m0) {
      const int64_t q = nq[jb[0] + j];
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[0][c] += part[c * n_runs + q];
    }
    if (m1) {
      const int64_t q = nq[jb[1] + j];
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[1][c] += part[c * n_runs + q];
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = i0 + h * kBlock;
    if (i >= n_nodes) continue;
    // one 24-byte store per node (three 8-byte stores wrote lines in pieces)
    Vec3<double> o;
#pragma unroll
    for (int c = 0; c < 3; ++c) o.v[c] = fixed && fixed[3 * i + c] ? 0.0 : acc[h][c];
    *reinterpret_cast<Vec3<double>*>(y + 3 * i) = o;
    if (state) {
      const Vec3<double> xv = *reinterpret_cast<const Vec3<double>*>(x + 3 * i);
#pragma unroll
      for (int c = 0; c < 3; ++c) dot += xv.v[c] * o.v[c];
    }
  }
  if (!state) return;
  double v1[1] = {dot};
  block_partials<1>(v1, partials);
}

// ---------------------------------------- matrix-free Tet10, chunked plan
// Elements are ordered along a Morton curve of their centroids and cut into
// chunks of kChunk consecutive elements, one chunk per workgroup.  The
// workgroup computes its elements' forces into LDS and sums them per node
// in fixed order: a node whose incident elements all lie in the chunk gets
// its y written directly; a node shared with other chunks gets one partial
// (3 doubles) per chunk, summed in chunk order by mf_shared_gather.  Against
// the two-pass operator this removes the 480 B per element force round trip
// through HBM; results are deterministic.
constexpr int kChunk = kBlock;   // elements per chunk = threads per block

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Morton key of each element's corner centroid (10 bits per axis)
__global__ void mfp_morton(int64_t m, const int32_t* __restrict__ conn,
                           const double* __restrict__ x, const double* __restrict__ y,
                           const double* __restrict__ z, double lx, double ly, double lz,
                           double sx, double sy, double sz, uint32_t* __restrict__ keys,
                           uint32_t* __restrict__ ids) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < m;
       e += int64_t(gridDim.x) * blockDim.x) {
    double c[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t nd = conn[e * 10 + a];
      c[0] += x[nd];
      c[1] += y[nd];
      c[2] += z[nd];
    }
    const double lo[3] = {lx, ly, lz}, sc[3] = {sx, sy, sz};
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double t = (0.25 * c[k] - lo[k]) * sc[k];
      t = t < 0.0 ? 0.0 : (t > 1023.0 ? 1023.0 : t);
      q[k] = uint32_t(t);
    }
    keys[e] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    ids[e] = uint32_t(e);
  }
}

__global__ void mfp_permute(int64_t m, const uint32_t* __restrict__ order,
                            const int32_t* __restrict__ conn, const double* __restrict__ geo,
                            int32_t* __restrict__ conn_c, double* __restrict__ geo_c) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < m * 10;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = t / 10, a = t - 10 * e;
    const int64_t src = int64_t(order[e]) * 10 + a;
    conn_c[t] = conn[src];
    geo_c[t] = geo[src];
  }
}

// slot t = 10 e + a of the chunk order -> key (chunk, node), value t
__global__ void mfp_pair_keys(int64_t m10, const int32_t* __restrict__ conn_c,
                              uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < m10;
       t += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t chunk = uint64_t(t / (10 * kChunk));
    keys[t] = (chunk << 32) | uint64_t(uint32_t(conn_c[t]));
    vals[t] = uint32_t(t);
  }
}

// head flags of runs of equal keys (or of keys below `limit` when not null)
__global__ void mfp_heads(int64_t n, const uint64_t* __restrict__ keys,
                          const int64_t* __restrict__ limit, int32_t* __restrict__ heads) {
  const int64_t lim = limit ? *limit : n;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const bool in = i < lim;
    const uint64_t k = keys[i] >> (limit ? 32 : 0);
    heads[i] = (in && (i == 0 || k != (keys[i - 1] >> (limit ? 32 : 0)))) ? 1 : 0;
  }
}

// pairs (chunk, node) from the sorted slot keys and their run index
__global__ void mfp_pairs(int64_t m10, int64_t n_chunks, const uint64_t* __restrict__ keys,
                          const uint32_t* __restrict__ vals, const int32_t* __restrict__ scan,
                          uint16_t* __restrict__ slot_list, int32_t* __restrict__ pair_node,
                          int32_t* __restrict__ pair_slot_off, int32_t* __restrict__ chunk_off,
                          int64_t* __restrict__ counts) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < m10;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    const int64_t chunk = int64_t(k >> 32);
    const int32_t p = scan[i] - 1;
    slot_list[i] = uint16_t(int64_t(vals[i]) - chunk * (10 * kChunk));
    if (i == 0 || k != keys[i - 1]) {
      pair_node[p] = int32_t(uint32_t(k));
      pair_slot_off[p] = int32_t(i);
      if (i == 0 || chunk != int64_t(keys[i - 1] >> 32)) chunk_off[chunk] = p;
    }
    if (i == m10 - 1) {
      counts[0] = p + 1;
      pair_slot_off[p + 1] = int32_t(m10);
      chunk_off[n_chunks] = p + 1;
    }
  }
}

// A pair is shared when its node has incident elements outside the chunk.
// keys = (node, chunk) for shared pairs, all-ones otherwise; pair_out = -1.
__global__ void mfp_shared_keys(int64_t m10, const int64_t* __restrict__ counts,
                                const uint64_t* __restrict__ slot_keys,
                                const int32_t* __restrict__ pair_node,
                                const int32_t* __restrict__ pair_slot_off,
                                const int32_t* __restrict__ offs, uint64_t* __restrict__ keys,
                                uint32_t* __restrict__ vals, int32_t* __restrict__ pair_out) {
  const int64_t n_pairs = counts[0];
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p < m10;
       p += int64_t(gridDim.x) * blockDim.x) {
    uint64_t key = ~uint64_t(0);
    if (p < n_pairs) {
      const int32_t node = pair_node[p];
      const int32_t s0 = pair_slot_off[p];
      const int32_t cnt = pair_slot_off[p + 1] - s0;
      const int32_t tot = offs[node + 1] - offs[node];
      if (cnt != tot) key = (uint64_t(uint32_t(node)) << 32) | (slot_keys[s0] >> 32);
      pair_out[p] = -1;
    }
    keys[p] = key;
    vals[p] = uint32_t(p);
  }
}

// counts[1] = number of shared pairs (keys sorted, all-ones last)
__global__ void mfp_count_shared(int64_t m10, const uint64_t* __restrict__ keys,
                                 int64_t* __restrict__ counts) {
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r < m10;
       r += int64_t(gridDim.x) * blockDim.x) {
    const bool in = keys[r] != ~uint64_t(0);
    if (in && (r == m10 - 1 || keys[r + 1] == ~uint64_t(0))) counts[1] = r + 1;
  }
}

__global__ void mfp_shared_fill(int64_t m10, const uint64_t* __restrict__ keys,
                                const uint32_t* __restrict__ vals,
                                const int32_t* __restrict__ scan,
                                int64_t* __restrict__ counts, int32_t* __restrict__ pair_out,
                                int32_t* __restrict__ shared_node,
                                int32_t* __restrict__ shared_off) {
  const int64_t n_sh = counts[1];
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r < n_sh;
       r += int64_t(gridDim.x) * blockDim.x) {
    pair_out[vals[r]] = int32_t(r);
    const int32_t j = scan[r] - 1;
    const uint32_t node = uint32_t(keys[r] >> 32);
    if (r == 0 || node != uint32_t(keys[r - 1] >> 32)) {
      shared_node[j] = int32_t(node);
      shared_off[j] = int32_t(r);
    }
    if (r == n_sh - 1) {
      counts[2] = j + 1;
      shared_off[j + 1] = int32_t(n_sh);
    }
  }
}

// y = K x for one chunk per block (see above); p.Ap partials of the direct
// nodes in partials[blockIdx].
__global__ void __launch_bounds__(kBlock)
mf_chunk_apply(int64_t m, const int32_t* __restrict__ conn_c, const double* __restrict__ geo_c,
               double lam, double mu, const double* __restrict__ x,
               const uint16_t* __restrict__ slot_list, const int32_t* __restrict__ chunk_off,
               const int32_t* __restrict__ pair_node, const int32_t* __restrict__ pair_slot_off,
               const int32_t* __restrict__ pair_out, const uint8_t* __restrict__ fixed,
               double* __restrict__ y, double* __restrict__ part_buf, PcgState* state,
               double* __restrict__ partials) {
  if (state && state->halt) return;
  // forces of the chunk, component-major: fl[(3 a + c) * kChunk + e_local]
  __shared__ double fl[kChunk * 30];
  const int64_t e = int64_t(blockIdx.x) * kChunk + threadIdx.x;
  if (e < m) {
    double f[10][3];
    t10_forces(conn_c + e * 10, geo_c + e * 10, lam, mu, x, f);
#pragma unroll
    for (int a = 0; a < 10; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) fl[(3 * a + c) * kChunk + threadIdx.x] = f[a][c];
  }
  __syncthreads();
  double dot = 0.0;
  const int32_t p1 = chunk_off[blockIdx.x + 1];
  for (int32_t p = chunk_off[blockIdx.x] + threadIdx.x; p < p1; p += kBlock) {
    double o[3] = {0.0, 0.0, 0.0};
    const int32_t s1 = pair_slot_off[p + 1];
    for (int32_t k = pair_slot_off[p]; k < s1; ++k) {
      const int sl = slot_list[k];  // e_local * 10 + a
      const int el = sl / 10, a = sl - 10 * el;
#pragma unroll
      for (int c = 0; c < 3; ++c) o[c] += fl[(3 * a + c) * kChunk + el];
    }
    const int32_t out = pair_out[p];
    if (out >= 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) part_buf[3 * int64_t(out) + c] = o[c];
    } else {
      const int64_t node = pair_node[p];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int64_t d = 3 * node + c;
        if (fixed && fixed[d]) o[c] = 0.0;
        y[d] = o[c];
        if (state) dot += x[d] * o[c];
      }
    }
  }
  if (!state) return;
  double v1[1] = {dot};
  block_partials<1>(v1, partials);
}

// y of the shared nodes: sum of their chunk partials in chunk order
__global__ void __launch_bounds__(kBlock)
mf_shared_gather(int64_t n_sn, const int32_t* __restrict__ shared_node,
                 const int32_t* __restrict__ shared_off, const double* __restrict__ part_buf,
                 const uint8_t* __restrict__ fixed, const double* __restrict__ x,
                 double* __restrict__ y, PcgState* state, double* __restrict__ partials) {
  if (state && state->halt) return;
  double dot = 0.0;
  const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (j < n_sn) {
    double o[3] = {0.0, 0.0, 0.0};
    for (int32_t r = shared_off[j]; r < shared_off[j + 1]; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) o[c] += part_buf[3 * int64_t(r) + c];
    const int64_t node = shared_node[j];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t d = 3 * node + c;
      if (fixed && fixed[d]) o[c] = 0.0;
      y[d] = o[c];
      if (state) dot += x[d] * o[c];
    }
  }
  if (!state) return;
  double v1[1] = {dot};
  block_partials<1>(v1, partials);
}

// ---------------------------------------- matrix-free Tet10, FP32 inner operator
// For the mixed solver's Chebyshev preconditioner: products with the
// Jacobi-scaled operator S K S in FP32 (geometry rounded once to FP32,
// forces and sums in FP32).  The element kernel gathers x = S d on the fly;
// the node kernel scales its force sum by S and applies the Chebyshev
// update of cheb_step (or, for the power iteration, just stores S K S d).


// Per-element inputs of the low-precision operator in tile-packed order:
// value a (0..9) of element e = t kBlock + l sits at ((t 10 + a) kBlock + l),
// so each of the 10 loads of a thread is one contiguous 64-lane access.
__device__ __forceinline__ int64_t tile_packed(int64_t t, int a) {
  return (t * 10 + a) * kTile + threadIdx.x;
}

template <class T>
__global__ void mf_pack_tiles(int64_t m, const int32_t* __restrict__ conn,
                              const double* __restrict__ geo, const uint16_t* __restrict__ rank,
                              int32_t* __restrict__ connT, T* __restrict__ geoT,
                              uint16_t* __restrict__ rankT) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < 10 * m;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = i / 10;
    const int a = int(i - 10 * e);
    const int64_t t = e / kTile;
    const int64_t o = (t * 10 + a) * kTile + (e - t * kTile);
    connT[o] = conn[i];
    rankT[o] = rank[i];
    geoT[o] = T(geo[i]);
  }
}


// PRE: d is already S-scaled (the Chebyshev vectors of the SQ form), so
// the kernel gathers it as is; otherwise it gathers S d.  Writes the run
// partials described above.  conn, geo and rank are tile-packed
// (mf_pack_tiles).
template <class T, bool PRE>
__global__ void __launch_bounds__(kTile)
mf_t10_elem_lp(int64_t m, const int32_t* __restrict__ conn, const T* __restrict__ geo,
               T lam, T mu, const T* __restrict__ sc, const T* __restrict__ d,
               const uint16_t* __restrict__ rank, const int32_t* __restrict__ toff,
               const uint16_t* __restrict__ tstart, T* __restrict__ part,
               const PcgState* state) {
  if (state && state->halt) return;
  // Forces are staged in two passes: (x, y) as one 8- or 16-byte pair per
  // item, then z in the same buffer.  Two components per item instead of
  // three keep the block at 20 KB of LDS (FP32), so six blocks fit a CU
  // and a 384,000-element mesh (1,500 tiles) runs in one round of
  // resident blocks instead of 1.17.
  __shared__ Vec2<T> fl[kTileItems];
  T* fz = reinterpret_cast<T*>(fl);
  const int64_t n_tiles = (m + kTile - 1) / kTile;
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < n_tiles; t += gridDim.x) {
    const int64_t e = t * kTile + threadIdx.x;
    T fzl[10];
    uint16_t rk[10];
    if (e < m) {
      T J[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[r][c] = geo[tile_packed(t, 3 * r + c)];
      const T V = geo[tile_packed(t, 9)];
      int32_t nd[10];
#pragma unroll
      for (int a = 0; a < 10; ++a) nd[a] = conn[tile_packed(t, a)];
      T xa[10][3];
#pragma unroll
      for (int a = 0; a < 10; ++a) {
        const Vec3<T> x = *reinterpret_cast<const Vec3<T>*>(d + 3 * int64_t(nd[a]));
        if constexpr (PRE) {
#pragma unroll
          for (int c = 0; c < 3; ++c) xa[a][c] = x.v[c];
        } else {
          const Vec3<T> sv = *reinterpret_cast<const Vec3<T>*>(sc + 3 * int64_t(nd[a]));
#pragma unroll
          for (int c = 0; c < 3; ++c) xa[a][c] = sv.v[c] * x.v[c];
        }
      }
      T f[10][3];
      t10_forces_4pt<T>(xa, J, V, lam, mu, f);
#pragma unroll
      for (int a = 0; a < 10; ++a) {
        rk[a] = rank[tile_packed(t, a)];
        Vec2<T> v;
        v.v[0] = f[a][0];
        v.v[1] = f[a][1];
        fl[rk[a]] = v;
        fzl[a] = f[a][2];
      }
    }
    __syncthreads();
    // one partial per run of the tile (runs cover the tile's items)
    const int64_t k0 = t * kTileItems;
    const int nt = int((k0 + kTileItems < 10 * m ? k0 + kTileItems : 10 * m) - k0);
    const int32_t q0 = toff[t], q1 = toff[t + 1];
    for (int32_t q = q0 + threadIdx.x; q < q1; q += kTile) {
      const int kb = tstart[q], ke = q + 1 < q1 ? int(tstart[q + 1]) : nt;
      T s0 = T(0), s1 = T(0);
      for (int k = kb; k < ke; ++k) {
        const Vec2<T> v = fl[k];
        s0 += v.v[0];
        s1 += v.v[1];
      }
      T* o = part + 3 * int64_t(q);
      o[0] = s0;
      o[1] = s1;
    }
    __syncthreads();
    if (e < m) {
#pragma unroll
      for (int a = 0; a < 10; ++a) fz[rk[a]] = fzl[a];
    }
    __syncthreads();
    for (int32_t q = q0 + threadIdx.x; q < q1; q += kTile) {
      const int kb = tstart[q], ke = q + 1 < q1 ? int(tstart[q + 1]) : nt;
      T s2 = T(0);
      for (int k = kb; k < ke; ++k) s2 += fz[k];
      part[3 * int64_t(q) + 2] = s2;
    }
    __syncthreads();
  }
}

// ---- FP32 element forces on the matrix cores.
// The two contractions of the affine Tet10 operator with the reference
// gradient table d[q][a][k] = dN_a/dxi_k at the 4 points of the exact rule
// (point q: L_q = kT4a, the other three kT4b) are batched GEMMs:
//   G_c[q][k] = sum_a d[q][a][k] u_c[a]          (gradients of u_c)
//   f_c[a]    = sum_{q,k} d[q][a][k] P_c[q][k]    (forces)
// with P = (V/4) sigma(U Jinv) Jinv^T per point, as in t10_forces_4pt.  A
// wave takes 16 elements at a time, four lanes per element: lane l serves
// element e = l % 16 and point / node group j = l / 16.  With
// v_mfma_f32_16x16x4f32 (A 16x4: lane l holds A[l%16][l/16]; B 4x16: lane l
// holds B[l/16][l%16]; D 16x16: lane l, register r holds D[4(l/16)+r][l%16])
//   stage 1: D1[4q+k][e] = sum_s A1_s[4q+k][i] B1_s[i][e], B1_s[i][e] =
//            u_c[4s+i] of element e (nodes 10, 11 and k = 3 padded with
//            zeros), three k-steps: lane (e, j) ends with G_c[j][0..2];
//   stage 2: D2[a][e] = sum_s A2_s[a][q] B2_s[q][e], B2_s[q][e] = P_c[q][s]
//            (lane (e, j) holds point j), three k-steps: lane (e, j) ends
//            with f_c[4j + r], r = 0..3 (a < 10).
// 18 MFMAs and about 60 VALU multiply-adds per lane per 16 elements, against
// about 580 multiply-adds per element on one lane in t10_forces_4pt; every
// lane gathers three nodes instead of ten.  Same run plan, LDS staging and
// run partials as mf_t10_elem_lp, so the node kernel is shared.
typedef float cf_float4 __attribute__((ext_vector_type(4)));

struct T10Grad4 {
  float d[4][10][3];
};

constexpr T10Grad4 build_t10_grad4() {
  T10Grad4 t{};
  for (int q = 0; q < 4; ++q) {
    double L[4] = {kT4b, kT4b, kT4b, kT4b};
    L[q] = kT4a;
    double dL[10][4] = {};
    for (int i = 0; i < 4; ++i) dL[i][i] = 4.0 * L[i] - 1.0;
    for (int s = 0; s < 6; ++s) {
      const int i = kTE[s][0], j = kTE[s][1];
      dL[4 + s][i] = 4.0 * L[j];
      dL[4 + s][j] = 4.0 * L[i];
    }
    for (int a = 0; a < 10; ++a)
      for (int k = 0; k < 3; ++k) t.d[q][a][k] = float(dL[a][k + 1] - dL[a][0]);
  }
  return t;
}

__constant__ T10Grad4 kGrad4 = build_t10_grad4();

__device__ __forceinline__ cf_float4 mfma16x16x4(float a, float b, cf_float4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <bool PRE>
__global__ void __launch_bounds__(kTile)
mf_t10_elem_mfma(int64_t m, const int32_t* __restrict__ conn, const float* __restrict__ geo,
                 float lam, float mu, const float* __restrict__ sc, const float* __restrict__ d,
                 const uint16_t* __restrict__ rank, const int32_t* __restrict__ toff,
                 const uint16_t* __restrict__ tstart, float* __restrict__ part,
                 const PcgState* state) {
  if (state && state->halt) return;
  __shared__ float fl[kTileItems * 3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int el0 = lane & 15, j = lane >> 4;
  // constant operands: A1_s (row 4q+k, column node 4s+j) and A2_s (row node
  // el0, column point j, direction s)
  float a1[3], a2[3];
  {
    const int q = el0 >> 2, k = el0 & 3;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const int a = 4 * s + j;
      a1[s] = (k < 3 && a < 10) ? kGrad4.d[q][a][k] : 0.0f;
      a2[s] = el0 < 10 ? kGrad4.d[j][el0][s] : 0.0f;
    }
  }
  const float lq = 0.25f * lam, mq = 0.25f * mu;
  const int64_t n_tiles = (m + kTile - 1) / kTile;
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < n_tiles; t += gridDim.x) {
    // All loads of the four groups first, with no branches: an element past
    // m reads element m - 1 (the padding of the last tile is not written)
    // and gets V = 0, node slots 10 and 11 read node 9 (their A1 entries are
    // zero), so every wait is on loads that are already in flight.
    const int64_t last = m - 1 - t * kTile;  // last valid element of the tile
    int64_t base[4];
    bool ok[4];
    int32_t nd[4][3];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int el = wave * 64 + g * 16 + el0;  // element within the tile
      ok[g] = el <= last;
      base[g] = t * 10 * kTile + (ok[g] ? el : last);
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int a = 4 * s + j < 10 ? 4 * s + j : 9;
        nd[g][s] = conn[base[g] + int64_t(a) * kTile];
      }
    }
    float xb[4][3][3];
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const Vec3<float> x = *reinterpret_cast<const Vec3<float>*>(d + 3 * int64_t(nd[g][s]));
        if constexpr (PRE) {
#pragma unroll
          for (int c = 0; c < 3; ++c) xb[g][s][c] = x.v[c];
        } else {
          const Vec3<float> sv =
              *reinterpret_cast<const Vec3<float>*>(sc + 3 * int64_t(nd[g][s]));
#pragma unroll
          for (int c = 0; c < 3; ++c) xb[g][s][c] = sv.v[c] * x.v[c];
        }
      }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      float J[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[r][c] = geo[base[g] + int64_t(3 * r + c) * kTile];
      const float V = ok[g] ? geo[base[g] + 9 * int64_t(kTile)] : 0.0f;
      // stage 1: reference gradients at point j
      cf_float4 G[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        G[c] = cf_float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 3; ++s) G[c] = mfma16x16x4(a1[s], xb[g][s][c], G[c]);
      }
      // point algebra: H = U Jinv, (V/4) sigma, P = (V/4) sigma Jinv^T
      float H[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < 3; ++i) H[c][i] = G[c][0] * J[0][i] + G[c][1] * J[1][i] + G[c][2] * J[2][i];
      const float lv = lq * V, mv = mq * V;
      const float ltr = lv * (H[0][0] + H[1][1] + H[2][2]);
      float sg[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sg[c][c] = 2.0f * mv * H[c][c] + ltr;
#pragma unroll
        for (int i = c + 1; i < 3; ++i) sg[c][i] = sg[i][c] = mv * (H[c][i] + H[i][c]);
      }
      // stage 2: forces of nodes 4j .. 4j + 3
      cf_float4 F[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        F[c] = cf_float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const float p = sg[c][0] * J[s][0] + sg[c][1] * J[s][1] + sg[c][2] * J[s][2];
          F[c] = mfma16x16x4(a2[s], p, F[c]);
        }
      }
      if (ok[g]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int a = 4 * j + r;
          if (a < 10) {
            const int rk = rank[base[g] + int64_t(a) * kTile];
            fl[3 * rk] = F[0][r];
            fl[3 * rk + 1] = F[1][r];
            fl[3 * rk + 2] = F[2][r];
          }
        }
      }
    }
    __syncthreads();
    // one partial per run of the tile (runs cover the tile's items)
    const int64_t k0 = t * kTileItems;
    const int nt = int((k0 + kTileItems < 10 * m ? k0 + kTileItems : 10 * m) - k0);
    const int32_t q0 = toff[t], q1 = toff[t + 1];
    for (int32_t q = q0 + threadIdx.x; q < q1; q += kTile) {
      const int kb = tstart[q], ke = q + 1 < q1 ? int(tstart[q + 1]) : nt;
      float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
      for (int k = kb; k < ke; ++k) {
        s0 += fl[3 * k];
        s1 += fl[3 * k + 1];
        s2 += fl[3 * k + 2];
      }
      float* o = part + 3 * int64_t(q);
      o[0] = s0;
      o[1] = s1;
      o[2] = s2;
    }
    __syncthreads();
  }
}

// Which FP32 element kernel the inner products use: CF_INNER_KERNEL=mfma
// (mf_t10_elem_mfma) or =valu (mf_t10_elem_lp<float>); read at each solve.
bool inner_mfma() {
  const char* v = std::getenv("CF_INNER_KERNEL");
  if (v && std::strcmp(v, "mfma") == 0) return true;
  if (v && std::strcmp(v, "valu") == 0) return false;
  return false;
}

// FP32 element forces of the Chebyshev / power products (PRE as in
// mf_t10_elem_lp) with the kernel chosen by inner_mfma().
template <class T, bool PRE>
void launch_elem_lp(bool mfma, int grid, hipStream_t s, int64_t m, const int32_t* conn,
                    const T* geo, T lam, T mu, const T* sc, const T* d, const uint16_t* rank,
                    const int32_t* toff, const uint16_t* tstart, T* part, const PcgState* state) {
  if constexpr (std::is_same_v<T, float>) {
    if (mfma) {
      mf_t10_elem_mfma<PRE><<<grid, kTile, 0, s>>>(m, conn, geo, lam, mu, sc, d, rank, toff,
                                                    tstart, part, state);
      return;
    }
  }
  mf_t10_elem_lp<T, PRE><<<grid, kTile, 0, s>>>(m, conn, geo, lam, mu, sc, d, rank, toff,
                                                 tstart, part, state);
}

// node sums of the FP32 forces (the node's run partials in tile order).
// mode 0 (power iteration): zn = S sum.  Modes 3 and 4 serve the two-level
// preconditioner (product K zc of its coarse iterate): mode 3 starts the
// Chebyshev iteration on the residual rin - K zc, w = S^2 (rin - sum),
// zn = cw w (cw = 1 / theta); mode 4 returns the residual zn = rin - sum.
// Modes 1 and 2: one Chebyshev step in three-term form on S-scaled vectors
// (z = S zhat, w = S^2 r fixed over the steps; the element kernel gathered
// zc, so sum = K zc and S^2 sum = S (S K S) zhat):
//   zn = zc + cd (zc - zp) + cw (w - S^2 sum),
// with zp = 0 in mode 1 (the first step).  Five FP32 vector streams per
// node (S, w, zc, zp read, zn written) instead of the seven of the
// residual-recursive form.  Mode 5 is mode 2 for the last step of the
// two-level form, which adds its iterate to zn (zn += ..., the coarse
// iterate) instead of storing it.  zn may alias zp: each entry is read and
// then written by the same thread.
// Each thread takes two nodes (i and i + kBlock of its block's 2 kBlock), so
// one launch fits in a single round of resident waves (one node per thread
// needs 1.03 rounds at 539,601 nodes: 14.8 vs 13.2 us), and it loads the
// node vectors before the dependent nro -> nq -> partial chain, whose loads
// it interleaves over the two nodes.
// R4 (every node has at most four runs, the usual case): the node's run ids
// come from one 16-byte load of run4[i] (-1 = none), so the partial loads
// wait on one load instead of the nro -> nq chain.
template <class T, bool R4 = false>
__global__ void __launch_bounds__(kBlock)
mf_gather_cheb(int64_t n_nodes, const int32_t* __restrict__ nro,
               const uint32_t* __restrict__ nq, const T* __restrict__ part,
               const T* __restrict__ sc, int mode, const T* __restrict__ zc,
               const T* zp, T* zn, T* __restrict__ w, T cd, T cw,
               const T* __restrict__ rin, const PcgState* state,
               const int4* __restrict__ run4 = nullptr) {
  if (state && state->halt) return;
  const int64_t i0 = tile_index(blockIdx.x) * (2 * kBlock) + threadIdx.x;
  // per node: S, then zc and w (modes 1, 2, 5) or rin (modes 3, 4), zp
  // (modes 2, 5), the sum zn is added to (mode 5)
  T acc[2][3] = {}, sv[2][3] = {}, zv[2][3] = {}, wv[2][3] = {}, pv[2][3] = {},
    uv[2][3] = {};
  int32_t jb[2] = {0, 0}, je[2] = {0, 0};
  int4 r4[2] = {make_int4(-1, -1, -1, -1), make_int4(-1, -1, -1, -1)};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = i0 + h * kBlock;
    if (i >= n_nodes) continue;
    if constexpr (R4) {
      r4[h] = run4[i];
    } else {
      jb[h] = nro[i];
      je[h] = nro[i + 1];
    }
    // one 12- or 24-byte access per node vector (three scalar accesses per
    // node wrote each line in pieces: twice the bytes of zn reached HBM)
    auto ld3 = [&](const T* v, T (&o)[3]) {
      const Vec3<T> t = *reinterpret_cast<const Vec3<T>*>(v + 3 * i);
      o[0] = t.v[0];
      o[1] = t.v[1];
      o[2] = t.v[2];
    };
    if (mode != 4) ld3(sc, sv[h]);
    if (mode == 3 || mode == 4) {
      ld3(rin, wv[h]);
    } else if (mode != 0) {
      ld3(zc, zv[h]);
      ld3(w, wv[h]);
    }
    if (mode == 2 || mode == 5) ld3(zp, pv[h]);
    if (mode == 5) ld3(zn, uv[h]);
  }
  if constexpr (R4) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q[4] = {r4[h].x, r4[h].y, r4[h].z, r4[h].w};
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (q[u] >= 0) {
          const Vec3<T> v = *reinterpret_cast<const Vec3<T>*>(part + 3 * int64_t(q[u]));
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[h][c] += v.v[c];
        }
    }
  }
  for (int32_t j = 0; !R4; ++j) {
    const bool m0 = jb[0] + j < je[0], m1 = jb[1] + j < je[1];
    if (!m0 && !m1) break;
    if (m0) {
      const T* v = part + 3 * int64_t(nq[jb[0] + j]);
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[0][c] += v[c];
    }
    if (m1) {
      const T* v = part + 3 * int64_t(nq[jb[1] + j]);
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[1][c] += v[c];
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = i0 + h * kBlock;
    if (i >= n_nodes) continue;
    Vec3<T> out, wout;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T s = sv[h][c], o = acc[h][c];
      if (mode == 0) {
        out.v[c] = s * o;
      } else if (mode == 3) {
        const T wk = s * s * (wv[h][c] - o);
        wout.v[c] = wk;
        out.v[c] = cw * wk;
      } else if (mode == 4) {
        out.v[c] = wv[h][c] - o;
      } else {
        const T zk = zv[h][c];
        const T dz = mode == 1 ? zk : zk - pv[h][c];
        const T znew = zk + cd * dz + cw * (wv[h][c] - s * s * o);
        out.v[c] = mode == 5 ? uv[h][c] + znew : znew;
      }
    }
    *reinterpret_cast<Vec3<T>*>(zn + 3 * i) = out;
    if (mode == 3) *reinterpret_cast<Vec3<T>*>(w + 3 * i) = wout;
  }
}

template <class T>
__global__ void to_type(int64_t n, const double* __restrict__ a, T* __restrict__ b) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    b[i] = T(a[i]);
}

// inv[perm[j]] = j
__global__ void invert_permutation(int64_t n, const uint32_t* __restrict__ perm,
                                   uint32_t* __restrict__ inv) {
  for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < n;
       j += int64_t(gridDim.x) * blockDim.x)
    inv[perm[j]] = uint32_t(j);
}


// Diagonal 3x3 node blocks of the Tet10 elasticity matrix without
// assembling it: block_i = sum_(e,a) V (lam M + mu M^T + mu tr(M) I),
// M = Jinv^T R_aa Jinv.
__global__ void mf_t10_diag(int64_t n_nodes, const int32_t* __restrict__ offs,
                            const uint32_t* __restrict__ perm, const double* __restrict__ geo,
                            double lam, double mu, double* __restrict__ blocks) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    double acc[3][3] = {};
    for (int32_t j = offs[i]; j < offs[i + 1]; ++j) {
      const uint32_t item = perm[j];
      const int64_t e = item / 10u;
      const int a = int(item - uint32_t(e) * 10u);
      const double* g = geo + e * 10;
      double T[3][3], M[3][3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          T[k][c] = kT10m.R[a][a][k][0] * g[c] + kT10m.R[a][a][k][1] * g[3 + c] +
                    kT10m.R[a][a][k][2] * g[6 + c];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) M[r][c] = g[r] * T[0][c] + g[3 + r] * T[1][c] + g[6 + r] * T[2][c];
      const double lv = lam * g[9], mv = mu * g[9];
      const double tr = mv * (M[0][0] + M[1][1] + M[2][2]);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[r][c] += lv * M[r][c] + mv * M[c][r] + (r == c ? tr : 0.0);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) blocks[9 * i + 3 * r + c] = acc[r][c];
  }
}

// Preconditioner from diagonal node blocks (same rules as diag_nb3).
__global__ void blocks_to_precond(int64_t n_nodes, const double* __restrict__ blocks,
                                  const uint8_t* __restrict__ fixed, int block3,
                                  double* __restrict__ dinv, int32_t* __restrict__ bad) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    double blk[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) blk[r][c] = blocks[9 * i + 3 * r + c];
    bool fx[3];
    for (int c = 0; c < 3; ++c) fx[c] = fixed && fixed[3 * i + c];
    if (!block3) {
      for (int c = 0; c < 3; ++c) {
        const double d = blk[c][c];
        if (fx[c]) {
          dinv[3 * i + c] = 0.0;
        } else if (d > 0.0) {
          dinv[3 * i + c] = 1.0 / d;
        } else {
          dinv[3 * i + c] = 0.0;
          atomicAdd(bad, 1);
        }
      }
      continue;
    }
    for (int c = 0; c < 3; ++c)
      if (fx[c])
        for (int t = 0; t < 3; ++t) blk[c][t] = blk[t][c] = (c == t) ? 1.0 : 0.0;
    double o[3][3];
    if (!inv3(blk, o)) {
      atomicAdd(bad, 1);
      for (int t = 0; t < 9; ++t) dinv[9 * i + t] = 0.0;
      continue;
    }
    for (int c = 0; c < 3; ++c)
      for (int cp = 0; cp < 3; ++cp) dinv[9 * i + 3 * c + cp] = (fx[c] || fx[cp]) ? 0.0 : o[c][cp];
  }
}

// ------------------------------------------------------------- sliced NB3

// Solver-internal copy of an NB3 matrix in a sliced layout (SELL-C-sigma over
// node rows with C = 16 rows per slice and sigma = 2048): node rows are stably
// sorted by decreasing length inside windows of 2048 rows and cut into slices
// of 16.  A wave owns one slice; its 64 lanes are 4 phases x 16 rows, phase h
// taking blocks k = 4g + h of its row, so a slice of longest row L costs
// ceil(L / 4) steps g.  Step g of slice s stores, for lane = 16h + row:
// the block column at cols[(koff[s] + g) * 64 + lane] and the 9 block values
// at vals[((koff[s] + g) * 9 + q) * 64 + lane], q = 3c + c'.  Every load of
// the product is a coalesced 64-lane access; the 4 phases of a row are summed
// with two lane swaps at the end; padding (rows shorter than the slice's
// longest, about 5 % on Tet10 box meshes) holds zero values and column -1.
// The FP64 product skips the loads of padding lanes: in the last step of a
// slice of equal-length rows whole phases are padding, so 16 lanes (one
// 128-byte line of each value plane) are never fetched.  The FP32/FP16
// copies load them (zeros, gathering x[0]): a padding phase is only half or
// a quarter of a line there, and the divergent masking cost more than it saved
// (cheb_step<float> 124 vs 117 us, <_Float16> 226 vs 179 us at 539,601 rows).
// Four phases give 4x more waves than one-row-per-lane slices, which keeps
// every SIMD busy on meshes of a few 10^5 nodes.
constexpr int kSliceRows = 16;
constexpr int kPhases = 64 / kSliceRows;
constexpr int kSigma = 2048;
constexpr int kSellUnroll = 4;
// FP64 values: 5 steps in flight per lane (207.2 vs 209.3 us per product at
// 539,601 node rows over a full 6,000-iteration solve; 3, 6 and 8 steps:
// 209.3, 224.8, 218.1 us)
constexpr int kSellUnroll64 = 5;
// consecutive slices of one wave: 1 measured fastest (4: 251 vs 218 us per
// FP64 product at 539,601 node rows), the CG update folds the partials
constexpr int kSlicesPerWave = 1;

// Slices per die: the product's blocks are dealt round-robin to the kXcd
// dies, so block b runs on die d = b % kXcd.  xsplit[d] .. xsplit[d + 1] is a
// contiguous range of slices with about 1/kXcd of the steps, and the blocks
// of die d walk it in order: the x entries a die gathers then come from one
// neighbourhood of the mesh and stay in its L2.  (Splitting by slice count
// instead puts the long corner-node rows on one die.)  FP64 product at
// 539,601 node rows: 208.0 us against 214.5 us with slices dealt round-robin.
// Returns the wave's first slice and the end of its die's range.
__device__ __forceinline__ int64_t sell_first_slice(const int32_t* __restrict__ xsplit,
                                                    int64_t* end) {
  const int d = int(blockIdx.x % kXcd);
  const int64_t i = blockIdx.x / kXcd;
  *end = xsplit[d + 1];
  return xsplit[d] + (i * kWaves + threadIdx.x / 64) * kSlicesPerWave;
}

// xsplit[d] = first slice s with koff[s] >= d W / kXcd (W = total steps);
// one thread per die boundary.
__global__ void sell_xcd_split(int64_t n_slices, const int32_t* __restrict__ koff,
                               int32_t* __restrict__ xsplit) {
  const int d = threadIdx.x;
  if (d > kXcd) return;
  const int64_t W = koff[n_slices];
  const int64_t target = (W * d + kXcd - 1) / kXcd;
  int64_t lo = 0, hi = n_slices;  // first s in [0, n_slices] with koff[s] >= target
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (koff[mid] >= target) hi = mid; else lo = mid + 1;
  }
  xsplit[d] = int32_t(d == 0 ? 0 : (d == kXcd ? n_slices : lo));
}

// Sort key: window index above, descending row length below (exact, so the
// sorted slice lengths never exceed the unsorted ones the workspace is sized by).
__global__ void sell_keys(int64_t n, const int32_t* __restrict__ ptr, uint64_t* __restrict__ keys,
                          uint32_t* __restrict__ ids) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t len = uint32_t(ptr[i + 1] - ptr[i]);
    keys[i] = (uint64_t(i / kSigma) << 32) | uint64_t(0xFFFFFFFFu - len);
    ids[i] = uint32_t(i);
  }
}

// steps[s] = ceil(longest row of slice s / kPhases), rows[] in sorted order
__global__ void sell_slice_len(int64_t n, int64_t n_slices, const int32_t* __restrict__ ptr,
                               const uint32_t* __restrict__ rows, int32_t* __restrict__ steps) {
  for (int64_t sl = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; sl < n_slices;
       sl += int64_t(gridDim.x) * blockDim.x) {
    int32_t mx = 0;
    for (int t = 0; t < kSliceRows; ++t) {
      const int64_t lin = sl * kSliceRows + t;
      if (lin >= n) break;
      const int64_t r = rows[lin];
      const int32_t len = ptr[r + 1] - ptr[r];
      mx = len > mx ? len : mx;
    }
    steps[sl] = (mx + kPhases - 1) / kPhases;
  }
}

// One thread per (slice, lane): fills the lane's column of every step.
template <class OT>
__global__ void sell_fill(int64_t n, int64_t n_slices, const int32_t* __restrict__ ptr,
                          const int32_t* __restrict__ idx, const double* __restrict__ val,
                          const uint32_t* __restrict__ rows, const int32_t* __restrict__ koff,
                          const double* __restrict__ scale, int32_t* __restrict__ scol,
                          OT* __restrict__ sval) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n_slices * 64;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t sl = t / 64;
    const int lane = int(t - sl * 64);
    const int h = lane / kSliceRows;
    const int64_t lin = sl * kSliceRows + (lane % kSliceRows);
    const int64_t r = lin < n ? int64_t(rows[lin]) : -1;
    const int32_t g0 = koff[sl], steps = koff[sl + 1] - g0;
    const int32_t base = r >= 0 ? ptr[r] : 0;
    const int32_t len = r >= 0 ? ptr[r + 1] - base : 0;
    for (int32_t g = 0; g < steps; ++g) {
      const int32_t k = g * kPhases + h;
      const int64_t at = int64_t(g0) + g;
      int32_t col = -1;
      double v[9] = {};
      if (k < len) {
        col = idx[base + k];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int cp = 0; cp < 3; ++cp) {
            double a = val[int64_t(9) * base + c * 3 * len + 3 * k + cp];
            if (scale) a = scale[3 * r + c] * a * scale[3 * int64_t(col) + cp];
            v[3 * c + cp] = a;
          }
      }
      scol[at * 64 + lane] = col;
#pragma unroll
      for (int q = 0; q < 9; ++q) sval[(at * 9 + q) * 64 + lane] = OT(v[q]);
    }
  }
}

// Row products of one slice: one slice per wave, kSellUnroll steps per
// iteration with the next steps' column indices loaded one iteration ahead,
// so the x gathers of a step never wait on its own index load.  Returns the
// row (or -1) whose 3 sums `o` this lane holds after the phase reduction.
template <class VT, class XT, class AT, bool ABS = false>
__device__ __forceinline__ int64_t sell_rows(int64_t n, int64_t sl,
                                             const uint32_t* __restrict__ rows,
                                             const int32_t* __restrict__ koff,
                                             const int32_t* __restrict__ scol,
                                             const VT* __restrict__ sval,
                                             const XT* __restrict__ x, AT (&o)[3]) {
  constexpr int U = sizeof(VT) == 8 ? kSellUnroll64 : kSellUnroll;
  constexpr bool kMask = sizeof(VT) == 8;
  const int lane = threadIdx.x & 63;
  const int64_t lin = sl * kSliceRows + (lane % kSliceRows);
  const int32_t g0 = koff[sl], g1 = koff[sl + 1];
  AT a0 = 0, a1 = 0, a2 = 0;
  int32_t c[U];
#pragma unroll
  for (int u = 0; u < U; ++u) c[u] = g0 + u < g1 ? scol[int64_t(g0 + u) * 64 + lane] : -1;
  for (int32_t g = g0; g < g1; g += U) {
    AT xv[U][3], v[U][9];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = g + u < g1 && (!kMask || c[u] >= 0);
      const int64_t c3 = c[u] >= 0 ? 3 * int64_t(c[u]) : 0;
      // the block column's three x entries as one access (FP64: 16 + 8 bytes
      // instead of three 8-byte gathers)
      Vec3<XT> xx{};
      if (ok) xx = *reinterpret_cast<const Vec3<XT>*>(x + c3);
      xv[u][0] = AT(xx.v[0]);
      xv[u][1] = AT(xx.v[1]);
      xv[u][2] = AT(xx.v[2]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const VT* p = sval + int64_t(g + u) * 9 * 64 + lane;
      const bool ok = g + u < g1 && (!kMask || c[u] >= 0);
#pragma unroll
      for (int q = 0; q < 9; ++q) v[u][q] = ok ? AT(ld(p + q * 64)) : AT(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      c[u] = g + U + u < g1 ? scol[int64_t(g + U + u) * 64 + lane] : -1;
    if constexpr (ABS) {
      // |A| |x|: absolute values at the use (FMA source modifiers), so the
      // loads keep their batched waits
#pragma unroll
      for (int u = 0; u < U; ++u) {
        a0 += fabs(v[u][0]) * fabs(xv[u][0]) + fabs(v[u][1]) * fabs(xv[u][1]) +
              fabs(v[u][2]) * fabs(xv[u][2]);
        a1 += fabs(v[u][3]) * fabs(xv[u][0]) + fabs(v[u][4]) * fabs(xv[u][1]) +
              fabs(v[u][5]) * fabs(xv[u][2]);
        a2 += fabs(v[u][6]) * fabs(xv[u][0]) + fabs(v[u][7]) * fabs(xv[u][1]) +
              fabs(v[u][8]) * fabs(xv[u][2]);
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        a0 += v[u][0] * xv[u][0] + v[u][1] * xv[u][1] + v[u][2] * xv[u][2];
        a1 += v[u][3] * xv[u][0] + v[u][4] * xv[u][1] + v[u][5] * xv[u][2];
        a2 += v[u][6] * xv[u][0] + v[u][7] * xv[u][1] + v[u][8] * xv[u][2];
      }
    }
  }
#pragma unroll
  for (int m = kSliceRows; m < 64; m <<= 1) {
    a0 += __shfl_xor(a0, m, 64);
    a1 += __shfl_xor(a1, m, 64);
    a2 += __shfl_xor(a2, m, 64);
  }
  o[0] = a0;
  o[1] = a1;
  o[2] = a2;
  return (lane < kSliceRows && lin < n) ? int64_t(rows[lin]) : -1;
}

// y = A x on the sliced copy, fixed rows zeroed; with `state`: p.Ap
// partials.  Grid: the product grid of build_sell (blocks of 4 waves, kXcd
// per-die slice ranges, see sell_first_slice).
template <class VT, class XT, class AT>
__global__ void __launch_bounds__(kBlock)
spmv_sell(int64_t n, int64_t n_slices, const uint32_t* __restrict__ rows,
          const int32_t* __restrict__ koff, const int32_t* __restrict__ scol,
          const VT* __restrict__ sval, const XT* __restrict__ x, XT* __restrict__ y,
          const uint8_t* __restrict__ fixed, PcgState* state, double* __restrict__ partials,
          const int32_t* __restrict__ xsplit) {
  if (state && state->halt) return;
  int64_t end;
  const int64_t sl0 = sell_first_slice(xsplit, &end);
  double dot = 0.0;
  for (int64_t sl = sl0; sl < sl0 + kSlicesPerWave && sl < end; ++sl) {
    // the row's fixed flags and x entries are loaded before the row products,
    // so the stores at the end wait on nothing (with the 24-byte x gathers of
    // sell_rows: 208.1 against 210.0 us per FP64 product at 539,601 node rows)
    const int lane = threadIdx.x & 63;
    const int64_t lin = sl * kSliceRows + (lane % kSliceRows);
    const int64_t r0 = (lane < kSliceRows && lin < n) ? int64_t(rows[lin]) : -1;
    bool fx[3] = {false, false, false};
    XT xd[3] = {XT(0), XT(0), XT(0)};
    if (r0 >= 0) {
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        fx[cc] = fixed && fixed[3 * r0 + cc];
        if (state) xd[cc] = x[3 * r0 + cc];
      }
    }
    AT o[3];
    sell_rows<VT, XT, AT>(n, sl, rows, koff, scol, sval, x, o);
    if (r0 >= 0) {
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const int64_t d = 3 * r0 + cc;
        if (fx[cc]) o[cc] = AT(0);
        y[d] = XT(o[cc]);
        if (state) dot += double(xd[cc]) * double(XT(o[cc]));
      }
    }
  }
  if (!state) return;
  double v1[1] = {dot};
  block_partials<1>(v1, partials);
}

// || |A_ff| |x_f| ||^2 partials on the sliced copy (xf: x with its fixed
// entries zeroed; fixed rows are dropped): the rounding floor of the FP64
// residual b_f - A_ff x_f is about eps || |A_ff| |x_f| || (Higham, Accuracy
// and Stability of Numerical Algorithms, 2002, sec. 3.5).  Grid: as spmv_sell.
__global__ void __launch_bounds__(kBlock)
sell_abs_norm(int64_t n, int64_t n_slices, const uint32_t* __restrict__ rows,
              const int32_t* __restrict__ koff, const int32_t* __restrict__ scol,
              const double* __restrict__ sval, const double* __restrict__ xf,
              const uint8_t* __restrict__ fixed, double* __restrict__ partials,
              const int32_t* __restrict__ xsplit) {
  int64_t end;
  const int64_t sl0 = sell_first_slice(xsplit, &end);
  double sq = 0.0;
  for (int64_t sl = sl0; sl < sl0 + kSlicesPerWave && sl < end; ++sl) {
    double o[3];
    const int64_t r = sell_rows<double, double, double, true>(n, sl, rows, koff, scol, sval, xf,
                                                              o);
    if (r >= 0) {
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        if (!(fixed && fixed[3 * r + cc])) sq += o[cc] * o[cc];
    }
  }
  double v1[1] = {sq};
  block_partials<1>(v1, partials);
}

// The same for a scalar CSR matrix (grid-stride rows).
__global__ void __launch_bounds__(kBlock)
csr_abs_norm(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
             const double* __restrict__ val, const double* __restrict__ xf,
             const uint8_t* __restrict__ fixed, double* __restrict__ partials) {
  double sq = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (fixed && fixed[i]) continue;
    double acc = 0.0;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) acc += fabs(val[k]) * fabs(xf[idx[k]]);
    sq += acc * acc;
  }
  double v1[1] = {sq};
  block_partials<1>(v1, partials);
}

// One step of the Chebyshev iteration on S A S (Jacobi-scaled, FP32 vectors,
// values stored as VT):  w -= S A S d;  d' = cd d + cw w;  z += d'.
// S zeroes fixed rows and columns, so no mask is needed.
template <class VT>
__global__ void __launch_bounds__(kBlock)
cheb_step(int64_t n, int64_t n_slices, const uint32_t* __restrict__ rows,
          const int32_t* __restrict__ koff, const int32_t* __restrict__ scol,
          const VT* __restrict__ sval, const float* __restrict__ d, float* __restrict__ dn,
          float* __restrict__ w, float* __restrict__ z, float cd, float cw,
          const PcgState* state, const int32_t* __restrict__ xsplit) {
  if (state->halt) return;
  int64_t end;
  const int64_t sl0 = sell_first_slice(xsplit, &end);
  for (int64_t sl = sl0; sl < sl0 + kSlicesPerWave && sl < end; ++sl) {
    float o[3];
    const int64_t r = sell_rows<VT, float, float>(n, sl, rows, koff, scol, sval, d, o);
    if (r < 0) continue;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const int64_t i = 3 * r + cc;
      const float wn = w[i] - o[cc];
      w[i] = wn;
      const float dd = cd * d[i] + cw * wn;
      dn[i] = dd;
      z[i] += dd;
    }
  }
}

// ------------------------------------------------------------- finalize

enum Fin : int {
  kFinAlpha = 0,      // pAp -> alpha (or breakdown)
  kFinCgUpdate = 1,   // (r.r, r.z) -> convergence test, beta
  kFinStore = 2,      // sum -> *out
  kFinStart = 3,      // (r.r, r.z) -> state
  kFinMxStart = 4,    // r.r of a fresh FP32 residual, new direction
  kFinMxReplace = 5,  // r.r of a replaced FP32 residual, kept direction
  kFinMxUpdate = 6,   // recursive FP32 r.r -> beta / replacement request
  kFinFcgCheck = 7,   // r.r after the update -> convergence test
  kFinFcgBeta = 8,    // (r.z, r.z_old) -> flexible (Polak-Ribiere) beta
  kFinMax = 9,        // max of the partials -> *out
};

// One block folds the per-block partials of the previous kernel (G of
// them per value, fixed order) and updates the CG scalars.
__global__ void __launch_bounds__(kBlock)
finalize(int mode, int G, const double* __restrict__ partials, PcgState* state,
         double* __restrict__ out, double aux) {
  const bool gated = (mode == kFinAlpha || mode == kFinCgUpdate || mode == kFinMxUpdate ||
                      mode == kFinFcgCheck || mode == kFinFcgBeta);
  if (gated && state->halt) return;
  if (mode == kFinMax) {
    __shared__ double red[kBlock / 64];
    double m = 0.0;
    for (int b = threadIdx.x; b < G; b += blockDim.x) m = fmax(m, partials[b]);
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double mm = red[0];
      for (int k = 1; k < kBlock / 64; ++k) mm = fmax(mm, red[k]);
      *out = mm;
    }
    return;
  }
  const int nv = (mode == kFinCgUpdate || mode == kFinStart || mode == kFinFcgBeta) ? 2 : 1;
  double v[2] = {0.0, 0.0};
  for (int k = 0; k < nv; ++k)
    for (int b = threadIdx.x; b < G; b += blockDim.x) v[k] += partials[k * G + b];
  block_sum<2>(v);
  if (threadIdx.x != 0) return;
  switch (mode) {
    case kFinAlpha:
      state->pAp = v[0];
      if (!(v[0] > 0.0) || !isfinite(v[0]))
        state->halt = kBreakdown;
      else
        state->alpha = state->rz / v[0];
      break;
    case kFinCgUpdate:
      state->iter += 1;
      if (!isfinite(v[0]) || !isfinite(v[1])) {
        state->halt = kBreakdown;
      } else if (v[0] <= state->stop2) {
        state->rr = v[0];
        state->halt = kConverged;
      } else if (state->iter >= state->maxit) {
        state->rr = v[0];
        state->halt = kMaxit;
      } else {
        state->beta = v[1] / state->rz;
        state->rz = v[1];
        state->rr = v[0];
      }
      break;
    case kFinStore:
      *out = v[0];
      break;
    case kFinStart:
      state->rr = v[0];
      state->rz = v[1];
      state->rzs[0] = v[1];
      break;
    case kFinMxStart:
      state->rz = v[0];
      state->rr = v[0];
      state->refresh2 = v[0] * aux;
      break;
    case kFinMxReplace:
      // keep the search direction: beta = new r.r / previous r.r
      state->beta = v[0] / state->rz;
      state->rz = v[0];
      state->rr = v[0];
      state->refresh2 = v[0] * aux;
      break;
    case kFinMxUpdate: {
      state->iter += 1;
      const double rn = v[0];
      if (!isfinite(rn)) {
        state->halt = kBreakdown;
      } else if (rn <= state->stop2 || rn <= state->refresh2 || state->iter >= state->maxit) {
        // leave rz at the previous r.r: after the replacement the direction
        // update uses beta = (S r_true).(S r_true) / previous r.r
        state->rr = rn;
        state->halt = kRefresh;
      } else {
        state->beta = rn / state->rz;
        state->rz = rn;
        state->rr = rn;
      }
      break;
    }
    case kFinFcgCheck:
      state->iter += 1;
      state->rr = v[0];
      if (!isfinite(v[0]))
        state->halt = kBreakdown;
      else if (v[0] <= state->stop2)
        state->halt = kConverged;
      else if (state->iter >= state->maxit)
        state->halt = kMaxit;
      break;
    case kFinFcgBeta:
      // v[0] = r.z (new), v[1] = r.z_old; rz holds the previous r.z
      if (!(v[0] > 0.0) || !isfinite(v[0]) || !isfinite(v[1])) {
        state->halt = kBreakdown;
      } else {
        state->beta = (v[0] - v[1]) / state->rz;
        state->rz = v[0];
      }
      break;
    default:
      break;
  }
}

// ------------------------------------------------------------- host side

// Optional timing of the CG product kernels: HIP events around every 8th
// SpMV launch of a solve, on the solve's stream.  Averages are kept per
// value type (FP64 / FP32) until cf_spmv_profile_reset().
struct SpmvProfile {
  bool enabled = false;
  double ms[2] = {0.0, 0.0};
  int64_t launches[2] = {0, 0};
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, int>> pending;  // (event index, precision slot)
  size_t used = 0;
};
SpmvProfile g_prof;
constexpr int kProfileEvery = 8;

hipEvent_t prof_event() {
  if (g_prof.used == g_prof.pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_prof.pool.push_back(e);
  }
  return g_prof.pool[g_prof.used++];
}

void prof_begin(int slot, int iter, hipStream_t s, int* token) {
  *token = -1;
  if (!g_prof.enabled || iter % kProfileEvery) return;
  hipEvent_t a = prof_event(), b = prof_event();
  if (!a || !b) return;
  hipEventRecord(a, s);
  *token = int(g_prof.used - 2);
  g_prof.pending.push_back({*token, slot});
}

void prof_end(int token, hipStream_t s) {
  if (token >= 0) hipEventRecord(g_prof.pool[token + 1], s);
}

// Called after the stream is synchronised: fold finished pairs in.
void prof_collect() {
  for (auto& pr : g_prof.pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, g_prof.pool[pr.first], g_prof.pool[pr.first + 1]) ==
        hipSuccess) {
      g_prof.ms[pr.second] += ms;
      g_prof.launches[pr.second] += 1;
    }
  }
  g_prof.pending.clear();
  g_prof.used = 0;
}

int n_dofs(const cf_matrix* A) { return A->format == CF_FMT_NB3 ? 3 * A->n_rows : A->n_rows; }

// Lanes per row for the SpMV kernels: NB3 rows of Tet10 meshes hold 10..65
// blocks (90..585 values), P1 scalar rows 5..15 entries.
constexpr int kLanesNb3 = 32;
constexpr int kLanesCsr = 8;

template <class VT, class XT, class AT>
int launch_spmv(const cf_matrix* A, const VT* vals, const XT* x, XT* y, const uint8_t* fixed,
                PcgState* state, double* partials, int grid, hipStream_t s) {
  if (A->format == CF_FMT_NB3) {
    spmv_nb3_wave<VT, XT, AT><<<grid, kBlock, 0, s>>>(A->n_rows, A->indptr, A->indices, vals, x,
                                                      y, fixed, state, partials);
  } else {
    spmv_csr<VT, XT, AT, kLanesCsr><<<grid, kBlock, 0, s>>>(A->n_rows, A->indptr, A->indices,
                                                            vals, x, y, fixed, state, partials);
  }
  CF_LAUNCHED();
  return CF_OK;
}

struct Layout {
  double* r;
  double* p;
  double* Ap;
  double* dinv;
  double* partials;
  double* partials_b;  // (r.r, r.z) partials of cg_update_fused
  PcgState* state;
  int32_t* bad;
  double* lz;  // Lanczos coefficients (two-level set-up)
  // mixed only
  double* s;
  float* rh;
  float* ph;
  float* Aph;
  float* yh;
  float* vals32;
  // sliced copy of an NB3 matrix (solver-internal product layout)
  int64_t n_slices;
  int32_t* xsplit;      // kXcd + 1 slice boundaries (sell_xcd_split)
  uint32_t* rows;
  int32_t* koff;
  int32_t* slice_len;
  int32_t* scol;
  double* sval64;
  float* sval32;
  _Float16* sval16;
  double* z0;  // flexible-CG preconditioned residuals (current / previous)
  double* z1;
  uint64_t* keys_in;
  uint64_t* keys_out;
  uint32_t* ids_in;
  void* tmp;
  size_t tmp_bytes;
  size_t total;
};

int pcg_grid() {
  int g = cf::cu_count() * 8;
  return g - g % kXcd;
}

// blocks for `tiles` tiles (see tile_index)
int64_t tile_grid(int64_t tiles) { return tiles < 1 ? 1 : tiles; }

int64_t sell_grid(int64_t n_slices) {
  const int64_t per_block = (kBlock / 64) * kSlicesPerWave;
  return tile_grid((n_slices + per_block - 1) / per_block);
}

// grid of the fused CG update / direction kernels: each of their blocks folds
// the previous reduction's partials, so fewer, fuller blocks
int pcg_vec_grid() {
  int g = cf::cu_count() * 4;
  return g - g % kXcd;
}

int bits_for(int64_t n) {
  int b = 1;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

// Upper bound on the number of steps of the sliced copy: the slice steps of
// the unsorted rows (sorting inside windows only lowers each slice maximum's
// rank, so the sorted total never exceeds this).
// Steps of the sliced layout before the window sort: per slice of
// kSliceRows rows, ceil(longest row / kPhases), summed over the slices with
// integer atomics (exact in any order).  One wave-reduced add per wave.
__device__ unsigned long long g_sell_steps;

__global__ void __launch_bounds__(kBlock)
sell_steps_count(int64_t n_rows, const int32_t* __restrict__ ptr,
                 unsigned long long* __restrict__ out) {
  const int64_t n_slices = (n_rows + kSliceRows - 1) / kSliceRows;
  unsigned long long sum = 0;
  for (int64_t sl = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; sl < n_slices;
       sl += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r1 = (sl + 1) * kSliceRows < n_rows ? (sl + 1) * kSliceRows : n_rows;
    int32_t mx = 0;
    for (int64_t r = sl * kSliceRows; r < r1; ++r) {
      const int32_t len = ptr[r + 1] - ptr[r];
      mx = len > mx ? len : mx;
    }
    sum += uint64_t((mx + kPhases - 1) / kPhases);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) sum += __shfl_xor(sum, m, 64);
  if ((threadIdx.x & 63) == 0 && sum) atomicAdd(out, sum);
}

__global__ void zero_u64(unsigned long long* p) { *p = 0; }

// Workspace bound of the sliced copy (total steps of the unsorted slices;
// the window sort never makes them longer), counted on the device so that
// only 8 bytes cross to the host.  Calls serialise on a process-wide lock
// around the per-device counter.  s: the caller's stream (0: null stream).
int sell_bound(const cf_matrix* A, int64_t* lsum, hipStream_t s = nullptr) {
  static std::mutex lock;
  std::lock_guard<std::mutex> guard(lock);
  unsigned long long* d = nullptr;
  CF_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&d), HIP_SYMBOL(g_sell_steps)));
  zero_u64<<<1, 1, 0, s>>>(d);
  CF_LAUNCHED();
  const int64_t n_slices = (int64_t(A->n_rows) + kSliceRows - 1) / kSliceRows;
  sell_steps_count<<<cf::stride_grid(n_slices, kBlock), kBlock, 0, s>>>(
      A->n_rows, static_cast<const int32_t*>(A->indptr), d);
  CF_LAUNCHED();
  unsigned long long total = 0;
  CF_HIP(hipMemcpyAsync(&total, d, sizeof(total), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  *lsum = int64_t(total);
  return CF_OK;
}

// Carves the sliced copy of an NB3 matrix (slice structure, FP64 values and,
// for the mixed precisions, the low-precision values) and its sort space.
void sell_layout(const cf_matrix* A, int precision, int64_t lsum, cf::Carve& c, Layout& L) {
  const int64_t N = A->n_rows;
  L.n_slices = (N + kSliceRows - 1) / kSliceRows;
  L.rows = c.take<uint32_t>(N);
  L.koff = c.take<int32_t>(L.n_slices + 1);
  L.slice_len = c.take<int32_t>(L.n_slices + 1);
  L.xsplit = c.take<int32_t>(kXcd + 1);
  L.scol = c.take<int32_t>(size_t(64) * lsum);
  L.sval64 = c.take<double>(size_t(9 * 64) * lsum);
  if (precision == CF_PREC_MIXED) L.sval32 = c.take<float>(size_t(9 * 64) * lsum);
  if (precision == CF_PREC_MIXED16) L.sval16 = c.take<_Float16>(size_t(9 * 64) * lsum);
  L.keys_in = c.take<uint64_t>(N);
  L.keys_out = c.take<uint64_t>(N);
  L.ids_in = c.take<uint32_t>(N);
  size_t sort_bytes = 0, scan_bytes = 0;
  const int end_bit = bits_for((N / kSigma + 1)) + 32;
  rocprim::radix_sort_pairs<SortConfig>(nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                            (uint32_t*)nullptr, (uint32_t*)nullptr, size_t(N), 0, end_bit);
  rocprim::exclusive_scan(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, 0,
                          size_t(L.n_slices + 1), rocprim::plus<int32_t>());
  L.tmp_bytes = std::max(sort_bytes, scan_bytes);
  L.tmp = c.take<char>(L.tmp_bytes);
}

// partial sums written by one sliced product (its grid, at most)
int64_t sell_parts(int64_t n_rows) {
  return kXcd * sell_grid((n_rows + kSliceRows - 1) / kSliceRows);
}

Layout layout(const cf_matrix* A, int precision, int precond, void* base, int64_t lsum) {
  const int64_t n = n_dofs(A);
  const bool nb3 = A->format == CF_FMT_NB3;
  const int64_t nval = nb3 ? 9 * A->nnz : A->nnz;
  cf::Carve c(base);
  Layout L{};
  L.r = c.take<double>(n);
  L.p = c.take<double>(n);
  L.Ap = c.take<double>(n);
  L.dinv = c.take<double>(precond == CF_PC_BLOCK3 ? 3 * n : n);
  const int64_t parts =
      nb3 ? std::max<int64_t>(pcg_grid(), sell_parts(A->n_rows)) : pcg_grid();
  L.partials = c.take<double>(4 * size_t(parts));
  L.partials_b = c.take<double>(2 * size_t(pcg_vec_grid()));
  L.state = c.take<PcgState>(1);
  L.bad = c.take<int32_t>(1);
  const bool mixed = precision != CF_PREC_FP64;
  if (mixed) {
    L.s = c.take<double>(n);
    L.rh = c.take<float>(n);
    L.ph = c.take<float>(n);
    L.Aph = c.take<float>(n);
    L.yh = c.take<float>(n);
    if (!nb3) L.vals32 = c.take<float>(nval);
  }
  if (mixed && nb3) {
    L.z0 = c.take<double>(n);
    L.z1 = c.take<double>(n);
  }
  if (nb3) sell_layout(A, precision, lsum, c, L);
  L.total = c.used;
  return L;
}

// Slice structure of the sliced copy: window sort, slice steps, offsets, and
// the per-die slice ranges; *product_grid = blocks of the sliced products.
// Grid of the sliced product: kXcd x the blocks of the longest per-die slice
// range (reads the die split back, so it synchronises with stream s).
int sell_product_grid(const Layout& L, hipStream_t s, int* product_grid) {
  int32_t xs[kXcd + 1];
  CF_HIP(hipMemcpyAsync(xs, L.xsplit, sizeof(xs), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  int64_t most = 0;
  for (int d = 0; d < kXcd; ++d) most = std::max<int64_t>(most, xs[d + 1] - xs[d]);
  *product_grid = int(kXcd * sell_grid(most));
  return CF_OK;
}

// product_grid null: the caller reads the grid later (sell_product_grid).
int build_sell(const cf_matrix* A, const Layout& L, hipStream_t s, int* product_grid) {
  const int64_t N = A->n_rows;
  const int grid = cf::stride_grid(N, kBlock);
  sell_keys<<<grid, kBlock, 0, s>>>(N, A->indptr, L.keys_in, L.ids_in);
  CF_LAUNCHED();
  size_t tb = L.tmp_bytes;
  const int end_bit = bits_for((N / kSigma + 1)) + 32;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(L.tmp, tb, L.keys_in, L.keys_out, L.ids_in, L.rows, size_t(N),
                                   0, end_bit, s));
  CF_HIP(hipMemsetAsync(L.slice_len + L.n_slices, 0, sizeof(int32_t), s));
  sell_slice_len<<<cf::stride_grid(L.n_slices, kBlock), kBlock, 0, s>>>(N, L.n_slices, A->indptr,
                                                                         L.rows, L.slice_len);
  CF_LAUNCHED();
  tb = L.tmp_bytes;
  CF_HIP(rocprim::exclusive_scan(L.tmp, tb, L.slice_len, L.koff, 0, size_t(L.n_slices + 1),
                                 rocprim::plus<int32_t>(), s));
  sell_xcd_split<<<1, 64, 0, s>>>(L.n_slices, L.koff, L.xsplit);
  CF_LAUNCHED();
  return product_grid ? sell_product_grid(L, s, product_grid) : CF_OK;
}

// Values of the sliced copy: A (scale null) or S A S, stored as OT.
template <class OT>
int fill_sell(const cf_matrix* A, const Layout& L, const double* scale, OT* out, hipStream_t s) {
  const int fgrid = cf::stride_grid(L.n_slices * 64, kBlock);
  sell_fill<OT><<<fgrid, kBlock, 0, s>>>(A->n_rows, L.n_slices, A->indptr, A->indices, A->values,
                                         L.rows, L.koff, scale, L.scol, out);
  CF_LAUNCHED();
  return CF_OK;
}

int check_matrix(const cf_matrix* A) {
  CF_REQUIRE(A && A->indptr && A->values && A->n_rows > 0 && A->nnz >= 0,
             "cf_matrix: bad matrix");
  CF_REQUIRE(A->format == CF_FMT_CSR || A->format == CF_FMT_NB3, "cf_matrix: bad format %d",
             A->format);
  return CF_OK;
}

int read_state(const PcgState* d, PcgState* h, hipStream_t s) {
  CF_HIP(hipMemcpyAsync(h, d, sizeof(PcgState), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  return CF_OK;
}

// Fold the partials of the kernel just launched (grid pcg_grid()).
// `parts`: number of block partials the preceding kernel wrote (its grid)
int fin(int mode, const Layout& L, hipStream_t s, double* out = nullptr, double aux = 0.0,
        int parts = 0) {
  finalize<<<1, kBlock, 0, s>>>(mode, parts > 0 ? parts : pcg_grid(), L.partials, L.state, out,
                                aux);
  CF_LAUNCHED();
  return CF_OK;
}

}  // namespace

extern "C" int cf_spmv(const cf_matrix* A, const double* x, double* y, const uint8_t* fixed,
                       void* stream) {
  int rc = check_matrix(A);
  if (rc) return rc;
  CF_REQUIRE(x && y, "cf_spmv: null vector");
  hipStream_t s = cf::as_stream(stream);
  return launch_spmv<double, double, double>(A, A->values, x, y, fixed, nullptr, nullptr,
                                             pcg_grid(), s);
}

namespace {

// y = |A| |x| (fixed rows 0), one thread per DOF row: the magnitude that
// bounds the FP64 rounding of A x (used once per solve, not on a hot path).
__global__ void spmv_abs(int32_t format, int64_t n_rows, const int32_t* __restrict__ ptr,
                         const int32_t* __restrict__ idx, const double* __restrict__ val,
                         const double* __restrict__ x, const uint8_t* __restrict__ fixed,
                         double* __restrict__ y) {
  const int64_t n = format == CF_FMT_NB3 ? 3 * n_rows : n_rows;
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x) {
    double acc = 0.0;
    if (format == CF_FMT_NB3) {
      const int64_t i = t / 3;
      const int c = int(t - 3 * i);
      const int32_t base = ptr[i], len = ptr[i + 1] - base;
      const double* v = val + int64_t(9) * base + int64_t(c) * 3 * len;
      for (int32_t k = 0; k < len; ++k) {
        const int64_t j3 = 3 * int64_t(idx[base + k]);
        acc += fabs(v[3 * k]) * fabs(x[j3]) + fabs(v[3 * k + 1]) * fabs(x[j3 + 1]) +
               fabs(v[3 * k + 2]) * fabs(x[j3 + 2]);
      }
    } else {
      for (int32_t k = ptr[t]; k < ptr[t + 1]; ++k) acc += fabs(val[k]) * fabs(x[idx[k]]);
    }
    y[t] = (fixed && fixed[t]) ? 0.0 : acc;
  }
}

}  // namespace

extern "C" int cf_spmv_abs(const cf_matrix* A, const double* x, double* y, const uint8_t* fixed,
                           void* stream) {
  int rc = check_matrix(A);
  if (rc) return rc;
  CF_REQUIRE(x && y, "cf_spmv_abs: null vector");
  hipStream_t s = cf::as_stream(stream);
  const int64_t n = n_dofs(A);
  spmv_abs<<<cf::stride_grid(n, kBlock), kBlock, 0, s>>>(A->format, A->n_rows, A->indptr,
                                                         A->indices, A->values, x, fixed, y);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_pcg_workspace_bytes(const cf_matrix* A, int32_t precision, int32_t precond,
                                      size_t* bytes) {
  int rc = check_matrix(A);
  if (rc) return rc;
  CF_REQUIRE(bytes, "cf_pcg_workspace_bytes: null output");
  int64_t lsum = 0;
  if (A->format == CF_FMT_NB3 && (rc = sell_bound(A, &lsum))) return rc;
  *bytes = layout(A, precision, precond, nullptr, lsum).total;
  return CF_OK;
}

namespace {

// FP64 PCG on the masked free system.  `apply(xin, y, state, partials)`
// launches y = A xin with fixed rows zeroed; when `state` is set it also
// writes `apply_parts` block partials of p.Ap (the CG product).  `dinv` holds the
// preconditioner (Jacobi 1/A_ii, or inverse 3x3 node blocks).
// Iterations to enqueue before the next host check.  The remaining count is
// predicted from the faster of two convergence rates (since the start, and
// over the last batch when the residual fell there), and 80 % of it is
// enqueued: iterations launched after convergence exit at once but still
// cost a launch each, and CG residuals are not monotone, so the prediction
// must not overshoot.  Without a usable rate the batch doubles.
int next_batch(const PcgState& h, double rr_start, double prev_rr, int prev_it, int batch,
               int lo, int hi) {
  if (h.halt != kRunning) return batch;
  double rate = 0.0;  // log residual decrease per iteration (< 0 when converging)
  if (h.iter > 0 && h.rr < rr_start) rate = std::log(h.rr / rr_start) / h.iter;
  const int done_it = h.iter - prev_it;
  if (done_it > 0 && h.rr < prev_rr) rate = std::min(rate, std::log(h.rr / prev_rr) / done_it);
  if (!(rate < 0.0)) return batch * 2 > hi ? hi : batch * 2;
  const double need = std::log(h.stop2 / h.rr) / rate;
  const int nb = int(0.8 * need) + 2;
  return nb < lo ? lo : (nb > hi ? hi : nb);
}

// Residual replacement at the end of a CG solve.  CG stops on its
// recursive residual, which drifts from the true one b - A x by rounding.
// When the true residual misses the bar, the solver first measures the FP64
// floor eps || |A_ff| |x_f| || (where the operator has stored values): a
// true residual within kFloorFactor of it is what FP64 can resolve for this
// system, so the solve stops there with floor_limited set.  Otherwise CG
// restarts from the true residual (it replaces the recursive one) for a
// segment that ends when the recursive residual has fallen 10x below the
// restart value (or to the bar); a restart that does not halve the true
// residual also ends the solve with floor_limited set.
struct TrueResidual {
  static constexpr int kMaxRestarts = 8;
  // a true residual within this factor of the FP64 floor is accepted
  static constexpr double kFloorFactor = 2.0;
  double bar2;            // (tol ||b_f||)^2
  double last2 = 0.0;     // last true r.r
  double prev2 = HUGE_VAL;// true r.r at the previous restart
  double floor_rr = 0.0;  // (kFloorFactor eps || |A_ff| |x_f| ||)^2, 0 = unknown
  int restarts = 0;
  int floor_limited = 0;
  explicit TrueResidual(double bar) : bar2(bar) {}
  // || |A_ff| |x_f| ||^2 measured at the current iterate
  void set_floor(double absnorm2) {
    const double e = kFloorFactor * std::numeric_limits<double>::epsilon();
    floor_rr = absnorm2 > 0.0 && std::isfinite(absnorm2) ? e * e * absnorm2 : 0.0;
  }
  // true r.r of the current iterate -> whether to restart from it
  bool restart(double rr2) {
    last2 = rr2;
    if (!(rr2 > bar2)) return false;  // met (or NaN: reported, no restart)
    if (rr2 <= floor_rr) {            // at the FP64 floor: nothing left to gain
      floor_limited = 1;
      return false;
    }
    if (restarts >= kMaxRestarts || rr2 > 0.25 * prev2) {
      floor_limited = 1;
      return false;
    }
    prev2 = rr2;
    ++restarts;
    return true;
  }
  double segment_stop2() const { return std::max(bar2, 0.01 * last2); }
  // A restart the caller requires whatever the residual (the outer operator
  // changed): the segment then reduces the new true residual by
  // sqrt(1 / kForcedStop2).  The residual of a matrix-free solution in the
  // assembled system is rounding-level; what the segment has to remove is
  // the low-mode part of the solution difference, which the two-level cycle
  // takes out within a few iterations.
  static constexpr double kForcedStop2 = 0.1;
  bool forced = false;
  bool force(double rr2) {
    last2 = rr2;
    forced = rr2 > 0.0;
    if (!forced) return false;
    prev2 = rr2;
    ++restarts;
    return true;
  }
  double stop2_after(bool was_forced) const {
    return was_forced ? kForcedStop2 * last2 : segment_stop2();
  }
};

// FP64 floor of a converged iterate whose true residual misses the bar:
// abs_norm(xf) enqueues || |A_ff| |xf| ||^2 into state->floor2 (returns
// CF_OK, < 0 when the current operator has no stored values, or an error).
// L.Ap is free at this point and holds x_f.
template <class AbsNorm>
int measure_floor(AbsNorm& abs_norm, int64_t n, const double* x, const uint8_t* fixed,
                  const Layout& L, TrueResidual& tr, hipStream_t s) {
  if constexpr (std::is_same_v<std::decay_t<AbsNorm>, std::nullptr_t>) {
    return CF_OK;
  } else {
    split_vector<<<pcg_grid(), kBlock, 0, s>>>(n, x, fixed, 0, L.Ap);
    CF_LAUNCHED();
    const int fr = abs_norm(static_cast<const double*>(L.Ap));
    if (fr < 0) return CF_OK;
    if (fr) return fr;
    PcgState t{};
    const int rc = read_state(L.state, &t, s);
    if (rc) return rc;
    tr.set_floor(t.floor2);
    return CF_OK;
  }
}

template <int PC, class Apply, class AbsNorm = std::nullptr_t>
int pcg_fp64_core(Apply&& apply, int apply_parts, int64_t n, int64_t units, const double* b,
                  double* x, const uint8_t* fixed, double tol, int maxit, const Layout& L,
                  cf_pcg_stats* st, hipStream_t s, AbsNorm&& abs_norm = nullptr) {
  const int G = pcg_grid();
  int rc;
  PcgState h{};
  // reference rhs: b_f - A_fF x_F
  split_vector<<<G, kBlock, 0, s>>>(n, x, fixed, 1, L.p);
  CF_LAUNCHED();
  if ((rc = apply(L.p, L.Ap, nullptr, nullptr))) return rc;
  residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials, &L.state->pad);
  CF_LAUNCHED();
  if ((rc = fin(kFinStore, L, s, &L.state->pad))) return rc;
  // r = b - A x  (full x, warm start included)
  if ((rc = apply(x, L.Ap, nullptr, nullptr))) return rc;
  residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials, &L.state->rr);
  CF_LAUNCHED();
  if ((rc = fin(kFinStore, L, s, &L.state->rr))) return rc;
  cg_start<PC><<<G, kBlock, 0, s>>>(units, L.r, L.p, L.dinv, L.state, L.partials);
  CF_LAUNCHED();
  if ((rc = fin(kFinStart, L, s))) return rc;
  if ((rc = read_state(L.state, &h, s))) return rc;
  const double bb = h.pad;
  st->b_norm = std::sqrt(bb);
  st->iterations = 0;
  st->residual_updates = 0;
  st->restarts = 0;
  st->floor_limited = 0;
  if (!std::isfinite(bb) || !std::isfinite(h.rr)) {
    cf::set_error("cf_pcg_solve: non-finite right-hand side");
    return CF_ERR_BREAKDOWN;
  }
  if (bb == 0.0) {
    // zero reduced right-hand side: the solution of the free block is 0
    split_vector<<<G, kBlock, 0, s>>>(n, x, fixed, 1, x);
    CF_LAUNCHED();
    CF_HIP(hipStreamSynchronize(s));
    st->rel_residual = 0.0;
    return CF_OK;
  }
  h.stop2 = tol * tol * bb;
  h.maxit = maxit;
  h.iter = 0;
  h.halt = h.rr <= h.stop2 ? kConverged : kRunning;
  h.ticket = 0;
  CF_HIP(hipMemcpyAsync(L.state, &h, sizeof(PcgState), hipMemcpyHostToDevice, s));
  int batch = 8;
  double prev_rr = h.rr;
  double rr_start = h.rr;
  int prev_it = 0;
  const int GV = pcg_vec_grid();
  int64_t enq = 0;
  TrueResidual tr(tol * tol * bb);
  // three launches per iteration: product (+ p.Ap partials), update (folds
  // p.Ap itself), direction (folds r.r, r.z itself)
  for (;;) {
    while (h.halt == kRunning) {
      for (int t = 0; t < batch; ++t, ++enq) {
        int tok;
        prof_begin(0, h.iter + t, s, &tok);
        if ((rc = apply(L.p, L.Ap, L.state, L.partials))) return rc;
        prof_end(tok, s);
        const int slot = int(enq & 1);
        cg_update_fused<PC><<<GV, kBlock, 0, s>>>(units, x, L.r, L.p, L.Ap, L.dinv, L.state,
                                                  L.partials, apply_parts, L.partials_b, slot);
        CF_LAUNCHED();
        cg_direction_fused<PC><<<GV, kBlock, 0, s>>>(units, L.r, L.p, L.dinv, L.state,
                                                     L.partials_b, GV, slot);
        CF_LAUNCHED();
      }
      if ((rc = read_state(L.state, &h, s))) return rc;
      prof_collect();
      // size the next batch from the observed convergence rate
      batch = next_batch(h, rr_start, prev_rr, prev_it, batch, 8, 512);
      prev_rr = h.rr;
      prev_it = h.iter;
    }
    // true FP64 residual of the free system (r = b - A x, its r.r into pad)
    if ((rc = apply(x, L.Ap, nullptr, nullptr))) return rc;
    residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials, &L.state->pad);
    CF_LAUNCHED();
    if ((rc = fin(kFinStore, L, s, &L.state->pad))) return rc;
    PcgState ts{};
    if ((rc = read_state(L.state, &ts, s))) return rc;
    tr.last2 = ts.pad;
    if (h.halt == kConverged && ts.pad > tr.bar2 && tr.floor_rr == 0.0 &&
        (rc = measure_floor(abs_norm, n, x, fixed, L, tr, s)))
      return rc;
    if (h.halt != kConverged || !tr.restart(ts.pad)) break;
    // residual replacement: restart CG from the true residual now in r
    cg_start<PC><<<G, kBlock, 0, s>>>(units, L.r, L.p, L.dinv, L.state, L.partials);
    CF_LAUNCHED();
    if ((rc = fin(kFinStart, L, s))) return rc;
    if ((rc = read_state(L.state, &h, s))) return rc;
    h.stop2 = tr.segment_stop2();
    h.halt = h.rr <= h.stop2 ? kConverged : (h.iter >= maxit ? kMaxit : kRunning);
    CF_HIP(hipMemcpyAsync(L.state, &h, sizeof(PcgState), hipMemcpyHostToDevice, s));
    enq = 0;
    batch = 8;
    prev_rr = rr_start = h.rr;
    prev_it = h.iter;
  }
  st->iterations = h.iter;
  st->rel_residual = std::sqrt(tr.last2 / bb);
  st->restarts = tr.restarts;
  st->floor_limited = tr.floor_limited;
  if (h.halt == kBreakdown) {
    cf::set_error("cf_pcg_solve: breakdown after %d iterations (NaN/Inf or p.Ap <= 0)", h.iter);
    return CF_ERR_BREAKDOWN;
  }
  if (h.halt == kMaxit) {
    cf::set_error("cf_pcg_solve: not converged after %d iterations (rel. residual %.3e)",
                  h.iter, st->rel_residual);
    return CF_ERR_NOT_CONVERGED;
  }
  return CF_OK;
}

template <int PC>
int pcg_fp64(const cf_matrix* A, const double* b, double* x, const uint8_t* fixed, double tol,
             int maxit, const Layout& L, cf_pcg_stats* st, hipStream_t s) {
  const int64_t n = n_dofs(A);
  const int64_t units = PC == CF_PC_BLOCK3 ? A->n_rows : n;
  const int G = pcg_grid();
  int rc;
  int GS = G;
  if (A->format == CF_FMT_NB3 &&
      ((rc = build_sell(A, L, s, &GS)) || (rc = fill_sell<double>(A, L, nullptr, L.sval64, s))))
    return rc;
  const int32_t* xsplit = L.xsplit;
  auto apply = [&](const double* xin, double* y, PcgState* state, double* partials) {
    if (A->format == CF_FMT_NB3) {
      spmv_sell<double, double, double><<<GS, kBlock, 0, s>>>(
          A->n_rows, L.n_slices, L.rows, L.koff, L.scol, L.sval64, xin, y, fixed, state, partials,
          xsplit);
      CF_LAUNCHED();
      return CF_OK;
    }
    return launch_spmv<double, double, double>(A, A->values, xin, y, fixed, state, partials, G,
                                               s);
  };
  auto abs_norm = [&](const double* xf) -> int {
    if (A->format == CF_FMT_NB3) {
      sell_abs_norm<<<GS, kBlock, 0, s>>>(A->n_rows, L.n_slices, L.rows, L.koff, L.scol,
                                          L.sval64, xf, fixed, L.partials, xsplit);
      CF_LAUNCHED();
      return fin(kFinStore, L, s, &L.state->floor2, 0.0, GS);
    }
    csr_abs_norm<<<G, kBlock, 0, s>>>(A->n_rows, A->indptr, A->indices, A->values, xf, fixed,
                                      L.partials);
    CF_LAUNCHED();
    return fin(kFinStore, L, s, &L.state->floor2, 0.0, G);
  };
  return pcg_fp64_core<PC>(apply, GS, n, units, b, x, fixed, tol, maxit, L, st, s, abs_norm);
}

// NB3 systems: FP64 flexible CG, preconditioned by S p(S A S) S with p the
// Chebyshev iteration (steps - 1 products) on [b / ratio, b] run in FP32 on a
// sliced copy of S A S stored as VT (float or _Float16), or matrix-free.  b is
// the smaller of the Gershgorin bound and 1.25x a 40-step power estimate;
// p(t) > 0 for every t < b + b / ratio, so the preconditioner is SPD even where
// rounding of the stored values shifts the lowest eigenvalues.  Polak-Ribiere
// beta keeps the outer iteration robust to the FP32 rounding inside p.
// Degree and ratio, measured on the 160x20x20 Tet10 cantilever (ms per
// solve, FP32 matrix-free / FP32 sliced / FP16 sliced): (8, 100) 579 / 1218 /
// 1685; (16, 3000) 502 / 1106 / 1652; (20, 1000) 496 / 1134 / 1709;
// (24, 3000) 493 / 1095 / 1864; (32, 10000) 500 / 1256 / 1974.  FP16
// rounding of the stored values makes high degrees lose there.
struct ChebShape {
  int steps;
  double ratio;
};
constexpr ChebShape kChebF32{24, 3000.0};
// inside the two-level preconditioner the coarse space takes the lowest modes
constexpr ChebShape kCheb2L{24, 3000.0};
constexpr ChebShape kChebF16{16, 3000.0};
constexpr int kChebMaxSteps = 32;
constexpr int kPowerSteps = 40;

// Enqueues the power iteration for the largest eigenvalue of S A S on
// stream s (power(v, y): y = S A S v on that stream); the square of the
// estimate ends in state->pad, read with read_state.
// Vectors of the low-precision part (Chebyshev, power iteration) are carved
// as raw bytes; T is float, or double for the FP64 twin of the solvers.
template <class T>
T* lp(float* p) { return reinterpret_cast<T*>(p); }

template <class T = float, class Power>
int power_launch(Power&& power, int64_t n, const Layout& L, hipStream_t s) {
  const int G = pcg_grid();
  int rc;
  T* ph = lp<T>(L.ph);
  T* Aph = lp<T>(L.Aph);
  power_start<T><<<G, kBlock, 0, s>>>(n, L.s, ph);
  CF_LAUNCHED();
  for (int k = 0; k < kPowerSteps; ++k) {
    if ((rc = power(ph, Aph))) return rc;
    norm2_partials<T><<<G, kBlock, 0, s>>>(n, Aph, L.partials);
    CF_LAUNCHED();
    if ((rc = fin(kFinStore, L, s, &L.state->pad))) return rc;
    power_scale<T><<<G, kBlock, 0, s>>>(n, Aph, ph, L.state);
    CF_LAUNCHED();
  }
  return CF_OK;
}

// Lanczos estimate of the largest eigenvalue of S K S (the largest Ritz
// value of the tridiagonal T_k; C. Lanczos 1950, C. C. Paige 1972): from the
// power iteration's start vector, 12 steps reach the estimate of its 40
// steps (0.968 against 0.965 of the largest eigenvalue on a Tet10 80x10x10
// cantilever), so the same Chebyshev interval costs 12 products instead of
// 40.  lz[2k] = alpha_k = v_k . (S K S v_k), lz[2k + 1] = beta_k^2.
constexpr int kLanczosSteps = 12;

template <class T>
__global__ void __launch_bounds__(kBlock)
lz_dot(int64_t n, const T* __restrict__ v, const T* __restrict__ y, double* __restrict__ partials) {
  double a = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    a += double(v[i]) * double(y[i]);
  double v1[1] = {a};
  block_partials<1>(v1, partials);
}

// w = y - alpha_k v - beta_{k-1} vp in place in y (no vp term at k = 0),
// partials of ||w||^2
template <class T>
__global__ void __launch_bounds__(kBlock)
lz_update(int64_t n, T* __restrict__ y, const T* __restrict__ v, const T* __restrict__ vp,
          const double* __restrict__ lz, int k, double* __restrict__ partials) {
  const T a = T(lz[2 * k]);
  double acc = 0.0;
  if (k == 0) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
      const T w = y[i] - a * v[i];
      y[i] = w;
      acc += double(w) * double(w);
    }
  } else {
    const T b = T(sqrt(lz[2 * k - 1]));
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
      const T w = y[i] - a * v[i] - b * vp[i];
      y[i] = w;
      acc += double(w) * double(w);
    }
  }
  double v1[1] = {acc};
  block_partials<1>(v1, partials);
}

// v_{k+1} = w / beta_k in place (0 after a breakdown, beta_k = 0)
template <class T>
__global__ void lz_scale(int64_t n, T* __restrict__ w, const double* __restrict__ lz, int k) {
  const double b2 = lz[2 * k + 1];
  const T inv = b2 > 0.0 ? T(1.0 / sqrt(b2)) : T(0);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    w[i] *= inv;
}

// kLanczosSteps steps on three rotating vectors (ph, Aph, yh); power(v, y):
// y = S K S v.  The coefficients stay in L.lz (lanczos_max reads them).
template <class T, class Power>
int lanczos_launch(Power&& power, int64_t n, const Layout& L, hipStream_t s) {
  const int G = pcg_grid();
  int rc;
  T* buf[3] = {lp<T>(L.ph), lp<T>(L.Aph), lp<T>(L.yh)};
  power_start<T><<<G, kBlock, 0, s>>>(n, L.s, buf[0]);
  CF_LAUNCHED();
  norm2_partials<T><<<G, kBlock, 0, s>>>(n, buf[0], L.partials);
  CF_LAUNCHED();
  if ((rc = fin(kFinStore, L, s, &L.state->pad))) return rc;
  power_scale<T><<<G, kBlock, 0, s>>>(n, buf[0], buf[0], L.state);
  CF_LAUNCHED();
  for (int k = 0; k < kLanczosSteps; ++k) {
    T* v = buf[k % 3];
    T* vp = buf[(k + 2) % 3];
    T* w = buf[(k + 1) % 3];
    if ((rc = power(v, w))) return rc;
    lz_dot<T><<<G, kBlock, 0, s>>>(n, v, w, L.partials);
    CF_LAUNCHED();
    if ((rc = fin(kFinStore, L, s, &L.lz[2 * k]))) return rc;
    lz_update<T><<<G, kBlock, 0, s>>>(n, w, v, vp, L.lz, k, L.partials);
    CF_LAUNCHED();
    if ((rc = fin(kFinStore, L, s, &L.lz[2 * k + 1]))) return rc;
    if (k + 1 < kLanczosSteps) {
      lz_scale<T><<<G, kBlock, 0, s>>>(n, w, L.lz, k);
      CF_LAUNCHED();
    }
  }
  return CF_OK;
}

// The Lanczos coefficients to the host (waits for the iteration on s).
int read_lanczos(const Layout& L, double* lz, hipStream_t s) {
  CF_HIP(hipMemcpyAsync(lz, L.lz, sizeof(double) * 2 * kLanczosSteps, hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  return CF_OK;
}

// Largest eigenvalue of the tridiagonal T (alpha_k, beta_k = sqrt(lz[2k + 1]),
// k < m): bisection on the Sturm count of the LDL^T pivots below x, inside
// the Gershgorin interval.
double lanczos_max(const double* lz, int m) {
  double lo = 0.0, hi = 0.0;
  for (int k = 0; k < m; ++k) {
    const double b0 = k > 0 ? std::sqrt(lz[2 * k - 1]) : 0.0;
    const double b1 = k + 1 < m ? std::sqrt(lz[2 * k + 1]) : 0.0;
    lo = k == 0 ? lz[0] - b0 - b1 : std::min(lo, lz[2 * k] - b0 - b1);
    hi = k == 0 ? lz[0] + b0 + b1 : std::max(hi, lz[2 * k] + b0 + b1);
  }
  auto below = [&](double x) {  // eigenvalues of T below x
    int c = 0;
    double d = 1.0;
    for (int k = 0; k < m; ++k) {
      const double b2 = k > 0 ? lz[2 * k - 1] : 0.0;
      d = lz[2 * k] - x - (k > 0 ? b2 / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0.0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(std::fabs(hi), 1.0); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (below(mid) >= m) hi = mid; else lo = mid;
  }
  return hi;
}

// The flexible CG + Chebyshev driver, independent of how the products are
// formed.  outer(xin, y, state): FP64 y = A xin with fixed rows zeroed
// (state set: p.Ap partials, outer_parts of them); power(v, y): FP32
// y = S A S v; cheb(k, cd, cw, d, dn): Chebyshev step k (w -= S A S d;
// dn = cd d + cw w; z += dn on rh / yh).  gersh: Gershgorin bound of S A S,
// or 0 when none is known (then b = 1.3 x the power estimate); shape: degree
// and interval ratio of the polynomial.  SQ: cheb works on S-scaled vectors
// in three-term form (d = z_k in, dn = z_{k-1} in and z_{k+1} out; see
// fcg_update and mf_gather_cheb32).
// Outer iterations enqueued between host checks of the flexible CG: an
// iteration is ~25 launches, and those enqueued after convergence still cost
// their launch on the host, so batches stay short (the check itself is one
// small read-back per batch).
constexpr int kFcgBatchMax = 16;
// Relative recursive residual at which a two-operator solve switches to its
// final outer operator (the assembled matrix).  Measured at 384,000 Tet10
// elements: switching at 1e-6 or 1e-7 gives 56 outer iterations, switching
// after convergence (with a restart from the final operator's residual) 60,
// switching at 1e-8 71 (the operators' rounding difference is no longer
// small against r).
constexpr double kSwitchRel = 1e-6;

template <bool SQ, class T, class Outer, class Power, class Cheb,
          class Pre2L = std::nullptr_t, class SwitchOuter = std::nullptr_t,
          class AbsNorm = std::nullptr_t>
int fcg_cheb_core(Outer&& outer, const int& outer_parts, Power&& power, Cheb&& cheb, double gersh,
                  ChebShape shape, int64_t n, const double* b, double* x, const uint8_t* fixed, double tol,
                  int maxit, const Layout& L, cf_pcg_stats* st, hipStream_t s,
                  Pre2L&& pre2l = nullptr, const T* zh2l = nullptr, int extra_products = 0,
                  double power_est = 0.0, SwitchOuter&& switch_outer = nullptr,
                  AbsNorm&& abs_norm = nullptr) {
  constexpr bool kTwoLevel = !std::is_same<std::decay_t<Pre2L>, std::nullptr_t>::value;
  constexpr bool kSwitch = !std::is_same<std::decay_t<SwitchOuter>, std::nullptr_t>::value;
  const int G = pcg_grid();
  int rc;
  PcgState h{};
  // power_est > 0: the caller ran power_launch already (state->pad = its square)
  double est = power_est;
  if (!(est > 0.0)) {
    if ((rc = power_launch<T>(power, n, L, s))) return rc;
    if ((rc = read_state(L.state, &h, s))) return rc;
    est = std::sqrt(h.pad);
  }
  const double bmax = gersh > 0.0 ? std::min(1.25 * est, gersh) : 1.3 * est;
  if (!(bmax > 0.0) || !std::isfinite(bmax)) {
    cf::set_error("cf_pcg_solve: no spectrum bound for the Chebyshev preconditioner "
                  "(power estimate %.3e, Gershgorin %.3e)", est, gersh);
    return CF_ERR_BREAKDOWN;
  }
  if (shape.steps < 2 || shape.steps > kChebMaxSteps || !(shape.ratio > 1.0)) {
    cf::set_error("cf_pcg_solve: bad Chebyshev shape (%d steps, ratio %g)", shape.steps,
                  shape.ratio);
    return CF_ERR_ARG;
  }
  const double amin = bmax / shape.ratio;
  const double theta = 0.5 * (bmax + amin), delta = 0.5 * (bmax - amin), sigma = theta / delta;
  T cd[kChebMaxSteps], cw[kChebMaxSteps];
  double rho0 = 1.0 / sigma;
  for (int k = 1; k < shape.steps; ++k) {
    const double rho1 = 1.0 / (2.0 * sigma - rho0);
    cd[k] = T(rho1 * rho0);
    cw[k] = T(2.0 * rho1 / delta);
    rho0 = rho1;
  }
  const T inv_theta = T(1.0 / theta);
  T* D[2] = {lp<T>(L.ph), lp<T>(L.Aph)};
  T* rh = lp<T>(L.rh);
  T* yh = lp<T>(L.yh);
  double* Z[2] = {L.z0, L.z1};
  // One outer iteration is: product (p.Ap partials), fcg_update_fused (x, r
  // and the Chebyshev start), the Chebyshev steps, fcg_zdots_fused (stop test,
  // z = S zh), fcg_direction_fused -- no finalize launches.  w = rh, zh = yh
  // (SQ: zh = the last three-term iterate).
  const int GV = pcg_vec_grid();
  auto precondition = [&](int apply_step, double* zc, const double* zo, int sample, int slot) {
    if constexpr (kTwoLevel) {
      // two-level: the hook updates x, r and forms the preconditioned residual in zh2l
      int r2 = pre2l(apply_step, slot, sample, cd, cw, inv_theta, shape.steps);
      if (r2) return r2;
      fcg_zdots_fused<SQ, T><<<GV, kBlock, 0, s>>>(n, L.r, L.s, zh2l, zc, zo, L.state,
                                                   L.partials_b, GV, apply_step, L.partials);
      CF_LAUNCHED();
      fcg_direction_fused<<<GV, kBlock, 0, s>>>(n, zc, L.p, L.state, L.partials, GV, slot);
      CF_LAUNCHED();
      return CF_OK;
    }
    if (apply_step) {
      fcg_update_fused<SQ, T><<<GV, kBlock, 0, s>>>(n, x, L.r, L.p, L.Ap, L.s, rh, D[0], yh,
                                                    inv_theta, L.state, L.partials, outer_parts,
                                                    L.partials_b, slot);
    } else {
      fcg_update<SQ, T><<<G, kBlock, 0, s>>>(n, x, L.r, L.p, L.Ap, L.s, rh, D[0], yh, inv_theta,
                                             0, L.state, L.partials);
    }
    CF_LAUNCHED();
    int r2;
    for (int k = 1; k < shape.steps; ++k) {
      int tok = -1;
      if (k == 1) prof_begin(1, sample, s, &tok);
      if ((r2 = cheb(k, cd[k], cw[k], D[(k - 1) & 1], D[k & 1]))) return r2;
      if (k == 1) prof_end(tok, s);
    }
    // SQ: the last three-term step left z in D[(steps - 1) & 1]
    const T* zh = SQ ? D[(shape.steps - 1) & 1] : yh;
    fcg_zdots_fused<SQ, T><<<GV, kBlock, 0, s>>>(n, L.r, L.s, zh, zc, zo, L.state, L.partials_b,
                                                 GV, apply_step, L.partials);
    CF_LAUNCHED();
    fcg_direction_fused<<<GV, kBlock, 0, s>>>(n, zc, L.p, L.state, L.partials, GV, slot);
    CF_LAUNCHED();
    return CF_OK;
  };
  // reference rhs b_f - A_fF x_F, then r = b - A x
  split_vector<<<G, kBlock, 0, s>>>(n, x, fixed, 1, L.p);
  CF_LAUNCHED();
  if ((rc = outer(L.p, L.Ap, nullptr))) return rc;
  residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials, &L.state->pad);
  CF_LAUNCHED();
  if ((rc = fin(kFinStore, L, s, &L.state->pad))) return rc;
  if ((rc = outer(x, L.Ap, nullptr))) return rc;
  residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials, &L.state->rr);
  CF_LAUNCHED();
  if ((rc = fin(kFinStore, L, s, &L.state->rr))) return rc;
  if ((rc = read_state(L.state, &h, s))) return rc;
  const double bb = h.pad;
  st->b_norm = std::sqrt(bb);
  st->iterations = 0;
  st->residual_updates = 0;
  st->restarts = 0;
  st->floor_limited = 0;
  if (!std::isfinite(bb) || !std::isfinite(h.rr)) {
    cf::set_error("cf_pcg_solve: non-finite right-hand side");
    return CF_ERR_BREAKDOWN;
  }
  if (bb == 0.0) {
    split_vector<<<G, kBlock, 0, s>>>(n, x, fixed, 1, x);
    CF_LAUNCHED();
    CF_HIP(hipStreamSynchronize(s));
    st->rel_residual = 0.0;
    return CF_OK;
  }
  h.stop2 = tol * tol * bb;
  h.maxit = maxit;
  h.iter = 0;
  h.rz = 1.0;
  h.rzs[0] = h.rzs[1] = 1.0;
  h.halt = h.rr <= h.stop2 ? kConverged : kRunning;
  CF_HIP(hipMemcpyAsync(L.state, &h, sizeof(PcgState), hipMemcpyHostToDevice, s));
  // first direction: z_old = 0 and p = 0, so p = z, and r.z lands in rzs[0]
  CF_HIP(hipMemsetAsync(Z[0], 0, size_t(n) * sizeof(double), s));
  CF_HIP(hipMemsetAsync(L.p, 0, size_t(n) * sizeof(double), s));
  int opened = 0;  // preconditioner applications that opened a CG segment
  if (h.halt == kRunning) {
    if ((rc = precondition(0, Z[1], Z[0], 1, 1))) return rc;
    ++opened;
  }
  int64_t enq = 0;
  int batch = 8;
  double prev_rr = h.rr;
  double rr_start = h.rr;
  int prev_it = 0;
  TrueResidual tr(tol * tol * bb);
  bool switched_mid = false;
  double switch2 = 0.0;  // switch the outer operator when r.r <= switch2
  if constexpr (kSwitch) switch2 = kSwitchRel * kSwitchRel * bb;
  for (;;) {
    while (h.halt == kRunning) {
      for (int t = 0; t < batch; ++t, ++enq) {
        int tok;
        prof_begin(0, int(enq), s, &tok);
        if ((rc = outer(L.p, L.Ap, L.state))) return rc;
        prof_end(tok, s);
        if ((rc = precondition(1, Z[enq & 1], Z[(enq + 1) & 1], int(enq), int(enq & 1))))
          return rc;
      }
      if ((rc = read_state(L.state, &h, s))) return rc;
      prof_collect();
      if constexpr (kSwitch) {
        // Switch the outer operator while the recursive residual is still far
        // above the rounding-level difference of the two operators: the true
        // residual of the final operator replaces r in place, p and the CG
        // scalars are kept (residual replacement, van der Vorst and Ye, SIAM
        // J. Sci. Comput. 22 (2000) 1635), so no Krylov information is lost.
        if (!switched_mid && h.halt == kRunning && h.rr <= switch2 && switch_outer()) {
          switched_mid = true;
          if ((rc = outer(x, L.Ap, nullptr))) return rc;
          residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials,
                                        &L.state->pad);
          CF_LAUNCHED();
        }
      }
      batch = next_batch(h, rr_start, prev_rr, prev_it, batch, 2, kFcgBatchMax);
      prev_rr = h.rr;
      prev_it = h.iter;
    }
    // converged with the first outer operator: the caller may switch to the
    // final one, which then gets at least one segment of its own
    bool switched = false;
    if constexpr (kSwitch)
      if (h.halt == kConverged && !switched_mid) switched = switch_outer();
    // true FP64 residual (r = b - A x, its r.r into pad)
    if ((rc = outer(x, L.Ap, nullptr))) return rc;
    residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials, &L.state->pad);
    CF_LAUNCHED();
    if ((rc = fin(kFinStore, L, s, &L.state->pad))) return rc;
    PcgState fin_state{};
    if ((rc = read_state(L.state, &fin_state, s))) return rc;
    tr.last2 = fin_state.pad;
    if (h.halt != kConverged) break;
    if (!switched && fin_state.pad > tr.bar2 && tr.floor_rr == 0.0 &&
        (rc = measure_floor(abs_norm, n, x, fixed, L, tr, s)))
      return rc;
    if (switched ? !tr.force(fin_state.pad) : !tr.restart(fin_state.pad)) break;
    // residual replacement: a fresh flexible CG from the true residual in r
    h.stop2 = tr.stop2_after(switched);
    h.rr = fin_state.pad;
    h.rz = 1.0;
    h.rzs[0] = h.rzs[1] = 1.0;
    h.halt = h.rr <= h.stop2 ? kConverged : (h.iter >= maxit ? kMaxit : kRunning);
    CF_HIP(hipMemcpyAsync(L.state, &h, sizeof(PcgState), hipMemcpyHostToDevice, s));
    CF_HIP(hipMemsetAsync(Z[0], 0, size_t(n) * sizeof(double), s));
    CF_HIP(hipMemsetAsync(L.p, 0, size_t(n) * sizeof(double), s));
    if (h.halt == kRunning) {
      if ((rc = precondition(0, Z[1], Z[0], 1, 1))) return rc;
      ++opened;
    }
    enq = 0;
    batch = 4;  // restart segments are short
    prev_rr = rr_start = h.rr;
    prev_it = h.iter;
  }
  st->iterations = h.iter;
  // low-precision products: every outer iteration and every segment start
  // applies the polynomial once (steps - 1 products, plus the two-level extras)
  st->residual_updates = (shape.steps - 1 + extra_products) * (h.iter + opened);
  st->rel_residual = std::sqrt(tr.last2 / bb);
  st->restarts = tr.restarts;
  st->floor_limited = tr.floor_limited;
  if (h.halt == kBreakdown) {
    cf::set_error("cf_pcg_solve: flexible CG breakdown after %d iterations (NaN/Inf, "
                  "p.Ap <= 0 or r.z <= 0)", h.iter);
    return CF_ERR_BREAKDOWN;
  }
  if (h.halt == kMaxit) {
    cf::set_error("cf_pcg_solve: not converged after %d iterations (rel. residual %.3e)",
                  h.iter, st->rel_residual);
    return CF_ERR_NOT_CONVERGED;
  }
  return CF_OK;
}

// Assembled NB3 systems: products on the sliced copies (FP64 outer, S A S
// stored as VT inner).
template <class VT>
int pcg_fcg_cheb(const cf_matrix* A, VT* sval, const double* b, double* x, const uint8_t* fixed,
                 double tol, int maxit, const Layout& L, cf_pcg_stats* st, hipStream_t s) {
  const int64_t n = n_dofs(A);
  const int G = pcg_grid();
  int GS = 0;
  int rc;
  PcgState h{};
  make_scale<<<G, kBlock, 0, s>>>(n, L.dinv, L.s);
  CF_LAUNCHED();
  const int32_t* xsplit = L.xsplit;
  if ((rc = build_sell(A, L, s, &GS)) || (rc = fill_sell<double>(A, L, nullptr, L.sval64, s)) ||
      (rc = fill_sell<VT>(A, L, L.s, sval, s)))
    return rc;
  gershgorin_nb3<<<G, kBlock, 0, s>>>(A->n_rows, A->indptr, A->indices, A->values, L.s,
                                      L.partials);
  CF_LAUNCHED();
  if ((rc = fin(kFinMax, L, s, &L.state->pad, 0.0, G))) return rc;
  if ((rc = read_state(L.state, &h, s))) return rc;
  const double gersh = h.pad;
  auto outer = [&](const double* xin, double* y, PcgState* state) {
    spmv_sell<double, double, double><<<GS, kBlock, 0, s>>>(
        A->n_rows, L.n_slices, L.rows, L.koff, L.scol, L.sval64, xin, y, fixed, state,
        L.partials, xsplit);
    CF_LAUNCHED();
    return CF_OK;
  };
  auto power = [&](const float* v, float* y) {
    spmv_sell<VT, float, float><<<GS, kBlock, 0, s>>>(A->n_rows, L.n_slices, L.rows, L.koff,
                                                       L.scol, sval, v, y, nullptr, nullptr,
                                                       nullptr, xsplit);
    CF_LAUNCHED();
    return CF_OK;
  };
  auto cheb = [&](int, float cdk, float cwk, const float* din, float* dout) {
    cheb_step<VT><<<GS, kBlock, 0, s>>>(A->n_rows, L.n_slices, L.rows, L.koff, L.scol, sval, din,
                                        dout, L.rh, L.yh, cdk, cwk, L.state, xsplit);
    CF_LAUNCHED();
    return CF_OK;
  };
  auto abs_norm = [&](const double* xf) -> int {
    sell_abs_norm<<<GS, kBlock, 0, s>>>(A->n_rows, L.n_slices, L.rows, L.koff, L.scol, L.sval64,
                                        xf, fixed, L.partials, xsplit);
    CF_LAUNCHED();
    return fin(kFinStore, L, s, &L.state->floor2, 0.0, GS);
  };
  return fcg_cheb_core<false, float>(outer, GS, power, cheb, gersh,
                                     sizeof(VT) == 2 ? kChebF16 : kChebF32, n, b, x, fixed, tol,
                                     maxit, L, st, s, nullptr, nullptr, 0, 0.0, nullptr,
                                     abs_norm);
}

// Scalar CSR systems (P1 heat, capacity-dominated theta-method steps): FP32
// inner CG on S A S with FP64 residual replacement.
int pcg_mixed(const cf_matrix* A, const double* b, double* x, const uint8_t* fixed, double tol,
              int maxit, const Layout& L, cf_pcg_stats* st, hipStream_t s) {
  const int64_t n = n_dofs(A);
  const int G = pcg_grid();
  // Replace the residual once the FP32 recursion has dropped it by 10x
  // since the last replacement; restart the direction when the replaced
  // residual is more than 2x the recursive one (the FP32 recursion has
  // drifted); give up after 8 replacements without a 2x gain.
  const double kRefresh2 = 1e-2;
  const double kRestart2 = 4.0;
  const int kStallLimit = 8;
  int rc;
  PcgState h{};
  make_scale<<<G, kBlock, 0, s>>>(n, L.dinv, L.s);
  CF_LAUNCHED();
  scale_csr_to_f32<<<G, kBlock, 0, s>>>(A->n_rows, A->indptr, A->indices, A->values, L.s,
                                        L.vals32);
  CF_LAUNCHED();
  cf_matrix A32 = *A;
  // reference norm ||b_f - A_fF x_F||
  split_vector<<<G, kBlock, 0, s>>>(n, x, fixed, 1, L.p);
  CF_LAUNCHED();
  if ((rc = launch_spmv<double, double, double>(A, A->values, L.p, L.Ap, fixed, nullptr,
                                                nullptr, G, s)))
    return rc;
  residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials, &L.state->pad);
  CF_LAUNCHED();
  if ((rc = fin(kFinStore, L, s, &L.state->pad))) return rc;
  // FP64 residual of the initial guess
  if ((rc = launch_spmv<double, double, double>(A, A->values, x, L.Ap, fixed, nullptr, nullptr,
                                                G, s)))
    return rc;
  residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials, &L.state->rr);
  CF_LAUNCHED();
  if ((rc = fin(kFinStore, L, s, &L.state->rr))) return rc;
  if ((rc = read_state(L.state, &h, s))) return rc;
  const double bb = h.pad;
  st->b_norm = std::sqrt(bb);
  st->iterations = 0;
  st->residual_updates = 0;
  st->restarts = 0;
  st->floor_limited = 0;
  if (!std::isfinite(bb) || !std::isfinite(h.rr)) {
    cf::set_error("cf_pcg_solve: non-finite right-hand side");
    return CF_ERR_BREAKDOWN;
  }
  if (bb == 0.0) {
    split_vector<<<G, kBlock, 0, s>>>(n, x, fixed, 1, x);
    CF_LAUNCHED();
    CF_HIP(hipStreamSynchronize(s));
    st->rel_residual = 0.0;
    return CF_OK;
  }
  const double stop_true2 = tol * tol * bb;
  double true_rr = h.rr;
  if (true_rr <= stop_true2) {
    st->rel_residual = std::sqrt(true_rr / bb);
    return CF_OK;
  }
  CF_HIP(hipMemsetAsync(L.yh, 0, sizeof(float) * n, s));
  mx_load_residual<<<G, kBlock, 0, s>>>(n, L.r, L.s, L.rh, L.ph, 1, L.state, L.partials,
                                        kRefresh2);
  CF_LAUNCHED();
  if ((rc = fin(kFinMxStart, L, s, nullptr, kRefresh2))) return rc;
  if ((rc = read_state(L.state, &h, s))) return rc;
  int iters = 0, updates = 0, stall = 0;
  double best_rr = true_rr;
  for (;;) {
    // Map the true stopping test onto the scaled FP32 residual through the
    // ratio observed at the last replacement.
    const double ratio = true_rr / h.rr;  // ||r||^2 / ||S r||^2
    h.stop2 = stop_true2 / ratio;
    h.maxit = maxit;
    h.iter = iters;
    h.halt = kRunning;
    h.ticket = 0;
    CF_HIP(hipMemcpyAsync(L.state, &h, sizeof(PcgState), hipMemcpyHostToDevice, s));
    int batch = 16;
    while (h.halt == kRunning) {
      for (int t = 0; t < batch; ++t) {
        int tok;
        prof_begin(1, h.iter + t, s, &tok);
        if ((rc = launch_spmv<float, float, float>(&A32, L.vals32, L.ph, L.Aph, nullptr, L.state,
                                                   L.partials, G, s)))
          return rc;
        prof_end(tok, s);
        if ((rc = fin(kFinAlpha, L, s))) return rc;
        mx_update<<<G, kBlock, 0, s>>>(n, L.yh, L.rh, L.ph, L.Aph, L.state, L.partials);
        CF_LAUNCHED();
        if ((rc = fin(kFinMxUpdate, L, s))) return rc;
        mx_direction<<<G, kBlock, 0, s>>>(n, L.rh, L.ph, L.state);
        CF_LAUNCHED();
      }
      if ((rc = read_state(L.state, &h, s))) return rc;
      prof_collect();
      batch = batch * 2 > 256 ? 256 : batch * 2;
    }
    iters = h.iter;
    if (h.halt == kBreakdown) {
      cf::set_error("cf_pcg_solve: breakdown after %d iterations (NaN/Inf or p.Ap <= 0)", iters);
      st->iterations = iters;
      return CF_ERR_BREAKDOWN;
    }
    // residual replacement: x += S y, r = b - A x in FP64
    mx_accumulate<<<G, kBlock, 0, s>>>(n, x, L.yh, L.s);
    CF_LAUNCHED();
    if ((rc = launch_spmv<double, double, double>(A, A->values, x, L.Ap, fixed, nullptr,
                                                  nullptr, G, s)))
      return rc;
    residual<<<G, kBlock, 0, s>>>(n, b, L.Ap, fixed, L.r, L.state, L.partials, &L.state->pad);
    CF_LAUNCHED();
    if ((rc = fin(kFinStore, L, s, &L.state->pad))) return rc;
    PcgState t{};
    if ((rc = read_state(L.state, &t, s))) return rc;
    true_rr = t.pad;
    ++updates;
    if (!std::isfinite(true_rr)) {
      cf::set_error("cf_pcg_solve: non-finite residual after %d iterations", iters);
      st->iterations = iters;
      return CF_ERR_BREAKDOWN;
    }
    if (true_rr <= stop_true2 || iters >= maxit) break;
    if (true_rr < 0.25 * best_rr) {
      best_rr = true_rr;
      stall = 0;
    } else if (++stall >= kStallLimit) {
      break;
    }
    const double rec_rr = h.rr;  // recursive FP32 r.r before the replacement
    mx_load_residual<<<G, kBlock, 0, s>>>(n, L.r, L.s, L.rh, L.ph, 0, L.state, L.partials,
                                          kRefresh2);
    CF_LAUNCHED();
    if ((rc = fin(kFinMxReplace, L, s, nullptr, kRefresh2))) return rc;
    if ((rc = read_state(L.state, &h, s))) return rc;
    if (h.rr > kRestart2 * rec_rr) {
      // restart: p = r (mx_load_residual with start = 1 resets p and rz)
      mx_load_residual<<<G, kBlock, 0, s>>>(n, L.r, L.s, L.rh, L.ph, 1, L.state, L.partials,
                                            kRefresh2);
      CF_LAUNCHED();
      if ((rc = fin(kFinMxStart, L, s, nullptr, kRefresh2))) return rc;
    } else {
      mx_direction_forced<<<G, kBlock, 0, s>>>(n, L.rh, L.ph, L.state);
      CF_LAUNCHED();
    }
    if ((rc = read_state(L.state, &h, s))) return rc;
  }
  st->iterations = iters;
  st->residual_updates = updates;
  st->rel_residual = std::sqrt(true_rr / bb);
  if (true_rr > stop_true2) {
    cf::set_error("cf_pcg_solve: mixed precision not converged after %d iterations and %d "
                  "FP64 residual replacements (rel. residual %.3e%s)",
                  iters, updates, st->rel_residual,
                  stall >= kStallLimit ? ", stagnated: FP32 cannot resolve this system" : "");
    return CF_ERR_NOT_CONVERGED;
  }
  return CF_OK;
}

}  // namespace

extern "C" int cf_pcg_solve(const cf_matrix* A, const double* b, double* x,
                            const uint8_t* fixed, int32_t precision, int32_t precond,
                            double tol, int32_t maxit, void* ws, size_t ws_bytes,
                            cf_pcg_stats* stats_host, void* stream) {
  int rc = check_matrix(A);
  if (rc) return rc;
  CF_REQUIRE(b && x && ws && stats_host, "cf_pcg_solve: null argument");
  CF_REQUIRE(precision == CF_PREC_FP64 || precision == CF_PREC_MIXED ||
                 precision == CF_PREC_MIXED16,
             "cf_pcg_solve: bad precision %d", precision);
  CF_REQUIRE(precond == CF_PC_JACOBI || precond == CF_PC_BLOCK3, "cf_pcg_solve: bad precond %d",
             precond);
  CF_REQUIRE(!(precond == CF_PC_BLOCK3 && A->format != CF_FMT_NB3),
             "cf_pcg_solve: block3 preconditioner needs an NB3 matrix");
  CF_REQUIRE(!(precond == CF_PC_BLOCK3 && precision != CF_PREC_FP64),
             "cf_pcg_solve: mixed precision uses Jacobi scaling");
  CF_REQUIRE(tol > 0.0 && maxit > 0, "cf_pcg_solve: tol and maxit must be positive");
  hipStream_t s = cf::as_stream(stream);
  int64_t lsum = 0;
  if (A->format == CF_FMT_NB3 && (rc = sell_bound(A, &lsum, s))) return rc;
  Layout L = layout(A, precision, precond, ws, lsum);
  if (ws_bytes < L.total) {
    cf::set_error("cf_pcg_solve: workspace %zu < %zu bytes", ws_bytes, L.total);
    return CF_ERR_WORKSPACE;
  }
  const int G = pcg_grid();
  *stats_host = cf_pcg_stats{};
  CF_HIP(hipMemsetAsync(L.state, 0, sizeof(PcgState), s));
  CF_HIP(hipMemsetAsync(L.bad, 0, sizeof(int32_t), s));
  if (A->format == CF_FMT_NB3) {
    diag_nb3<<<G, kBlock, 0, s>>>(A->n_rows, A->indptr, A->indices, A->values, fixed,
                                  precond == CF_PC_BLOCK3, L.dinv, L.bad);
  } else {
    diag_csr<<<G, kBlock, 0, s>>>(A->n_rows, A->indptr, A->indices, A->values, fixed, L.dinv,
                                  L.bad);
  }
  CF_LAUNCHED();
  int32_t bad = 0;
  CF_HIP(hipMemcpyAsync(&bad, L.bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  if (bad) {
    cf::set_error("cf_pcg_solve: %d free diagonal entries/blocks are not positive "
                  "(matrix not SPD on the free DOFs)",
                  bad);
    return CF_ERR_BREAKDOWN;
  }
  if (precision != CF_PREC_FP64 && A->format == CF_FMT_NB3) {
    if (precision == CF_PREC_MIXED16)
      return pcg_fcg_cheb<_Float16>(A, L.sval16, b, x, fixed, tol, maxit, L, stats_host, s);
    return pcg_fcg_cheb<float>(A, L.sval32, b, x, fixed, tol, maxit, L, stats_host, s);
  }
  if (precision != CF_PREC_FP64)
    return pcg_mixed(A, b, x, fixed, tol, maxit, L, stats_host, s);
  if (precond == CF_PC_BLOCK3)
    return pcg_fp64<CF_PC_BLOCK3>(A, b, x, fixed, tol, maxit, L, stats_host, s);
  return pcg_fp64<CF_PC_JACOBI>(A, b, x, fixed, tol, maxit, L, stats_host, s);
}

extern "C" int cf_spmv_profile(int32_t enable, double* avg_ms_fp64, int64_t* n_fp64,
                               double* avg_ms_fp32, int64_t* n_fp32) {
  if (avg_ms_fp64) *avg_ms_fp64 = g_prof.launches[0] ? g_prof.ms[0] / g_prof.launches[0] : 0.0;
  if (n_fp64) *n_fp64 = g_prof.launches[0];
  if (avg_ms_fp32) *avg_ms_fp32 = g_prof.launches[1] ? g_prof.ms[1] / g_prof.launches[1] : 0.0;
  if (n_fp32) *n_fp32 = g_prof.launches[1];
  if (enable >= 0) {
    g_prof.enabled = enable != 0;
    g_prof.ms[0] = g_prof.ms[1] = 0.0;
    g_prof.launches[0] = g_prof.launches[1] = 0;
  }
  return CF_OK;
}

// ------------------------------------------------------------- matrix-free C ABI

namespace {

struct MfArgs {
  int64_t n_nodes, n_elem;
  const int32_t* conn;
  const double* geo;
  double lam, mu;
  const int32_t* offs;
  const uint32_t* perm;
};

int mf_check(const MfArgs& a) {
  CF_REQUIRE(a.n_nodes > 0 && a.n_elem > 0 && a.conn && a.geo && a.offs && a.perm,
             "matrix-free Tet10: bad arguments");
  return CF_OK;
}

// Run plan of the matrix-free operator (see mf_t10_elem): built per solve
// from the node-sorted incidence list, in the solver's workspace.
struct MfRuns {
  int64_t n_runs;    // set by mf_runs_build (host)
  uint16_t* rank;    // rank of each item in its tile
  int32_t* toff;     // first run of each tile
  uint16_t* tstart;  // start of each run inside its tile
  uint32_t* nq;      // runs of each node, tile order
  int32_t* nro;      // first entry of each node in nq
  int4* run4;        // the same, four per node (-1 padded), when use_run4
  int32_t* over;     // nodes with more than four runs (device)
  bool use_run4;     // set by mf_runs_build (host)
  // build temporaries
  uint32_t* keys_in;
  uint32_t* keys_out;
  uint32_t* ids_in;
  uint32_t* tslot;
  uint32_t* rk;
  uint16_t* rl;
  void* tmp;
  size_t tmp_bytes;
};

void mf_runs_layout(int64_t n_nodes, int64_t n_elem, cf::Carve& c, MfRuns* R) {
  const int64_t items = 10 * n_elem;
  const int64_t n_tiles = (n_elem + kTile - 1) / kTile;
  R->n_runs = 0;
  R->rank = c.take<uint16_t>(items);
  R->toff = c.take<int32_t>(n_tiles + 1);
  R->tstart = c.take<uint16_t>(items);
  R->nq = c.take<uint32_t>(items);
  R->nro = c.take<int32_t>(n_nodes + 1);
  R->run4 = c.take<int4>(n_nodes);
  R->over = c.take<int32_t>(1);
  R->use_run4 = false;
  R->rk = c.take<uint32_t>(items);
  R->rl = c.take<uint16_t>(items);
  R->keys_in = c.take<uint32_t>(items + 1);
  R->keys_out = c.take<uint32_t>(items + 1);
  R->ids_in = c.take<uint32_t>(items);
  R->tslot = c.take<uint32_t>(items);
  size_t sort_bytes = 0, scan_bytes = 0;
  rocprim::radix_sort_pairs<SortConfig>(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                            (uint32_t*)nullptr, (uint32_t*)nullptr, size_t(items), 0, 32);
  rocprim::exclusive_scan(nullptr, scan_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u,
                          size_t(items + 1), rocprim::plus<uint32_t>());
  R->tmp_bytes = std::max(sort_bytes, scan_bytes);
  R->tmp = c.take<char>(R->tmp_bytes);
}

// Run plan from the node-sorted incidence list perm and its node offsets
// offs (one host synchronisation, for the run count).
int mf_runs_build(int64_t n_nodes, int64_t n_elem, const int32_t* offs, const uint32_t* perm,
                  MfRuns* R, hipStream_t s) {
  const MfRuns& M = *R;
  const int64_t items = 10 * n_elem;
  const int grid = cf::stride_grid(items + 1, kBlock);
  mf_store_keys<<<grid, kBlock, 0, s>>>(items, perm, M.keys_in, M.ids_in);
  CF_LAUNCHED();
  size_t tb = M.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(M.tmp, tb, M.keys_in, M.keys_out, M.ids_in, M.tslot,
                                   size_t(items), 0,
                                   bits_for((n_elem + kTile - 1) / kTile + 1), s));
  mf_store_rank<<<grid, kBlock, 0, s>>>(items, perm, M.tslot, M.rank);
  CF_LAUNCHED();
  uint32_t* head = M.keys_in;
  uint32_t* ridx = M.keys_out;
  uint32_t* rs = M.ids_in;
  mf_tile_heads<<<grid, kBlock, 0, s>>>(items, perm, head);
  CF_LAUNCHED();
  mf_node_heads<<<cf::stride_grid(n_nodes, kBlock), kBlock, 0, s>>>(n_nodes, offs, head);
  CF_LAUNCHED();
  tb = M.tmp_bytes;
  CF_HIP(rocprim::exclusive_scan(M.tmp, tb, head, ridx, 0u, size_t(items + 1),
                                 rocprim::plus<uint32_t>(), s));
  mf_run_fill<<<grid, kBlock, 0, s>>>(items, perm, M.rank, head, ridx, M.rk, rs);
  CF_LAUNCHED();
  mf_run_len<<<grid, kBlock, 0, s>>>(items, ridx, rs, M.rl, n_nodes, offs, M.nro);
  CF_LAUNCHED();
  // number the runs tile by tile: sort the run starts rk
  uint32_t n_runs = 0;
  CF_HIP(hipMemcpyAsync(&n_runs, ridx + items, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  CF_REQUIRE(n_runs > 0 && int64_t(n_runs) <= items, "matrix-free plan: bad run count %u",
             n_runs);
  R->n_runs = n_runs;
  iota_u32<<<grid, kBlock, 0, s>>>(n_runs, M.ids_in);
  CF_LAUNCHED();
  tb = M.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(M.tmp, tb, M.rk, M.keys_out, M.ids_in, M.tslot,
                                   size_t(n_runs), 0, bits_for(items + 1), s));
  mf_run_tile_order<<<grid, kBlock, 0, s>>>(n_runs, M.keys_out, M.tslot, M.nq, M.tstart);
  CF_LAUNCHED();
  const int64_t n_tiles = (n_elem + kTile - 1) / kTile;
  mf_tile_run_off<<<cf::stride_grid(n_tiles + 1, kBlock), kBlock, 0, s>>>(n_tiles, n_runs,
                                                                          M.keys_out, M.toff);
  CF_LAUNCHED();
  CF_HIP(hipMemsetAsync(M.over, 0, sizeof(int32_t), s));
  mf_run4_fill<<<cf::stride_grid(n_nodes, kBlock), kBlock, 0, s>>>(n_nodes, M.nro, M.nq, M.run4,
                                                                   M.over);
  CF_LAUNCHED();
  int32_t over = 0;
  CF_HIP(hipMemcpyAsync(&over, M.over, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  R->use_run4 = over == 0;
  return CF_OK;
}

// blocks of the run-sum node kernels (mf_gather_runs, mf_gather_cheb32):
// two nodes per thread, 2 kBlock nodes per block
int64_t mf_gather_grid(int64_t n_nodes) {
  return tile_grid((n_nodes + 2 * kBlock - 1) / (2 * kBlock));
}

// fe_doubles: element forces (30 per element) or, for the chunked plan,
// shared-node partials (3 per shared pair); parts: most block partials a
// product writes.
size_t mf_layout(int64_t n_nodes, int64_t fe_doubles, int64_t parts, int precond, void* base,
                 Layout* L, double** fe, double** blocks, MfRuns* R = nullptr,
                 int64_t n_elem = 0) {
  cf::Carve c(base);
  const int64_t n = 3 * n_nodes;
  L->r = c.take<double>(n);
  L->p = c.take<double>(n);
  L->Ap = c.take<double>(n);
  L->dinv = c.take<double>(precond == CF_PC_BLOCK3 ? 3 * n : n);
  L->partials = c.take<double>(4 * size_t(std::max<int64_t>(pcg_grid(), parts)));
  L->partials_b = c.take<double>(2 * size_t(pcg_vec_grid()));
  L->state = c.take<PcgState>(1);
  L->bad = c.take<int32_t>(1);
  *fe = c.take<double>(std::max<int64_t>(fe_doubles, 1));
  *blocks = c.take<double>(9 * n_nodes);
  if (R) mf_runs_layout(n_nodes, n_elem, c, R);
  L->total = c.used;
  return c.used;
}

// two-pass operator: fe holds the FP64 run partials (at most 3 per item)
size_t mf_layout(int64_t n_nodes, int64_t n_elem, int precond, void* base, Layout* L,
                 double** fe, double** blocks, MfRuns* R) {
  return mf_layout(n_nodes, 30 * n_elem, mf_gather_grid(n_nodes), precond, base, L, fe, blocks,
                   R, n_elem);
}

int mf_apply(const MfArgs& a, const MfRuns& R, const double* xin, double* y,
             const uint8_t* fixed, double* part, PcgState* state, double* partials,
             hipStream_t s) {
  mf_t10_elem<<<elem_grid(a.n_elem), kTile, 0, s>>>(
      a.n_elem, a.conn, a.geo, a.lam, a.mu, xin, R.rank, R.toff, R.tstart, R.n_runs, part, state);
  CF_LAUNCHED();
  mf_gather_runs<<<int(mf_gather_grid(a.n_nodes)), kBlock, 0, s>>>(
      a.n_nodes, R.nro, R.nq, R.n_runs, part, fixed, xin, y, state, partials);
  CF_LAUNCHED();
  return CF_OK;
}

// ---- chunked plan (see mf_chunk_apply)

struct MfPlan {
  int64_t n_chunks, m10;
  int32_t* conn_c;
  double* geo_c;
  uint16_t* slot_list;
  int32_t* chunk_off;
  int32_t* pair_node;
  int32_t* pair_slot_off;
  int32_t* pair_out;
  int32_t* shared_node;
  int32_t* shared_off;
  int64_t* counts;  // n_pairs, n_shared_pairs, n_shared_nodes
  uint64_t* keys_a;
  uint64_t* keys_b;
  uint32_t* vals_a;
  uint32_t* vals_b;
  int32_t* flags;
  int32_t* scan;
  void* tmp;
  size_t tmp_bytes;
  size_t total;
};

MfPlan mf_plan_layout(int64_t n_elem, void* base) {
  cf::Carve c(base);
  MfPlan P{};
  P.n_chunks = (n_elem + kChunk - 1) / kChunk;
  P.m10 = 10 * n_elem;
  P.counts = c.take<int64_t>(4);
  P.conn_c = c.take<int32_t>(P.m10);
  P.geo_c = c.take<double>(P.m10);
  P.slot_list = c.take<uint16_t>(P.m10);
  P.chunk_off = c.take<int32_t>(P.n_chunks + 1);
  P.pair_node = c.take<int32_t>(P.m10);
  P.pair_slot_off = c.take<int32_t>(P.m10 + 1);
  P.pair_out = c.take<int32_t>(P.m10);
  P.shared_node = c.take<int32_t>(P.m10);
  P.shared_off = c.take<int32_t>(P.m10 + 1);
  P.keys_a = c.take<uint64_t>(P.m10);
  P.keys_b = c.take<uint64_t>(P.m10);
  P.vals_a = c.take<uint32_t>(P.m10);
  P.vals_b = c.take<uint32_t>(P.m10);
  P.flags = c.take<int32_t>(P.m10);
  P.scan = c.take<int32_t>(P.m10);
  size_t a = 0, b = 0, d = 0;
  rocprim::radix_sort_pairs<SortConfig>(nullptr, a, (uint64_t*)nullptr, (uint64_t*)nullptr,
                            (uint32_t*)nullptr, (uint32_t*)nullptr, size_t(P.m10), 0, 64);
  rocprim::radix_sort_pairs<SortConfig>(nullptr, b, (uint32_t*)nullptr, (uint32_t*)nullptr,
                            (uint32_t*)nullptr, (uint32_t*)nullptr, size_t(n_elem), 0, 30);
  rocprim::inclusive_scan(nullptr, d, (int32_t*)nullptr, (int32_t*)nullptr, size_t(P.m10),
                          rocprim::plus<int32_t>());
  P.tmp_bytes = std::max(a, std::max(b, d));
  P.tmp = c.take<char>(P.tmp_bytes);
  P.total = c.used;
  return P;
}

int mf_plan_build(const MfArgs& A, const double* x, const double* y, const double* z,
                  const double* bbox, const MfPlan& P, int64_t counts_host[3], hipStream_t s) {
  const int64_t m = A.n_elem, m10 = P.m10;
  const int gm = cf::stride_grid(m, kBlock), g10 = cf::stride_grid(m10, kBlock);
  double sc[3];
  for (int k = 0; k < 3; ++k) {
    const double ext = bbox[3 + k] - bbox[k];
    sc[k] = ext > 0.0 ? 1023.0 / ext : 0.0;
  }
  CF_HIP(hipMemsetAsync(P.counts, 0, 4 * sizeof(int64_t), s));
  CF_HIP(hipMemsetAsync(P.shared_off, 0, sizeof(int32_t), s));
  // 1. Morton order of the elements, permuted connectivity and geometry
  mfp_morton<<<gm, kBlock, 0, s>>>(m, A.conn, x, y, z, bbox[0], bbox[1], bbox[2], sc[0], sc[1],
                                   sc[2], P.vals_a, P.vals_b);
  CF_LAUNCHED();
  uint32_t* mkeys = reinterpret_cast<uint32_t*>(P.keys_a);
  uint32_t* morder = reinterpret_cast<uint32_t*>(P.keys_b);
  size_t tb = P.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(P.tmp, tb, P.vals_a, mkeys, P.vals_b, morder, size_t(m), 0, 30,
                                   s));
  mfp_permute<<<g10, kBlock, 0, s>>>(m, morder, A.conn, A.geo, P.conn_c, P.geo_c);
  CF_LAUNCHED();
  // 2. (chunk, node) pairs of the slots
  mfp_pair_keys<<<g10, kBlock, 0, s>>>(m10, P.conn_c, P.keys_a, P.vals_a);
  CF_LAUNCHED();
  int chunk_bits = 1;
  while ((int64_t(1) << chunk_bits) < P.n_chunks + 1) ++chunk_bits;
  tb = P.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(P.tmp, tb, P.keys_a, P.keys_b, P.vals_a, P.vals_b,
                                   size_t(m10), 0, 32 + chunk_bits, s));
  mfp_heads<<<g10, kBlock, 0, s>>>(m10, P.keys_b, nullptr, P.flags);
  CF_LAUNCHED();
  tb = P.tmp_bytes;
  CF_HIP(rocprim::inclusive_scan(P.tmp, tb, P.flags, P.scan, size_t(m10),
                                 rocprim::plus<int32_t>(), s));
  mfp_pairs<<<g10, kBlock, 0, s>>>(m10, P.n_chunks, P.keys_b, P.vals_b, P.scan, P.slot_list,
                                   P.pair_node, P.pair_slot_off, P.chunk_off, P.counts);
  CF_LAUNCHED();
  // 3. shared pairs in (node, chunk) order -> partial ids, shared node lists
  mfp_shared_keys<<<g10, kBlock, 0, s>>>(m10, P.counts, P.keys_b, P.pair_node, P.pair_slot_off,
                                         A.offs, P.keys_a, P.vals_a, P.pair_out);
  CF_LAUNCHED();
  tb = P.tmp_bytes;
  CF_HIP(rocprim::radix_sort_pairs<SortConfig>(P.tmp, tb, P.keys_a, P.keys_b, P.vals_a, P.vals_b,
                                   size_t(m10), 0, 64, s));
  mfp_count_shared<<<g10, kBlock, 0, s>>>(m10, P.keys_b, P.counts);
  CF_LAUNCHED();
  mfp_heads<<<g10, kBlock, 0, s>>>(m10, P.keys_b, P.counts + 1, P.flags);
  CF_LAUNCHED();
  tb = P.tmp_bytes;
  CF_HIP(rocprim::inclusive_scan(P.tmp, tb, P.flags, P.scan, size_t(m10),
                                 rocprim::plus<int32_t>(), s));
  mfp_shared_fill<<<g10, kBlock, 0, s>>>(m10, P.keys_b, P.vals_b, P.scan, P.counts, P.pair_out,
                                         P.shared_node, P.shared_off);
  CF_LAUNCHED();
  CF_HIP(hipMemcpyAsync(counts_host, P.counts, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  return CF_OK;
}

int64_t mf_shared_grid(int64_t n_sn) { return tile_grid((n_sn + kBlock - 1) / kBlock); }

int mf_plan_apply(const MfArgs& A, const MfPlan& P, const int64_t counts[3], const double* xin,
                  double* y, const uint8_t* fixed, double* part_buf, PcgState* state,
                  double* partials, hipStream_t s) {
  mf_chunk_apply<<<int(P.n_chunks), kBlock, 0, s>>>(
      A.n_elem, P.conn_c, P.geo_c, A.lam, A.mu, xin, P.slot_list, P.chunk_off, P.pair_node,
      P.pair_slot_off, P.pair_out, fixed, y, part_buf, state, partials);
  CF_LAUNCHED();
  if (counts[2] > 0) {
    mf_shared_gather<<<int(mf_shared_grid(counts[2])), kBlock, 0, s>>>(
        counts[2], P.shared_node, P.shared_off, part_buf, fixed, xin, y, state,
        partials ? partials + P.n_chunks : nullptr);
    CF_LAUNCHED();
  }
  return CF_OK;
}

}  // namespace

extern "C" int cf_mf_t10_workspace_bytes(int64_t n_nodes, int64_t n_elem, int32_t precond,
                                         size_t* bytes) {
  CF_REQUIRE(bytes && n_nodes > 0 && n_elem > 0, "cf_mf_t10_workspace_bytes: bad arguments");
  Layout L{};
  double *fe, *blocks;
  MfRuns R{};
  *bytes = mf_layout(n_nodes, n_elem, precond, nullptr, &L, &fe, &blocks, &R);
  return CF_OK;
}

extern "C" int cf_mf_t10_apply(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                               const double* geo, double lam, double mu, const int32_t* offs,
                               const uint32_t* perm, const double* x, double* y,
                               const uint8_t* fixed, void* ws, size_t ws_bytes, void* stream) {
  MfArgs a{n_nodes, n_elem, conn, geo, lam, mu, offs, perm};
  int rc = mf_check(a);
  if (rc) return rc;
  CF_REQUIRE(x && y && ws, "cf_mf_t10_apply: null argument");
  CF_REQUIRE(10 * n_elem < (int64_t(1) << 31), "cf_mf_t10_apply: too many elements");
  Layout L{};
  double *fe, *blocks;
  MfRuns R{};
  const size_t need = mf_layout(n_nodes, n_elem, CF_PC_JACOBI, ws, &L, &fe, &blocks, &R);
  if (ws_bytes < need) {
    cf::set_error("cf_mf_t10_apply: workspace %zu < %zu bytes", ws_bytes, need);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  if ((rc = mf_runs_build(n_nodes, n_elem, offs, perm, &R, s))) return rc;
  return mf_apply(a, R, x, y, fixed, fe, nullptr, nullptr, s);
}

extern "C" int cf_pcg_solve_mf_t10(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                                   const double* geo, double lam, double mu,
                                   const int32_t* offs, const uint32_t* perm,
                                   const double* b, double* x, const uint8_t* fixed,
                                   int32_t precond, double tol, int32_t maxit, void* ws,
                                   size_t ws_bytes, cf_pcg_stats* stats_host, void* stream) {
  MfArgs a{n_nodes, n_elem, conn, geo, lam, mu, offs, perm};
  int rc = mf_check(a);
  if (rc) return rc;
  CF_REQUIRE(b && x && ws && stats_host, "cf_pcg_solve_mf_t10: null argument");
  CF_REQUIRE(precond == CF_PC_JACOBI || precond == CF_PC_BLOCK3,
             "cf_pcg_solve_mf_t10: bad precond %d", precond);
  CF_REQUIRE(tol > 0.0 && maxit > 0, "cf_pcg_solve_mf_t10: tol and maxit must be positive");
  CF_REQUIRE(10 * n_elem < (int64_t(1) << 31), "cf_pcg_solve_mf_t10: too many elements");
  Layout L{};
  double *fe, *blocks;
  MfRuns R{};
  const size_t need = mf_layout(n_nodes, n_elem, precond, ws, &L, &fe, &blocks, &R);
  if (ws_bytes < need) {
    cf::set_error("cf_pcg_solve_mf_t10: workspace %zu < %zu bytes", ws_bytes, need);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  *stats_host = cf_pcg_stats{};
  if ((rc = mf_runs_build(n_nodes, n_elem, offs, perm, &R, s))) return rc;
  CF_HIP(hipMemsetAsync(L.state, 0, sizeof(PcgState), s));
  CF_HIP(hipMemsetAsync(L.bad, 0, sizeof(int32_t), s));
  mf_t10_diag<<<cf::stride_grid(n_nodes, kBlock), kBlock, 0, s>>>(n_nodes, offs, perm, geo, lam,
                                                                  mu, blocks);
  CF_LAUNCHED();
  blocks_to_precond<<<cf::stride_grid(n_nodes, kBlock), kBlock, 0, s>>>(
      n_nodes, blocks, fixed, precond == CF_PC_BLOCK3, L.dinv, L.bad);
  CF_LAUNCHED();
  int32_t bad = 0;
  CF_HIP(hipMemcpyAsync(&bad, L.bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  if (bad) {
    cf::set_error("cf_pcg_solve_mf_t10: %d free diagonal entries/blocks are not positive", bad);
    return CF_ERR_BREAKDOWN;
  }
  auto apply = [&](const double* xin, double* y, PcgState* state, double* partials) {
    return mf_apply(a, R, xin, y, fixed, fe, state, partials, s);
  };
  const int64_t n = 3 * n_nodes;
  const int parts = int(mf_gather_grid(n_nodes));
  if (precond == CF_PC_BLOCK3)
    return pcg_fp64_core<CF_PC_BLOCK3>(apply, parts, n, n_nodes, b, x, fixed, tol, maxit, L,
                                       stats_host, s);
  return pcg_fp64_core<CF_PC_JACOBI>(apply, parts, n, n, b, x, fixed, tol, maxit, L, stats_host,
                                     s);
}

extern "C" int cf_mf_t10_plan_bytes(int64_t n_elem, size_t* bytes) {
  CF_REQUIRE(n_elem > 0 && bytes, "cf_mf_t10_plan_bytes: bad arguments");
  CF_REQUIRE(10 * n_elem < (int64_t(1) << 31), "cf_mf_t10_plan_bytes: too many elements");
  *bytes = mf_plan_layout(n_elem, nullptr).total;
  return CF_OK;
}

extern "C" int cf_mf_t10_plan_build(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                                    const double* geo, const double* x, const double* y,
                                    const double* z, const int32_t* offs, const double* bbox,
                                    void* plan, size_t plan_bytes, int64_t* counts,
                                    void* stream) {
  MfArgs a{n_nodes, n_elem, conn, geo, 0.0, 0.0, offs, nullptr};
  CF_REQUIRE(n_nodes > 0 && n_elem > 0 && conn && geo && offs && x && y && z && bbox && plan &&
                 counts,
             "cf_mf_t10_plan_build: bad arguments");
  CF_REQUIRE(10 * n_elem < (int64_t(1) << 31), "cf_mf_t10_plan_build: too many elements");
  MfPlan P = mf_plan_layout(n_elem, plan);
  if (plan_bytes < P.total) {
    cf::set_error("cf_mf_t10_plan_build: plan buffer %zu < %zu bytes", plan_bytes, P.total);
    return CF_ERR_WORKSPACE;
  }
  return mf_plan_build(a, x, y, z, bbox, P, counts, cf::as_stream(stream));
}

extern "C" int cf_mf_t10_plan_workspace_bytes(int64_t n_nodes, int64_t n_elem,
                                              const int64_t* counts, int32_t precond,
                                              size_t* bytes) {
  CF_REQUIRE(n_nodes > 0 && n_elem > 0 && counts && bytes,
             "cf_mf_t10_plan_workspace_bytes: bad arguments");
  Layout L{};
  double *fe, *blocks;
  const int64_t parts = (n_elem + kChunk - 1) / kChunk + mf_shared_grid(counts[2]);
  *bytes = mf_layout(n_nodes, 3 * counts[1], parts, precond, nullptr, &L, &fe, &blocks);
  return CF_OK;
}

extern "C" int cf_mf_t10_plan_apply(int64_t n_nodes, int64_t n_elem, const void* plan,
                                    const int64_t* counts, double lam, double mu,
                                    const double* x, double* y, const uint8_t* fixed, void* ws,
                                    size_t ws_bytes, void* stream) {
  CF_REQUIRE(n_nodes > 0 && n_elem > 0 && plan && counts && x && y && ws,
             "cf_mf_t10_plan_apply: bad arguments");
  MfArgs a{n_nodes, n_elem, nullptr, nullptr, lam, mu, nullptr, nullptr};
  MfPlan P = mf_plan_layout(n_elem, const_cast<void*>(plan));
  Layout L{};
  double *fe, *blocks;
  const int64_t parts = P.n_chunks + mf_shared_grid(counts[2]);
  const size_t need = mf_layout(n_nodes, 3 * counts[1], parts, CF_PC_JACOBI, ws, &L, &fe, &blocks);
  if (ws_bytes < need) {
    cf::set_error("cf_mf_t10_plan_apply: workspace %zu < %zu bytes", ws_bytes, need);
    return CF_ERR_WORKSPACE;
  }
  return mf_plan_apply(a, P, counts, x, y, fixed, fe, nullptr, nullptr, cf::as_stream(stream));
}

extern "C" int cf_pcg_solve_mf_t10_plan(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                                        const double* geo, const int32_t* offs,
                                        const uint32_t* perm, const void* plan,
                                        const int64_t* counts, double lam, double mu,
                                        const double* b, double* x, const uint8_t* fixed,
                                        int32_t precond, double tol, int32_t maxit, void* ws,
                                        size_t ws_bytes, cf_pcg_stats* stats_host,
                                        void* stream) {
  MfArgs a{n_nodes, n_elem, conn, geo, lam, mu, offs, perm};
  int rc = mf_check(a);
  if (rc) return rc;
  CF_REQUIRE(plan && counts && b && x && ws && stats_host,
             "cf_pcg_solve_mf_t10_plan: null argument");
  CF_REQUIRE(precond == CF_PC_JACOBI || precond == CF_PC_BLOCK3,
             "cf_pcg_solve_mf_t10_plan: bad precond %d", precond);
  CF_REQUIRE(tol > 0.0 && maxit > 0, "cf_pcg_solve_mf_t10_plan: tol and maxit must be positive");
  MfPlan P = mf_plan_layout(n_elem, const_cast<void*>(plan));
  Layout L{};
  double *fe, *blocks;
  const int64_t parts = P.n_chunks + mf_shared_grid(counts[2]);
  const size_t need = mf_layout(n_nodes, 3 * counts[1], parts, precond, ws, &L, &fe, &blocks);
  if (ws_bytes < need) {
    cf::set_error("cf_pcg_solve_mf_t10_plan: workspace %zu < %zu bytes", ws_bytes, need);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  *stats_host = cf_pcg_stats{};
  CF_HIP(hipMemsetAsync(L.state, 0, sizeof(PcgState), s));
  CF_HIP(hipMemsetAsync(L.bad, 0, sizeof(int32_t), s));
  mf_t10_diag<<<cf::stride_grid(n_nodes, kBlock), kBlock, 0, s>>>(n_nodes, offs, perm, geo, lam,
                                                                  mu, blocks);
  CF_LAUNCHED();
  blocks_to_precond<<<cf::stride_grid(n_nodes, kBlock), kBlock, 0, s>>>(
      n_nodes, blocks, fixed, precond == CF_PC_BLOCK3, L.dinv, L.bad);
  CF_LAUNCHED();
  int32_t bad = 0;
  CF_HIP(hipMemcpyAsync(&bad, L.bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  if (bad) {
    cf::set_error("cf_pcg_solve_mf_t10_plan: %d free diagonal entries/blocks are not positive",
                  bad);
    return CF_ERR_BREAKDOWN;
  }
  int64_t cnt[3] = {counts[0], counts[1], counts[2]};
  auto apply = [&](const double* xin, double* yout, PcgState* state, double* partials) {
    return mf_plan_apply(a, P, cnt, xin, yout, fixed, fe, state, partials, s);
  };
  const int64_t n = 3 * n_nodes;
  const int apply_parts = int(P.n_chunks + (cnt[2] > 0 ? mf_shared_grid(cnt[2]) : 0));
  if (precond == CF_PC_BLOCK3)
    return pcg_fp64_core<CF_PC_BLOCK3>(apply, apply_parts, n, n_nodes, b, x, fixed, tol, maxit,
                                       L, stats_host, s);
  return pcg_fp64_core<CF_PC_JACOBI>(apply, apply_parts, n, n, b, x, fixed, tol, maxit, L,
                                     stats_host, s);
}

namespace {

// Workspace of the Chebyshev / two-level solvers.  The vectors of the
// low-precision part (Chebyshev iterates, power iteration, inner geometry,
// scaling and run partials) hold lpb bytes per value: 4 (FP32) or 8 (the
// FP64 twin of the same solver).
struct MfMixed {
  double* fe;      // FP64 run partials of the outer product (component planes)
  double* blocks;  // diagonal 3x3 blocks
  void* geoL;      // geometry in the inner precision, tile-packed
  int32_t* connT;  // connectivity, tile-packed
  uint16_t* rankT; // run-plan ranks, tile-packed
  void* feL;       // inner run partials (AoS, 3 per run)
  void* sL;        // Jacobi scaling in the inner precision
  MfRuns R;          // run plan shared by the FP64 and inner products
};

void mf_mixed_layout_base(int64_t n_nodes, int64_t n_elem, cf::Carve& c, Layout* L, MfMixed* M,
                          int lpb, const cf_matrix* A) {
  const int64_t n = 3 * n_nodes;
  L->r = c.take<double>(n);
  L->p = c.take<double>(n);
  L->Ap = c.take<double>(n);
  L->dinv = c.take<double>(n);
  int64_t parts = std::max<int64_t>(pcg_grid(), mf_gather_grid(n_nodes));
  if (A) parts = std::max<int64_t>(parts, sell_parts(A->n_rows));
  L->partials = c.take<double>(4 * size_t(parts));
  L->partials_b = c.take<double>(2 * size_t(pcg_vec_grid()));
  L->state = c.take<PcgState>(1);
  L->bad = c.take<int32_t>(1);
  L->lz = c.take<double>(2 * kLanczosSteps);
  L->s = c.take<double>(n);
  L->rh = reinterpret_cast<float*>(c.take<char>(n * lpb));
  L->ph = reinterpret_cast<float*>(c.take<char>(n * lpb));
  L->Aph = reinterpret_cast<float*>(c.take<char>(n * lpb));
  L->yh = reinterpret_cast<float*>(c.take<char>(n * lpb));
  L->z0 = c.take<double>(n);
  L->z1 = c.take<double>(n);
  M->fe = c.take<double>(30 * n_elem);
  M->blocks = c.take<double>(9 * n_nodes);
  const int64_t packed = 10 * ((n_elem + kTile - 1) / kTile) * kTile;
  M->geoL = c.take<char>(packed * lpb);
  M->connT = c.take<int32_t>(packed);
  M->rankT = c.take<uint16_t>(packed);
  M->feL = c.take<char>(30 * n_elem * lpb);
  M->sL = c.take<char>(n * lpb);
  mf_runs_layout(n_nodes, n_elem, c, &M->R);
}

}  // namespace

namespace {

// ------------------------------------------------------------- two-level preconditioner
// Coarse space of the mixed matrix-free solver: the six rigid-body modes of
// node aggregates.  Aggregate of a node = the group of agg_tiles_for(tiles)
// consecutive element tiles (in the caller's element order) that holds the
// first element touching it.  Basis rows of node i in aggregate g:
// translations e_k and rotations (e_k x (x_i - c_g)) / L_g (c_g centroid, L_g
// RMS node distance from it), zero on fixed DOFs.  Coarse matrix
// A_c = P^T K_ff P by probing: aggregates get a distance-2 colouring, the
// modes of all aggregates of one colour are applied together with the FP64
// matrix-free operator, and because no two aggregates of a colour share a
// neighbour each coarse block is read off exactly (no atomics).  A_c is
// factored and inverted with rocSOLVER; the preconditioner is
//   z1 = C r,  z2 = z1 + Cheb(r - K z1),  z = z2 + C (r - K z2),
// C = P A_c^-1 P^T, Cheb the FP32 Chebyshev polynomial of S K S.
constexpr int kAggTiles = 6;  // the most tiles per aggregate (1536 elements)
constexpr int kAggTarget = 250;  // aggregates aimed at (a 1,500-unknown coarse problem)

// Tiles per aggregate for a mesh of n_tiles element tiles: about kAggTarget
// aggregates, 2 to kAggTiles tiles each.  Measured in full Tet10 cantilever
// solves: at 384,000 elements (1,500 tiles) 6 tiles per aggregate are best
// (57.9 ms; 4: 60.1, 3: 62.6, 2: 78.9 ms, the coarse factorisation growing
// with the coarse size; 8 to 24 tiles were slower too), at 48,000 elements
// (188 tiles) 1 to 3 are (22.3 ms, 45 to 50 outer iterations; 6: 30.2 ms,
// 71: too coarse a coarse space).
int agg_tiles_for(int64_t n_tiles) {
  const int64_t at = (n_tiles + kAggTarget - 1) / kAggTarget;
  return int(std::min<int64_t>(kAggTiles, std::max<int64_t>(2, at)));
}
constexpr int kMaxColors = 64;
constexpr int kPairsPerElem = 45;  // distinct aggregate pairs of 10 nodes

__device__ __forceinline__ void rigid_row(int m, double dx, double dy, double dz, double invl,
                                          double (&o)[3]) {
  o[0] = o[1] = o[2] = 0.0;
  if (m < 3) {
    o[m] = 1.0;
  } else if (m == 3) {  // e_x x d
    o[1] = -dz * invl;
    o[2] = dy * invl;
  } else if (m == 4) {  // e_y x d
    o[0] = dz * invl;
    o[2] = -dx * invl;
  } else {              // e_z x d
    o[0] = -dy * invl;
    o[1] = dx * invl;
  }
}

// group of the first element touching node i (-1: none); used[g] = 1
__global__ void agg_first(int64_t n_nodes, const int32_t* __restrict__ offs,
                          const uint32_t* __restrict__ perm, int agg_tiles,
                          int32_t* __restrict__ graw, uint32_t* __restrict__ used) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    int32_t g = -1;
    if (offs[i] < offs[i + 1]) {
      g = int32_t(perm[offs[i]] / uint32_t(kAggUnitItems * agg_tiles));
      used[g] = 1u;
    }
    graw[i] = g;
  }
}

// compact aggregate ids; sort keys (aggregate, or n_agg for untouched nodes)
__global__ void agg_compact(int64_t n_nodes, const int32_t* __restrict__ graw,
                            const uint32_t* __restrict__ gid, int64_t n_groups,
                            int32_t* __restrict__ agg, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ ids) {
  const uint32_t n_agg = gid[n_groups];
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t g = graw[i] >= 0 ? int32_t(gid[graw[i]]) : -1;
    agg[i] = g;
    keys[i] = g >= 0 ? uint32_t(g) : n_agg;
    ids[i] = uint32_t(i);
  }
}

// agg_off[g] = first position of aggregate g in the sorted keys (g = 0..n_agg)
__global__ void agg_offsets(int64_t n_nodes, int64_t n_agg, const uint32_t* __restrict__ keys,
                            int32_t* __restrict__ agg_off) {
  for (int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; g <= n_agg;
       g += int64_t(gridDim.x) * blockDim.x) {
    int64_t lo = 0, hi = n_nodes;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (int64_t(keys[mid]) >= g) hi = mid; else lo = mid + 1;
    }
    agg_off[g] = int32_t(lo);
  }
}

// Fallback aggregates (a node numbering that scatters the element-order
// groups): the nodes that belong to an element, sorted by their coordinate
// along the longest side of the node bounding box (stable: ties keep the
// node order) and cut into equal runs -- slabs across that axis whatever
// the numbering, so aggregates that share an element lie a few ids apart.

// per-block minimum and maximum of x, y, z:
// part[6 b .. 6 b + 5] = min x, min y, min z, max x, max y, max z
__global__ void __launch_bounds__(kBlock)
node_bbox(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
          const double* __restrict__ z, double* __restrict__ part) {
  __shared__ double sh[6][kBlock];
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double c[3] = {x[i], y[i], z[i]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      v[k] = fmin(v[k], c[k]);
      v[3 + k] = fmax(v[3 + k], c[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
#pragma unroll
      for (int k = 0; k < 6; ++k)
        sh[k][threadIdx.x] = k < 3 ? fmin(sh[k][threadIdx.x], sh[k][threadIdx.x + w])
                                   : fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 6; ++k) part[6 * int64_t(blockIdx.x) + k] = sh[k][0];
}

// sort keys: the coordinate as an order-preserving uint32 (UINT32_MAX for a
// node in no element, which sorts last), node ids
__global__ void slab_keys(int64_t n_nodes, const int32_t* __restrict__ offs,
                          const double* __restrict__ coord, uint32_t* __restrict__ keys,
                          uint32_t* __restrict__ ids) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    uint32_t k = 0xffffffffu;
    if (offs[i] < offs[i + 1]) {
      const uint32_t b = __float_as_uint(float(coord[i]));
      k = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
      if (k == 0xffffffffu) k = 0xfffffffeu;
    }
    keys[i] = k;
    ids[i] = uint32_t(i);
  }
}

// count[0] = number of sorted keys below UINT32_MAX (the nodes in elements)
__global__ void slab_count(int64_t n_nodes, const uint32_t* __restrict__ keys,
                           uint32_t* __restrict__ count) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t lo = 0, hi = n_nodes;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (keys[mid] == 0xffffffffu) hi = mid; else lo = mid + 1;
  }
  count[0] = uint32_t(lo);
}

// agg[node] = position / npa for the sorted nodes in elements (agg is -1
// beforehand); agg_off[g] = g npa capped at n_in (g = 0..n_agg)
__global__ void slab_assign(int64_t n_in, int64_t npa, int64_t n_agg,
                            const uint32_t* __restrict__ sorted_ids, int32_t* __restrict__ agg,
                            int32_t* __restrict__ agg_off) {
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p < n_in;
       p += int64_t(gridDim.x) * blockDim.x)
    agg[sorted_ids[p]] = int32_t(p / npa);
  for (int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; g <= n_agg;
       g += int64_t(gridDim.x) * blockDim.x)
    agg_off[g] = int32_t(std::min<int64_t>(g * npa, n_in));
}

// centroid and 1 / RMS radius of each aggregate; one workgroup per
// aggregate, fixed node split and reduction tree
__global__ void __launch_bounds__(kBlock)
agg_frame(int64_t n_agg, const int32_t* __restrict__ agg_off,
          const uint32_t* __restrict__ agg_nodes, const double* __restrict__ x,
          const double* __restrict__ y, const double* __restrict__ z,
          double* __restrict__ frame) {
  for (int64_t g = blockIdx.x; g < n_agg; g += gridDim.x) {
    const int32_t b = agg_off[g], e = agg_off[g + 1];
    double c[3] = {0.0, 0.0, 0.0};
    for (int32_t k = b + threadIdx.x; k < e; k += blockDim.x) {
      const uint32_t i = agg_nodes[k];
      c[0] += x[i];
      c[1] += y[i];
      c[2] += z[i];
    }
    block_sum<3>(c);
    const double inv_n = e > b ? 1.0 / double(e - b) : 0.0;
    c[0] *= inv_n;
    c[1] *= inv_n;
    c[2] *= inv_n;
    double r2[1] = {0.0};
    for (int32_t k = b + threadIdx.x; k < e; k += blockDim.x) {
      const uint32_t i = agg_nodes[k];
      const double dx = x[i] - c[0], dy = y[i] - c[1], dz = z[i] - c[2];
      r2[0] += dx * dx + dy * dy + dz * dz;
    }
    block_sum<1>(r2);
    if (threadIdx.x == 0) {
      const double rr = r2[0] * inv_n;
      frame[4 * g] = c[0];
      frame[4 * g + 1] = c[1];
      frame[4 * g + 2] = c[2];
      frame[4 * g + 3] = rr > 0.0 ? 1.0 / sqrt(rr) : 1.0;
    }
    __syncthreads();
  }
}

// distinct aggregates of element e's nodes, ascending; returns their count
__device__ __forceinline__ int elem_aggs(const int32_t* __restrict__ conn, int64_t e,
                                         const int32_t* __restrict__ agg, int32_t (&g)[10]) {
  int n = 0;
  for (int a = 0; a < 10; ++a) {
    const int32_t v = agg[conn[e * 10 + a]];
    if (v < 0) continue;
    int p = n;
    bool dup = false;
    for (int q = 0; q < n; ++q) dup |= g[q] == v;
    if (dup) continue;
    while (p > 0 && g[p - 1] > v) {
      g[p] = g[p - 1];
      --p;
    }
    g[p] = v;
    ++n;
  }
  return n;
}

// number of aggregate pairs (g < h) of each element
__global__ void agg_pair_count(int64_t m, const int32_t* __restrict__ conn,
                               const int32_t* __restrict__ agg, uint32_t* __restrict__ cnt) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e <= m;
       e += int64_t(gridDim.x) * blockDim.x) {
    if (e == m) {
      cnt[e] = 0;
      continue;
    }
    int32_t g[10];
    const int n = elem_aggs(conn, e, agg, g);
    cnt[e] = uint32_t(n * (n - 1) / 2);
  }
}

__global__ void agg_pair_fill(int64_t m, const int32_t* __restrict__ conn,
                              const int32_t* __restrict__ agg, const uint32_t* __restrict__ off,
                              int64_t n_agg, uint64_t* __restrict__ keys) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < m;
       e += int64_t(gridDim.x) * blockDim.x) {
    int32_t g[10];
    const int n = elem_aggs(conn, e, agg, g);
    uint32_t o = off[e];
    for (int a = 0; a < n; ++a)
      for (int b = a + 1; b < n; ++b) keys[o++] = uint64_t(g[a]) * uint64_t(n_agg) + uint64_t(g[b]);
  }
}

// probe vectors of colour c: v_m = mode m of every aggregate of colour c
__global__ void probe_build(int64_t n_nodes, const int32_t* __restrict__ agg,
                            const int32_t* __restrict__ color, int c,
                            const double* __restrict__ frame, const double* __restrict__ x,
                            const double* __restrict__ y, const double* __restrict__ z,
                            const uint8_t* __restrict__ fixed, double* __restrict__ v) {
  const int64_t n = 3 * n_nodes;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t g = agg[i];
    const bool on = g >= 0 && color[g] == c;
    const double* f = frame + 4 * (g >= 0 ? g : 0);
    for (int mm = 0; mm < 6; ++mm) {
      double o[3] = {0.0, 0.0, 0.0};
      if (on) rigid_row(mm, x[i] - f[0], y[i] - f[1], z[i] - f[2], f[3], o);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int64_t d = 3 * i + k;
        v[mm * n + d] = (fixed && fixed[d]) ? 0.0 : o[k];
      }
    }
  }
}

// A_c[(h, mp), (g, mm)] = sum over the nodes i of h of mode mp(i) . y_mm(i), g =
// the colour-c neighbour of h (if any); one workgroup per aggregate h, fixed
// node split and reduction tree.  Column-major, leading dimension 6 n_agg.
__global__ void __launch_bounds__(kBlock)
probe_extract(int64_t n_agg, int n_col, int c, const int32_t* __restrict__ nbc,
              const int32_t* __restrict__ agg_off, const uint32_t* __restrict__ agg_nodes,
              const double* __restrict__ frame, const double* __restrict__ x,
              const double* __restrict__ y, const double* __restrict__ z, int64_t n,
              const double* __restrict__ yv, double* __restrict__ A) {
  const int64_t nc = 6 * n_agg;
  for (int64_t h = blockIdx.x; h < n_agg; h += gridDim.x) {
    const int32_t g = nbc[h * n_col + c];
    if (g < 0) continue;  // uniform over the block
    const double* f = frame + 4 * h;
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; ++q) acc[q] = 0.0;
    for (int32_t k = agg_off[h] + threadIdx.x; k < agg_off[h + 1]; k += blockDim.x) {
      const uint32_t i = agg_nodes[k];
      const double dx = x[i] - f[0], dy = y[i] - f[1], dz = z[i] - f[2];
      double yy[6][3];
#pragma unroll
      for (int mm = 0; mm < 6; ++mm)
#pragma unroll
        for (int q = 0; q < 3; ++q) yy[mm][q] = yv[mm * n + 3 * int64_t(i) + q];
#pragma unroll
      for (int mp = 0; mp < 6; ++mp) {
        double o[3];
        rigid_row(mp, dx, dy, dz, f[3], o);
#pragma unroll
        for (int mm = 0; mm < 6; ++mm)
          acc[6 * mp + mm] += o[0] * yy[mm][0] + o[1] * yy[mm][1] + o[2] * yy[mm][2];
      }
    }
    block_sum<36>(acc);
    if (threadIdx.x == 0)
#pragma unroll
      for (int q = 0; q < 36; ++q) A[(6 * h + q / 6) + (6 * int64_t(g) + q % 6) * nc] = acc[q];
    __syncthreads();
  }
}


// A vanishing diagonal (an aggregate without free DOFs, or a degenerate mode)
// becomes an identity row and column of the lower triangle rocSOLVER reads.
// The probed blocks A[h, g] and A[g, h] agree to rounding, and potrf uses the
// lower one, so no symmetrisation is needed.
// Coarse modes that the aggregate's free nodes do not support independently:
// one thread per aggregate factors its 6x6 diagonal block (Cholesky in mode
// order) and flags every mode whose pivot falls to 1e-10 of the block's
// largest diagonal entry or below -- a vanishing mode (no free node), or a
// dependent one (two free nodes, say, leave the rotation about the line
// through them without stiffness although every diagonal entry is positive).
__global__ void coarse_dependent_modes(int64_t n_agg, int64_t nc, const double* __restrict__ A,
                                       int32_t* __restrict__ dep) {
  for (int64_t g = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; g < n_agg;
       g += int64_t(gridDim.x) * blockDim.x) {
    const int64_t o = 6 * g;
    double L[6][6] = {};
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) mx = fmax(mx, A[(o + i) + (o + i) * nc]);
    const double tol = 1e-10 * mx;
    bool ok[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      double d = A[(o + k) + (o + k) * nc];
      for (int m = 0; m < k; ++m)
        if (ok[m]) d -= L[k][m] * L[k][m];
      ok[k] = d > tol;
      dep[o + k] = ok[k] ? 0 : 1;
      if (!ok[k]) continue;
      L[k][k] = sqrt(d);
      for (int i = k + 1; i < 6; ++i) {
        double v = A[(o + i) + (o + k) * nc];
        for (int m = 0; m < k; ++m)
          if (ok[m]) v -= L[i][m] * L[k][m];
        L[i][k] = v / L[k][k];
      }
    }
  }
}

// Flagged modes get identity rows and columns for the factorisation (lower
// triangle, column-major nc x nc); coarse_drop_modes then clears them in the
// inverse, so the coarse correction leaves them out.  Two flagged modes j1 < j2
// both clear entry (j2, j1): the same value, so the overlap is benign.
__global__ void coarse_regularize(int64_t nc, double* __restrict__ A,
                                  const int32_t* __restrict__ dep) {
  for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < nc;
       j += int64_t(gridDim.x) * blockDim.x) {
    if (!dep[j]) continue;
    for (int64_t i = j + 1; i < nc; ++i) A[i + j * nc] = 0.0;  // column j below the diagonal
    for (int64_t k = 0; k < j; ++k) A[j + k * nc] = 0.0;       // row j left of it
    A[j + j * nc] = 1.0;
  }
}

// upper triangle := lower triangle (after potri), 32 x 32 tiles through LDS
__global__ void __launch_bounds__(kBlock)
coarse_fill_upper(int64_t nc, double* __restrict__ A) {
  __shared__ double t[32][33];
  const int64_t nt = (nc + 31) / 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int64_t b = blockIdx.x; b < nt * nt; b += gridDim.x) {
    const int64_t bi = b / nt, bj = b - bi * nt;  // source tile: rows bi, columns bj
    if (bj > bi) continue;                        // lower tiles only (uniform per block)
    // t[c][rr] = A[bi*32 + rr, bj*32 + c], loaded down the columns (coalesced)
    for (int r = ty; r < 32; r += 8) {
      const int64_t i = bi * 32 + tx, j = bj * 32 + r;
      t[r][tx] = (i < nc && j < nc) ? A[i + j * nc] : 0.0;
    }
    __syncthreads();
    // A[j, i] = A[i, j] for j < i, i = bi*32 + r, j = bj*32 + tx (coalesced)
    for (int r = ty; r < 32; r += 8) {
      const int64_t i = bi * 32 + r, j = bj * 32 + tx;
      if (i < nc && j < nc && j < i) A[j + i * nc] = t[tx][r];
    }
    __syncthreads();
  }
}

// rc[(g, m)] = sum over the nodes of g of mode m . r (r FP64 or FP32); one
// workgroup of kRestrictBlock threads per aggregate (a few hundred
// aggregates: wide groups keep each thread's chain of gathers short), fixed
// node split and reduction tree.
constexpr int kRestrictBlock = 1024;

template <class RT>
__global__ void __launch_bounds__(kRestrictBlock)
coarse_restrict(int64_t n_agg, const int32_t* __restrict__ agg_off,
                const uint32_t* __restrict__ agg_nodes, const double* __restrict__ frame,
                const double* __restrict__ x, const double* __restrict__ y,
                const double* __restrict__ z, const RT* __restrict__ r, double* __restrict__ rc,
                const PcgState* state) {
  if (state && state->halt) return;
  for (int64_t g = blockIdx.x; g < n_agg; g += gridDim.x) {
    const double* f = frame + 4 * g;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int32_t k = agg_off[g] + threadIdx.x; k < agg_off[g + 1]; k += blockDim.x) {
      const uint32_t i = agg_nodes[k];
      const double dx = x[i] - f[0], dy = y[i] - f[1], dz = z[i] - f[2];
      const double r0 = double(r[3 * int64_t(i)]), r1 = double(r[3 * int64_t(i) + 1]),
                   r2 = double(r[3 * int64_t(i) + 2]);
#pragma unroll
      for (int mm = 0; mm < 6; ++mm) {
        double o[3];
        rigid_row(mm, dx, dy, dz, f[3], o);
        acc[mm] += o[0] * r0 + o[1] * r1 + o[2] * r2;
      }
    }
    block_sum<6, kRestrictBlock / 64>(acc);
    if (threadIdx.x == 0)
#pragma unroll
      for (int q = 0; q < 6; ++q) rc[6 * g + q] = acc[q];
    __syncthreads();
  }
}

// zc = Ainv rc with the FP32 copy of the symmetric inverse (column i read as
// row i; half the bytes of the FP64 one, products summed in FP64): one wave
// per entry, two entries per lane and load (nc = 6 x aggregates is even),
// four loads in flight
__global__ void __launch_bounds__(kBlock)
coarse_gemv(int64_t nc, const float* __restrict__ Ainv, const double* __restrict__ rc,
            double* __restrict__ zc, const PcgState* state) {
  if (state && state->halt) return;
  const int lane = threadIdx.x & 63;
  const int64_t half = nc / 2;
  for (int64_t i = (blockIdx.x * int64_t(blockDim.x) + threadIdx.x) / 64; i < nc;
       i += int64_t(gridDim.x) * blockDim.x / 64) {
    const float2* col = reinterpret_cast<const float2*>(Ainv + i * nc);
    const double2* r2 = reinterpret_cast<const double2*>(rc);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t j = lane;
    for (; j + 192 < half; j += 256) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float2 a = col[j + 64 * u];
        const double2 r = r2[j + 64 * u];
        acc[u] += double(a.x) * r.x + double(a.y) * r.y;
      }
    }
    for (; j < half; j += 64) {
      const float2 a = col[j];
      const double2 r = r2[j];
      acc[0] += double(a.x) * r.x + double(a.y) * r.y;
    }
    const double sum = wave_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
    if (lane == 0) zc[i] = sum;
  }
}

// u (inner precision T) = [ADD ? u : 0] + P zc (zero on fixed DOFs)
template <bool ADD, class T>
__global__ void coarse_prolong(int64_t n_nodes, const int32_t* __restrict__ agg,
                               const double* __restrict__ frame, const double* __restrict__ x,
                               const double* __restrict__ y, const double* __restrict__ z,
                               const uint8_t* __restrict__ fixed, const double* __restrict__ zc,
                               T* __restrict__ u, const PcgState* state) {
  if (state && state->halt) return;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t g = agg[i];
    double s[3] = {0.0, 0.0, 0.0};
    if (g >= 0) {
      const double* f = frame + 4 * g;
      const double dx = x[i] - f[0], dy = y[i] - f[1], dz = z[i] - f[2];
      for (int mm = 0; mm < 6; ++mm) {
        double o[3];
        rigid_row(mm, dx, dy, dz, f[3], o);
        const double c = zc[6 * int64_t(g) + mm];
        s[0] += o[0] * c;
        s[1] += o[1] * c;
        s[2] += o[2] * c;
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int64_t d = 3 * i + k;
      const T v = (fixed && fixed[d]) ? T(0) : T(s[k]);
      u[d] = ADD ? u[d] + v : v;
    }
  }
}

// ------------------------------------------ banded coarse factorisation
// Aggregates are numbered in the solver's element order, so two aggregates
// that share an element lie a few ids apart and the coarse matrix is block
// banded (half-bandwidth 6 * 2 + 5 = 17 at 250 aggregates for a 160x20x20
// beam).  band_cholesky factors it in one workgroup, band_forward forms
// L^-1 column by column and one rocBLAS SYRK the dense inverse L^-T L^-1;
// together they replace rocSOLVER's dense potrf + potri (6.4 ms at 1,500
// unknowns, mostly panel kernels and launch gaps) whenever the
// half-bandwidth is at most kBandW - 2.
constexpr int kBandW = 64;         // LDS window of the factorisation, ring of the substitution
constexpr int kBandChunk = 32;     // band rows fetched per chunk of factorisation steps
constexpr int kBandThreads = 256;  // threads of band_cholesky
constexpr int kBandNV = kBandChunk * kBandW / kBandThreads;

// v <- rows r0 .. r0 + kBandChunk - 1 of the band, v[q] = A(r, r - bw + t) for
// the (r, t) of slot threadIdx.x + q kBandThreads (zero outside the band)
__device__ __forceinline__ void band_fetch(int nc, int bw, const double* __restrict__ A,
                                           int64_t r0, double (&v)[kBandNV]) {
#pragma unroll
  for (int q = 0; q < kBandNV; ++q) {
    const int idx = int(threadIdx.x) + q * kBandThreads;
    const int64_t r = r0 + idx / kBandW;
    const int t = idx % kBandW;
    v[q] = (t <= bw && r < nc) ? A[r + (r - bw + t) * int64_t(nc)] : 0.0;
  }
}

// Right-looking Cholesky A = L L^T of the SPD matrix A (column-major nc x nc,
// lower triangle read, zero beyond half-bandwidth bw <= kBandW - 2), one
// column per step: column k is scaled by its pivot, then the trailing
// bw x bw triangle is updated.  The active rows live in an LDS window
// (entry (i, j) at win[i % kBandW][j % kBandW]); row k + 1 + bw enters at
// step k from a staging buffer that is filled one chunk of kBandChunk steps
// ahead, so the HBM latency of the band rows stays hidden.  Writes the
// factor by columns (Lc[k kBandW + t] = L(k + t, k)) and by rows
// (Lr[i kBandW + t] = L(i, i - t)), dinv[k] = 1 / L(k, k), and info = 1 +
// the first non-positive pivot (0 if none).
__global__ void __launch_bounds__(kBandThreads)
band_cholesky(int nc, int bw, const double* __restrict__ A, double* __restrict__ Lc,
              double* __restrict__ Lr, double* __restrict__ dinv, int32_t* __restrict__ info) {
  constexpr int W = kBandW;
  __shared__ double win[W][W];
  __shared__ double stg[kBandChunk][W];
  // the trailing triangle's entries (a, b), b <= a < bw, as a | b << 8
  __shared__ uint16_t pab[(W - 2) * (W - 1) / 2];
  const int th = threadIdx.x;
  const int np = bw * (bw + 1) / 2;
  for (int p = th; p < np; p += kBandThreads) {
    int a = int((sqrtf(8.0f * float(p) + 1.0f) - 1.0f) * 0.5f);
    while ((a + 1) * (a + 2) / 2 <= p) ++a;
    while (a * (a + 1) / 2 > p) --a;
    pab[p] = uint16_t(a | ((p - a * (a + 1) / 2) << 8));
  }
  for (int idx = th; idx < W * W; idx += kBandThreads) {
    const int i = idx / W, j = idx % W;
    if (i <= bw && i < nc && j <= i) win[i][j] = A[i + int64_t(j) * nc];
  }
  double v[kBandNV];
  band_fetch(nc, bw, A, 1 + bw, v);
#pragma unroll
  for (int q = 0; q < kBandNV; ++q) {
    const int idx = th + q * kBandThreads;
    stg[idx / W][idx % W] = v[q];
  }
  __syncthreads();
  band_fetch(nc, bw, A, 1 + bw + kBandChunk, v);
  int32_t bad = 0;
  // this thread's first trailing entry (a0, b0), kept in registers
  const int a0 = th < np ? (pab[th] & 255) : 0, b0 = th < np ? (pab[th] >> 8) : 0;
  for (int k0 = 0; k0 < nc; k0 += kBandChunk) {
    for (int s = 0; s < kBandChunk && k0 + s < nc; ++s) {
      const int k = k0 + s, ks = k % W;
      // One phase per column: the column's entries are read unscaled and
      // scaled by 1 / l where they are used (the same products as storing
      // the scaled column first), so the update needs no barrier of its own
      // and its reads go out beside the pivot's (1,500 unknowns, half-
      // bandwidth 17: 853 against 936 us).
      const double wout = (th > 0 && th <= bw && k + th < nc) ? win[(k + th) % W][ks] : 0.0;
      double wa = 0.0, wb = 0.0, wab = 0.0;
      const bool upd = th < np && k + 1 + a0 < nc;
      const int ia = (k + 1 + a0) % W, jb = (k + 1 + b0) % W;
      if (upd) {
        wa = win[ia][ks];
        wb = win[jb][ks];
        wab = win[ia][jb];
      }
      double d = win[ks][ks];
      if (!(d > 0.0)) {
        if (bad == 0) bad = k + 1;
        d = 1.0;
      }
      const double l = sqrt(d), rl = 1.0 / l;
      if (th <= bw) {
        const int i = k + th;
        if (i < nc) {
          const double lik = th == 0 ? l : wout * rl;
          Lc[int64_t(k) * W + th] = lik;
          Lr[int64_t(i) * W + th] = lik;
        } else {
          Lc[int64_t(k) * W + th] = 0.0;
        }
        if (th == 0) dinv[k] = rl;
      }
      if (upd) win[ia][jb] = wab - (wa * rl) * (wb * rl);
      for (int p = th + kBandThreads; p < np; p += kBandThreads) {
        const int a = pab[p] & 255, b = pab[p] >> 8;  // entry (k + 1 + a, k + 1 + b)
        if (k + 1 + a < nc) {
          const int i = (k + 1 + a) % W, j = (k + 1 + b) % W;
          win[i][j] -= (win[i][ks] * rl) * (win[j][ks] * rl);
        }
      }
      const int r = k + 1 + bw;
      if (r < nc && th <= bw) win[r % W][(r - bw + th) % W] = stg[s][th];
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kBandNV; ++q) {
      const int idx = th + q * kBandThreads;
      stg[idx / W][idx % W] = v[q];
    }
    __syncthreads();
    band_fetch(nc, bw, A, 1 + bw + int64_t(k0) + 2 * kBandChunk, v);
  }
  if (th == 0) *info = bad;
}

// Inverse factor Y = L^-1 from the band factor, one column per lane: lane j
// solves L y = e_j forward over rows j0 .. nc - 1 (j0 = first column of its
// wave; y = 0 above row j) and stores Y(i, j) at Y[i nc + j] (rows above j0
// are left to the caller's zero fill).  The inverse of A = L L^T is then
// Y^T Y, one rocBLAS SYRK.  The lane's last kBandW values stay in an LDS
// ring (lane-major, no bank conflicts); the factor entries are wave-uniform
// loads.  Fixed summation order: bitwise reproducible.
__global__ void __launch_bounds__(64)
band_forward(int nc, int bw, const double* __restrict__ Lr, const double* __restrict__ dinv,
             double* __restrict__ Y) {
  constexpr int W = kBandW;
  __shared__ double ring[W][64];
  const int lane = threadIdx.x;
  const int j0 = blockIdx.x * 64, j = j0 + lane;
  const bool live = j < nc;
  for (int i = j0; i < nc; ++i) {
    const double* lr = Lr + int64_t(i) * W;
    const int tmax = bw < i - j0 ? bw : i - j0;
    double s0 = i == j ? 1.0 : 0.0, s1 = 0.0;
    int t = 1;
    for (; t + 1 <= tmax; t += 2) {
      s0 -= lr[t] * ring[(i - t) % W][lane];
      s1 -= lr[t + 1] * ring[(i - t - 1) % W][lane];
    }
    if (t <= tmax) s0 -= lr[t] * ring[(i - t) % W][lane];
    const double y = (s0 + s1) * dinv[i];
    ring[i % W][lane] = y;
    if (live) Y[int64_t(i) * nc + j] = y;
  }
}

// band_forward for half-bandwidths below BWM, with the lane's last BWM - 1
// values in registers: row r sits in slot (r - j0) mod BWM and the step loop
// is unrolled BWM times, so every slot index is a constant.  The factor rows
// (BWM entries each, dinv beside them) are staged in LDS a chunk of 2 BWM
// rows at a time, the next chunk's loads in flight while the wave works
// through the current one (wave-uniform global loads per step left each
// step waiting on the L2).  The factor arrays are cleared before the
// factorisation, so the entries beyond the band (and past the last row) are
// zero and every step runs the same BWM - 1 multiply-adds with no bounds.
template <int BWM>
__global__ void __launch_bounds__(64)
band_forward_reg(int nc, const double* __restrict__ Lr, const double* __restrict__ dinv,
                 double* __restrict__ Y) {
  constexpr int CH = 2 * BWM;           // rows per staged chunk
  constexpr int PER = CH * BWM / 64;    // factor entries each lane fetches per chunk
  static_assert((CH * BWM) % 64 == 0 && CH <= 64 && BWM <= kBandW, "chunk shape");
  __shared__ double st[2][CH * BWM];
  __shared__ double sd[2][CH];
  const int lane = threadIdx.x;
  const int j0 = blockIdx.x * 64, j = j0 + lane;
  const bool live = j < nc;
  double v[PER], dv = 0.0;
  auto fetch = [&](int r0) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int q = lane + 64 * m;
      const int r = r0 + q / BWM;
      v[m] = r < nc ? Lr[int64_t(r) * kBandW + q % BWM] : 0.0;
    }
    dv = (lane < CH && r0 + lane < nc) ? dinv[r0 + lane] : 0.0;
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int m = 0; m < PER; ++m) st[buf][lane + 64 * m] = v[m];
    if (lane < CH) sd[buf][lane] = dv;
  };
  fetch(j0);
  stage(0);
  __syncthreads();
  if (j0 + CH < nc) fetch(j0 + CH);
  double h[BWM];
#pragma unroll
  for (int t = 0; t < BWM; ++t) h[t] = 0.0;
  int buf = 0;
  for (int ib = j0; ib < nc; ib += CH) {
#pragma unroll
    for (int u2 = 0; u2 < CH; ++u2) {
      const int u = u2 % BWM;
      const int i = ib + u2;
      if (i >= nc) break;
      const double* lr = st[buf] + u2 * BWM;
      // the older terms first, the previous row's last: from one row's
      // value to the next the chain is one multiply-add and one multiply
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int t = BWM - 1; t >= 2; --t) {
        const double x = lr[t] * h[(u - t + BWM) % BWM];
        if (t & 1) s0 += x; else s1 += x;
      }
      const double rhs = (i == j ? 1.0 : 0.0) - (s0 + s1);
      const double y = (rhs - lr[1] * h[(u - 1 + BWM) % BWM]) * sd[buf][u2];
      h[u] = y;
      if (live) Y[int64_t(i) * nc + j] = y;
    }
    // the next chunk (in flight since the last stage) to the other buffer,
    // then the one after it in flight
    __syncthreads();
    stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
    if (ib + 2 * CH < nc) fetch(ib + 2 * CH);
  }
}
constexpr int kBandReg = 24;  // band_forward_reg<kBandReg> serves half-bandwidths below it

}  // namespace

namespace {

struct Coarse {
  int64_t n_groups = 0, n_agg = 0, nc = 0;  // host
  int agg_tiles = kAggTiles;                 // element tiles per aggregate
  int64_t bw = 0;                            // half-bandwidth of the coarse matrix
  int n_col = 0;
  int32_t* graw;
  uint32_t* used;
  uint32_t* gid;
  int32_t* agg;        // node -> aggregate (-1: no element)
  int32_t* agg_off;    // aggregate -> positions in agg_nodes
  uint32_t* keys_in;
  uint32_t* keys_out;
  uint32_t* ids_in;
  uint32_t* agg_nodes; // nodes sorted by aggregate
  double* frame;       // centroid, 1 / RMS radius
  uint32_t* pcnt;
  uint32_t* poff;
  uint64_t* pkeys_in;
  uint64_t* pkeys_out;
  int32_t* color;
  int32_t* nbc;        // n_agg x n_col: neighbour (or self) of each colour, -1
  double* A;           // coarse matrix, then its inverse (column-major nc x nc)
  float* A32;          // FP32 copy of the symmetric inverse
  double* rc;
  double* zc;
  double* probe;       // 6 probe vectors, 6 products
  void* u;             // preconditioned residual being built (inner precision)
  void* rl;            // copy of the outer residual in the inner precision
  void* r2;            // residual after the smoothing step (inner precision)
  int32_t* info;
  int32_t* dep;        // coarse modes dropped as dependent (1) or kept (0)
  void* tmp;
  size_t tmp_bytes = 0;
};

void coarse_layout(int64_t n_nodes, int64_t n_elem, cf::Carve& c, Coarse* C, int lpb) {
  const int64_t n = 3 * n_nodes, n_tiles = (10 * n_elem + kAggUnitItems - 1) / kAggUnitItems;
  const int at = agg_tiles_for(n_tiles);
  const int64_t G = (n_tiles + at - 1) / at, ncmax = 6 * G;
  C->agg_tiles = at;
  const int64_t maxp = kPairsPerElem * n_elem;
  C->n_groups = G;
  C->graw = c.take<int32_t>(n_nodes);
  C->used = c.take<uint32_t>(G + 1);
  C->gid = c.take<uint32_t>(G + 1);
  C->agg = c.take<int32_t>(n_nodes);
  C->agg_off = c.take<int32_t>(G + 1);
  C->keys_in = c.take<uint32_t>(n_nodes);
  C->keys_out = c.take<uint32_t>(n_nodes);
  C->ids_in = c.take<uint32_t>(n_nodes);
  C->agg_nodes = c.take<uint32_t>(n_nodes);
  C->frame = c.take<double>(4 * G);
  C->pcnt = c.take<uint32_t>(n_elem + 1);
  C->poff = c.take<uint32_t>(n_elem + 1);
  C->pkeys_in = c.take<uint64_t>(maxp);
  C->pkeys_out = c.take<uint64_t>(maxp);
  C->color = c.take<int32_t>(G);
  C->nbc = c.take<int32_t>(G * kMaxColors);
  C->A = c.take<double>(ncmax * ncmax);
  C->A32 = c.take<float>(ncmax * ncmax);
  C->rc = c.take<double>(ncmax);
  C->zc = c.take<double>(ncmax);
  C->probe = c.take<double>(12 * n);
  C->u = c.take<char>(n * lpb);
  C->rl = c.take<char>(n * lpb);
  C->r2 = c.take<char>(n * lpb);
  C->info = c.take<int32_t>(2);
  C->dep = c.take<int32_t>(ncmax);
  size_t b1 = 0, b2 = 0, b3 = 0, b4 = 0;
  rocprim::radix_sort_pairs<SortConfig>(nullptr, b1, (uint32_t*)nullptr, (uint32_t*)nullptr,
                            (uint32_t*)nullptr, (uint32_t*)nullptr, size_t(n_nodes), 0, 32);
  rocprim::exclusive_scan(nullptr, b2, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u,
                          size_t(G + 1), rocprim::plus<uint32_t>());
  rocprim::exclusive_scan(nullptr, b3, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u,
                          size_t(n_elem + 1), rocprim::plus<uint32_t>());
  rocprim::radix_sort_keys<SortConfig>(nullptr, b4, (uint64_t*)nullptr, (uint64_t*)nullptr, size_t(maxp), 0,
                           64);
  size_t b5 = 0;
  rocprim::unique(nullptr, b5, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr,
                  size_t(maxp), rocprim::equal_to<uint64_t>());
  C->tmp_bytes = std::max(std::max(std::max(b1, b2), std::max(b3, b4)), b5);
  C->tmp = c.take<char>(C->tmp_bytes);
}

// workspace of the Chebyshev / two-level solvers; C (two-level only) and A
// (assembled outer operator: its sliced copy, lsum slice steps) may be null
size_t mf_mixed_layout(int64_t n_nodes, int64_t n_elem, void* base, Layout* L, MfMixed* M,
                       Coarse* C, int lpb = 4, const cf_matrix* A = nullptr,
                       int64_t lsum = 0) {
  cf::Carve c(base);
  mf_mixed_layout_base(n_nodes, n_elem, c, L, M, lpb, A);
  if (C) coarse_layout(n_nodes, n_elem, c, C, lpb);
  if (A) sell_layout(A, CF_PREC_FP64, lsum, c, *L);
  L->total = c.used;
  return c.used;
}

// rocBLAS handle of the current device for the rocSOLVER factorisation (one
// per device for the life of the process; not thread-safe).
rocblas_handle solver_handle() {
  static rocblas_handle handles[64] = {};
  static std::mutex mu;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!handles[dev] && rocblas_create_handle(&handles[dev]) != rocblas_status_success)
    handles[dev] = nullptr;
  return handles[dev];
}

// Aggregates, colouring, coarse matrix by probing, and its inverse.
int coarse_setup(const MfArgs& a, const MfRuns& R, const double* xs, const double* ys,
                 const double* zs, const uint8_t* fixed, double* part, Coarse* C,
                 hipStream_t s) {
  const int64_t n_nodes = a.n_nodes, n_elem = a.n_elem, n = 3 * n_nodes, G = C->n_groups;
  const int gn = cf::stride_grid(n_nodes, kBlock), ge = cf::stride_grid(n_elem + 1, kBlock);
  size_t tb;
  // Aggregates from the element order; when their graph needs more than
  // kMaxColors colours (a node numbering that scatters the groups), slabs
  // across the longest side of the bounding box instead.
  uint32_t n_agg = 0;
  int64_t agg_band = 0;  // largest id distance of two coupled aggregates
  int n_col = 0;
  std::vector<int32_t> color;
  std::vector<std::vector<int32_t>> adj;
  for (int attempt = 0;; ++attempt) {
    if (attempt == 0) {
      // first-touching element group, compacted ids
      CF_HIP(hipMemsetAsync(C->used, 0, sizeof(uint32_t) * (G + 1), s));
      agg_first<<<gn, kBlock, 0, s>>>(n_nodes, a.offs, a.perm, C->agg_tiles, C->graw, C->used);
      CF_LAUNCHED();
      tb = C->tmp_bytes;
      CF_HIP(rocprim::exclusive_scan(C->tmp, tb, C->used, C->gid, 0u, size_t(G + 1),
                                     rocprim::plus<uint32_t>(), s));
      agg_compact<<<gn, kBlock, 0, s>>>(n_nodes, C->graw, C->gid, G, C->agg, C->keys_in,
                                        C->ids_in);
      CF_LAUNCHED();
      CF_HIP(hipMemcpyAsync(&n_agg, C->gid + G, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      CF_HIP(hipStreamSynchronize(s));
      CF_REQUIRE(n_agg > 0 && int64_t(n_agg) <= G, "two-level: bad aggregate count %u", n_agg);
      tb = C->tmp_bytes;
      CF_HIP(rocprim::radix_sort_pairs<SortConfig>(C->tmp, tb, C->keys_in, C->keys_out, C->ids_in,
                                       C->agg_nodes, size_t(n_nodes), 0,
                                       bits_for(int64_t(n_agg) + 1), s));
      agg_offsets<<<cf::stride_grid(n_agg + 1, kBlock), kBlock, 0, s>>>(n_nodes, n_agg,
                                                                        C->keys_out, C->agg_off);
      CF_LAUNCHED();
    } else {
      // slabs: nodes sorted along the longest side of the bounding box
      std::vector<double> box(6 * size_t(gn));
      node_bbox<<<gn, kBlock, 0, s>>>(n_nodes, xs, ys, zs, C->probe);
      CF_LAUNCHED();
      CF_HIP(hipMemcpyAsync(box.data(), C->probe, sizeof(double) * box.size(),
                            hipMemcpyDeviceToHost, s));
      CF_HIP(hipStreamSynchronize(s));
      double ext[3];
      for (int k = 0; k < 3; ++k) {
        double lo = INFINITY, hi = -INFINITY;
        for (int blk = 0; blk < gn; ++blk) {
          lo = std::min(lo, box[6 * size_t(blk) + k]);
          hi = std::max(hi, box[6 * size_t(blk) + 3 + k]);
        }
        ext[k] = hi - lo;
      }
      const int axis = ext[0] >= ext[1] && ext[0] >= ext[2] ? 0 : (ext[1] >= ext[2] ? 1 : 2);
      slab_keys<<<gn, kBlock, 0, s>>>(n_nodes, a.offs, axis == 0 ? xs : axis == 1 ? ys : zs,
                                      C->keys_in, C->ids_in);
      CF_LAUNCHED();
      tb = C->tmp_bytes;
      CF_HIP(rocprim::radix_sort_pairs<SortConfig>(C->tmp, tb, C->keys_in, C->keys_out, C->ids_in,
                                       C->agg_nodes, size_t(n_nodes), 0, 32, s));
      slab_count<<<1, 64, 0, s>>>(n_nodes, C->keys_out, C->used);
      CF_LAUNCHED();
      uint32_t n_in = 0;
      CF_HIP(hipMemcpyAsync(&n_in, C->used, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      CF_HIP(hipStreamSynchronize(s));
      CF_REQUIRE(n_in > 0, "two-level: no node belongs to an element");
      const int64_t npa = (int64_t(n_in) + G - 1) / G;  // nodes per aggregate
      n_agg = uint32_t((int64_t(n_in) + npa - 1) / npa);
      CF_HIP(hipMemsetAsync(C->agg, 0xff, sizeof(int32_t) * n_nodes, s));
      slab_assign<<<gn, kBlock, 0, s>>>(n_in, npa, n_agg, C->agg_nodes, C->agg, C->agg_off);
      CF_LAUNCHED();
    }
    C->n_agg = n_agg;
    C->nc = 6 * int64_t(n_agg);
    agg_frame<<<int(std::min<int64_t>(n_agg, 65535)), kBlock, 0, s>>>(n_agg, C->agg_off,
                                                                      C->agg_nodes, xs, ys, zs,
                                                                      C->frame);
    CF_LAUNCHED();
    // aggregate adjacency through elements
    agg_pair_count<<<ge, kBlock, 0, s>>>(n_elem, a.conn, C->agg, C->pcnt);
    CF_LAUNCHED();
    tb = C->tmp_bytes;
    CF_HIP(rocprim::exclusive_scan(C->tmp, tb, C->pcnt, C->poff, 0u, size_t(n_elem + 1),
                                   rocprim::plus<uint32_t>(), s));
    uint32_t n_pairs = 0;
    CF_HIP(hipMemcpyAsync(&n_pairs, C->poff + n_elem, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          s));
    CF_HIP(hipStreamSynchronize(s));
    CF_REQUIRE(int64_t(n_pairs) <= kPairsPerElem * n_elem, "two-level: %u aggregate pairs",
               n_pairs);
    std::vector<uint64_t> pk;
    if (n_pairs) {
      agg_pair_fill<<<ge, kBlock, 0, s>>>(n_elem, a.conn, C->agg, C->poff, n_agg, C->pkeys_in);
      CF_LAUNCHED();
      tb = C->tmp_bytes;
      CF_HIP(rocprim::radix_sort_keys<SortConfig>(C->tmp, tb, C->pkeys_in, C->pkeys_out, size_t(n_pairs), 0,
                                      bits_for(int64_t(n_agg) * n_agg + 1), s));
      // distinct pairs only (about 3 per aggregate instead of one per element
      // on a slab boundary) cross to the host
      tb = C->tmp_bytes;
      CF_HIP(rocprim::unique(C->tmp, tb, C->pkeys_out, C->pkeys_in, C->used, size_t(n_pairs),
                             rocprim::equal_to<uint64_t>(), s));
      uint32_t n_unique = 0;
      CF_HIP(hipMemcpyAsync(&n_unique, C->used, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      CF_HIP(hipStreamSynchronize(s));
      CF_REQUIRE(n_unique > 0 && n_unique <= n_pairs, "two-level: %u distinct pairs", n_unique);
      pk.resize(n_unique);
      CF_HIP(hipMemcpyAsync(pk.data(), C->pkeys_in, sizeof(uint64_t) * n_unique,
                            hipMemcpyDeviceToHost, s));
      CF_HIP(hipStreamSynchronize(s));
    }
    // distance-2 greedy colouring on the host (n_agg is small)
    adj.assign(n_agg, {});
    for (uint32_t g = 0; g < n_agg; ++g) adj[g].push_back(int32_t(g));
    uint64_t last = ~uint64_t(0);
    agg_band = 0;
    for (uint64_t key : pk) {
      if (key == last) continue;
      last = key;
      const int32_t g = int32_t(key / n_agg), h = int32_t(key % n_agg);
      adj[g].push_back(h);
      adj[h].push_back(g);
      agg_band = std::max<int64_t>(agg_band, g > h ? g - h : h - g);
    }
    color.assign(n_agg, -1);
    std::vector<uint32_t> stamp(kMaxColors, 0);
    n_col = 0;
    bool fits = true;
    for (uint32_t g = 0; g < n_agg && fits; ++g) {
      for (int32_t h : adj[g])
        for (int32_t k : adj[h])
          if (color[k] >= 0) stamp[color[k]] = g + 1;
      int c = 0;
      while (c < kMaxColors && stamp[c] == g + 1) ++c;
      fits = c < kMaxColors;
      color[g] = c;
      n_col = std::max(n_col, c + 1);
    }
    if (fits) break;
    CF_REQUIRE(attempt == 0, "two-level: aggregate graph needs more than %d colours",
               kMaxColors);
  }
  std::vector<int32_t> nbc(size_t(n_agg) * n_col, -1);
  for (uint32_t h = 0; h < n_agg; ++h)
    for (int32_t g : adj[h]) nbc[size_t(h) * n_col + color[g]] = g;
  C->n_col = n_col;
  CF_HIP(hipMemcpyAsync(C->color, color.data(), sizeof(int32_t) * n_agg, hipMemcpyHostToDevice,
                        s));
  CF_HIP(hipMemcpyAsync(C->nbc, nbc.data(), sizeof(int32_t) * nbc.size(), hipMemcpyHostToDevice,
                        s));
  // coarse matrix: 6 products per colour
  const int64_t nc = C->nc;
  CF_HIP(hipMemsetAsync(C->A, 0, sizeof(double) * nc * nc, s));
  double* v = C->probe;
  double* yv = C->probe + 6 * n;
  int rc;
  for (int c = 0; c < n_col; ++c) {
    probe_build<<<gn, kBlock, 0, s>>>(n_nodes, C->agg, C->color, c, C->frame, xs, ys, zs, fixed, v);
    CF_LAUNCHED();
    for (int mm = 0; mm < 6; ++mm)
      if ((rc = mf_apply(a, R, v + mm * n, yv + mm * n, fixed, part, nullptr, nullptr, s)))
        return rc;
    probe_extract<<<int(std::min<int64_t>(n_agg, 65535)), kBlock, 0, s>>>(
        n_agg, n_col, c, C->nbc, C->agg_off, C->agg_nodes, C->frame, xs, ys, zs, n, yv, C->A);
    CF_LAUNCHED();
  }
  const int gc = cf::stride_grid(nc, 64);
  coarse_dependent_modes<<<cf::stride_grid(n_agg, 64), 64, 0, s>>>(n_agg, nc, C->A, C->dep);
  CF_LAUNCHED();
  coarse_regularize<<<gc, 64, 0, s>>>(nc, C->A, C->dep);
  CF_LAUNCHED();
  C->bw = 6 * agg_band + 5;  // scalar half-bandwidth
  return CF_OK;
}

// Rows and columns of the dropped coarse modes of the inverse set to zero
// (one thread per mode clears both).
__global__ void coarse_drop_modes(int64_t nc, double* __restrict__ Ainv,
                                  const int32_t* __restrict__ dep) {
  for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < nc;
       j += int64_t(gridDim.x) * blockDim.x) {
    if (!dep[j]) continue;
    for (int64_t i = 0; i < nc; ++i) {
      Ainv[i + j * nc] = 0.0;
      Ainv[j + i * nc] = 0.0;
    }
  }
}

// Factor and invert the coarse matrix on stream s, with no host
// synchronisation (coarse_factor_check reads the pivot flags later).
// Narrow bands use the banded kernels, wide ones rocSOLVER's dense
// potrf/potri.  The environment variable CF_COARSE_FACTOR=dense (rocSOLVER)
// or =ring (LDS-ring inverse) pins a path, so that each one can be checked
// on narrow-band problems.
// started (optional): recorded on s once the factor arrays are cleared
int coarse_factor(Coarse* C, int64_t n, hipStream_t s, hipEvent_t started = nullptr) {
  const int64_t nc = C->nc, bw = C->bw;
  const char* pin = std::getenv("CF_COARSE_FACTOR");
  const bool pin_dense = pin && std::strcmp(pin, "dense") == 0;
  const bool pin_ring = pin && std::strcmp(pin, "ring") == 0;
  CF_HIP(hipMemsetAsync(C->info, 0, 2 * sizeof(int32_t), s));
  if (!pin_dense && bw <= kBandW - 2 && 2 * nc * kBandW + nc + nc * nc <= 12 * n) {
    // banded factor, Y = L^-1 by forward substitution, A^-1 = Y^T Y; the
    // probe vectors' space holds the factor and Y
    double* Lc = C->probe;
    double* Lr = Lc + nc * kBandW;
    double* dinv = Lr + nc * kBandW;
    double* Y = dinv + nc;
    CF_HIP(hipMemsetAsync(Lc, 0, sizeof(double) * 2 * nc * kBandW, s));
    CF_HIP(hipMemsetAsync(Y, 0, sizeof(double) * nc * nc, s));
    if (started) CF_HIP(hipEventRecord(started, s));
    band_cholesky<<<1, kBandThreads, 0, s>>>(int(nc), int(bw), C->A, Lc, Lr, dinv, C->info);
    CF_LAUNCHED();
    const int gi = int((nc + 63) / 64);
    if (bw < kBandReg && !pin_ring)
      band_forward_reg<kBandReg><<<gi, 64, 0, s>>>(int(nc), Lr, dinv, Y);
    else
      band_forward<<<gi, 64, 0, s>>>(int(nc), int(bw), Lr, dinv, Y);
    CF_LAUNCHED();
    // Y[i nc + j] = Y(i, j) is Y^T column-major, so (Y^T)(Y^T)^T = Y^T Y is
    // A^-1: one DGEMM (rocBLAS's SYRK splits into several GEMM launches and
    // took 0.75 ms here against one GEMM's fraction of that)
    rocblas_handle hnd = solver_handle();
    CF_REQUIRE(hnd != nullptr, "two-level: no rocBLAS handle");
    CF_REQUIRE(rocblas_set_stream(hnd, s) == rocblas_status_success,
               "two-level: rocblas_set_stream");
    const double one = 1.0, zero = 0.0;
    CF_REQUIRE(rocblas_dgemm(hnd, rocblas_operation_none, rocblas_operation_transpose,
                             rocblas_int(nc), rocblas_int(nc), rocblas_int(nc), &one, Y,
                             rocblas_int(nc), Y, rocblas_int(nc), &zero, C->A,
                             rocblas_int(nc)) == rocblas_status_success,
               "two-level: rocblas_dgemm failed");
  } else {
    rocblas_handle hnd = solver_handle();
    CF_REQUIRE(hnd != nullptr, "two-level: no rocBLAS handle");
    CF_REQUIRE(rocblas_set_stream(hnd, s) == rocblas_status_success,
               "two-level: rocblas_set_stream");
    CF_REQUIRE(rocsolver_dpotrf(hnd, rocblas_fill_lower, rocblas_int(nc), C->A, rocblas_int(nc),
                                C->info) == rocblas_status_success,
               "two-level: rocsolver_dpotrf failed");
    CF_REQUIRE(rocsolver_dpotri(hnd, rocblas_fill_lower, rocblas_int(nc), C->A, rocblas_int(nc),
                                C->info + 1) == rocblas_status_success,
               "two-level: rocsolver_dpotri failed");
  }
  // exact symmetry: the lower triangle (column-major) mirrored onto the upper
  const int64_t nt = (nc + 31) / 32;
  coarse_fill_upper<<<int(std::min<int64_t>(nt * nt, 65535)), kBlock, 0, s>>>(nc, C->A);
  CF_LAUNCHED();
  coarse_drop_modes<<<cf::stride_grid(nc, 64), 64, 0, s>>>(nc, C->A, C->dep);
  CF_LAUNCHED();
  to_type<float><<<cf::stride_grid(nc * nc, kBlock), kBlock, 0, s>>>(nc * nc, C->A, C->A32);
  CF_LAUNCHED();
  return CF_OK;
}

// Pivot flags of coarse_factor (waits for stream s).
int coarse_factor_check(const Coarse& C, hipStream_t s) {
  int32_t info[2] = {0, 0};
  CF_HIP(hipMemcpyAsync(info, C.info, sizeof(info), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  CF_REQUIRE(info[0] == 0, "two-level: coarse matrix not positive definite (pivot %d)", info[0]);
  CF_REQUIRE(info[1] == 0, "two-level: coarse inverse failed (potri info %d)", info[1]);
  return CF_OK;
}

// Side stream (default priority, non-blocking) and two events per device
// for the Chebyshev bound and the sliced copy of A, which run beside the
// coarse set-up;
// created on first use and kept until the process ends.  Solves that
// overlap on one device share them.
struct Side {
  hipStream_t s = nullptr;
  hipEvent_t ready = nullptr, done = nullptr;
};

// Per device: serialises two-level solves (they share the side stream and
// its events) and guards the lazy creation below.
std::mutex& device_mutex(int dev) {
  static std::mutex mu[64];
  return mu[dev];
}

const Side* side_stream() {
  static Side side[64];
  static std::mutex mu;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  Side& S = side[dev];
  if (!S.s && hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking) != hipSuccess) S.s = nullptr;
  if (S.s && !S.ready && hipEventCreateWithFlags(&S.ready, hipEventDisableTiming) != hipSuccess)
    S.ready = nullptr;
  if (S.ready && !S.done && hipEventCreateWithFlags(&S.done, hipEventDisableTiming) != hipSuccess)
    S.done = nullptr;
  return S.done ? &S : nullptr;
}

// Makes stream s wait for the side stream's work when it goes out of scope,
// on every return path, so that the caller's stream-ordered frees of the
// workspace come after it.
struct SideJoin {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  ~SideJoin() {
    if (done) (void)hipStreamWaitEvent(s, done, 0);
  }
};

// Runs fn() (launches only) on a second host thread with the caller's
// device current, so that its launch calls overlap the caller's host work;
// joined by join() or when the object goes out of scope.  join() returns
// fn's status and moves its error text to the calling thread.  If no thread
// can be started, fn runs inline (no exception crosses the C ABI).
class HostTask {
 public:
  template <class F>
  explicit HostTask(F&& fn) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    try {
      t_ = std::thread([this, dev, fn]() mutable {
        if (hipSetDevice(dev) == hipSuccess) {
          rc_ = fn();
        } else {
          cf::set_error("hipSetDevice(%d) failed on the launch thread", dev);
          rc_ = CF_ERR_HIP;
        }
        if (rc_) msg_ = cf_last_error();
      });
    } catch (...) {
      rc_ = fn();
    }
  }
  HostTask(const HostTask&) = delete;
  HostTask& operator=(const HostTask&) = delete;
  ~HostTask() { join(); }
  int join() {
    if (t_.joinable()) {
      t_.join();
      if (rc_) cf::set_error("%s", msg_.c_str());
    }
    return rc_;
  }

 private:
  int rc_ = CF_OK;
  std::string msg_;
  std::thread t_;
};

}  // namespace

namespace {

// Flexible CG preconditioned by the matrix-free Chebyshev polynomial in
// precision T (float: the mixed solvers; double: their FP64 twins), alone or,
// with xyz, inside the two-level cycle.  The outer FP64 products are
// matrix-free, or, when A is given, products with the assembled NB3 matrix A
// (its sliced FP64 copy), so that the system solved is A's own.
template <class T>
int mf_mixed_solve(const char* who, const cf_matrix* A, int64_t n_nodes, int64_t n_elem,
                   const int32_t* conn, const double* geo, double lam, double mu,
                   const int32_t* offs, const uint32_t* perm, const double* xs, const double* ys,
                   const double* zs, bool two_level, const double* b, double* x,
                   const uint8_t* fixed, double tol, int32_t maxit, void* ws, size_t ws_bytes,
                   cf_pcg_stats* stats_host, void* stream) {
  MfArgs a{n_nodes, n_elem, conn, geo, lam, mu, offs, perm};
  int rc = mf_check(a);
  if (rc) return rc;
  CF_REQUIRE(b && x && ws && stats_host, "%s: null argument", who);
  CF_REQUIRE(!two_level || (xs && ys && zs), "%s: node coordinates required", who);
  CF_REQUIRE(10 * n_elem < (int64_t(1) << 31), "%s: too many elements", who);
  CF_REQUIRE(tol > 0.0 && maxit > 0, "%s: tol and maxit must be positive", who);
  int64_t lsum = 0;
  if (A) {
    if ((rc = check_matrix(A))) return rc;
    CF_REQUIRE(A->format == CF_FMT_NB3 && A->n_rows == n_nodes,
               "%s: the assembled operator must be the NB3 matrix of the mesh (%d node rows, "
               "%lld nodes)", who, int(A->n_rows), (long long)n_nodes);
    if ((rc = sell_bound(A, &lsum, cf::as_stream(stream)))) return rc;
  }
  Layout L{};
  MfMixed M{};
  Coarse C{};
  const size_t need = mf_mixed_layout(n_nodes, n_elem, ws, &L, &M, two_level ? &C : nullptr,
                                      int(sizeof(T)), A, lsum);
  if (ws_bytes < need) {
    cf::set_error("%s: workspace %zu < %zu bytes", who, ws_bytes, need);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  const int G = pcg_grid();
  const int64_t n = 3 * n_nodes;
  T* geoL = static_cast<T*>(M.geoL);
  T* feL = static_cast<T*>(M.feL);
  T* sL = static_cast<T*>(M.sL);
  *stats_host = cf_pcg_stats{};
  CF_HIP(hipMemsetAsync(L.state, 0, sizeof(PcgState), s));
  CF_HIP(hipMemsetAsync(L.bad, 0, sizeof(int32_t), s));
  mf_t10_diag<<<cf::stride_grid(n_nodes, kBlock), kBlock, 0, s>>>(n_nodes, offs, perm, geo, lam,
                                                                  mu, M.blocks);
  CF_LAUNCHED();
  blocks_to_precond<<<cf::stride_grid(n_nodes, kBlock), kBlock, 0, s>>>(n_nodes, M.blocks, fixed,
                                                                        0, L.dinv, L.bad);
  CF_LAUNCHED();
  int32_t bad = 0;
  CF_HIP(hipMemcpyAsync(&bad, L.bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  if (bad) {
    cf::set_error("%s: %d free diagonal entries are not positive", who, bad);
    return CF_ERR_BREAKDOWN;
  }
  make_scale<<<G, kBlock, 0, s>>>(n, L.dinv, L.s);
  CF_LAUNCHED();
  to_type<T><<<G, kBlock, 0, s>>>(n, L.s, sL);
  CF_LAUNCHED();
  if ((rc = mf_runs_build(n_nodes, n_elem, offs, perm, &M.R, s))) return rc;
  mf_pack_tiles<T><<<cf::stride_grid(10 * n_elem, kBlock), kBlock, 0, s>>>(
      n_elem, conn, geo, M.R.rank, M.connT, geoL, M.rankT);
  CF_LAUNCHED();
  // the sliced product's grid is read back when the outer products switch to
  // A (a synchronisation point anyway), not here.  Two-level solves build
  // the sliced copy on the side stream, after the Lanczos estimate and beside
  // the coarse factorisation: the outer products switch to A only many
  // iterations later, and the factorisation leaves most of the GPU idle.
  int GS = 0;
  const bool sell_on_side = two_level && A;
  if (A && !sell_on_side &&
      ((rc = build_sell(A, L, s, nullptr)) ||
       (rc = fill_sell<double>(A, L, nullptr, L.sval64, s))))
    return rc;
  const int ge = elem_grid(n_elem);
  const int gg = int(mf_gather_grid(n_nodes));
  const int32_t* xsplit = L.xsplit;
  // With A the outer products start matrix-free (FP64, cheap) and switch to
  // A once that converges: the solver then restarts from A's true residual
  // and ends on A's own system (the matrix-free operator equals A only up
  // to rounding, which the conditioning of a clamped beam magnifies).
  bool use_A = false;
  int outer_parts = gg;
  auto outer = [&](const double* xin, double* y, PcgState* state) {
    if (use_A) {
      spmv_sell<double, double, double><<<GS, kBlock, 0, s>>>(
          A->n_rows, L.n_slices, L.rows, L.koff, L.scol, L.sval64, xin, y, fixed, state,
          L.partials, xsplit);
      CF_LAUNCHED();
      return CF_OK;
    }
    return mf_apply(a, M.R, xin, y, fixed, M.fe, state, state ? L.partials : nullptr, s);
  };
  int switch_rc = CF_OK;
  auto switch_outer = [&]() {
    if (!A || use_A) return false;
    if ((switch_rc = sell_product_grid(L, s, &GS))) return false;
    use_A = true;
    outer_parts = GS;
    return true;
  };
  T* rl = static_cast<T*>(C.rl);
  T* u = static_cast<T*>(C.u);
  T* r2v = static_cast<T*>(C.r2);
  // the power iteration multiplies plain vectors (S K S v); the Chebyshev
  // steps run on S-scaled vectors, so their element kernel gathers no S.
  // Step k reads z_k (din) and z_{k-1} (dout, none at k = 1) and overwrites
  // dout with z_{k+1}.  Modes 3 and 4: products of the two-level form.
  // zout: where mode 5 adds its step (else the step overwrites dout); ps:
  // the stream (default s)
  const bool use_mfma = inner_mfma();
  auto product = [&](const T* din, T* dout, int mode, T cdk, T cwk, const PcgState* state,
                     T* zout = nullptr, hipStream_t ps = nullptr) {
    if (!ps) ps = s;
    if (mode == 0)
      launch_elem_lp<T, false>(use_mfma, ge, ps, n_elem, M.connT, geoL, T(lam), T(mu), sL, din,
                               M.rankT, M.R.toff, M.R.tstart, feL, state);
    else
      launch_elem_lp<T, true>(use_mfma, ge, ps, n_elem, M.connT, geoL, T(lam), T(mu), sL, din,
                              M.rankT, M.R.toff, M.R.tstart, feL, state);
    CF_LAUNCHED();
    if (M.R.use_run4)
      mf_gather_cheb<T, true><<<gg, kBlock, 0, ps>>>(n_nodes, M.R.nro, M.R.nq, feL, sL, mode, din,
                                                     dout, zout ? zout : dout, lp<T>(L.rh), cdk,
                                                     cwk, two_level ? rl : nullptr, state,
                                                     M.R.run4);
    else
      mf_gather_cheb<T><<<gg, kBlock, 0, ps>>>(n_nodes, M.R.nro, M.R.nq, feL, sL, mode, din,
                                               dout, zout ? zout : dout, lp<T>(L.rh), cdk, cwk,
                                               two_level ? rl : nullptr, state);
    CF_LAUNCHED();
    return CF_OK;
  };
  auto power = [&](const T* v, T* y) { return product(v, y, 0, T(0), T(0), nullptr); };
  auto cheb = [&](int k, T cdk, T cwk, const T* din, T* dout) {
    return product(din, dout, k == 1 ? 1 : 2, cdk, cwk, L.state);
  };
  // Two-level: the Lanczos estimate of the Chebyshev bound (and, with A, the
  // sliced copy of A) runs on a side stream while this stream builds, factors
  // and inverts the coarse matrix (whose host colouring and one-workgroup
  // factorisation leave most of the GPU idle).
  double power_est = 0.0;
  // two-level solves on one device run one at a time (shared side stream,
  // events and rocBLAS handle); the lock is held until the solve returns
  std::unique_lock<std::mutex> device_lock;
  if (two_level) {
    int dev = 0;
    CF_HIP(hipGetDevice(&dev));
    CF_REQUIRE(dev >= 0 && dev < 64, "two-level: device %d out of range", dev);
    device_lock = std::unique_lock<std::mutex>(device_mutex(dev));
  }
  SideJoin join;
  if (two_level) {
    const Side* side = side_stream();
    const hipStream_t sp = side ? side->s : s;
    if (side) {
      CF_HIP(hipEventRecord(side->ready, s));
      CF_HIP(hipStreamWaitEvent(sp, side->ready, 0));
      join.s = s;
      join.done = side->done;
    }
    auto power_side = [&](const T* v, T* y) {
      return product(v, y, 0, T(0), T(0), nullptr, nullptr, sp);
    };
    // its ~200 launches go out from a second host thread
    double lz[2 * kLanczosSteps] = {};
    HostTask launches([&]() {
      int r = lanczos_launch<T>(power_side, n, L, sp);
      if (r == CF_OK) r = read_lanczos(L, lz, sp);
      if (r == CF_OK && sell_on_side) r = build_sell(A, L, sp, nullptr);
      // recorded here as well, so that an early return's join still orders
      // the caller's stream after this side work
      if (side && hipEventRecord(side->done, sp) != hipSuccess && r == CF_OK) {
        cf::set_error("two-level: hipEventRecord on the side stream failed");
        r = CF_ERR_HIP;
      }
      return r;
    });
    if ((rc = coarse_setup(a, M.R, xs, ys, zs, fixed, M.fe, &C, s))) return rc;
    if ((rc = launches.join())) return rc;
    // the values of the sliced copy (a bandwidth-bound pass over A) go beside
    // the banded factorisation, whose one-workgroup kernels leave the GPU
    // idle, rather than beside the probing products or the clearing of the
    // factor arrays
    const bool fill_after = side && sell_on_side;
    if ((rc = coarse_factor(&C, n, s, fill_after ? side->ready : nullptr))) return rc;
    if (fill_after) CF_HIP(hipStreamWaitEvent(sp, side->ready, 0));
    if (sell_on_side && (rc = fill_sell<double>(A, L, nullptr, L.sval64, sp))) return rc;
    if (side) CF_HIP(hipEventRecord(side->done, sp));
    if ((rc = coarse_factor_check(C, s))) return rc;
    power_est = lanczos_max(lz, kLanczosSteps);
    if (side) CF_HIP(hipStreamWaitEvent(s, side->done, 0));
  }
  // FP64 floor of the final operator when it is A (matrix-free: none stored)
  auto abs_norm = [&](const double* xf) -> int {
    if (!use_A) return -1;
    sell_abs_norm<<<GS, kBlock, 0, s>>>(A->n_rows, L.n_slices, L.rows, L.koff, L.scol, L.sval64,
                                        xf, fixed, L.partials, xsplit);
    CF_LAUNCHED();
    return fin(kFinStore, L, s, &L.state->floor2, 0.0, GS);
  };
  if (!two_level) {
    rc = fcg_cheb_core<true, T>(outer, outer_parts, power, cheb, 0.0, kChebF32, n, b, x, fixed,
                                tol, maxit, L, stats_host, s, nullptr, nullptr, 0, 0.0,
                                switch_outer, abs_norm);
    return rc ? rc : switch_rc;
  }
  // two-level: z1 = C r; z2 = z1 + Cheb(r - K z1); z = z2 + C (r - K z2)
  const int GV = pcg_vec_grid();
  const int ga = int(std::min<int64_t>(C.n_agg, 65535));
  const int gc = int(std::min<int64_t>((C.nc + kWaves - 1) / kWaves, 65535));
  auto coarse_apply = [&](auto* rin, bool add) {
    coarse_restrict<<<ga, kRestrictBlock, 0, s>>>(C.n_agg, C.agg_off, C.agg_nodes, C.frame, xs,
                                                  ys, zs, rin, C.rc, L.state);
    CF_LAUNCHED();
    coarse_gemv<<<gc, kBlock, 0, s>>>(C.nc, C.A32, C.rc, C.zc, L.state);
    CF_LAUNCHED();
    if (add)
      coarse_prolong<true, T><<<G, kBlock, 0, s>>>(n_nodes, C.agg, C.frame, xs, ys, zs, fixed,
                                                   C.zc, u, L.state);
    else
      coarse_prolong<false, T><<<G, kBlock, 0, s>>>(n_nodes, C.agg, C.frame, xs, ys, zs, fixed,
                                                    C.zc, u, L.state);
    CF_LAUNCHED();
    return CF_OK;
  };
  T* D[2] = {lp<T>(L.ph), lp<T>(L.Aph)};
  auto pre2l = [&](int apply_step, int slot, int sample, const T* cd, const T* cw, T inv_theta,
                   int steps) {
    if (apply_step)
      fcg_update2l<true, T><<<GV, kBlock, 0, s>>>(n, x, L.r, L.p, L.Ap, rl, L.state, L.partials,
                                                  outer_parts, L.partials_b, slot);
    else
      fcg_update2l<false, T><<<GV, kBlock, 0, s>>>(n, x, L.r, L.p, L.Ap, rl, L.state, nullptr, 0,
                                                   L.partials_b, slot);
    CF_LAUNCHED();
    int r2;
    if ((r2 = coarse_apply(L.r, false))) return r2;                       // u = C r
    if ((r2 = product(u, D[0], 3, T(0), inv_theta, L.state))) return r2;  // Cheb start
    for (int k = 1; k < steps; ++k) {
      int tok = -1;
      if (k == 1) prof_begin(1, sample, s, &tok);
      if (k > 1 && k == steps - 1)  // the last step adds its iterate to u = z1 (mode 5)
        r2 = product(D[(k - 1) & 1], D[k & 1], 5, cd[k], cw[k], L.state, u);
      else
        r2 = cheb(k, cd[k], cw[k], D[(k - 1) & 1], D[k & 1]);
      if (r2) return r2;
      if (k == 1) prof_end(tok, s);
    }
    if (steps == 2) {
      vec_add<T><<<G, kBlock, 0, s>>>(n, u, D[1], L.state);  // u = z1 + Cheb
      CF_LAUNCHED();
    }
    if ((r2 = product(u, r2v, 4, T(0), T(0), L.state))) return r2;  // r2 = r - K u
    return coarse_apply(r2v, true);                                 // u += C r2
  };
  rc = fcg_cheb_core<true, T>(outer, outer_parts, power, cheb, 0.0, kCheb2L, n, b, x, fixed,
                              tol, maxit, L, stats_host, s, pre2l, u, 2, power_est,
                              switch_outer, abs_norm);
  return rc ? rc : switch_rc;
}

}  // namespace

extern "C" int cf_mf_t10_mixed_workspace_bytes(int64_t n_nodes, int64_t n_elem, size_t* bytes) {
  CF_REQUIRE(n_nodes > 0 && n_elem > 0 && bytes, "cf_mf_t10_mixed_workspace_bytes: bad arguments");
  Layout L{};
  MfMixed M{};
  *bytes = mf_mixed_layout(n_nodes, n_elem, nullptr, &L, &M, nullptr);
  return CF_OK;
}

extern "C" int cf_pcg_solve_mf_t10_mixed(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                                         const double* geo, double lam, double mu,
                                         const int32_t* offs, const uint32_t* perm,
                                         const double* b, double* x, const uint8_t* fixed,
                                         double tol, int32_t maxit, void* ws, size_t ws_bytes,
                                         cf_pcg_stats* stats_host, void* stream) {
  return mf_mixed_solve<float>("cf_pcg_solve_mf_t10_mixed", nullptr, n_nodes, n_elem, conn, geo,
                               lam, mu, offs, perm, nullptr, nullptr, nullptr, false, b, x, fixed,
                               tol, maxit, ws, ws_bytes, stats_host, stream);
}

extern "C" int cf_mf_t10_twolevel_workspace_bytes(int64_t n_nodes, int64_t n_elem,
                                                  size_t* bytes) {
  CF_REQUIRE(n_nodes > 0 && n_elem > 0 && bytes,
             "cf_mf_t10_twolevel_workspace_bytes: bad arguments");
  Layout L{};
  MfMixed M{};
  Coarse C{};
  *bytes = mf_mixed_layout(n_nodes, n_elem, nullptr, &L, &M, &C);
  return CF_OK;
}

extern "C" int cf_pcg_solve_mf_t10_twolevel(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                                            const double* geo, double lam, double mu,
                                            const int32_t* offs, const uint32_t* perm,
                                            const double* x_nodes, const double* y_nodes,
                                            const double* z_nodes, const double* b, double* x,
                                            const uint8_t* fixed, double tol, int32_t maxit,
                                            void* ws, size_t ws_bytes, cf_pcg_stats* stats_host,
                                            void* stream) {
  return mf_mixed_solve<float>("cf_pcg_solve_mf_t10_twolevel", nullptr, n_nodes, n_elem, conn,
                               geo, lam, mu, offs, perm, x_nodes, y_nodes, z_nodes, true, b, x,
                               fixed, tol, maxit, ws, ws_bytes, stats_host, stream);
}

extern "C" int cf_t10_twolevel_workspace_bytes(const cf_matrix* A, int64_t n_nodes,
                                               int64_t n_elem, int32_t precision,
                                               size_t* bytes) {
  CF_REQUIRE(n_nodes > 0 && n_elem > 0 && bytes, "cf_t10_twolevel_workspace_bytes: bad arguments");
  CF_REQUIRE(precision == CF_PREC_FP64 || precision == CF_PREC_MIXED,
             "cf_t10_twolevel_workspace_bytes: precision must be CF_PREC_FP64 or CF_PREC_MIXED");
  int64_t lsum = 0;
  int rc;
  if (A) {
    if ((rc = check_matrix(A))) return rc;
    CF_REQUIRE(A->format == CF_FMT_NB3 && A->n_rows == n_nodes,
               "cf_t10_twolevel_workspace_bytes: A must be the NB3 matrix of the mesh");
    if ((rc = sell_bound(A, &lsum))) return rc;
  }
  Layout L{};
  MfMixed M{};
  Coarse C{};
  *bytes = mf_mixed_layout(n_nodes, n_elem, nullptr, &L, &M, &C,
                           precision == CF_PREC_FP64 ? 8 : 4, A, lsum);
  return CF_OK;
}

extern "C" int cf_pcg_solve_t10_twolevel(const cf_matrix* A, int64_t n_nodes, int64_t n_elem,
                                         const int32_t* conn, const double* geo, double lam,
                                         double mu, const int32_t* offs, const uint32_t* perm,
                                         const double* x_nodes, const double* y_nodes,
                                         const double* z_nodes, const double* b, double* x,
                                         const uint8_t* fixed, int32_t precision, double tol,
                                         int32_t maxit, void* ws, size_t ws_bytes,
                                         cf_pcg_stats* stats_host, void* stream) {
  const char* who = "cf_pcg_solve_t10_twolevel";
  if (precision == CF_PREC_MIXED)
    return mf_mixed_solve<float>(who, A, n_nodes, n_elem, conn, geo, lam, mu, offs, perm,
                                 x_nodes, y_nodes, z_nodes, true, b, x, fixed, tol, maxit, ws,
                                 ws_bytes, stats_host, stream);
  if (precision == CF_PREC_FP64)
    return mf_mixed_solve<double>(who, A, n_nodes, n_elem, conn, geo, lam, mu, offs, perm,
                                  x_nodes, y_nodes, z_nodes, true, b, x, fixed, tol, maxit, ws,
                                  ws_bytes, stats_host, stream);
  cf::set_error("%s: precision must be CF_PREC_FP64 or CF_PREC_MIXED", who);
  return CF_ERR_ARG;
}

extern "C" int cf_invert_permutation(int64_t n, const uint32_t* perm, uint32_t* inv,
                                     void* stream) {
  CF_REQUIRE(n >= 0 && (n == 0 || (perm && inv)), "cf_invert_permutation: bad arguments");
  if (n == 0) return CF_OK;
  invert_permutation<<<cf::stride_grid(n, kBlock), kBlock, 0, cf::as_stream(stream)>>>(n, perm,
                                                                                         inv);
  CF_LAUNCHED();
  return CF_OK;
}

// ------------------------------------------------------------- no-alloc solve
// cf_pcg_solve_noalloc: cf_pcg_solve for scalar CSR systems in FP64 with
// Jacobi preconditioning, whose workspace (CG vectors, block partials, solver
// state) lives in a file-scope device array instead of a caller buffer.

namespace {

constexpr int64_t kNoAllocMaxDofs = int64_t(1) << 18;  // 262,144 unknowns
// four FP64 vectors of kNoAllocMaxDofs, partials for grids of up to 4096
// blocks, the state and the flag, each carved at 256-byte alignment
constexpr size_t kNoAllocBytes = 4 * (size_t(kNoAllocMaxDofs) * 8) + 4 * 4096 * 8 +
                                 2 * 4096 * 8 + 4 * 256;

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// per device: the device address of `scratch` (the host-side name is only a
// stand-in for it)
void* scratch_address() {
  static std::mutex mu;
  static void* addr[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    addr[dev] = p;
  }
  return addr[dev];
}

__global__ void zero_words(int64_t n, uint32_t* __restrict__ p) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = 0u;
}

}  // namespace

extern "C" int cf_pcg_solve_noalloc(const cf_matrix* A, const double* b, double* x,
                                    const uint8_t* fixed, double tol, int32_t maxit,
                                    cf_pcg_stats* stats_host, void* stream) {
  int rc = check_matrix(A);
  if (rc) return rc;
  CF_REQUIRE(b && x && stats_host, "cf_pcg_solve_noalloc: null argument");
  CF_REQUIRE(A->format == CF_FMT_CSR,
             "cf_pcg_solve_noalloc: scalar CSR matrices only (use cf_pcg_solve for NB3)");
  CF_REQUIRE(tol > 0.0 && maxit > 0, "cf_pcg_solve_noalloc: tol and maxit must be positive");
  if (int64_t(A->n_rows) > kNoAllocMaxDofs) {
    cf::set_error("cf_pcg_solve_noalloc: %d unknowns exceed the limit of %lld",
                  int(A->n_rows), (long long)kNoAllocMaxDofs);
    return CF_ERR_ARG;
  }
  void* base = scratch_address();
  if (!base) {
    cf::set_error("cf_pcg_solve_noalloc: hipGetSymbolAddress of the scratch array failed");
    return CF_ERR_HIP;
  }
  Layout L = layout(A, CF_PREC_FP64, CF_PC_JACOBI, base, 0);
  if (L.total > kNoAllocBytes) {
    cf::set_error("cf_pcg_solve_noalloc: needs %zu bytes of scratch, has %zu (grid of %d blocks)",
                  L.total, kNoAllocBytes, pcg_grid());
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  const int G = pcg_grid();
  *stats_host = cf_pcg_stats{};
  // the array keeps the previous call's contents: clear the state and flag
  zero_words<<<1, 64, 0, s>>>(int64_t(sizeof(PcgState) / 4), reinterpret_cast<uint32_t*>(L.state));
  CF_LAUNCHED();
  zero_words<<<1, 64, 0, s>>>(1, reinterpret_cast<uint32_t*>(L.bad));
  CF_LAUNCHED();
  diag_csr<<<G, kBlock, 0, s>>>(A->n_rows, A->indptr, A->indices, A->values, fixed, L.dinv,
                                L.bad);
  CF_LAUNCHED();
  int32_t bad = 0;
  CF_HIP(hipMemcpyAsync(&bad, L.bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  if (bad) {
    cf::set_error("cf_pcg_solve_noalloc: %d free diagonal entries are not positive "
                  "(matrix not SPD on the free DOFs)", bad);
    return CF_ERR_BREAKDOWN;
  }
  return pcg_fp64<CF_PC_JACOBI>(A, b, x, fixed, tol, maxit, L, stats_host, s);
}

// ------------------------------------------------------------- inner product alone
namespace {

template <class T>
__global__ void fill_value(int64_t n, T v, T* __restrict__ p) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

template <class T>
int mf_lp_apply(int64_t n_nodes, int64_t n_elem, const int32_t* conn, const double* geo,
                double lam, double mu, const int32_t* offs, const uint32_t* perm,
                const double* scale, const T* x, T* y, int32_t reps, float* avg_ms, void* ws,
                size_t ws_bytes, hipStream_t s) {
  MfArgs a{n_nodes, n_elem, conn, geo, lam, mu, offs, perm};
  int rc = mf_check(a);
  if (rc) return rc;
  CF_REQUIRE(x && y && ws, "cf_mf_t10_lp_apply: null argument");
  CF_REQUIRE(10 * n_elem < (int64_t(1) << 31), "cf_mf_t10_lp_apply: too many elements");
  Layout L{};
  MfMixed M{};
  const size_t need = mf_mixed_layout(n_nodes, n_elem, ws, &L, &M, nullptr, int(sizeof(T)));
  if (ws_bytes < need) {
    cf::set_error("cf_mf_t10_lp_apply: workspace %zu < %zu bytes", ws_bytes, need);
    return CF_ERR_WORKSPACE;
  }
  const int64_t n = 3 * n_nodes;
  T* sL = static_cast<T*>(M.sL);
  if (scale)
    to_type<T><<<pcg_grid(), kBlock, 0, s>>>(n, scale, sL);
  else
    fill_value<T><<<pcg_grid(), kBlock, 0, s>>>(n, T(1), sL);
  CF_LAUNCHED();
  if ((rc = mf_runs_build(n_nodes, n_elem, offs, perm, &M.R, s))) return rc;
  T* geoL = static_cast<T*>(M.geoL);
  mf_pack_tiles<T><<<cf::stride_grid(10 * n_elem, kBlock), kBlock, 0, s>>>(
      n_elem, conn, geo, M.R.rank, M.connT, geoL, M.rankT);
  CF_LAUNCHED();
  const int ge = elem_grid(n_elem);
  const int gg = int(mf_gather_grid(n_nodes));
  const bool use_mfma = inner_mfma();
  auto product = [&]() {
    launch_elem_lp<T, false>(use_mfma, ge, s, n_elem, M.connT, geoL, T(lam), T(mu), sL, x,
                             M.rankT, M.R.toff, M.R.tstart, static_cast<T*>(M.feL), nullptr);
    CF_LAUNCHED();
    if (M.R.use_run4)
      mf_gather_cheb<T, true><<<gg, kBlock, 0, s>>>(n_nodes, M.R.nro, M.R.nq,
                                                    static_cast<T*>(M.feL), sL, 0, nullptr,
                                                    nullptr, y, nullptr, T(0), T(0), nullptr,
                                                    nullptr, M.R.run4);
    else
      mf_gather_cheb<T><<<gg, kBlock, 0, s>>>(n_nodes, M.R.nro, M.R.nq, static_cast<T*>(M.feL),
                                              sL, 0, nullptr, nullptr, y, nullptr, T(0), T(0),
                                              nullptr, nullptr);
    CF_LAUNCHED();
    return CF_OK;
  };
  if ((rc = product())) return rc;
  if (reps > 0) {
    hipEvent_t e0, e1;
    CF_HIP(hipEventCreate(&e0));
    CF_HIP(hipEventCreate(&e1));
    CF_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps && rc == CF_OK; ++r) rc = product();
    CF_HIP(hipEventRecord(e1, s));
    CF_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    CF_HIP(hipEventElapsedTime(&ms, e0, e1));
    CF_HIP(hipEventDestroy(e0));
    CF_HIP(hipEventDestroy(e1));
    if (avg_ms) *avg_ms = ms / float(reps);
  }
  return rc;
}

}  // namespace

extern "C" int cf_mf_t10_lp_workspace_bytes(int64_t n_nodes, int64_t n_elem, int32_t precision,
                                            size_t* bytes) {
  CF_REQUIRE(n_nodes > 0 && n_elem > 0 && bytes, "cf_mf_t10_lp_workspace_bytes: bad arguments");
  CF_REQUIRE(precision == CF_PREC_MIXED || precision == CF_PREC_FP64,
             "cf_mf_t10_lp_workspace_bytes: precision must be CF_PREC_MIXED or CF_PREC_FP64");
  Layout L{};
  MfMixed M{};
  *bytes = mf_mixed_layout(n_nodes, n_elem, nullptr, &L, &M, nullptr,
                           precision == CF_PREC_FP64 ? 8 : 4);
  return CF_OK;
}

extern "C" int cf_mf_t10_lp_apply(int64_t n_nodes, int64_t n_elem, const int32_t* conn,
                                  const double* geo, double lam, double mu, const int32_t* offs,
                                  const uint32_t* perm, int32_t precision, const double* scale,
                                  const void* x, void* y, int32_t reps, float* avg_ms_host,
                                  void* ws, size_t ws_bytes, void* stream) {
  hipStream_t s = cf::as_stream(stream);
  if (precision == CF_PREC_MIXED)
    return mf_lp_apply<float>(n_nodes, n_elem, conn, geo, lam, mu, offs, perm, scale,
                              static_cast<const float*>(x), static_cast<float*>(y), reps,
                              avg_ms_host, ws, ws_bytes, s);
  if (precision == CF_PREC_FP64)
    return mf_lp_apply<double>(n_nodes, n_elem, conn, geo, lam, mu, offs, perm, scale,
                               static_cast<const double*>(x), static_cast<double*>(y), reps,
                               avg_ms_host, ws, ws_bytes, s);
  cf::set_error("cf_mf_t10_lp_apply: precision must be CF_PREC_MIXED or CF_PREC_FP64");
  return CF_ERR_ARG;
}
