// Batched theta-method heat solves: many independent scalar systems (one per
// mesh) stored as one block-diagonal CSR, stepped together.
//
// Replaces, for a batch of meshes, the per-mesh loop of tet_heat.py:234-274
// ((C + th h K) T+ = (C - (1-th) h K) T + h F with one LU) by warm-started
// CG per step with per-system scalars: every launch covers all systems,
// every CG scalar (alpha, beta, residual, stop test) is per system, and a
// system that has converged skips its tiles until the next step.
//
// Work is cut into tiles of consecutive rows that never straddle two
// systems; one workgroup owns one tile and writes one partial per dot
// product, and a fold kernel (one workgroup per system) sums the tiles of
// its system in a fixed order, so results are deterministic.  Each step's
// solve starts from the linear extrapolation 2 T_n - T_{n-1} of the last two
// states.  A whole time step (right-hand side + K CG iterations, K from the
// second step, the first that starts from an extrapolated state) is captured
// once as a hipGraph and replayed per step; the host reads the per-system
// flags after each replay and replays a shorter iteration graph for the rare
// system that needs more.
//
// Precision: FP64 CG, or FP32 CG on the Jacobi-scaled matrix S A S inside
// FP64 iterative refinement (residual r = b - A x and the solution in FP64;
// each sweep solves S A S d = S r to a relative 1e-5 in FP32 and adds S d).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cf_common.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// Per-system CG scalars; level 0 = FP64 (plain CG or refinement), level 1 =
// the FP32 inner CG of the mixed mode.
struct SysState {
  double rz[2], rr[2], stop2[2], alpha[2], beta[2];
  double bb;          // ||b||^2 of the step
  int32_t iters[2];   // iterations of the current step / sweep
  int32_t halt[2];    // 0 running, 1 converged, 2 iteration cap, 3 breakdown
  int32_t total;      // iterations over all steps (level 1 in mixed mode)
  int32_t worst;      // largest iteration count of one step
  int32_t failed;     // steps that ended without convergence
  int32_t sweeps;     // refinement sweeps over all steps (mixed)
  int32_t worst_inner;// largest FP32 iteration count of one sweep (mixed)
};

enum Halt : int32_t { kRun = 0, kDone = 1, kCap = 2, kBroken = 3 };

struct Tiles {
  const int32_t* sys;  // tile -> system
  const int32_t* r0;   // first row
  const int32_t* r1;   // one past the last row
  int32_t n;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sums of NV values; every thread gets the totals.
template <int NV>
__device__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][kWaves];
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
    for (int w = 0; w < kWaves; ++w) s += sh[k][w];
    v[k] = s;
  }
}

template <int NV>
__device__ void tile_partials(double (&v)[NV], double* __restrict__ partials, int32_t n_tiles) {
  block_sum<NV>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * int64_t(n_tiles) + blockIdx.x] = v[k];
}

// ------------------------------------------------------------------ kernels

enum SpmvMode : int { kPlain = 0, kDot = 1, kRhs = 2 };

// y = A x over the tile's rows, L lanes per row.  kDot: gated on the
// system's level-LV flag, writes the tile partial of x.y.  kRhs: y = A x + h F.
template <class VT, class XT, int L, int MODE, int LV>
__global__ void __launch_bounds__(kBlock)
bt_spmv(Tiles t, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
        const VT* __restrict__ val, const XT* __restrict__ x, XT* __restrict__ y,
        const double* __restrict__ F, double h, const SysState* __restrict__ st,
        double* __restrict__ partials) {
  const int32_t tile = blockIdx.x;
  const int32_t s = t.sys[tile];
  if (MODE == kDot && st[s].halt[LV] != kRun) return;
  const int lane = threadIdx.x % L;
  double dot = 0.0;
  for (int32_t r = t.r0[tile] + threadIdx.x / L; r < t.r1[tile]; r += kBlock / L) {
    const int32_t b = ptr[r], e = ptr[r + 1];
    double a = 0.0;
    for (int32_t k = b + lane; k < e; k += L) a += double(val[k]) * double(x[idx[k]]);
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) {
      if (MODE == kRhs) a += h * F[r];
      y[r] = XT(a);
      if (MODE == kDot) dot += double(x[r]) * double(XT(a));
    }
  }
  if (MODE != kDot) return;
  double v[1] = {dot};
  tile_partials<1>(v, partials, t.n);
}

// FP64 CG start: r = b - Ax, p = z = dinv r; partials r.r, r.z, b.b.
__global__ void __launch_bounds__(kBlock)
bt_start(Tiles t, const double* __restrict__ b, const double* __restrict__ Ax,
         const double* __restrict__ dinv, double* __restrict__ r, double* __restrict__ p,
         double* __restrict__ partials) {
  const int32_t tile = blockIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  for (int32_t i = t.r0[tile] + threadIdx.x; i < t.r1[tile]; i += kBlock) {
    const double ri = b[i] - Ax[i];
    const double zi = dinv[i] * ri;
    r[i] = ri;
    p[i] = zi;
    v[0] += ri * ri;
    v[1] += ri * zi;
    v[2] += b[i] * b[i];
  }
  tile_partials<3>(v, partials, t.n);
}

// x += alpha p, r -= alpha Ap, z = PRE ? dinv r : r; partials r.r, r.z
template <class T, bool PRE, int LV>
__global__ void __launch_bounds__(kBlock)
bt_update(Tiles t, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
          const T* __restrict__ Ap, const double* __restrict__ dinv,
          const SysState* __restrict__ st, double* __restrict__ partials) {
  const int32_t tile = blockIdx.x;
  const int32_t s = t.sys[tile];
  if (st[s].halt[LV] != kRun) return;
  const T alpha = T(st[s].alpha[LV]);
  double v[2] = {0.0, 0.0};
  for (int32_t i = t.r0[tile] + threadIdx.x; i < t.r1[tile]; i += kBlock) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    const double zi = PRE ? dinv[i] * double(ri) : double(ri);
    v[0] += double(ri) * double(ri);
    v[1] += double(ri) * zi;
  }
  tile_partials<2>(v, partials, t.n);
}

// p = z + beta p
template <class T, bool PRE, int LV>
__global__ void __launch_bounds__(kBlock)
bt_direction(Tiles t, const T* __restrict__ r, T* __restrict__ p,
             const double* __restrict__ dinv, const SysState* __restrict__ st) {
  const int32_t tile = blockIdx.x;
  const int32_t s = t.sys[tile];
  if (st[s].halt[LV] != kRun) return;
  const T beta = T(st[s].beta[LV]);
  for (int32_t i = t.r0[tile] + threadIdx.x; i < t.r1[tile]; i += kBlock) {
    const T zi = PRE ? T(dinv[i] * double(r[i])) : r[i];
    p[i] = zi + beta * p[i];
  }
}

// Mixed: r = b - Ax (FP64), partials r.r and b.b (gated on level 0)
__global__ void __launch_bounds__(kBlock)
bt_residual(Tiles t, const double* __restrict__ b, const double* __restrict__ Ax,
            double* __restrict__ r, const SysState* __restrict__ st, int gate,
            double* __restrict__ partials) {
  const int32_t tile = blockIdx.x;
  if (gate && st[t.sys[tile]].halt[0] != kRun) return;
  double v[2] = {0.0, 0.0};
  for (int32_t i = t.r0[tile] + threadIdx.x; i < t.r1[tile]; i += kBlock) {
    const double ri = b[i] - Ax[i];
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += b[i] * b[i];
  }
  tile_partials<2>(v, partials, t.n);
}

// Mixed inner start: rh = ph = fp32(S r), dh = 0; partial rh.rh
__global__ void __launch_bounds__(kBlock)
bt_inner_start(Tiles t, const double* __restrict__ r, const double* __restrict__ sc,
               float* __restrict__ rh, float* __restrict__ ph, float* __restrict__ dh,
               const SysState* __restrict__ st, double* __restrict__ partials) {
  const int32_t tile = blockIdx.x;
  if (st[t.sys[tile]].halt[0] != kRun) return;
  double v[1] = {0.0};
  for (int32_t i = t.r0[tile] + threadIdx.x; i < t.r1[tile]; i += kBlock) {
    const float ri = float(sc[i] * r[i]);
    rh[i] = ri;
    ph[i] = ri;
    dh[i] = 0.0f;
    v[0] += double(ri) * double(ri);
  }
  tile_partials<1>(v, partials, t.n);
}

// Mixed correction: x += S dh
__global__ void __launch_bounds__(kBlock)
bt_correct(Tiles t, double* __restrict__ x, const double* __restrict__ sc,
           const float* __restrict__ dh, SysState* __restrict__ st) {
  const int32_t tile = blockIdx.x;
  if (st[t.sys[tile]].halt[0] != kRun) return;
  for (int32_t i = t.r0[tile] + threadIdx.x; i < t.r1[tile]; i += kBlock)
    x[i] += sc[i] * double(dh[i]);
}

enum FoldMode : int {
  kFoldStart = 0,      // (r.r, r.z, b.b) -> level 0 CG start
  kFoldAlpha = 1,      // p.Ap -> alpha
  kFoldBeta = 2,       // (r.r, r.z) -> stop test, beta
  kFoldEnd = 3,        // step bookkeeping
  kFoldCheckFirst = 4, // mixed: (r.r, b.b) of the first sweep -> stop2, halt
  kFoldCheck = 5,      // mixed: r.r -> halt
  kFoldInnerStart = 6, // mixed: rh.rh -> level 1 start
};

// One workgroup per system folds its tiles' partials in tile order.
template <int MODE, int LV>
__global__ void __launch_bounds__(kBlock)
bt_fold(const int32_t* __restrict__ sys_t0, int32_t n_tiles, const double* __restrict__ partials,
        SysState* __restrict__ st, double tol2, double inner_tol2, int32_t maxit) {
  const int32_t s = blockIdx.x;
  SysState& S = st[s];
  if ((MODE == kFoldAlpha || MODE == kFoldBeta || MODE == kFoldCheck) && S.halt[LV] != kRun)
    return;
  if (MODE == kFoldInnerStart && S.halt[0] != kRun) {
    if (threadIdx.x == 0) S.halt[1] = kDone;
    return;
  }
  if (MODE == kFoldEnd) {
    if (threadIdx.x == 0) {
      S.worst = max(S.worst, S.iters[0]);
      if (S.halt[0] != kDone) S.failed += 1;
    }
    return;
  }
  constexpr int NV = MODE == kFoldStart ? 3 : (MODE == kFoldBeta || MODE == kFoldCheckFirst) ? 2 : 1;
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int32_t tl = sys_t0[s] + threadIdx.x; tl < sys_t0[s + 1]; tl += kBlock)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += partials[k * int64_t(n_tiles) + tl];
  block_sum<NV>(v);
  if (threadIdx.x != 0) return;
  switch (MODE) {
    case kFoldStart:
      S.bb = v[2];
      S.stop2[0] = tol2 * v[2];
      S.rr[0] = v[0];
      S.rz[0] = v[1];
      S.iters[0] = 0;
      S.halt[0] = !(isfinite(v[0]) && isfinite(v[2])) ? kBroken
                  : (v[0] <= S.stop2[0]) ? kDone : kRun;
      break;
    case kFoldAlpha:
      if (!(v[0] > 0.0) || !isfinite(v[0])) {
        S.halt[LV] = kBroken;
      } else {
        S.alpha[LV] = S.rz[LV] / v[0];
      }
      break;
    case kFoldBeta: {
      const double rr = v[0], rz = v[1];
      S.iters[LV] += 1;
      S.total += 1;
      S.rr[LV] = rr;
      if (!isfinite(rr) || !isfinite(rz)) {
        S.halt[LV] = kBroken;
      } else if (rr <= S.stop2[LV]) {
        S.halt[LV] = kDone;
      } else if (S.iters[LV] >= maxit) {
        S.halt[LV] = kCap;
      } else {
        S.beta[LV] = rz / S.rz[LV];
        S.rz[LV] = rz;
      }
      break;
    }
    case kFoldCheckFirst:
      S.bb = v[1];
      S.stop2[0] = tol2 * v[1];
      S.rr[0] = v[0];
      S.iters[0] = 0;
      S.halt[0] = !isfinite(v[0]) ? kBroken : (v[0] <= S.stop2[0]) ? kDone : kRun;
      break;
    case kFoldCheck:
      S.rr[0] = v[0];
      S.sweeps += 1;
      S.worst_inner = max(S.worst_inner, S.iters[1]);
      S.iters[0] += 1;
      if (!isfinite(v[0]))
        S.halt[0] = kBroken;
      else if (v[0] <= S.stop2[0])
        S.halt[0] = kDone;
      else if (S.iters[0] >= maxit)
        S.halt[0] = kCap;
      break;
    case kFoldInnerStart:
      S.rr[1] = v[0];
      S.rz[1] = v[0];
      S.stop2[1] = inner_tol2 * v[0];
      S.iters[1] = 0;
      S.halt[1] = (v[0] > 0.0 && isfinite(v[0])) ? kRun : kDone;
      break;
    default:
      break;
  }
}

__global__ void bt_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    y[i] = x[i];
}

// Start of a step's solve: the linear extrapolation 2 T_n - T_{n-1} of the
// last two states (T holds T_n, Tprev T_{n-1}; Tprev = T_0 before the first
// step, which then starts from T_0); Tprev becomes T_n.  The step's
// right-hand side R T_n + h F is formed before this.
__global__ void bt_extrapolate(int64_t n, double* __restrict__ T, double* __restrict__ Tprev) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double t = T[i];
    T[i] = 2.0 * t - Tprev[i];
    Tprev[i] = t;
  }
}

// dinv = 1 / A_ii (Jacobi), sc = 1 / sqrt(A_ii); bad counts non-positive diagonals
__global__ void bt_diag(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                        const double* __restrict__ val, double* __restrict__ dinv,
                        double* __restrict__ sc, int32_t* __restrict__ bad) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double d = 0.0;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k)
      if (idx[k] == i) d += val[k];
    if (d > 0.0) {
      dinv[i] = 1.0 / d;
      sc[i] = 1.0 / sqrt(d);
    } else {
      dinv[i] = 0.0;
      sc[i] = 0.0;
      atomicAdd(bad, 1);
    }
  }
}

// vals32 = fp32(S A S)
__global__ void bt_scale_f32(int64_t n, const int32_t* __restrict__ ptr,
                             const int32_t* __restrict__ idx, const double* __restrict__ val,
                             const double* __restrict__ sc, float* __restrict__ out) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) out[k] = float(sc[i] * val[k] * sc[idx[k]]);
}

__global__ void bt_init_state(int32_t n_sys, SysState* st) {
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n_sys; s += gridDim.x * blockDim.x)
    st[s] = SysState{};
}

// ------------------------------------------------------------------ host side

constexpr int kLaneChoices[3] = {4, 8, 16};
constexpr int kTileMult[3] = {1, 2, 4};  // rows per tile = kBlock / L * mult

struct BatchLayout {
  double *b, *r, *p, *Ap, *dinv, *sc, *Tprev;
  float *rh, *ph, *Aph, *dh, *vals32;
  double* partials;
  SysState* st;
  int32_t *tsys, *tr0, *tr1, *sys_t0, *bad;
  size_t total;
};

int64_t max_tiles(int32_t n_sys, const int32_t* rows_host, int32_t rows_per_tile) {
  int64_t nt = 0;
  for (int32_t s = 0; s < n_sys; ++s)
    nt += cf::ceil_div(int64_t(rows_host[s + 1]) - rows_host[s], rows_per_tile);
  return nt;
}

// smallest tile of the autotuner's choices (most tiles)
constexpr int kMinTileRows = kBlock / 16;

BatchLayout batch_layout(int32_t n_sys, const int32_t* rows_host, int64_t nnz, int precision,
                         void* base) {
  const int64_t n = rows_host[n_sys];
  const int64_t nt = max_tiles(n_sys, rows_host, kMinTileRows);
  cf::Carve c(base);
  BatchLayout L{};
  L.b = c.take<double>(n);
  L.r = c.take<double>(n);
  L.p = c.take<double>(n);
  L.Ap = c.take<double>(n);
  L.dinv = c.take<double>(n);
  L.sc = c.take<double>(n);
  L.Tprev = c.take<double>(n);
  if (precision != CF_PREC_FP64) {
    L.rh = c.take<float>(n);
    L.ph = c.take<float>(n);
    L.Aph = c.take<float>(n);
    L.dh = c.take<float>(n);
    L.vals32 = c.take<float>(nnz);
  }
  L.partials = c.take<double>(3 * nt);
  L.st = c.take<SysState>(n_sys);
  L.tsys = c.take<int32_t>(nt);
  L.tr0 = c.take<int32_t>(nt);
  L.tr1 = c.take<int32_t>(nt);
  L.sys_t0 = c.take<int32_t>(n_sys + 1);
  L.bad = c.take<int32_t>(1);
  L.total = c.used;
  return L;
}

// Build the tiles on the host and upload them (outside any graph capture).
int upload_tiles(int32_t n_sys, const int32_t* rows_host, int32_t rows_per_tile,
                 const BatchLayout& L, Tiles* t, hipStream_t s) {
  std::vector<int32_t> sys, r0, r1, t0(size_t(n_sys) + 1);
  for (int32_t k = 0; k < n_sys; ++k) {
    t0[k] = int32_t(sys.size());
    for (int32_t a = rows_host[k]; a < rows_host[k + 1]; a += rows_per_tile) {
      sys.push_back(k);
      r0.push_back(a);
      r1.push_back(std::min(a + rows_per_tile, rows_host[k + 1]));
    }
  }
  t0[n_sys] = int32_t(sys.size());
  const size_t nt = sys.size();
  if (nt) {
    CF_HIP(hipMemcpyAsync(L.tsys, sys.data(), nt * 4, hipMemcpyHostToDevice, s));
    CF_HIP(hipMemcpyAsync(L.tr0, r0.data(), nt * 4, hipMemcpyHostToDevice, s));
    CF_HIP(hipMemcpyAsync(L.tr1, r1.data(), nt * 4, hipMemcpyHostToDevice, s));
  }
  CF_HIP(hipMemcpyAsync(L.sys_t0, t0.data(), t0.size() * 4, hipMemcpyHostToDevice, s));
  // the host vectors die at return: finish the copies first
  CF_HIP(hipStreamSynchronize(s));
  *t = Tiles{L.tsys, L.tr0, L.tr1, int32_t(nt)};
  return CF_OK;
}

struct Problem {
  int32_t n_sys;
  const int32_t* ptr;
  const int32_t* idx;
  const double* A;
  const double* R;
  const double* F;
  double h;
  double* T;
  double tol2, inner_tol2;
  int32_t maxit;
};

template <class VT, class XT, int MODE, int LV>
int launch_spmv(int lanes, const Tiles& t, const Problem& P, const VT* val, const XT* x, XT* y,
                const BatchLayout& L, hipStream_t s) {
  if (t.n == 0) return CF_OK;
  switch (lanes) {
    case 4:
      bt_spmv<VT, XT, 4, MODE, LV><<<t.n, kBlock, 0, s>>>(t, P.ptr, P.idx, val, x, y, P.F, P.h,
                                                         L.st, L.partials);
      break;
    case 8:
      bt_spmv<VT, XT, 8, MODE, LV><<<t.n, kBlock, 0, s>>>(t, P.ptr, P.idx, val, x, y, P.F, P.h,
                                                         L.st, L.partials);
      break;
    default:
      bt_spmv<VT, XT, 16, MODE, LV><<<t.n, kBlock, 0, s>>>(t, P.ptr, P.idx, val, x, y, P.F, P.h,
                                                          L.st, L.partials);
      break;
  }
  CF_LAUNCHED();
  return CF_OK;
}

template <int MODE, int LV>
int fold(const Tiles& t, const Problem& P, const BatchLayout& L, hipStream_t s) {
  bt_fold<MODE, LV><<<P.n_sys, kBlock, 0, s>>>(L.sys_t0, t.n, L.partials, L.st, P.tol2,
                                               P.inner_tol2, P.maxit);
  CF_LAUNCHED();
  return CF_OK;
}

// ---- FP64 CG
int enqueue_cg_iterations(int k, int lanes, const Tiles& t, const Problem& P,
                          const BatchLayout& L, hipStream_t s) {
  int rc;
  for (int it = 0; it < k; ++it) {
    if ((rc = launch_spmv<double, double, kDot, 0>(lanes, t, P, P.A, L.p, L.Ap, L, s)) ||
        (rc = fold<kFoldAlpha, 0>(t, P, L, s)))
      return rc;
    bt_update<double, true, 0><<<t.n, kBlock, 0, s>>>(t, P.T, L.r, L.p, L.Ap, L.dinv, L.st,
                                                      L.partials);
    CF_LAUNCHED();
    if ((rc = fold<kFoldBeta, 0>(t, P, L, s))) return rc;
    bt_direction<double, true, 0><<<t.n, kBlock, 0, s>>>(t, L.r, L.p, L.dinv, L.st);
    CF_LAUNCHED();
  }
  return CF_OK;
}

// b = R T_n + h F, then the extrapolated start (bt_extrapolate)
int enqueue_rhs_and_start(int lanes, const Tiles& t, const Problem& P, const BatchLayout& L,
                          int64_t n, hipStream_t s) {
  int rc;
  if ((rc = launch_spmv<double, double, kRhs, 0>(lanes, t, P, P.R, P.T, L.b, L, s))) return rc;
  bt_extrapolate<<<cf::stride_grid(n, kBlock), kBlock, 0, s>>>(n, P.T, L.Tprev);
  CF_LAUNCHED();
  return CF_OK;
}

int enqueue_fp64_step(int k, int lanes, const Tiles& t, const Problem& P, const BatchLayout& L,
                      int64_t n, hipStream_t s) {
  int rc;
  if ((rc = enqueue_rhs_and_start(lanes, t, P, L, n, s)) ||
      (rc = launch_spmv<double, double, kPlain, 0>(lanes, t, P, P.A, P.T, L.Ap, L, s)))
    return rc;
  bt_start<<<t.n, kBlock, 0, s>>>(t, L.b, L.Ap, L.dinv, L.r, L.p, L.partials);
  CF_LAUNCHED();
  if ((rc = fold<kFoldStart, 0>(t, P, L, s))) return rc;
  return enqueue_cg_iterations(k, lanes, t, P, L, s);
}

// ---- mixed: FP64 refinement sweeps around FP32 inner CG
int enqueue_sweep(int k_in, int lanes, const Tiles& t, const Problem& P, const BatchLayout& L,
                  hipStream_t s) {
  int rc;
  bt_inner_start<<<t.n, kBlock, 0, s>>>(t, L.r, L.sc, L.rh, L.ph, L.dh, L.st, L.partials);
  CF_LAUNCHED();
  if ((rc = fold<kFoldInnerStart, 1>(t, P, L, s))) return rc;
  for (int it = 0; it < k_in; ++it) {
    if ((rc = launch_spmv<float, float, kDot, 1>(lanes, t, P, L.vals32, L.ph, L.Aph, L, s)) ||
        (rc = fold<kFoldAlpha, 1>(t, P, L, s)))
      return rc;
    bt_update<float, false, 1><<<t.n, kBlock, 0, s>>>(t, L.dh, L.rh, L.ph, L.Aph, nullptr, L.st,
                                                      L.partials);
    CF_LAUNCHED();
    if ((rc = fold<kFoldBeta, 1>(t, P, L, s))) return rc;
    bt_direction<float, false, 1><<<t.n, kBlock, 0, s>>>(t, L.rh, L.ph, nullptr, L.st);
    CF_LAUNCHED();
  }
  bt_correct<<<t.n, kBlock, 0, s>>>(t, P.T, L.sc, L.dh, L.st);
  CF_LAUNCHED();
  // new FP64 residual and stop test
  if ((rc = launch_spmv<double, double, kPlain, 0>(lanes, t, P, P.A, P.T, L.Ap, L, s))) return rc;
  bt_residual<<<t.n, kBlock, 0, s>>>(t, L.b, L.Ap, L.r, L.st, 1, L.partials);
  CF_LAUNCHED();
  return fold<kFoldCheck, 0>(t, P, L, s);
}

int enqueue_mixed_step(int sweeps, int k_in, int lanes, const Tiles& t, const Problem& P,
                       const BatchLayout& L, int64_t n, hipStream_t s) {
  int rc;
  if ((rc = enqueue_rhs_and_start(lanes, t, P, L, n, s)) ||
      (rc = launch_spmv<double, double, kPlain, 0>(lanes, t, P, P.A, P.T, L.Ap, L, s)))
    return rc;
  bt_residual<<<t.n, kBlock, 0, s>>>(t, L.b, L.Ap, L.r, L.st, 0, L.partials);
  CF_LAUNCHED();
  if ((rc = fold<kFoldCheckFirst, 0>(t, P, L, s))) return rc;
  for (int k = 0; k < sweeps; ++k)
    if ((rc = enqueue_sweep(k_in, lanes, t, P, L, s))) return rc;
  return CF_OK;
}

int post_step(double* hist_row, int64_t n, const Tiles& t, const Problem& P, const BatchLayout& L,
              hipStream_t s) {
  int rc;
  if ((rc = fold<kFoldEnd, 0>(t, P, L, s))) return rc;
  if (hist_row) {
    bt_copy<<<cf::stride_grid(n, kBlock), kBlock, 0, s>>>(n, P.T, hist_row);
    CF_LAUNCHED();
  }
  return CF_OK;
}

// Graph of a stream-ordered sequence, captured on a private stream.
struct Graph {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  ~Graph() {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
  }
};

template <class F>
int capture(hipStream_t cs, Graph* out, F&& enqueue) {
  CF_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue(cs);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(cs, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    cf::set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
    return CF_ERR_HIP;
  }
  out->g = g;
  CF_HIP(hipGraphInstantiate(&out->e, g, nullptr, nullptr, 0));
  return CF_OK;
}

int read_flags(const BatchLayout& L, int32_t n_sys, std::vector<SysState>* h, hipStream_t s) {
  h->resize(n_sys);
  CF_HIP(hipMemcpyAsync(h->data(), L.st, sizeof(SysState) * n_sys, hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  return CF_OK;
}

int check_args(int32_t n_sys, const int32_t* rows_host, int64_t nnz) {
  CF_REQUIRE(n_sys > 0 && rows_host, "batch: need at least one system");
  CF_REQUIRE(rows_host[0] == 0, "batch: row offsets must start at 0");
  for (int32_t s = 0; s < n_sys; ++s)
    CF_REQUIRE(rows_host[s + 1] > rows_host[s], "batch: system %d has no rows", s);
  CF_REQUIRE(nnz > 0, "batch: empty matrix");
  return CF_OK;
}

}  // namespace

extern "C" int cf_batch_workspace_bytes(int32_t n_sys, const int32_t* sys_rows, int64_t nnz,
                                        int32_t precision, size_t* bytes) {
  int rc = check_args(n_sys, sys_rows, nnz);
  if (rc) return rc;
  CF_REQUIRE(bytes, "cf_batch_workspace_bytes: null output");
  *bytes = batch_layout(n_sys, sys_rows, nnz, precision, nullptr).total;
  return CF_OK;
}

extern "C" int cf_batch_autotune(int32_t n_sys, const int32_t* sys_rows, const int32_t* indptr,
                                 const int32_t* indices, const double* values, int64_t nnz,
                                 int32_t reps, int32_t* lanes_out, int32_t* rows_out,
                                 float* us_out, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_args(n_sys, sys_rows, nnz);
  if (rc) return rc;
  CF_REQUIRE(indptr && indices && values && ws && lanes_out && rows_out,
             "cf_batch_autotune: null argument");
  CF_REQUIRE(reps > 0, "cf_batch_autotune: reps must be positive");
  BatchLayout L = batch_layout(n_sys, sys_rows, nnz, CF_PREC_FP64, ws);
  if (ws_bytes < L.total) {
    cf::set_error("cf_batch_autotune: workspace %zu < %zu bytes", ws_bytes, L.total);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  const int64_t n = sys_rows[n_sys];
  bt_copy<<<cf::stride_grid(n, kBlock), kBlock, 0, s>>>(n, values, L.p);  // any x: reuse a buffer
  CF_LAUNCHED();
  Problem P{n_sys, indptr, indices, values, values, nullptr, 0.0, L.p, 0.0, 0.0, 1};
  hipEvent_t e0, e1;
  CF_HIP(hipEventCreate(&e0));
  CF_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int k = 0;
  for (int li = 0; li < 3; ++li)
    for (int mi = 0; mi < 3; ++mi, ++k) {
      const int lanes = kLaneChoices[li];
      const int rows = kBlock / lanes * kTileMult[mi];
      Tiles t;
      if ((rc = upload_tiles(n_sys, sys_rows, rows, L, &t, s))) return rc;
      // warm, then time `reps` products
      if ((rc = launch_spmv<double, double, kPlain, 0>(lanes, t, P, values, L.p, L.Ap, L, s)))
        return rc;
      CF_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r)
        if ((rc = launch_spmv<double, double, kPlain, 0>(lanes, t, P, values, L.p, L.Ap, L, s)))
          return rc;
      CF_HIP(hipEventRecord(e1, s));
      CF_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      CF_HIP(hipEventElapsedTime(&ms, e0, e1));
      const float us = 1e3f * ms / float(reps);
      if (us_out) us_out[k] = us;
      if (us < best) {
        best = us;
        *lanes_out = lanes;
        *rows_out = rows;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return CF_OK;
}

extern "C" int cf_batch_theta_run(int32_t n_sys, const int32_t* sys_rows, const int32_t* indptr,
                                  const int32_t* indices, int64_t nnz, const double* A_vals,
                                  const double* R_vals, const double* F, double h, double* T,
                                  double* hist, int32_t n_steps, int32_t precision, double tol,
                                  int32_t maxit, int32_t lanes, int32_t rows_per_tile,
                                  void* ws, size_t ws_bytes, cf_batch_stats* stats,
                                  void* stream) {
  int rc = check_args(n_sys, sys_rows, nnz);
  if (rc) return rc;
  CF_REQUIRE(indptr && indices && A_vals && R_vals && F && T && ws && stats,
             "cf_batch_theta_run: null argument");
  CF_REQUIRE(n_steps >= 0 && tol > 0.0 && maxit > 0, "cf_batch_theta_run: bad steps/tol/maxit");
  CF_REQUIRE(precision == CF_PREC_FP64 || precision == CF_PREC_MIXED,
             "cf_batch_theta_run: bad precision %d", precision);
  CF_REQUIRE(lanes == 4 || lanes == 8 || lanes == 16, "cf_batch_theta_run: lanes must be 4/8/16");
  CF_REQUIRE(rows_per_tile >= kMinTileRows && rows_per_tile % (kBlock / lanes) == 0,
             "cf_batch_theta_run: rows_per_tile must be a multiple of %d and >= %d",
             kBlock / lanes, kMinTileRows);
  BatchLayout L = batch_layout(n_sys, sys_rows, nnz, precision, ws);
  if (ws_bytes < L.total) {
    cf::set_error("cf_batch_theta_run: workspace %zu < %zu bytes", ws_bytes, L.total);
    return CF_ERR_WORKSPACE;
  }
  hipStream_t s = cf::as_stream(stream);
  const int64_t n = sys_rows[n_sys];
  *stats = cf_batch_stats{};
  stats->lanes = lanes;
  stats->rows_per_tile = rows_per_tile;
  Tiles t;
  if ((rc = upload_tiles(n_sys, sys_rows, rows_per_tile, L, &t, s))) return rc;
  CF_HIP(hipMemsetAsync(L.bad, 0, sizeof(int32_t), s));
  bt_diag<<<cf::stride_grid(n, kBlock), kBlock, 0, s>>>(n, indptr, indices, A_vals, L.dinv, L.sc,
                                                        L.bad);
  CF_LAUNCHED();
  if (precision == CF_PREC_MIXED) {
    bt_scale_f32<<<cf::stride_grid(n, kBlock), kBlock, 0, s>>>(n, indptr, indices, A_vals, L.sc,
                                                               L.vals32);
    CF_LAUNCHED();
  }
  bt_init_state<<<cf::stride_grid(n_sys, kBlock), kBlock, 0, s>>>(n_sys, L.st);
  CF_LAUNCHED();
  int32_t bad = 0;
  CF_HIP(hipMemcpyAsync(&bad, L.bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CF_HIP(hipStreamSynchronize(s));
  if (bad) {
    cf::set_error("cf_batch_theta_run: %d non-positive diagonal entries", bad);
    return CF_ERR_BREAKDOWN;
  }
  if (hist) {
    bt_copy<<<cf::stride_grid(n, kBlock), kBlock, 0, s>>>(n, T, hist);
    CF_LAUNCHED();
  }
  stats->n_steps = n_steps;
  if (n_steps == 0) {
    CF_HIP(hipStreamSynchronize(s));
    return CF_OK;
  }
  const double inner_tol = 1e-5;
  // mixed: maxit caps the refinement sweeps of a step (at most 200)
  const int32_t cap = precision == CF_PREC_MIXED ? std::min<int32_t>(maxit, 200) : maxit;
  Problem P{n_sys, indptr, indices, A_vals, R_vals, F, h, T, tol * tol, inner_tol * inner_tol,
            cap};
  // private capture stream, ordered after / before the caller's stream
  hipStream_t cs;
  CF_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{cs};
  hipEvent_t ev;
  CF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  struct EventGuard {
    hipEvent_t e;
    ~EventGuard() { (void)hipEventDestroy(e); }
  } eguard{ev};
  CF_HIP(hipEventRecord(ev, s));
  CF_HIP(hipStreamWaitEvent(cs, ev, 0));

  std::vector<SysState> hs;
  const bool mixed = precision == CF_PREC_MIXED;
  // iteration graphs for the eager steps and for systems the step graph
  // leaves unconverged
  const int kProbe = mixed ? 4 : 8;
  const int kInner = 24;   // FP32 iterations per sweep (inner stop at 1e-5 usually first)
  Graph probe;
  if (mixed)
    rc = capture(cs, &probe, [&](hipStream_t q) { return enqueue_sweep(kInner, lanes, t, P, L, q); });
  else
    rc = capture(cs, &probe,
                 [&](hipStream_t q) { return enqueue_cg_iterations(kProbe, lanes, t, P, L, q); });
  if (rc) return rc;
  // T_{-1} = T_0: the first step starts from T_0
  bt_copy<<<cf::stride_grid(n, kBlock), kBlock, 0, cs>>>(n, T, L.Tprev);
  CF_LAUNCHED();
  // The first two steps run eagerly, with small iteration graphs until every
  // system has converged; the step graph is sized from the second, the first
  // that starts from an extrapolated state (the first step starts from T_0
  // and usually needs more iterations than the later ones).
  const int n_eager = n_steps < 2 ? n_steps : 2;
  int worst = 0, worst_inner = 0;
  for (int32_t k = 0; k < n_eager; ++k) {
    if (mixed)
      rc = enqueue_mixed_step(0, kInner, lanes, t, P, L, n, cs);
    else
      rc = enqueue_fp64_step(0, lanes, t, P, L, n, cs);
    if (rc) return rc;
    for (;;) {
      if ((rc = read_flags(L, n_sys, &hs, cs))) return rc;
      bool running = false;
      for (const auto& x : hs) running |= (x.halt[0] == kRun);
      if (!running) break;
      CF_HIP(hipGraphLaunch(probe.e, cs));
      ++stats->graph_replays;
    }
    worst = worst_inner = 0;
    for (const auto& x : hs) {
      worst = std::max(worst, x.iters[0]);
      worst_inner = std::max(worst_inner, x.worst_inner);
    }
    if ((rc = post_step(hist ? hist + (k + 1) * n : nullptr, n, t, P, L, cs))) return rc;
  }
  // FP32 iterations per sweep in the step graph: the worst so far + 2
  const int k_in = std::min(kInner, worst_inner + 2);
  // step graph: rhs + enough iterations for the second step's worst system
  const int k_step = mixed ? std::max(1, worst) : worst + 2;
  Graph step;
  if (n_steps > n_eager) {
    if (mixed)
      rc = capture(cs, &step, [&](hipStream_t q) {
        return enqueue_mixed_step(k_step, k_in, lanes, t, P, L, n, q);
      });
    else
      rc = capture(cs, &step,
                   [&](hipStream_t q) { return enqueue_fp64_step(k_step, lanes, t, P, L, n, q); });
    if (rc) return rc;
  }
  stats->step_iterations = k_step;
  for (int32_t k = n_eager; k < n_steps; ++k) {
    CF_HIP(hipGraphLaunch(step.e, cs));
    ++stats->graph_replays;
    for (;;) {
      if ((rc = read_flags(L, n_sys, &hs, cs))) return rc;
      bool running = false;
      for (const auto& x : hs) running |= (x.halt[0] == kRun);
      if (!running) break;
      CF_HIP(hipGraphLaunch(probe.e, cs));
      ++stats->graph_replays;
      ++stats->extra_replays;
    }
    if ((rc = post_step(hist ? hist + (k + 1) * n : nullptr, n, t, P, L, cs))) return rc;
  }
  if ((rc = read_flags(L, n_sys, &hs, cs))) return rc;
  CF_HIP(hipEventRecord(ev, cs));
  CF_HIP(hipStreamWaitEvent(s, ev, 0));
  for (const auto& x : hs) {
    stats->iterations += x.total;
    stats->sweeps += x.sweeps;
    stats->worst_step = std::max(stats->worst_step, x.worst);
    stats->failed += x.failed;
    const double rel = x.bb > 0.0 ? std::sqrt(x.rr[0] / x.bb) : 0.0;
    stats->max_rel_residual = std::max(stats->max_rel_residual, rel);
  }
  if (stats->failed) {
    cf::set_error("cf_batch_theta_run: %d system-steps did not converge (tol %.1e, maxit %d)",
                  stats->failed, tol, maxit);
    return CF_ERR_NOT_CONVERGED;
  }
  return CF_OK;
}
