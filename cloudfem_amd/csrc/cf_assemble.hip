// Element geometry, matrix values, right-hand sides and face terms.
//
// Design: the reference builds a dense local matrix per element and lets
// scipy merge the triplets (e.g. elastic_tet10.py:164-173, 208-210).  Here
// nothing per element is materialised except the 10-double geometry record
// (inverse Jacobian + volume).  Each stored node pair of the CSR pattern is
// owned by one thread, which adds its contributors (the element node pairs
// that map to it, ascending element order); the contributors themselves are
// computed one per thread, in rounds, from the geometry record and constant
// reference tables.  For affine
// elements every local stiffness is a contraction of the element's inverse
// Jacobian with a reference tensor R_ab = sum_q w_q dN_a/dxi (dN_b/dxi)^T:
//   Tet10 elasticity   K_ab = V (lam M + mu M^T + mu tr(M) I),  M = Jinv^T R_ab Jinv
//   scalar diffusion   K_ab = V sum_q w_q k_q g_aq . g_bq
// so the kernel's HBM traffic is the CSR value write plus the contributor
// list, and the sums are deterministic (no atomics).
#include "cf_common.hpp"
#include "cf_tables.hpp"

namespace {

__constant__ cf::Tet10Tables kT10 = cf::build_tet10_tables();
__constant__ cf::FaceTables kFace = cf::build_face_tables();

constexpr int kBlock = 256;

// ------------------------------------------------------------- geometry

__global__ void tet_geometry(const double* __restrict__ x,
                             const double* __restrict__ y,
                             const double* __restrict__ z,
                             const int32_t* __restrict__ conn, int64_t m,
                             int k, double* __restrict__ geo,
                             int32_t* __restrict__ n_bad) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < m;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int32_t* c = conn + e * k;
    int n0 = c[0], n1 = c[1], n2 = c[2], n3 = c[3];
    double x0 = x[n0], y0 = y[n0], z0 = z[n0];
    // J[i][m] = (p_{m+1} - p_0)_i  (tet_poisson.py:23, tet10.py:115)
    double J[3][3] = {{x[n1] - x0, x[n2] - x0, x[n3] - x0},
                      {y[n1] - y0, y[n2] - y0, y[n3] - y0},
                      {z[n1] - z0, z[n2] - z0, z[n3] - z0}};
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double* g = geo + e * 10;
    if (det == 0.0 || !isfinite(det)) {
      atomicAdd(n_bad, 1);
      for (int t = 0; t < 10; ++t) g[t] = 0.0;
      continue;
    }
    double inv = 1.0 / det;
    // Jinv[m][i] = adj(J)[m][i] / det
    g[0] = c00 * inv;
    g[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * inv;
    g[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * inv;
    g[3] = c01 * inv;
    g[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * inv;
    g[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * inv;
    g[6] = c02 * inv;
    g[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * inv;
    g[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * inv;
    g[9] = fabs(det) / 6.0;
  }
}

struct Geo {
  double j[3][3];  // Jinv[m][i]
  double v;
};

__device__ __forceinline__ Geo load_geo(const double* __restrict__ geo, int64_t e) {
  Geo g;
  const double* p = geo + e * 10;
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int i = 0; i < 3; ++i) g.j[m][i] = p[m * 3 + i];
  g.v = p[9];
  return g;
}

// Physical gradient of P1 function a: g_a,i = sum_m ghat_a,m Jinv[m][i],
// ghat_0 = (-1,-1,-1), ghat_{1,2,3} = unit vectors (tet_poisson.py:27-31).
__device__ __forceinline__ void p1_grad(const Geo& g, int a, double out[3]) {
  if (a == 0) {
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = -(g.j[0][i] + g.j[1][i] + g.j[2][i]);
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = g.j[a - 1][i];
  }
}

// Physical gradient of Tet10 function a at Keast point q.
__device__ __forceinline__ void t10_grad(const Geo& g, int q, int a, double out[3]) {
  const double d0 = kT10.D[q][a][0], d1 = kT10.D[q][a][1], d2 = kT10.D[q][a][2];
#pragma unroll
  for (int i = 0; i < 3; ++i) out[i] = d0 * g.j[0][i] + d1 * g.j[1][i] + d2 * g.j[2][i];
}

// ------------------------------------------------------------- matrices

// Nodes per element of operator OP (a compile-time divisor for the
// contributor decode).
template <int OP>
constexpr int op_nodes() {
  return (OP == CF_OP_P1_STIFF || OP == CF_OP_P1_MASS || OP == CF_OP_P1_ELAST ||
          OP == CF_OP_P1_VMASS) ? 4 : 10;
}

// The part of element e's geometry record operator OP reads (mass
// operators: the volume only).
template <int OP>
__device__ __forceinline__ Geo load_geo_for(const double* __restrict__ geo, int64_t e) {
  if constexpr (OP == CF_OP_P1_MASS || OP == CF_OP_P1_VMASS || OP == CF_OP_T10_MASS) {
    Geo g{};
    g.v = geo[e * 10 + 9];
    return g;
  } else {
    return load_geo(geo, e);
  }
}

// Value of one contributor pid = e k^2 + a k + b (element e, local pair
// (a, b)) to its stored entry: a 3x3 block for the vector operators, v[0]
// for the scalar ones.  g: the geometry record of e, already loaded.
template <int OP>
__device__ __forceinline__ void contributor_value(uint32_t pid, int k, const Geo& g,
                                                  const double* __restrict__ geo,
                                                  const double* __restrict__ coef, double lam,
                                                  double mu, const double* __restrict__ Rs,
                                                  double (&v)[9]) {
  constexpr bool kVec9 = (OP == CF_OP_T10_ELAST || OP == CF_OP_P1_ELAST);
  constexpr int K = op_nodes<OP>();
  constexpr uint32_t kk = uint32_t(K * K);
  (void)k;
  const int64_t e = pid / kk;
  const int ab = int(pid - uint32_t(e) * kk);
  const int a = ab / K, b = ab - (ab / K) * K;
  (void)geo;
  if constexpr (OP == CF_OP_P1_MASS || OP == CF_OP_P1_VMASS) {
    // rho V (1 + d_ab)/20  (tet_heat.py:72-94, modal3d.py:64-66)
    v[0] = (coef[e] * g.v) * ((a == b ? 2.0 : 1.0) / 20.0);
  } else if constexpr (OP == CF_OP_T10_MASS) {
    v[0] = (coef[e] * g.v) * kT10.M[a][b];
  } else if constexpr (OP == CF_OP_P1_STIFF) {
    double ga[3], gb[3];
    p1_grad(g, a, ga);
    p1_grad(g, b, gb);
    v[0] = (coef[e] * g.v) * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
  } else if constexpr (OP == CF_OP_T10_STIFF) {
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      double ga[3], gb[3];
      t10_grad(g, q, a, ga);
      t10_grad(g, q, b, gb);
      sum += (kT10.w[q] * coef[e * 5 + q]) * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
    }
    v[0] = g.v * sum;
  } else if constexpr (kVec9) {
    double M[3][3];
    if constexpr (OP == CF_OP_T10_ELAST) {
      // M = Jinv^T R_ab Jinv
      const double* R = Rs + (a * 10 + b) * 9;
      double T[3][3];
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i)
          T[m][i] = R[3 * m] * g.j[0][i] + R[3 * m + 1] * g.j[1][i] + R[3 * m + 2] * g.j[2][i];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
          M[i][jj] = g.j[0][i] * T[0][jj] + g.j[1][i] * T[1][jj] + g.j[2][i] * T[2][jj];
    } else {
      double ga[3], gb[3];
      p1_grad(g, a, ga);
      p1_grad(g, b, gb);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) M[i][jj] = ga[i] * gb[jj];
    }
    const double lv = lam * g.v, mv = mu * g.v;
    const double tr = mv * (M[0][0] + M[1][1] + M[2][2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
        v[3 * i + jj] = lv * M[i][jj] + mv * M[jj][i] + (i == jj ? tr : 0.0);
  }
}

// A block owns kBlock consecutive stored entries (node pairs) s0 .. s0 + 255
// and their contributors contrib[seg[s0] .. seg[s0 + 256]) (element node
// pairs mapping to them, ascending element order within an entry).  The
// contributors are computed in rounds of kBlock, one per thread, so every
// lane works whatever the entries' contributor counts (1 to ~24 for Tet10):
// a round stages its values in LDS, then each entry thread adds the staged
// contributors of its own entry, in order.  The sum of an entry is therefore
// its contributors' sum in ascending element order (scipy's coo->csr sums
// the reference's triplets in that order too), bitwise reproducible, with no
// atomics.
//
// Vector operators store a 3x3 block per entry at its place in the DOF-level
// CSR value array: row node i = conn[e][a] of the first contributor (no
// search over ptr), DOF row 3i+c holding block(i, j)[c][0..2] at 3 (s - ptr[i]).
// The block's 3x3 results are staged in LDS and written row segment by row
// segment, consecutive lanes on consecutive addresses (direct per-thread
// stores touch each 128-byte line in three 8-byte pieces).
template <int OP>
__global__ void __launch_bounds__(kBlock)
assemble_values(const int32_t* __restrict__ ptr, const int32_t* __restrict__ seg,
                const uint32_t* __restrict__ contrib, const int32_t* __restrict__ conn,
                int64_t nnz, int k, const double* __restrict__ geo,
                const double* __restrict__ coef, double lam, double mu,
                double* __restrict__ out) {
  constexpr bool kRT = (OP == CF_OP_T10_ELAST);
  constexpr bool kVec9 = (OP == CF_OP_T10_ELAST || OP == CF_OP_P1_ELAST);
  constexpr int NV = kVec9 ? 9 : 1;
  // staged contributor values, contributor-major, padded to 10 doubles (80 B)
  // so that one contributor is five aligned 16-byte LDS reads
  constexpr int NS = kVec9 ? 10 : 1;
  __shared__ double Rs[kRT ? 900 : 1];
  __shared__ __align__(16) double stage[kBlock * NS];
  __shared__ int32_t segl[kBlock + 1];
  __shared__ int64_t rbase[kVec9 ? kBlock : 1];
  __shared__ int32_t rlen[kVec9 ? kBlock : 1];
  if constexpr (kRT) {
    const double* R = &kT10.R[0][0][0][0];
    for (int t = threadIdx.x; t < 900; t += blockDim.x) Rs[t] = R[t];
  }
  for (int64_t s0 = int64_t(blockIdx.x) * kBlock; s0 < nnz; s0 += int64_t(gridDim.x) * kBlock) {
    const int64_t m = nnz - s0 < kBlock ? nnz - s0 : kBlock;  // entries of this chunk
    const int64_t s = s0 + threadIdx.x;
    const bool own = threadIdx.x < m;
    segl[threadIdx.x] = seg[own ? s : s0 + m];
    if (threadIdx.x == 0) segl[kBlock] = seg[s0 + m];
    __syncthreads();
    const int32_t c0 = segl[0], c1 = segl[kBlock];
    const int32_t jb = segl[threadIdx.x], je = own ? segl[threadIdx.x + 1] : jb;
    // row of the entry from its first contributor
    int64_t row = 0;
    if constexpr (kVec9 || OP == CF_OP_P1_VMASS) {
      if (own) {
        const uint32_t pid = contrib[jb];
        constexpr int K = op_nodes<OP>();
        const uint32_t e = pid / uint32_t(K * K);
        const int a = int(pid - e * uint32_t(K * K)) / K;
        row = conn[int64_t(e) * K + a];
      }
    }
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int32_t base = c0; base < c1; base += kBlock) {
      const int32_t j = base + int32_t(threadIdx.x);
      if (j < c1) {
        const uint32_t pid = contrib[j];
        const Geo g = load_geo_for<OP>(geo, int64_t(pid / uint32_t(op_nodes<OP>() * op_nodes<OP>())));
        double v[9];
        contributor_value<OP>(pid, k, g, geo, coef, lam, mu, Rs, v);
        if constexpr (kVec9) {
          double2* o = reinterpret_cast<double2*>(stage + threadIdx.x * NS);
#pragma unroll
          for (int q = 0; q < 4; ++q) o[q] = make_double2(v[2 * q], v[2 * q + 1]);
          o[4] = make_double2(v[8], 0.0);
        } else {
          stage[threadIdx.x] = v[0];
        }
      }
      __syncthreads();
      const int32_t lo = jb > base ? jb : base;
      const int32_t hi = je < base + kBlock ? je : base + kBlock;
      for (int32_t jj = lo; jj < hi; ++jj) {
        if constexpr (kVec9) {
          const double2* o = reinterpret_cast<const double2*>(stage + (jj - base) * NS);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double2 t = o[q];
            acc[2 * q] += t.x;
            acc[2 * q + 1] += t.y;
          }
          acc[8] += o[4].x;
        } else {
          acc[0] += stage[jj - base];
        }
      }
      __syncthreads();
    }
    if constexpr (kVec9) {
      // stage the blocks, then write the chunk row segment by row segment
      if (own) {
        const int32_t b0 = ptr[row];
        rbase[threadIdx.x] = b0;
        rlen[threadIdx.x] = ptr[row + 1] - b0;
#pragma unroll
        for (int q = 0; q < 9; ++q) stage[threadIdx.x * 9 + q] = acc[q];
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 3; ++c)
        for (int t = threadIdx.x; t < 3 * m; t += kBlock) {
          const int q = t / 3, d = t - 3 * (t / 3);
          const int64_t b0 = rbase[q];
          out[9 * b0 + int64_t(c) * 3 * rlen[q] + 3 * (s0 + q - b0) + d] = stage[q * 9 + c * 3 + d];
        }
      __syncthreads();
    } else if constexpr (OP == CF_OP_P1_VMASS) {
      if (own) {
        const int64_t b0 = ptr[row], len = ptr[row + 1] - b0, kpos = s - b0;
        double* o = out + 3 * b0 + kpos;
        o[0] = acc[0];
        o[len] = acc[0];
        o[2 * len] = acc[0];
      }
    } else {
      if (own) out[s] = acc[0];
    }
  }
}

// ------------------------------------------------------------- rhs

template <int OP>
__global__ void assemble_rhs(const int32_t* __restrict__ offs,
                             const uint32_t* __restrict__ perm,
                             int64_t n_nodes, int k,
                             const double* __restrict__ geo,
                             const double* __restrict__ coef, double lam,
                             double mu, double* __restrict__ rhs) {
  constexpr bool kVec = !(OP == CF_RHS_P1_SOURCE || OP == CF_RHS_T10_SOURCE);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int32_t j = offs[i]; j < offs[i + 1]; ++j) {
      const uint32_t item = perm[j];
      const int64_t e = item / uint32_t(k);
      const int a = int(item - uint32_t(e) * uint32_t(k));
      if constexpr (OP == CF_RHS_P1_SOURCE) {
        acc[0] += (coef[e] * geo[e * 10 + 9]) / 4.0;
      } else if constexpr (OP == CF_RHS_T10_SOURCE) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < 5; ++q) sum += kT10.N[q][a] * (kT10.w[q] * coef[e * 5 + q]);
        acc[0] += geo[e * 10 + 9] * sum;
      } else if constexpr (OP == CF_RHS_P1_BODY) {
        const double v = geo[e * 10 + 9];
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += (coef[e * 3 + c] * v) / 4.0;
      } else if constexpr (OP == CF_RHS_T10_BODY) {
        const double v = geo[e * 10 + 9];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double sum = 0.0;
#pragma unroll
          for (int q = 0; q < 5; ++q) sum += (kT10.w[q] * kT10.N[q][a]) * coef[(e * 5 + q) * 3 + c];
          acc[c] += v * sum;
        }
      } else if constexpr (OP == CF_RHS_P1_THERMAL) {
        // V B_a^T D (s,s,s,0,0,0) = V (3 lam + 2 mu) s g_a
        Geo g = load_geo(geo, e);
        double ga[3];
        p1_grad(g, a, ga);
        const double sig = (3.0 * lam + 2.0 * mu) * coef[e];
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += (sig * ga[c]) * g.v;
      } else if constexpr (OP == CF_RHS_T10_THERMAL) {
        Geo g = load_geo(geo, e);
        double sum[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          double ga[3];
          t10_grad(g, q, a, ga);
          const double sig = kT10.w[q] * (3.0 * lam + 2.0 * mu) * coef[e * 5 + q];
#pragma unroll
          for (int c = 0; c < 3; ++c) sum[c] += sig * ga[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += g.v * sum[c];
      }
    }
    if constexpr (kVec) {
#pragma unroll
      for (int c = 0; c < 3; ++c) rhs[3 * i + c] += acc[c];
    } else {
      rhs[i] += acc[0];
    }
  }
}

// ------------------------------------------------------------- faces

__device__ __forceinline__ double face_area(const double* __restrict__ x,
                                            const double* __restrict__ y,
                                            const double* __restrict__ z,
                                            const int32_t* f) {
  const int a = f[0], b = f[1], c = f[2];
  const double ux = x[b] - x[a], uy = y[b] - y[a], uz = z[b] - z[a];
  const double vx = x[c] - x[a], vy = y[c] - y[a], vz = z[c] - z[a];
  const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
  return 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

template <int OP>
__global__ void face_rhs(const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ z,
                         const int32_t* __restrict__ faces, int nf,
                         const int32_t* __restrict__ offs,
                         const uint32_t* __restrict__ perm, int64_t n_nodes,
                         const double* __restrict__ coef,
                         double* __restrict__ rhs) {
  constexpr bool kVec = (OP == CF_FACE_P1_TRACTION || OP == CF_FACE_T10_TRACTION);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int32_t j = offs[i]; j < offs[i + 1]; ++j) {
      const uint32_t item = perm[j];
      const int64_t f = item / uint32_t(nf);
      const int a = int(item - uint32_t(f) * uint32_t(nf));
      const double area = face_area(x, y, z, faces + f * nf);
      if constexpr (OP == CF_FACE_P1_FLUX) {
        acc[0] += (coef[f] * area) / 3.0;
      } else if constexpr (OP == CF_FACE_T10_FLUX) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) sum += kFace.N6[q][a] * (kFace.w6[q] * coef[f * 6 + q]);
        acc[0] += area * sum;
      } else if constexpr (OP == CF_FACE_P1_TRACTION) {
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += (coef[f * 3 + c] * area) / 3.0;
      } else if constexpr (OP == CF_FACE_T10_TRACTION) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double sum = 0.0;
#pragma unroll
          for (int q = 0; q < 3; ++q) sum += (kFace.w3[q] * kFace.N3[q][a]) * coef[(f * 3 + q) * 3 + c];
          acc[c] += area * sum;
        }
      }
    }
    if constexpr (kVec) {
#pragma unroll
      for (int c = 0; c < 3; ++c) rhs[3 * i + c] += acc[c];
    } else {
      rhs[i] += acc[0];
    }
  }
}

template <int OP>
__global__ void face_values(const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z,
                            const int32_t* __restrict__ faces, int nf,
                            const int32_t* __restrict__ seg,
                            const uint32_t* __restrict__ contrib, int64_t nnz,
                            const double* __restrict__ coef,
                            double* __restrict__ out) {
  const int kk = nf * nf;
  for (int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; s < nnz;
       s += int64_t(gridDim.x) * blockDim.x) {
    double acc = 0.0;
    for (int32_t j = seg[s]; j < seg[s + 1]; ++j) {
      const uint32_t pid = contrib[j];
      const int64_t f = pid / uint32_t(kk);
      const int ab = int(pid - uint32_t(f) * uint32_t(kk));
      const int a = ab / nf, b = ab - (ab / nf) * nf;
      const double area = face_area(x, y, z, faces + f * nf);
      if constexpr (OP == CF_FACE_P1_ROBIN_K) {
        acc += (coef[f] * area) * ((a == b ? 2.0 : 1.0) / 12.0);
      } else {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q)
          sum += (kFace.N6[q][a] * (kFace.w6[q] * coef[f * 6 + q])) * kFace.N6[q][b];
        acc += area * sum;
      }
    }
    out[s] = acc;
  }
}

__global__ void add_subpattern(int64_t n_rows, const int32_t* __restrict__ ap,
                               const int32_t* __restrict__ ai,
                               double* __restrict__ av,
                               const int32_t* __restrict__ bp,
                               const int32_t* __restrict__ bi,
                               const double* __restrict__ bv, double alpha,
                               int32_t* __restrict__ missing) {
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r < n_rows;
       r += int64_t(gridDim.x) * blockDim.x) {
    for (int32_t t = bp[r]; t < bp[r + 1]; ++t) {
      const int32_t col = bi[t];
      int32_t lo = ap[r], hi = ap[r + 1];
      while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (ai[mid] < col)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < ap[r + 1] && ai[lo] == col)
        av[lo] += alpha * bv[t];
      else
        atomicAdd(missing, 1);
    }
  }
}

__global__ void expand_pattern3(int64_t n_nodes, const int32_t* __restrict__ ptr,
                                const int32_t* __restrict__ idx, int dense3,
                                int32_t* __restrict__ dptr,
                                int32_t* __restrict__ didx) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n_nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t base = ptr[i], len = ptr[i + 1] - base;
    for (int c = 0; c < 3; ++c) {
      if (dense3) {
        const int32_t rb = 9 * base + 3 * c * len;
        dptr[3 * i + c] = rb;
        for (int32_t q = 0; q < len; ++q) {
          const int32_t col = idx[base + q];
          didx[rb + 3 * q + 0] = 3 * col + 0;
          didx[rb + 3 * q + 1] = 3 * col + 1;
          didx[rb + 3 * q + 2] = 3 * col + 2;
        }
      } else {
        const int32_t rb = 3 * base + c * len;
        dptr[3 * i + c] = rb;
        for (int32_t q = 0; q < len; ++q) didx[rb + q] = 3 * idx[base + q] + c;
      }
    }
    if (i == n_nodes - 1) dptr[3 * n_nodes] = (dense3 ? 9 : 3) * ptr[n_nodes];
  }
}


// ------------------------------------------------------------- stresses

// Centroid strain (Voigt [xx,yy,zz,yz,xz,xy], engineering shear) and stress
// sigma = D (eps - s (1,1,1,0,0,0)) per element; P1 gradients are constant,
// Tet10 ones are taken at the centroid (elastic3d.py:276-289,
// elastic_tet10.py:263-278).  s[e] = alpha (T_c - T_ref) or null.
template <int K>
__global__ void element_stress(const int32_t* __restrict__ conn, int64_t m,
                               const double* __restrict__ geo,
                               const double* __restrict__ u,
                               const double* __restrict__ s_th, double lam, double mu,
                               double* __restrict__ sig, double* __restrict__ eps) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < m;
       e += int64_t(gridDim.x) * blockDim.x) {
    Geo g = load_geo(geo, e);
    double H[3][3] = {};  // H[c][i] = d u_c / d x_i
    for (int a = 0; a < K; ++a) {
      double ga[3];
      if constexpr (K == 4) {
        p1_grad(g, a, ga);
      } else {
        const double d0 = kT10.Dc[a][0], d1 = kT10.Dc[a][1], d2 = kT10.Dc[a][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) ga[i] = d0 * g.j[0][i] + d1 * g.j[1][i] + d2 * g.j[2][i];
      }
      const int64_t nd = conn[e * K + a];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double uc = u[3 * nd + c];
#pragma unroll
        for (int i = 0; i < 3; ++i) H[c][i] += ga[i] * uc;
      }
    }
    double ev[6] = {H[0][0], H[1][1], H[2][2], H[1][2] + H[2][1], H[0][2] + H[2][0],
                    H[0][1] + H[1][0]};
    const double th = s_th ? s_th[e] : 0.0;
    double d[6] = {ev[0] - th, ev[1] - th, ev[2] - th, ev[3], ev[4], ev[5]};
    const double l2m = lam + 2.0 * mu;
    double* so = sig + e * 6;
    double* eo = eps + e * 6;
    so[0] = l2m * d[0] + lam * d[1] + lam * d[2];
    so[1] = lam * d[0] + l2m * d[1] + lam * d[2];
    so[2] = lam * d[0] + lam * d[1] + l2m * d[2];
    so[3] = mu * d[3];
    so[4] = mu * d[4];
    so[5] = mu * d[5];
#pragma unroll
    for (int t = 0; t < 6; ++t) eo[t] = ev[t];
  }
}


// ------------------------------------------------------------- adjoint

// dJ/dk_e = -lambda_e^T (dK_e/dk_e) T_e for a diffusion coefficient that is
// constant per element: P1 -V grad(lambda).grad(T); Tet10
// -V sum_q w_q grad(lambda)_q.grad(T)_q (K_e is linear in k_e).
template <int K>
__global__ void adjoint_dk(const int32_t* __restrict__ conn, int64_t m,
                           const double* __restrict__ geo, const double* __restrict__ T,
                           const double* __restrict__ lam, double* __restrict__ out) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < m;
       e += int64_t(gridDim.x) * blockDim.x) {
    Geo g = load_geo(geo, e);
    double tv[K], lv[K];
#pragma unroll
    for (int a = 0; a < K; ++a) {
      const int64_t nd = conn[e * K + a];
      tv[a] = T[nd];
      lv[a] = lam[nd];
    }
    double acc = 0.0;
    constexpr int NQ = K == 4 ? 1 : 5;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double gt[3] = {0.0, 0.0, 0.0}, gl[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int a = 0; a < K; ++a) {
        double ga[3];
        if constexpr (K == 4)
          p1_grad(g, a, ga);
        else
          t10_grad(g, q, a, ga);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          gt[i] += ga[i] * tv[a];
          gl[i] += ga[i] * lv[a];
        }
      }
      const double w = K == 4 ? 1.0 : kT10.w[q];
      acc += w * (gl[0] * gt[0] + gl[1] * gt[1] + gl[2] * gt[2]);
    }
    out[e] = -g.v * acc;
  }
}

}  // namespace

// ------------------------------------------------------------- C ABI

extern "C" int cf_tet_geometry(const double* x, const double* y, const double* z,
                               const int32_t* conn, int64_t n_elem, int32_t k,
                               double* geo, int32_t* n_bad_dev, void* stream) {
  CF_REQUIRE(x && y && z && geo && n_bad_dev && (conn || n_elem == 0) && k >= 4,
             "cf_tet_geometry: bad arguments");
  if (n_elem == 0) return CF_OK;
  hipStream_t s = cf::as_stream(stream);
  tet_geometry<<<cf::stride_grid(n_elem, kBlock), kBlock, 0, s>>>(x, y, z, conn, n_elem, k,
                                                                  geo, n_bad_dev);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_assemble_matrix(int32_t op, const int32_t* indptr,
                                  const int32_t* indices, const int32_t* seg,
                                  const uint32_t* contrib, const int32_t* conn,
                                  int64_t n_nodes, int64_t nnz, int32_t k, const double* geo,
                                  const double* coef, double lam, double mu,
                                  double* values, void* stream) {
  (void)indices;
  CF_REQUIRE(indptr && seg && values && n_nodes > 0 && nnz >= 0,
             "cf_assemble_matrix: bad arguments");
  CF_REQUIRE(nnz == 0 || (contrib && geo && conn), "cf_assemble_matrix: null inputs");
  const bool p1 = (op == CF_OP_P1_STIFF || op == CF_OP_P1_MASS || op == CF_OP_P1_ELAST ||
                   op == CF_OP_P1_VMASS);
  CF_REQUIRE(k == (p1 ? 4 : 10), "cf_assemble_matrix: op %d needs k=%d, got %d", op,
             p1 ? 4 : 10, k);
  const bool needs_coef = !(op == CF_OP_T10_ELAST || op == CF_OP_P1_ELAST);
  CF_REQUIRE(!needs_coef || coef || nnz == 0, "cf_assemble_matrix: op %d needs coef", op);
  if (nnz == 0) return CF_OK;
  hipStream_t s = cf::as_stream(stream);
  // one round of resident blocks, each walking its share of the 256-entry
  // chunks: with the generic cap (8 blocks per CU) the 31 KB of LDS per block
  // left 768 of 2,048 blocks for a second round (Tet10 elasticity at
  // 15.0 M stored blocks: 690 against 746 us)
  const int64_t n_chunks = (nnz + kBlock - 1) / kBlock;
  int grid = 1;
#define CF_CASE(OPV)                                                                  \
  case OPV: {                                                                         \
    int per = 0;                                                                      \
    CF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, assemble_values<OPV>, kBlock, 0)); \
    const int64_t cap = int64_t(cf::cu_count()) * (per > 0 ? per : 1);               \
    grid = int(n_chunks < cap ? n_chunks : cap);                                      \
    assemble_values<OPV><<<grid, kBlock, 0, s>>>(indptr, seg, contrib, conn, nnz, k, geo, \
                                                 coef, lam, mu, values);              \
    break;                                                                            \
  }
  switch (op) {
    CF_CASE(CF_OP_P1_STIFF)
    CF_CASE(CF_OP_P1_MASS)
    CF_CASE(CF_OP_T10_STIFF)
    CF_CASE(CF_OP_T10_MASS)
    CF_CASE(CF_OP_T10_ELAST)
    CF_CASE(CF_OP_P1_ELAST)
    CF_CASE(CF_OP_P1_VMASS)
    default:
      cf::set_error("cf_assemble_matrix: unknown op %d", op);
      return CF_ERR_ARG;
  }
#undef CF_CASE
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_assemble_rhs(int32_t op, const int32_t* offs, const uint32_t* perm,
                               int64_t n_nodes, int32_t k, const double* geo,
                               const double* coef, double lam, double mu, double* rhs,
                               void* stream) {
  CF_REQUIRE(offs && rhs && geo && coef && n_nodes > 0, "cf_assemble_rhs: bad arguments");
  const bool p1 = (op == CF_RHS_P1_SOURCE || op == CF_RHS_P1_BODY || op == CF_RHS_P1_THERMAL);
  CF_REQUIRE(k == (p1 ? 4 : 10), "cf_assemble_rhs: op %d needs k=%d, got %d", op,
             p1 ? 4 : 10, k);
  hipStream_t s = cf::as_stream(stream);
  const int grid = cf::stride_grid(n_nodes, kBlock);
#define CF_CASE(OPV)                                                                       \
  case OPV:                                                                                \
    assemble_rhs<OPV><<<grid, kBlock, 0, s>>>(offs, perm, n_nodes, k, geo, coef, lam, mu, \
                                              rhs);                                        \
    break;
  switch (op) {
    CF_CASE(CF_RHS_P1_SOURCE)
    CF_CASE(CF_RHS_T10_SOURCE)
    CF_CASE(CF_RHS_P1_BODY)
    CF_CASE(CF_RHS_T10_BODY)
    CF_CASE(CF_RHS_P1_THERMAL)
    CF_CASE(CF_RHS_T10_THERMAL)
    default:
      cf::set_error("cf_assemble_rhs: unknown op %d", op);
      return CF_ERR_ARG;
  }
#undef CF_CASE
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_face_rhs(int32_t op, const double* x, const double* y, const double* z,
                           const int32_t* faces, int64_t n_faces, int32_t nf,
                           const int32_t* offs, const uint32_t* perm, int64_t n_nodes,
                           const double* coef, double* rhs, void* stream) {
  CF_REQUIRE(x && y && z && offs && rhs && n_nodes > 0, "cf_face_rhs: bad arguments");
  if (n_faces == 0) return CF_OK;
  CF_REQUIRE(faces && perm && coef, "cf_face_rhs: null face data");
  const bool p1 = (op == CF_FACE_P1_FLUX || op == CF_FACE_P1_TRACTION);
  CF_REQUIRE(nf == (p1 ? 3 : 6), "cf_face_rhs: op %d needs nf=%d", op, p1 ? 3 : 6);
  hipStream_t s = cf::as_stream(stream);
  const int grid = cf::stride_grid(n_nodes, kBlock);
#define CF_CASE(OPV)                                                                      \
  case OPV:                                                                               \
    face_rhs<OPV><<<grid, kBlock, 0, s>>>(x, y, z, faces, nf, offs, perm, n_nodes, coef, \
                                          rhs);                                           \
    break;
  switch (op) {
    CF_CASE(CF_FACE_P1_FLUX)
    CF_CASE(CF_FACE_T10_FLUX)
    CF_CASE(CF_FACE_P1_TRACTION)
    CF_CASE(CF_FACE_T10_TRACTION)
    default:
      cf::set_error("cf_face_rhs: unknown op %d", op);
      return CF_ERR_ARG;
  }
#undef CF_CASE
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_face_matrix(int32_t op, const double* x, const double* y, const double* z,
                              const int32_t* faces, int32_t nf, const int32_t* indptr,
                              const int32_t* indices, const int32_t* seg,
                              const uint32_t* contrib, int64_t n_nodes, int64_t nnz,
                              const double* coef, double* values, void* stream) {
  (void)indptr;
  (void)indices;
  (void)n_nodes;
  CF_REQUIRE(x && y && z && seg && values, "cf_face_matrix: bad arguments");
  if (nnz == 0) return CF_OK;
  CF_REQUIRE(faces && contrib && coef, "cf_face_matrix: null face data");
  hipStream_t s = cf::as_stream(stream);
  const int grid = cf::stride_grid(nnz, kBlock);
  if (op == CF_FACE_P1_ROBIN_K) {
    CF_REQUIRE(nf == 3, "cf_face_matrix: P1 Robin needs nf=3");
    face_values<CF_FACE_P1_ROBIN_K><<<grid, kBlock, 0, s>>>(x, y, z, faces, nf, seg, contrib,
                                                            nnz, coef, values);
  } else if (op == CF_FACE_T10_ROBIN_K) {
    CF_REQUIRE(nf == 6, "cf_face_matrix: Tet10 Robin needs nf=6");
    face_values<CF_FACE_T10_ROBIN_K><<<grid, kBlock, 0, s>>>(x, y, z, faces, nf, seg, contrib,
                                                             nnz, coef, values);
  } else {
    cf::set_error("cf_face_matrix: unknown op %d", op);
    return CF_ERR_ARG;
  }
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_csr_add_subpattern(int64_t n_rows, const int32_t* a_ptr,
                                     const int32_t* a_idx, double* a_val,
                                     const int32_t* b_ptr, const int32_t* b_idx,
                                     const double* b_val, double alpha,
                                     int32_t* n_missing_dev, void* stream) {
  CF_REQUIRE(a_ptr && a_idx && a_val && b_ptr && b_idx && b_val && n_missing_dev && n_rows > 0,
             "cf_csr_add_subpattern: bad arguments");
  hipStream_t s = cf::as_stream(stream);
  add_subpattern<<<cf::stride_grid(n_rows, kBlock), kBlock, 0, s>>>(
      n_rows, a_ptr, a_idx, a_val, b_ptr, b_idx, b_val, alpha, n_missing_dev);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_expand_pattern3(int64_t n_nodes, const int32_t* ptr, const int32_t* idx,
                                  int32_t dense3, int32_t* dof_ptr, int32_t* dof_idx,
                                  void* stream) {
  CF_REQUIRE(ptr && dof_ptr && n_nodes > 0, "cf_expand_pattern3: bad arguments");
  hipStream_t s = cf::as_stream(stream);
  expand_pattern3<<<cf::stride_grid(n_nodes, kBlock), kBlock, 0, s>>>(n_nodes, ptr, idx,
                                                                      dense3, dof_ptr, dof_idx);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_element_stress(const int32_t* conn, int64_t n_elem, int32_t k,
                                 const double* geo, const double* u, const double* s_thermal,
                                 double lam, double mu, double* sig, double* eps,
                                 void* stream) {
  CF_REQUIRE(geo && u && sig && eps && (conn || n_elem == 0) && (k == 4 || k == 10),
             "cf_element_stress: bad arguments");
  if (n_elem == 0) return CF_OK;
  hipStream_t s = cf::as_stream(stream);
  const int grid = cf::stride_grid(n_elem, kBlock);
  if (k == 4)
    element_stress<4><<<grid, kBlock, 0, s>>>(conn, n_elem, geo, u, s_thermal, lam, mu, sig, eps);
  else
    element_stress<10><<<grid, kBlock, 0, s>>>(conn, n_elem, geo, u, s_thermal, lam, mu, sig,
                                               eps);
  CF_LAUNCHED();
  return CF_OK;
}

extern "C" int cf_adjoint_dk(const int32_t* conn, int64_t n_elem, int32_t k, const double* geo,
                             const double* T, const double* lam, double* dk, void* stream) {
  CF_REQUIRE(geo && T && lam && dk && (conn || n_elem == 0) && (k == 4 || k == 10),
             "cf_adjoint_dk: bad arguments");
  if (n_elem == 0) return CF_OK;
  hipStream_t s = cf::as_stream(stream);
  const int grid = cf::stride_grid(n_elem, kBlock);
  if (k == 4)
    adjoint_dk<4><<<grid, kBlock, 0, s>>>(conn, n_elem, geo, T, lam, dk);
  else
    adjoint_dk<10><<<grid, kBlock, 0, s>>>(conn, n_elem, geo, T, lam, dk);
  CF_LAUNCHED();
  return CF_OK;
}
