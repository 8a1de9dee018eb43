// Shared host/device helpers of libcloudfem_hip (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "cloudfem_hip.h"

namespace cf {

void set_error(const char* fmt, ...);

// Number of compute units of the current device (cached per device).
int cu_count();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Round a byte count up to 256 so sub-buffers carved from a workspace stay
// aligned for 16-byte vector loads and rocPRIM's temporaries.
inline size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Bump allocator over a caller-owned workspace.
struct Carve {
  char* base;
  size_t used = 0;
  explicit Carve(void* p) : base(static_cast<char*>(p)) {}
  template <class T>
  T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base + used);
    used += align256(count * sizeof(T));
    return p;
  }
};

// Grid size of the grid-stride kernels: enough resident waves to cover HBM
// latency on every CU, and a fixed number of per-block partial sums so the
// deterministic reductions stay cheap.
inline int stride_grid(int64_t work_items, int block) {
  int64_t want = ceil_div(work_items, block);
  int64_t cap = int64_t(cu_count()) * 8;
  if (want < 1) want = 1;
  return int(want < cap ? want : cap);
}

}  // namespace cf

#define CF_HIP(expr)                                                     \
  do {                                                                   \
    hipError_t _e = (expr);                                              \
    if (_e != hipSuccess) {                                              \
      cf::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                    __FILE__, __LINE__);                                 \
      return CF_ERR_HIP;                                                 \
    }                                                                    \
  } while (0)

#define CF_LAUNCHED()                                                    \
  do {                                                                   \
    hipError_t _e = hipGetLastError();                                   \
    if (_e != hipSuccess) {                                              \
      cf::set_error("kernel launch failed: %s (%s:%d)",                  \
                    hipGetErrorString(_e), __FILE__, __LINE__);          \
      return CF_ERR_HIP;                                                 \
    }                                                                    \
  } while (0)

#define CF_REQUIRE(cond, ...)                                            \
  do {                                                                   \
    if (!(cond)) {                                                       \
      cf::set_error(__VA_ARGS__);                                        \
      return CF_ERR_ARG;                                                 \
    }                                                                    \
  } while (0)
