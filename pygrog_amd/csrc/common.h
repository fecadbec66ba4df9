// Shared helpers of libpygrog_hip.so: error reporting, complex arithmetic on
// float2/double2, launch-size helpers.  gfx950 only (64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/pygrog_hip.h"

namespace pg {

// Thread-local message of the last failure; read back by pg_last_error().
void set_error(const char *fmt, ...);

inline hipStream_t as_stream(pg_stream_t s) {
  return reinterpret_cast<hipStream_t>(s);
}

constexpr int kWave = 64;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Blocks for a grid-stride loop over n items.
inline unsigned grid_for(int64_t n, int threads, int64_t cap = 65536) {
  int64_t b = ceil_div(n, threads);
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

template <class T> struct V2;
template <> struct V2<float> { using type = float2; };
template <> struct V2<double> { using type = double2; };
template <class T> using v2_t = typename V2<T>::type;

template <class T>
__device__ __forceinline__ v2_t<T> cmake(T re, T im) {
  v2_t<T> r;
  r.x = re;
  r.y = im;
  return r;
}
// acc += a * b (complex), in the (re, im) operation order of c10::complex.
template <class T>
__device__ __forceinline__ void cmac(v2_t<T> &acc, v2_t<T> a, v2_t<T> b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
}

}  // namespace pg

#define PG_CHECK(cond, ...)        \
  do {                             \
    if (!(cond)) {                 \
      pg::set_error(__VA_ARGS__);  \
      return 1;                    \
    }                              \
  } while (0)

#define PG_HIP(expr)                                                    \
  do {                                                                  \
    hipError_t pg_e_ = (expr);                                          \
    if (pg_e_ != hipSuccess) {                                          \
      pg::set_error("%s failed: %s", #expr, hipGetErrorString(pg_e_));  \
      return 2;                                                         \
    }                                                                   \
  } while (0)

#define PG_LAUNCHED() PG_HIP(hipGetLastError())
