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

// Loads and stores through a global-address-space pointer.  Pointers held in
// device structs (the frame tables) are generic, so plain dereferences
// compile to flat accesses; an integer round trip into address space 1 makes
// them global_load / global_store with a scalar base where possible.
#define PG_GAS __attribute__((address_space(1)))
typedef float pg_f2n __attribute__((ext_vector_type(2)));
typedef float pg_f4n __attribute__((ext_vector_type(4)));
template <class T> __device__ __forceinline__ T gld(const T *p) {
  return *(const PG_GAS T *)(uintptr_t)p;
}
template <> __device__ __forceinline__ float2 gld<float2>(const float2 *p) {
  const pg_f2n v = *(const PG_GAS pg_f2n *)(uintptr_t)p;
  return make_float2(v.x, v.y);
}
template <> __device__ __forceinline__ float4 gld<float4>(const float4 *p) {
  const pg_f4n v = *(const PG_GAS pg_f4n *)(uintptr_t)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
typedef unsigned int pg_u2n __attribute__((ext_vector_type(2)));
template <> __device__ __forceinline__ uint2 gld<uint2>(const uint2 *p) {
  const pg_u2n v = *(const PG_GAS pg_u2n *)(uintptr_t)p;
  return make_uint2(v.x, v.y);
}
typedef unsigned int pg_u4n __attribute__((ext_vector_type(4)));
template <> __device__ __forceinline__ uint4 gld<uint4>(const uint4 *p) {
  const pg_u4n v = *(const PG_GAS pg_u4n *)(uintptr_t)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
template <class T> __device__ __forceinline__ void gst(T *p, T v) {
  *(PG_GAS T *)(uintptr_t)p = v;
}
template <> __device__ __forceinline__ void gst<float2>(float2 *p, float2 v) {
  pg_f2n n;
  n.x = v.x;
  n.y = v.y;
  *(PG_GAS pg_f2n *)(uintptr_t)p = n;
}
template <> __device__ __forceinline__ void gst<float4>(float4 *p, float4 v) {
  pg_f4n n;
  n.x = v.x;
  n.y = v.y;
  n.z = v.z;
  n.w = v.w;
  *(PG_GAS pg_f4n *)(uintptr_t)p = n;
}
template <> __device__ __forceinline__ void gst<uint4>(uint4 *p, uint4 v) {
  pg_u4n n;
  n.x = v.x;
  n.y = v.y;
  n.z = v.z;
  n.w = v.w;
  *(PG_GAS pg_u4n *)(uintptr_t)p = n;
}

// Streaming (non-temporal) variants for data read or written once: they do
// not displace what should stay in the caches (the operator table).
template <class T> __device__ __forceinline__ T gld_nt(const T *p) {
  return __builtin_nontemporal_load((const PG_GAS T *)(uintptr_t)p);
}
template <> __device__ __forceinline__ float2 gld_nt<float2>(const float2 *p) {
  const pg_f2n v = __builtin_nontemporal_load((const PG_GAS pg_f2n *)(uintptr_t)p);
  return make_float2(v.x, v.y);
}
template <> __device__ __forceinline__ float4 gld_nt<float4>(const float4 *p) {
  const pg_f4n v = __builtin_nontemporal_load((const PG_GAS pg_f4n *)(uintptr_t)p);
  return make_float4(v.x, v.y, v.z, v.w);
}
template <> __device__ __forceinline__ uint4 gld_nt<uint4>(const uint4 *p) {
  const pg_u4n v = __builtin_nontemporal_load((const PG_GAS pg_u4n *)(uintptr_t)p);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst_nt(float2 *p, float2 v) {
  pg_f2n n;
  n.x = v.x;
  n.y = v.y;
  __builtin_nontemporal_store(n, (PG_GAS pg_f2n *)(uintptr_t)p);
}
__device__ __forceinline__ void gst_nt(float4 *p, float4 v) {
  pg_f4n n;
  n.x = v.x;
  n.y = v.y;
  n.z = v.z;
  n.w = v.w;
  __builtin_nontemporal_store(n, (PG_GAS pg_f4n *)(uintptr_t)p);
}

// 24-bit sample format of the BF16 path's sample-major rows: the top three
// bytes of the fp32 value rounded to nearest even at bit 8, i.e. 16
// significant bits (NaN stays NaN, overflow rounds to infinity), packed
// little-endian 3 bytes per value; a sample's row of 2C values (re, im per
// coil) starts on a 16-byte boundary.  16 bits is what the split-BF16
// products use of an input anyway: x = hi + lo with hi = bf16(x) and
// lo = x - hi exactly representable in bf16.
typedef unsigned int pg_u3n __attribute__((ext_vector_type(3)));
__host__ __device__ constexpr int64_t s24_row_bytes(int64_t C) { return (6 * C + 15) / 16 * 16; }
__device__ __forceinline__ uint32_t s24_round(float x) {
  const uint32_t b = __float_as_uint(x);
  if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x7fffffu) != 0u) return (b >> 8) | 0x4000u;
  return (b + 0x7fu + ((b >> 8) & 1u)) >> 8;
}
// values 0..3 of three packed dwords (d0 = v0 | v1 << 24, d1 = v1 >> 8 |
// v2 << 16, d2 = v2 >> 16 | v3 << 8) as fp32 bit patterns
__device__ __forceinline__ void s24_unpack4(uint32_t d0, uint32_t d1, uint32_t d2, float *o) {
  o[0] = __uint_as_float(d0 << 8);
  o[1] = __uint_as_float((d1 << 16) | ((d0 >> 16) & 0xff00u));
  o[2] = __uint_as_float((d2 << 24) | ((d1 >> 8) & 0xffff00u));
  o[3] = __uint_as_float(d2 & 0xffffff00u);
}
__device__ __forceinline__ pg_u3n s24_pack4(float a, float b, float c, float d) {
  const uint32_t v0 = s24_round(a), v1 = s24_round(b), v2 = s24_round(c), v3 = s24_round(d);
  pg_u3n r;
  r.x = v0 | (v1 << 24);
  r.y = (v1 >> 8) | (v2 << 16);
  r.z = (v2 >> 16) | (v3 << 8);
  return r;
}
__device__ __forceinline__ float s24_load(const unsigned char *p) {
  const PG_GAS unsigned char *q = (const PG_GAS unsigned char *)(uintptr_t)p;
  return __uint_as_float(((uint32_t)q[0] << 8) | ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 24));
}

// Sample-row formats of the grouped path's sample-major copy.
constexpr int kRowC64 = 0;  // complex64, 8 C bytes a row
constexpr int kRowS24 = 1;  // 24-bit values (above), 16 * ceil(6 C / 16) bytes
constexpr int kRowBfp = 2;  // block floating point (below), 16 * ceil(C / 4) bytes

// Block-floating-point rows ("bfp15"): a row's 2C values (re, im per coil)
// in groups of 8 (4 coils; the last group zero-padded), 16 bytes a group
// (a 32-coil row is one 128-byte line).  A group holds 8 little-endian
// 16-bit fields f_i; bits 1..15 of f_i are a two's-complement mantissa
// m_i in [-16384, 16383] and bit 0 of f_i is bit i of the group's 8-bit
// exponent E; value_i = m_i * 2^(E - 140).  Encoding: M = max |x_i|; the
// scale s = max(floor(log2 M) - 13, -140), plus 1 if M * 2^-s rounds past
// 16383 (s = -140 when M = 0); m_i = x_i * 2^-s rounded to nearest even.
// So |error| <= 2^(s-1) <= M * 2^-14 for every value of a group.  Every
// finite fp32 input has such an s in [-140, 115] (inputs within 2^-15 of
// FLT_MAX round to infinity); a group holding a NaN or an infinity is
// stored as E = 255, m_i = -16384 and decodes to -infinity.  A split-BF16
// product reads x = hi + lo exactly from such a value (|m| < 2^14: hi the
// top 8 significant bits, lo the rest).
__host__ __device__ constexpr int64_t bfp_row_bytes(int64_t C) { return (C + 3) / 4 * 16; }
constexpr int kBfpBias = 140;
__device__ __forceinline__ uint4 bfp_encode8(const float (&x)[8]) {
  // M's bit pattern: the largest |x| as an unsigned integer (|x| orders as
  // its pattern does, and an infinity or a NaN is above every finite value)
  uint32_t mb = 0u;
#pragma unroll
  for (int i = 0; i < 8; ++i) mb = max(mb, __float_as_uint(x[i]) & 0x7fffffffu);
  // s = floor(log2 M) - 13 from the exponent field (zero and subnormal M:
  // field 0, s = -140), plus 1 when M * 2^-s rounds past 16383: the
  // fraction bits reach 16383.5 / 8192 (normal) or 16383.5 / 16384
  // (subnormal)
  const uint32_t ef = mb >> 23;
  const int s = (int)ef - kBfpBias + ((mb & 0x7fffffu) >= (ef != 0u ? 0x7ffe00u : 0x7fff00u) ? 1 : 0);
  const uint32_t E = (uint32_t)(s + kBfpBias);
  // m_i: x_i 2^-s (exact) plus 1.5 2^23 rounds to an integer, ties to even
  // (the sum lies in [2^23, 2^24), whose spacing is 1, for |m| < 2^22), and
  // the sum's low 16 bits are then m_i in two's complement (1.5 2^23 has
  // none set)
  uint32_t b[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) b[i] = __float_as_uint(__builtin_amdgcn_ldexpf(x[i], -s) + 12582912.0f);
  // dword k: the low halves of b[2k] and b[2k+1] (one byte permute), each
  // moved up one bit within its half, then E bit 2k at bit 0 and E bit
  // 2k + 1 at bit 16 (from E | E << 15)
  const uint32_t eb = E | (E << 15);
  uint32_t d[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t pk = __builtin_amdgcn_perm(b[2 * k + 1], b[2 * k], 0x05040100u);
    d[k] = ((pk << 1) & 0xFFFEFFFEu) | ((eb >> (2 * k)) & 0x10001u);
  }
  uint4 w = make_uint4(d[0], d[1], d[2], d[3]);
  if (mb >= 0x7f800000u) w = make_uint4(0x80018001u, 0x80018001u, 0x80018001u, 0x80018001u);  // E = 255, m = -16384
  return w;
}
// E - 140 + adj of a group (adj: a constant folded into the sum for free)
__device__ __forceinline__ int bfp_scale(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3, int adj = 0) {
  // the fields' low bytes (bytes 0 and 2 of each dword), four to a dword,
  // their bit 0 kept: byte k of lo holds bit k of E, byte k of hi bit 4 + k;
  // E - 140 = sum_k lo_k 2^k + sum_k hi_k 2^(4 + k) - 140 as two byte dot
  // products, the second accumulating onto the first, the first onto -140
  const uint32_t lo = __builtin_amdgcn_perm(d1, d0, 0x06040200u) & 0x01010101u;
  const uint32_t hi = __builtin_amdgcn_perm(d3, d2, 0x06040200u) & 0x01010101u;
  const uint32_t e = __builtin_amdgcn_udot4(lo, 0x08040201u, (uint32_t)(adj - kBfpBias), false);
  return (int)__builtin_amdgcn_udot4(hi, 0x80402010u, e, false);
}
// Decoding: with the exponent bits (bit 0 of each field) cleared, a field
// read as a signed 16-bit integer is 2 m exactly, so a value is
// ldexp(2 m, s - 1) = m 2^s -- one mask per two fields, then a convert of
// each 16-bit half (sub-dword operand select, no shifts) and an ldexp.
// (re, im) of coil k (0..3) of a group
__device__ __forceinline__ float2 bfp_decode_coil(uint4 g, int k) {
  const int s = bfp_scale(g.x, g.y, g.z, g.w, -1);
  const uint32_t d = (k == 0 ? g.x : k == 1 ? g.y : k == 2 ? g.z : g.w) & 0xFFFEFFFEu;
  return make_float2(__builtin_amdgcn_ldexpf((float)(short)(d & 0xFFFFu), s),
                     __builtin_amdgcn_ldexpf((float)(short)(d >> 16), s));
}
__device__ __forceinline__ void bfp_decode8(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3, float *o) {
  const int s = bfp_scale(d0, d1, d2, d3, -1);
  const uint32_t d[4] = {d0 & 0xFFFEFFFEu, d1 & 0xFFFEFFFEu, d2 & 0xFFFEFFFEu, d3 & 0xFFFEFFFEu};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    o[2 * k] = __builtin_amdgcn_ldexpf((float)(short)(d[k] & 0xFFFFu), s);
    o[2 * k + 1] = __builtin_amdgcn_ldexpf((float)(short)(d[k] >> 16), s);
  }
}

// Orders one wavefront's LDS writes before its later LDS reads (other lanes
// of the same wavefront); no other wavefront is involved.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
