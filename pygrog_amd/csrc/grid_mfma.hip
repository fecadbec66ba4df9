// Grouped complex GEMM on gfx950 BF16 MFMA (v_mfma_f32_16x16x32_bf16) with
// FP32 accumulation: the first pass of the grouped gridding modes.
//
// Replicas are grouped by quantised table entry e, so each operator K[e] is
// read once per group instead of once per replica (the plain path re-reads
// 8*C*C bytes per replica, which makes it table-bound once the table exceeds
// the on-chip caches: 565 MB at C=32, q=40 in 3D).
//
// Complex -> real form, interleaved so that a sample-major row of y_t is the
// K-dimension vector as stored:
//   rows 2co, 2co+1 = Re, Im of output coil co; cols 2ci, 2ci+1 = Re, Im of
//   input coil ci;  A = [[Kr, -Ki], [Ki, Kr]] per (co, ci) 2x2 block.
// One wavefront computes D (16 rows x 16 replicas) tiles for all MT row
// tiles of one entry; KS K-steps of 32.  Split-BF16 (n_terms = 3):
//   A B ~= Ahi Bhi + Ahi Blo + Alo Bhi, with x = hi + lo, hi = bf16(x),
//   lo = bf16(x - hi): error O(2^-16) relative instead of O(2^-8).
//
// MFMA 16x16x32 bf16 lane layout (64 lanes, q = lane / 16, j = lane % 16):
//   A operand: row j, k = 8q .. 8q+7;  B operand: column j, k = 8q .. 8q+7;
//   D: column j, rows 4q .. 4q+3.
#include "common.h"

namespace pg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float(((uint32_t)h) << 16);
}

// packed[e][part][mt][ks][lane][i], lane = (row % 16) + 16 * ((col % 32) / 8),
// i = col % 8, row = 16 mt + row % 16, col = 32 ks + col % 32.
__global__ void k_table_pack_bf16(const float2 *__restrict__ K, int64_t nK,
                                  int C, int MT, int KS, int n_parts,
                                  unsigned short *__restrict__ packed) {
  const int64_t per_part = (int64_t)MT * KS * 64 * 8;
  const int64_t total = nK * per_part;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / per_part;
    int64_t rem = t - e * per_part;
    const int i = (int)(rem % 8);
    rem /= 8;
    const int lane = (int)(rem % 64);
    rem /= 64;
    const int ks = (int)(rem % KS);
    const int mt = (int)(rem / KS);
    const int row = 16 * mt + (lane % 16);
    const int col = 32 * ks + 8 * (lane / 16) + i;
    const int co = row / 2, ci = col / 2;
    float a = 0.0f;
    if (co < C && ci < C) {
      const float2 k = K[(e * C + co) * C + ci];
      const bool re_row = (row % 2) == 0, re_col = (col % 2) == 0;
      a = re_row ? (re_col ? k.x : -k.y) : (re_col ? k.y : k.x);
    }
    const unsigned short hi = f2bf(a);
    const int64_t base = e * n_parts * per_part;
    packed[base + (t - e * per_part)] = hi;
    if (n_parts > 1) {
      const unsigned short lo = f2bf(a - bf2f(hi));
      packed[base + per_part + (t - e * per_part)] = lo;
    }
  }
}

constexpr int kTilesPerWave = 8;  // 16-replica N tiles per wavefront

template <int MT, int KS, int NTERMS>
__global__ __launch_bounds__(256) void k_apply_grouped(
    const float *__restrict__ y_t, const unsigned short *__restrict__ tbl,
    const uint32_t *__restrict__ ent_ptr, const uint32_t *__restrict__ ent_rep,
    const uint32_t *__restrict__ rep_src, const float *__restrict__ rep_s,
    const uint32_t *__restrict__ items, int64_t n_items, int64_t nK, int C,
    float *__restrict__ vbuf) {
  constexpr int NP = (NTERMS > 1) ? 2 : 1;
  const int wave = threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  const int64_t item = (int64_t)blockIdx.x * 4 + wave;
  if (item >= n_items) return;
  const int64_t e = items[2 * item];
  if (e >= nK) return;
  const uint32_t b1 = ent_ptr[e + 1];
  const uint32_t start = ent_ptr[e] + items[2 * item + 1];
  if (start >= b1) return;
  uint32_t stop = start + 16 * kTilesPerWave;
  if (stop > b1) stop = b1;
  const int twoC = 2 * C;
  const int64_t per_part = (int64_t)MT * KS * 64 * 8;

  bf16x8_t ahi[MT][KS];
  bf16x8_t alo[MT][KS];
  const unsigned short *tb = tbl + e * NP * per_part;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int64_t off = ((int64_t)(mt * KS + ks) * 64 + lane) * 8;
      u16x8_t h = *reinterpret_cast<const u16x8_t *>(tb + off);
      ahi[mt][ks] = __builtin_bit_cast(bf16x8_t, h);
      if (NP > 1) {
        u16x8_t l = *reinterpret_cast<const u16x8_t *>(tb + per_part + off);
        alo[mt][ks] = __builtin_bit_cast(bf16x8_t, l);
      }
    }

  const int j = lane % 16, q = lane / 16;
  for (uint32_t n0 = start; n0 < stop; n0 += 16) {
    const uint32_t idx = n0 + j;
    const bool valid = idx < stop;
    uint32_t r = 0, src = 0;
    float s = 0.0f;
    if (valid) {
      r = ent_rep[idx];
      src = rep_src[r];
      s = rep_s[r];
    }
    floatx4_t acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] = floatx4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int col0 = 32 * ks + 8 * q;
      float f[8];
      const float *yr = y_t + (int64_t)src * twoC + col0;
      // rows are 16-byte aligned only when 2C is a multiple of 4
      if (valid && col0 + 8 <= twoC && (twoC & 3) == 0) {
        const float4 a0 = *reinterpret_cast<const float4 *>(yr);
        const float4 a1 = *reinterpret_cast<const float4 *>(yr + 4);
        f[0] = a0.x; f[1] = a0.y; f[2] = a0.z; f[3] = a0.w;
        f[4] = a1.x; f[5] = a1.y; f[6] = a1.z; f[7] = a1.w;
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i)
          f[i] = (valid && col0 + i < twoC) ? yr[i] : 0.0f;
      }
      u16x8_t bh, bl;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const unsigned short h = f2bf(f[i]);
        bh[i] = h;
        bl[i] = f2bf(f[i] - bf2f(h));
      }
      const bf16x8_t bhi = __builtin_bit_cast(bf16x8_t, bh);
      const bf16x8_t blo = __builtin_bit_cast(bf16x8_t, bl);
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[mt][ks], bhi,
                                                          acc[mt], 0, 0, 0);
        if (NTERMS > 1) {
          acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[mt][ks], blo,
                                                            acc[mt], 0, 0, 0);
          acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[mt][ks], bhi,
                                                            acc[mt], 0, 0, 0);
        }
      }
    }
    if (valid) {
      float *vr = vbuf + (int64_t)r * twoC;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const int row = 16 * mt + 4 * q;
        float o[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = (acc[mt][i] * s) * s;
        if (row + 4 <= twoC && (twoC & 3) == 0) {
          *reinterpret_cast<float4 *>(vr + row) = make_float4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (row + i < twoC) vr[row + i] = o[i];
        }
      }
    }
  }
}

static void mfma_shape(int64_t C, int &MT, int &KS) {
  const int twoC = (int)(2 * C);
  MT = (twoC + 15) / 16;
  KS = (twoC + 31) / 32;
}

template <int MT, int KS>
static int launch_grouped(const void *y_t, const void *tbl,
                          const uint32_t *ent_ptr, const uint32_t *ent_rep,
                          const uint32_t *rep_src, const float *rep_s,
                          const uint32_t *items, int64_t n_items, int64_t K,
                          int64_t C, int32_t n_terms, void *vbuf,
                          hipStream_t st) {
  if (n_items == 0) return 0;
  PG_CHECK(items != nullptr, "grouped apply: null work list");
  const dim3 grid((unsigned)ceil_div(n_items, 4));
  if (n_terms == 3)
    hipLaunchKernelGGL((k_apply_grouped<MT, KS, 3>), grid, dim3(256), 0, st,
                       static_cast<const float *>(y_t),
                       static_cast<const unsigned short *>(tbl), ent_ptr,
                       ent_rep, rep_src, rep_s, items, n_items, K, (int)C,
                       static_cast<float *>(vbuf));
  else
    hipLaunchKernelGGL((k_apply_grouped<MT, KS, 1>), grid, dim3(256), 0, st,
                       static_cast<const float *>(y_t),
                       static_cast<const unsigned short *>(tbl), ent_ptr,
                       ent_rep, rep_src, rep_s, items, n_items, K, (int)C,
                       static_cast<float *>(vbuf));
  PG_LAUNCHED();
  return 0;
}

}  // namespace pg

using namespace pg;

extern "C" {

int pg_table_pack_bf16(const void *table, int64_t K, int64_t C,
                       int32_t n_parts, void *packed, pg_stream_t stream) {
  PG_CHECK(K >= 1 && C >= 1 && C <= 64, "table_pack_bf16: bad sizes");
  PG_CHECK(n_parts == 1 || n_parts == 2, "table_pack_bf16: n_parts 1 or 2");
  int MT, KS;
  mfma_shape(C, MT, KS);
  const int64_t total = K * (int64_t)MT * KS * 64 * 8;
  const int threads = 256;
  hipLaunchKernelGGL(k_table_pack_bf16, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<const float2 *>(table), K, (int)C, MT, KS,
                     (int)n_parts, static_cast<unsigned short *>(packed));
  PG_LAUNCHED();
  return 0;
}

int64_t pg_table_bf16_elems(int64_t K, int64_t C, int32_t n_parts) {
  if (C < 1 || C > 64) return -1;
  int MT, KS;
  mfma_shape(C, MT, KS);
  return K * n_parts * (int64_t)MT * KS * 64 * 8;
}

int pg_apply_grouped_bf16(const void *y_t, const void *table_bf16,
                          const uint32_t *ent_ptr, const uint32_t *ent_rep,
                          const uint32_t *rep_src, const float *rep_s,
                          const uint32_t *items, int64_t n_items,
                          int64_t K, int64_t C, int32_t n_terms, void *vbuf,
                          pg_stream_t stream) {
  PG_CHECK(C >= 1 && C <= 64, "grouped apply: coil count must be in [1, 64]");
  PG_CHECK(n_terms == 1 || n_terms == 3, "grouped apply: n_terms 1 or 3");
  PG_CHECK(K >= 1, "grouped apply: empty table");
  const hipStream_t st = as_stream(stream);
  int MT, KS;
  mfma_shape(C, MT, KS);
#define PG_GROUPED_CASE(MTV, KSV)                                            \
  if (MT == MTV && KS == KSV)                                                \
    return launch_grouped<MTV, KSV>(y_t, table_bf16, ent_ptr, ent_rep,       \
                                    rep_src, rep_s, items, n_items, K, C,     \
                                    n_terms, vbuf, st);
  PG_GROUPED_CASE(1, 1)
  PG_GROUPED_CASE(2, 1)
  PG_GROUPED_CASE(3, 2)
  PG_GROUPED_CASE(4, 2)
  PG_GROUPED_CASE(5, 3)
  PG_GROUPED_CASE(6, 3)
  PG_GROUPED_CASE(7, 4)
  PG_GROUPED_CASE(8, 4)
#undef PG_GROUPED_CASE
  pg::set_error("grouped apply: unsupported shape MT=%d KS=%d", MT, KS);
  return 1;
}

}  // extern "C"
