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
// A lane's B values of one K-step are 8 consecutive values of one sample
// row, i.e. exactly one 16-byte group of the block-floating-point rows
// (RF = kRowBfp), so every lane decodes its own group.  Sample rows come as
// complex64 (kRowC64), 24-bit (kRowS24) or bfp15 (kRowBfp) rows; V rows are
// written as complex64, or as bfp15 for kRowBfp.
#include "common.h"

namespace pg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float float2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float(((uint32_t)h) << 16);
}

// Compact complex table: packed[e][part][co][ci] = (re, im) as two bf16,
// co < CO = 8 MT, ci < CI = 16 KS (zero padded).  part 0 = hi = bf16(x),
// part 1 = lo = bf16(x - hi).  A lane's MFMA A fragment (row 16 mt + j,
// k = 32 ks + 8q .. +7) is the 4 complex values K[co][16 ks + 4q .. +3] of
// co = 8 mt + j/2, expanded in registers: even rows (Re of the output)
// take (re, -im) pairs, odd rows (Im) take (im, re) pairs.  Negating a bf16
// is exact and both rounding splits are odd functions, so these are the
// same bf16 values as a table stored in real 2C x 2C form, at half the bytes.
__global__ void k_table_pack_bf16(const float2 *__restrict__ K, int64_t nK,
                                  int C, int CO, int CI, int n_parts,
                                  unsigned short *__restrict__ packed) {
  const int64_t per_part = (int64_t)CO * CI * 2;
  const int64_t total = nK * (int64_t)CO * CI;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / ((int64_t)CO * CI);
    const int rem = (int)(t - e * CO * CI);
    const int co = rem / CI, ci = rem - (rem / CI) * CI;
    float2 k = make_float2(0.0f, 0.0f);
    if (co < C && ci < C) k = K[(e * C + co) * C + ci];
    const unsigned short hr = f2bf(k.x), hi_ = f2bf(k.y);
    unsigned short *dst = packed + e * n_parts * per_part + (int64_t)rem * 2;
    dst[0] = hr;
    dst[1] = hi_;
    if (n_parts > 1) {
      dst[per_part] = f2bf(k.x - bf2f(hr));
      dst[per_part + 1] = f2bf(k.y - bf2f(hi_));
    }
  }
}

#ifndef PG_APPLY_TILES
#define PG_APPLY_TILES 8
#endif
constexpr int kTilesPerWave = PG_APPLY_TILES;  // 16-replica N tiles per wavefront
// One wavefront (one work item) per workgroup: items run 1-8 tiles, and a
// workgroup of four holds its slot until its longest item ends (measured on
// MI355X, kooshball 192^3 x 32 coils, bf16x3c: apply 0.79 -> 0.68 ms).
#ifndef PG_APPLY_WPB
#define PG_APPLY_WPB 1
#endif
constexpr int kApplyWaves = PG_APPLY_WPB;  // wavefronts (work items) per workgroup

template <int MT, int KS, int NTERMS, int RF>
__global__ __launch_bounds__(64 * kApplyWaves) void k_apply_grouped(
    const void *__restrict__ y_t, const unsigned short *__restrict__ tbl,
    const uint32_t *__restrict__ ent_ptr, const uint32_t *__restrict__ ent_rep,
    const uint32_t *__restrict__ ent_src, const float *__restrict__ ent_s,
    const uint32_t *__restrict__ items, int64_t n_items, int64_t nK, int C,
    float *__restrict__ vbuf) {
  constexpr int NP = (NTERMS > 1) ? 2 : 1;
  const int wave = threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  const int64_t item = (int64_t)blockIdx.x * kApplyWaves + wave;
  if (item >= n_items) return;
  const int64_t e = items[2 * item];
  if (e >= nK) return;
  const uint32_t b1 = ent_ptr[e + 1];
  const uint32_t start = ent_ptr[e] + items[2 * item + 1];
  if (start >= b1) return;
  uint32_t stop = start + 16 * kTilesPerWave;
  if (stop > b1) stop = b1;
  const int twoC = 2 * C;
  constexpr int CO = 8 * MT, CI = 16 * KS;
  constexpr int64_t per_part = (int64_t)CO * CI * 2;
  // per-wavefront staging of one 16-replica output tile (rows of 16 MT
  // floats, padded): the tile is written to LDS in the MFMA layout and read
  // back so that 16 consecutive lanes store one replica's whole V row
  constexpr int RLP = 16 * MT + 4;
  __shared__ __align__(16) float s_out[kApplyWaves][16 * RLP];
  float *vt = s_out[threadIdx.x / 64];

  bf16x8_t ahi[MT][KS];
  bf16x8_t alo[MT][KS];
  const unsigned short *tb = tbl + e * NP * per_part;
  {
    const int jj = lane % 16, qq = lane / 16;
    const bool im_row = (jj & 1) != 0;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int co = 8 * mt + (jj >> 1);
        const int64_t off = ((int64_t)co * CI + 16 * ks + 4 * qq) * 2;
#pragma unroll
        for (int part = 0; part < NP; ++part) {
          const u16x8_t c = *reinterpret_cast<const u16x8_t *>(tb + part * per_part + off);
          u16x8_t a;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const unsigned short re = c[2 * i], im = c[2 * i + 1];
            a[2 * i] = im_row ? im : re;
            a[2 * i + 1] = im_row ? re : (unsigned short)(im ^ 0x8000u);
          }
          if (part == 0) ahi[mt][ks] = __builtin_bit_cast(bf16x8_t, a);
          else alo[mt][ks] = __builtin_bit_cast(bf16x8_t, a);
        }
      }
  }

  const int j = lane % 16, q = lane / 16;
  // every tile's replica, source and weight for this lane's column j: one
  // round of independent, contiguous loads for the whole work item (the
  // records are stored in entry order next to ent_rep)
  uint32_t rr[kTilesPerWave], src[kTilesPerWave];
  float sw[kTilesPerWave];
#pragma unroll
  for (int t = 0; t < kTilesPerWave; ++t) {
    const uint32_t idx = start + 16 * t + j;
    const bool v = idx < stop;
    rr[t] = v ? ent_rep[idx] : 0u;
    src[t] = v ? ent_src[idx] : 0u;
    sw[t] = v ? ent_s[idx] : 0.0f;
  }
  const bool vec = (twoC & 3) == 0;
  // sample-row fragments of one tile: KS x 8 floats per lane
  // 24-bit rows stay packed until their tile is used: 8 values = 24 bytes
  // = 6 dwords at byte 3 col0 (8-byte aligned); past the row's 2C values
  // the bytes are zero (value 0)
  auto load_packed = [&](int t, uint32_t (&d)[KS][6]) {
    const bool v = start + 16 * t + j < stop;
    const unsigned char *yr = static_cast<const unsigned char *>(y_t) + (int64_t)src[t] * s24_row_bytes(C);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int col0 = 32 * ks + 8 * q;
      if (v && col0 + 8 <= twoC) {
        const uint2 *p = reinterpret_cast<const uint2 *>(yr + 3 * col0);
        const uint2 a = gld(p), b = gld(p + 1), c = gld(p + 2);
        d[ks][0] = a.x; d[ks][1] = a.y; d[ks][2] = b.x;
        d[ks][3] = b.y; d[ks][4] = c.x; d[ks][5] = c.y;
      } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          uint32_t w = 0;
#pragma unroll
          for (int by = 0; by < 4; ++by) {
            const int o = 3 * col0 + 4 * k + by;
            if (v && o < 3 * twoC) w |= (uint32_t)gld(yr + o) << (8 * by);
          }
          d[ks][k] = w;
        }
      }
    }
  };
  // block-floating-point rows: 8 values = one 16-byte group at byte
  // 2 col0; groups past the row's 2C values do not exist (value 0)
  auto load_bfp = [&](int t, uint32_t (&d)[KS][4]) {
    const bool v = start + 16 * t + j < stop;
    const unsigned char *yr = static_cast<const unsigned char *>(y_t) + (int64_t)src[t] * bfp_row_bytes(C);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int col0 = 32 * ks + 8 * q;
      if (v && col0 < twoC) {
        const uint4 a = gld(reinterpret_cast<const uint4 *>(yr + 2 * col0));
        d[ks][0] = a.x; d[ks][1] = a.y; d[ks][2] = a.z; d[ks][3] = a.w;
      } else {
        d[ks][0] = d[ks][1] = d[ks][2] = d[ks][3] = 0u;  // E = 0, m = 0
      }
    }
  };
  auto load_rows = [&](int t, float (&f)[KS][8]) {
    const bool v = start + 16 * t + j < stop;
    const float *yr = static_cast<const float *>(y_t) + (int64_t)src[t] * twoC;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int col0 = 32 * ks + 8 * q;
      if (v && vec && col0 + 8 <= twoC) {
        const float4 a0 = *reinterpret_cast<const float4 *>(yr + col0);
        const float4 a1 = *reinterpret_cast<const float4 *>(yr + col0 + 4);
        f[ks][0] = a0.x; f[ks][1] = a0.y; f[ks][2] = a0.z; f[ks][3] = a0.w;
        f[ks][4] = a1.x; f[ks][5] = a1.y; f[ks][6] = a1.z; f[ks][7] = a1.w;
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i)
          f[ks][i] = (v && col0 + i < twoC) ? yr[col0 + i] : 0.0f;
      }
    }
  };
  const int n_tiles = (int)((stop - start + 15) / 16);
  // sample-row fragments DEPTH tiles ahead (DEPTH + 1 buffers in rotation;
  // two ahead up to 32 coils, where the registers allow it)
#ifndef PG_APPLY_DEPTH
#define PG_APPLY_DEPTH 2
#endif
#ifndef PG_APPLY_DEPTH24
#define PG_APPLY_DEPTH24 2
#endif
#ifndef PG_APPLY_DEPTHB
#define PG_APPLY_DEPTHB 3
#endif
  constexpr int DEPTH = KS <= 2 ? (RF == kRowBfp ? PG_APPLY_DEPTHB : RF == kRowS24 ? PG_APPLY_DEPTH24 : PG_APPLY_DEPTH) : 1;
  constexpr int NBUF = DEPTH + 1;
  float fr_[NBUF][KS][8];     // fp32 rows
  uint32_t fs_[NBUF][KS][6];  // 24-bit rows, packed
  uint32_t fb_[NBUF][KS][4];  // block-floating-point rows, packed
  auto load = [&](int t, int buf) {
    if constexpr (RF == kRowS24) load_packed(t, fs_[buf]);
    else if constexpr (RF == kRowBfp) load_bfp(t, fb_[buf]);
    else load_rows(t, fr_[buf]);
  };
#pragma unroll
  for (int d = 0; d < DEPTH; ++d)
    if (d < n_tiles) load(d, d);
#pragma unroll
  for (int t = 0; t < kTilesPerWave; ++t) {
    if (t >= n_tiles) break;
    float cur[KS][8];
    if constexpr (RF == kRowBfp) {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const uint32_t(&d)[4] = fb_[t % NBUF][ks];
        bfp_decode8(d[0], d[1], d[2], d[3], &cur[ks][0]);
      }
    } else if constexpr (RF == kRowS24) {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const uint32_t(&d)[6] = fs_[t % NBUF][ks];
        s24_unpack4(d[0], d[1], d[2], &cur[ks][0]);
        s24_unpack4(d[3], d[4], d[5], &cur[ks][4]);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int i = 0; i < 8; ++i) cur[ks][i] = fr_[t % NBUF][ks][i];
    }
    if (t + DEPTH < n_tiles) load(t + DEPTH, (t + DEPTH) % NBUF);
    floatx4_t acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] = floatx4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      // hi = bf16(x), lo = bf16(x - hi), both round to nearest even: the
      // hardware conversion (v_cvt_pk_bf16_f32) gives the bits of f2bf.
      // One packed conversion per pair for hi; the pair's fp32 hi values
      // are its two halves moved to the top of a word
      uint32_t hp[4], lp[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float x0 = cur[ks][2 * i], x1 = cur[ks][2 * i + 1];
        hp[i] = __builtin_bit_cast(uint32_t, __builtin_convertvector((float2_t){x0, x1}, bf16x2_t));
        const float h0 = __uint_as_float(hp[i] << 16), h1 = __uint_as_float(hp[i] & 0xffff0000u);
        lp[i] = __builtin_bit_cast(uint32_t, __builtin_convertvector((float2_t){x0 - h0, x1 - h1}, bf16x2_t));
      }
      const bf16x8_t bhi = __builtin_bit_cast(bf16x8_t, (u32x4_t){hp[0], hp[1], hp[2], hp[3]});
      const bf16x8_t blo = __builtin_bit_cast(bf16x8_t, (u32x4_t){lp[0], lp[1], lp[2], lp[3]});
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
    if constexpr (RF == kRowBfp) {
      // block-floating-point V rows (any C): through LDS, then one lane per
      // (replica, 4-coil group) encodes 8 values into its 16 bytes;
      // consecutive lanes write consecutive groups of a row
      const float s = sw[t];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        float4 o;
        o.x = (acc[mt][0] * s) * s;
        o.y = (acc[mt][1] * s) * s;
        o.z = (acc[mt][2] * s) * s;
        o.w = (acc[mt][3] * s) * s;
        *reinterpret_cast<float4 *>(&vt[j * RLP + 16 * mt + 4 * q]) = o;
      }
      wave_lds_sync();
      constexpr int NGR = 2 * MT;  // groups per (padded) row
      const int64_t vrb = bfp_row_bytes(C);
#pragma unroll
      for (int k = 0; k < (16 * NGR + 63) / 64; ++k) {
        const int idx = 64 * k + lane;
        const int jr = idx / NGR, gi = idx % NGR;
        const uint32_t rj = __shfl(rr[t], jr < 16 ? jr : 0);
        if (idx < 16 * NGR && 8 * gi < twoC && start + 16 * t + jr < stop) {
          // values past 2C are zero (the table's padded rows)
          const float4 lo4 = *reinterpret_cast<const float4 *>(&vt[jr * RLP + 8 * gi]);
          const float4 hi4 = *reinterpret_cast<const float4 *>(&vt[jr * RLP + 8 * gi + 4]);
          const float x[8] = {lo4.x, lo4.y, lo4.z, lo4.w, hi4.x, hi4.y, hi4.z, hi4.w};
          const uint4 w = bfp_encode8(x);
          gst(reinterpret_cast<uint4 *>(static_cast<unsigned char *>(static_cast<void *>(vbuf)) +
                                        (int64_t)rj * vrb + 16 * gi), w);
        }
      }
      wave_lds_sync();
    } else if (vec) {
      // even C: through LDS, then whole rows (16 lanes x 16 bytes per
      // replica row piece, consecutive lanes on consecutive bytes)
      const float s = sw[t];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        float4 o;
        o.x = (acc[mt][0] * s) * s;
        o.y = (acc[mt][1] * s) * s;
        o.z = (acc[mt][2] * s) * s;
        o.w = (acc[mt][3] * s) * s;
        *reinterpret_cast<float4 *>(&vt[j * RLP + 16 * mt + 4 * q]) = o;
      }
      wave_lds_sync();
      constexpr int CH = 4 * MT;  // float4 pieces per (padded) row
#pragma unroll
      for (int k = 0; k < CH / 4; ++k) {
        const int idx = 64 * k + lane;
        const int jr = idx / CH, ch = idx % CH;
        const uint32_t rj = __shfl(rr[t], jr);
        if (4 * ch < twoC && start + 16 * t + jr < stop)
          gst(reinterpret_cast<float4 *>(vbuf + (int64_t)rj * twoC + 4 * ch),
              *reinterpret_cast<const float4 *>(&vt[jr * RLP + 4 * ch]));
      }
      wave_lds_sync();
    } else if (start + 16 * t + j < stop) {
      const float s = sw[t];
      float *vr = vbuf + (int64_t)rr[t] * twoC;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const int row = 16 * mt + 4 * q;
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (row + i < twoC) vr[row + i] = (acc[mt][i] * s) * s;
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
                          const uint32_t *ent_src, const float *ent_s,
                          const uint32_t *items, int64_t n_items, int64_t K,
                          int64_t C, int32_t n_terms, int rf, void *vbuf,
                          hipStream_t st) {
  if (n_items == 0) return 0;
  PG_CHECK(items != nullptr, "grouped apply: null work list");
  PG_CHECK(rf == kRowC64 || ((uintptr_t)y_t & 15) == 0, "grouped apply: packed rows not 16-byte aligned");
  PG_CHECK(rf != kRowBfp || ((uintptr_t)vbuf & 15) == 0, "grouped apply: packed V rows not 16-byte aligned");
  const dim3 grid((unsigned)ceil_div(n_items, kApplyWaves));
#define PG_APPLY_LAUNCH(NT, S)                                                   \
  hipLaunchKernelGGL((k_apply_grouped<MT, KS, NT, S>), grid, dim3(64 * kApplyWaves), 0, st, \
                     y_t, static_cast<const unsigned short *>(tbl), ent_ptr,      \
                     ent_rep, ent_src, ent_s, items, n_items, K, (int)C,          \
                     static_cast<float *>(vbuf))
  if (n_terms == 3) {
    if (rf == kRowS24) PG_APPLY_LAUNCH(3, kRowS24);
    else if (rf == kRowBfp) PG_APPLY_LAUNCH(3, kRowBfp);
    else PG_APPLY_LAUNCH(3, kRowC64);
  } else {
    if (rf == kRowS24) PG_APPLY_LAUNCH(1, kRowS24);
    else if (rf == kRowBfp) PG_APPLY_LAUNCH(1, kRowBfp);
    else PG_APPLY_LAUNCH(1, kRowC64);
  }
#undef PG_APPLY_LAUNCH
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
  const int64_t total = K * (int64_t)(8 * MT) * (16 * KS);
  const int threads = 256;
  hipLaunchKernelGGL(k_table_pack_bf16, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<const float2 *>(table), K, (int)C, 8 * MT,
                     16 * KS, (int)n_parts, static_cast<unsigned short *>(packed));
  PG_LAUNCHED();
  return 0;
}

int64_t pg_table_bf16_elems(int64_t K, int64_t C, int32_t n_parts) {
  if (C < 1 || C > 64) return -1;
  int MT, KS;
  mfma_shape(C, MT, KS);
  return K * n_parts * (int64_t)(8 * MT) * (16 * KS) * 2;
}

static int apply_grouped(const void *y_t, const void *table_bf16,
                         const uint32_t *ent_ptr, const uint32_t *ent_rep,
                         const uint32_t *ent_src, const float *ent_s,
                         const uint32_t *items, int64_t n_items, int64_t K,
                         int64_t C, int32_t n_terms, int rf, void *vbuf,
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
                                    ent_src, ent_s, items, n_items, K, C,     \
                                    n_terms, rf, vbuf, st);
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

int pg_apply_grouped_bf16(const void *y_t, const void *table_bf16,
                          const uint32_t *ent_ptr, const uint32_t *ent_rep,
                          const uint32_t *ent_src, const float *ent_s,
                          const uint32_t *items, int64_t n_items,
                          int64_t K, int64_t C, int32_t n_terms, void *vbuf,
                          pg_stream_t stream) {
  return apply_grouped(y_t, table_bf16, ent_ptr, ent_rep, ent_src, ent_s, items,
                       n_items, K, C, n_terms, kRowC64, vbuf, stream);
}

int pg_apply_grouped_bf16_s24(const void *y_s, const void *table_bf16,
                              const uint32_t *ent_ptr, const uint32_t *ent_rep,
                              const uint32_t *ent_src, const float *ent_s,
                              const uint32_t *items, int64_t n_items,
                              int64_t K, int64_t C, int32_t n_terms, void *vbuf,
                              pg_stream_t stream) {
  return apply_grouped(y_s, table_bf16, ent_ptr, ent_rep, ent_src, ent_s, items,
                       n_items, K, C, n_terms, kRowS24, vbuf, stream);
}

int pg_apply_grouped_bf16_bfp(const void *y_b, const void *table_bf16,
                              const uint32_t *ent_ptr, const uint32_t *ent_rep,
                              const uint32_t *ent_src, const float *ent_s,
                              const uint32_t *items, int64_t n_items,
                              int64_t K, int64_t C, int32_t n_terms, void *vbuf,
                              pg_stream_t stream) {
  return apply_grouped(y_b, table_bf16, ent_ptr, ent_rep, ent_src, ent_s, items,
                       n_items, K, C, n_terms, kRowBfp, vbuf, stream);
}

}  // extern "C"
