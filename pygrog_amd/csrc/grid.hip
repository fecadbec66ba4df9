// Fused GROG gridding on gfx950: operator apply + deterministic cell-ordered
// segmented reduction, written as a dense coil-major grid.
//
// Data layout (all device-resident, built by the plan compiler):
//   replicas r of a frame are sorted by target cell (stable in (sample,
//   offset) order) and split into "units": runs of at most `chunk` replicas
//   of one cell.  cell_unit[c] is the first unit of cell c, unit_rep[u] the
//   first replica of unit u, unit_cell[u] its cell.
//
// Canonical arithmetic (independent of tile size, grid size, frame mix):
//   v_r       = sum_ci K[e_r][co][ci] * y[ci][src_r], ci ascending, then
//               (v_r * s_r) * s_r             (reference: pre_weights, then
//               the scatter weight, _grog.py:763-821)
//   p_u       = ((0 + v_{r0}) + v_{r0+1}) + ...        over the unit
//   grid[c]   = ((0 + p_{u0}) + p_{u0+1}) + ...        over the cell's units
// so results are bitwise reproducible run to run and alone vs batched.
//
// Work decomposition: a workgroup (256 threads) owns a tile of TC cells.  Its
// lanes form groups of GS >= C lanes (lane = output coil); each group folds
// one unit at a time, the unit partials are staged in LDS, folded per cell
// into an LDS tile accumulator and the tile is written coil-row by coil-row
// (coalesced, zeros included, no read-modify-write, no atomics).  The tile
// loop is persistent and XCD-aware: the 8 XCDs each walk a contiguous range
// of tiles, so neighbouring cells (which share source samples) stay in one
// XCD's L2.
#include "common.h"

namespace pg {

constexpr int kGridThreads = 256;
constexpr int kLdsAcc = 4096;   // GS * TC complex accumulators (32 KiB)
constexpr int kLdsPart = 2048;  // GS * NB complex unit partials (16 KiB)

template <int GS> struct TileCfg {
  static constexpr int TC = kLdsAcc / GS;
  static constexpr int NB = kLdsPart / GS;
};

__device__ __forceinline__ int find_frame(const uint32_t *__restrict__ tp,
                                          int F, uint32_t L) {
  int lo = 0, hi = F - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tp[mid] <= L) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// MODE 0: v = (K[e] y[:, src]) * s * s        (plain FP32 GROG apply)
// MODE 1: v = vbuf[r]                          (grouped modes, second pass)
// MODE 2: v = y[:, src] * s                    (SparseFFT scatter, identity op)
template <int GS, int MODE>
__device__ __forceinline__ float2 replica_value(const pg_frame &fr, uint32_t r,
                                                int lane, int C, int P) {
  float vx = 0.0f, vy = 0.0f;
  if (MODE == 1) {
    if (lane < C)
      return reinterpret_cast<const float2 *>(fr.vbuf)[(int64_t)r * C + lane];
    return make_float2(0.0f, 0.0f);
  }
  const uint32_t src = fr.rep_src[r];
  const float s = fr.rep_s[r];
  const float2 *y = reinterpret_cast<const float2 *>(fr.y);
  if (MODE == 2) {
    if (lane < C) {
      const float2 a = y[(int64_t)lane * fr.n_src + src];
      vx = a.x * s;
      vy = a.y * s;
    }
    return make_float2(vx, vy);
  }
  const uint32_t e = fr.rep_ent[r];
  const float4 *tb =
      reinterpret_cast<const float4 *>(fr.table) + (int64_t)e * P * C + lane;
  if (lane < C) {
#pragma unroll
    for (int p = 0; p < (GS + 1) / 2; ++p) {
      if (p < P) {
        const float4 k = tb[p * C];
        const float2 y0 = y[(int64_t)(2 * p) * fr.n_src + src];
        vx += k.x * y0.x - k.y * y0.y;
        vy += k.x * y0.y + k.y * y0.x;
        if (2 * p + 1 < C) {
          const float2 y1 = y[(int64_t)(2 * p + 1) * fr.n_src + src];
          vx += k.z * y1.x - k.w * y1.y;
          vy += k.z * y1.y + k.w * y1.x;
        }
      }
    }
  }
  vx = vx * s;
  vy = vy * s;
  return make_float2(vx * s, vy * s);
}

template <int GS, int MODE>
__global__ __launch_bounds__(kGridThreads) void k_grid(
    const pg_frame *__restrict__ frames, const uint32_t *__restrict__ tile_ptr,
    int F, int C, uint32_t n_tiles) {
  constexpr int TC = TileCfg<GS>::TC;
  constexpr int NB = TileCfg<GS>::NB;
  constexpr int NG = kGridThreads / GS;
  __shared__ float2 acc[GS * TC];
  __shared__ float2 part[NB * GS];
  const int lane = threadIdx.x % GS;
  const int grp = threadIdx.x / GS;
  const int P = (C + 1) / 2;

  // XCD-aware persistent order: iteration `it` of block b runs on XCD
  // it % 8 (gridDim.x is a multiple of 8), which owns tiles
  // [xcd * per_xcd, (xcd + 1) * per_xcd).
  const uint32_t per_xcd = (n_tiles + 7) / 8;
  const uint32_t n_it = per_xcd * 8;
  for (uint32_t it = blockIdx.x; it < n_it; it += gridDim.x) {
    const uint32_t L = (it % 8) * per_xcd + it / 8;
    if (L >= n_tiles) continue;
    const int f = find_frame(tile_ptr, F, L);
    const pg_frame fr = frames[f];
    const int64_t cell0 = (int64_t)(L - tile_ptr[f]) * TC;
    const int64_t rem = fr.n_cell - cell0;
    const int ncell = rem < TC ? (int)rem : TC;
    const uint32_t u0 = fr.cell_unit[cell0];
    const uint32_t u1 = fr.cell_unit[cell0 + ncell];

    for (int i = threadIdx.x; i < GS * TC; i += kGridThreads)
      acc[i] = make_float2(0.0f, 0.0f);
    __syncthreads();

    for (uint32_t ub = u0; ub < u1; ub += NB) {
      const int nbat = (u1 - ub) < (uint32_t)NB ? (int)(u1 - ub) : NB;
      for (int j = grp; j < nbat; j += NG) {
        const uint32_t u = ub + j;
        const uint32_t r0 = fr.unit_rep[u], r1 = fr.unit_rep[u + 1];
        float px = 0.0f, py = 0.0f;
        for (uint32_t r = r0; r < r1; ++r) {
          const float2 v = replica_value<GS, MODE>(fr, r, lane, C, P);
          px = px + v.x;
          py = py + v.y;
        }
        part[j * GS + lane] = make_float2(px, py);
      }
      __syncthreads();
      // Fold unit partials into the cell accumulators, units of a cell in
      // order (one thread per (run head, coil)).
      for (int i = threadIdx.x; i < nbat * GS; i += kGridThreads) {
        const int j = i / GS;
        const int co = i % GS;
        if (co >= C) continue;
        const uint32_t cell = fr.unit_cell[ub + j];
        if (j > 0 && fr.unit_cell[ub + j - 1] == cell) continue;
        const int t = (int)((int64_t)cell - cell0);
        float2 a = acc[co * TC + t];
        for (int jj = j; jj < nbat && fr.unit_cell[ub + jj] == cell; ++jj) {
          const float2 p = part[jj * GS + co];
          a.x = a.x + p.x;
          a.y = a.y + p.y;
        }
        acc[co * TC + t] = a;
      }
      __syncthreads();
    }
    float2 *grid = reinterpret_cast<float2 *>(fr.grid);
    for (int i = threadIdx.x; i < C * TC; i += kGridThreads) {
      const int co = i / TC;
      const int t = i - co * TC;
      if (t < ncell) grid[(int64_t)co * fr.n_cell + cell0 + t] = acc[co * TC + t];
    }
    __syncthreads();
  }
}

// packed[e][p][co] = (K[e][co][2p], K[e][co][2p+1])
__global__ void k_table_pack(const float2 *__restrict__ K, int64_t nK, int C,
                             float4 *__restrict__ packed) {
  const int P = (C + 1) / 2;
  const int64_t total = nK * P * C;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / ((int64_t)P * C);
    const int rem = (int)(t - e * P * C);
    const int p = rem / C;
    const int co = rem - p * C;
    const float2 *row = K + (e * C + co) * C;
    const float2 a = row[2 * p];
    const float2 b = (2 * p + 1 < C) ? row[2 * p + 1] : make_float2(0.f, 0.f);
    packed[t] = make_float4(a.x, a.y, b.x, b.y);
  }
}

// (C, N) -> (N, C) through a 32x33 LDS tile.
__global__ void k_transpose_cn(const float2 *__restrict__ y, int64_t C,
                               int64_t N, float2 *__restrict__ yt) {
  __shared__ float2 tile[32][33];
  const int64_t n0 = (int64_t)blockIdx.x * 32;
  const int64_t c0 = (int64_t)blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    const int64_t c = c0 + k, n = n0 + threadIdx.x;
    if (c < C && n < N) tile[k][threadIdx.x] = y[c * N + n];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    const int64_t n = n0 + k, c = c0 + threadIdx.x;
    if (c < C && n < N) yt[n * C + c] = tile[threadIdx.x][k];
  }
}

// out[b, co, m] = K[idx[m]][co] . y[b, :, m / kw]
__global__ void k_sparse_apply(const float2 *__restrict__ y,
                               const int64_t *__restrict__ idx,
                               const float2 *__restrict__ K,
                               float2 *__restrict__ out, int64_t B, int64_t C,
                               int64_t N, int64_t kw, int64_t nK) {
  const int64_t M = N * kw;
  const int64_t total = B * C * M;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t bc = t / M;
    const int64_t m = t - bc * M;
    const int64_t b = bc / C;
    const int64_t co = bc - b * C;
    const int64_t n = m / kw;
    const int64_t e = idx[m];
    float2 acc = make_float2(0.f, 0.f);
    if (e >= 0 && e < nK) {
      const float2 *row = K + (e * C + co) * C;
      const float2 *ys = y + b * C * N + n;
      for (int64_t ci = 0; ci < C; ++ci) cmac<float>(acc, row[ci], ys[ci * N]);
    }
    out[t] = acc;
  }
}

// out[b, m] = w[inv[m]] * grid[b, idx[inv[m]]]  (gather + unsort fused)
__global__ void k_gather_perm(const float2 *__restrict__ grid, int64_t G,
                              const int64_t *__restrict__ idx,
                              const float *__restrict__ w,
                              const int64_t *__restrict__ inv, int64_t n,
                              int64_t B, float2 *__restrict__ out) {
  const int64_t total = B * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / n;
    const int64_t m = t - b * n;
    const int64_t i = inv[m];
    const int64_t k = idx[i];
    float2 r = make_float2(0.f, 0.f);
    if (k >= 0 && k < G) {
      const float wi = w[i];
      const float2 g = grid[b * G + k];
      r = make_float2(wi * g.x, wi * g.y);
    }
    out[t] = r;
  }
}

__global__ void k_unit_count(const uint32_t *__restrict__ cp, int64_t S,
                             int64_t G, int chunk,
                             uint32_t *__restrict__ count) {
  const int64_t total = S * G;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / G, c = t - s * G;
    const uint32_t *p = cp + s * (G + 1);
    count[t] = (p[c + 1] - p[c] + chunk - 1) / chunk;
  }
}

template <int MODE>
static int launch_grid(const pg_frame *frames, const uint32_t *tile_ptr,
                       int64_t F, int64_t n_tiles, int64_t C,
                       int64_t tile_cells, int64_t n_blocks, pg_stream_t s) {
  PG_CHECK(C >= 1 && C <= 64, "grid: coil count must be in [1, 64], got %lld",
           (long long)C);
  PG_CHECK(F >= 1 && frames && tile_ptr, "grid: empty frame list");
  PG_CHECK(n_tiles >= 0 && n_tiles < (int64_t)UINT32_MAX, "grid: bad tiles");
  if (n_tiles == 0) return 0;
  int64_t nb = n_blocks > 0 ? n_blocks : 2048;
  if (nb > n_tiles) nb = n_tiles;
  nb = ((nb + 7) / 8) * 8;
  const hipStream_t st = as_stream(s);
  const dim3 grid((unsigned)nb), block(kGridThreads);
#define PG_GRID_CASE(GSV)                                                  \
  if (C <= GSV) {                                                          \
    PG_CHECK(tile_cells == TileCfg<GSV>::TC,                               \
             "grid: tile_cells must be %d for C=%lld", TileCfg<GSV>::TC,   \
             (long long)C);                                                \
    hipLaunchKernelGGL((k_grid<GSV, MODE>), grid, block, 0, st, frames,    \
                       tile_ptr, (int)F, (int)C, (uint32_t)n_tiles);       \
    PG_LAUNCHED();                                                         \
    return 0;                                                              \
  }
  PG_GRID_CASE(4)
  PG_GRID_CASE(8)
  PG_GRID_CASE(16)
  PG_GRID_CASE(32)
  PG_GRID_CASE(64)
#undef PG_GRID_CASE
  return 1;
}

}  // namespace pg

using namespace pg;

extern "C" {

int pg_grid_tile_cells(int64_t C) {
  if (C <= 4) return TileCfg<4>::TC;
  if (C <= 8) return TileCfg<8>::TC;
  if (C <= 16) return TileCfg<16>::TC;
  if (C <= 32) return TileCfg<32>::TC;
  if (C <= 64) return TileCfg<64>::TC;
  return 0;
}

int pg_table_pack_c64(const void *table, int64_t K, int64_t C, void *packed,
                      pg_stream_t stream) {
  PG_CHECK(K >= 1 && C >= 1 && C <= 64, "table_pack: bad sizes");
  const int64_t total = K * ((C + 1) / 2) * C;
  const int threads = 256;
  hipLaunchKernelGGL(k_table_pack, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<const float2 *>(table), K, (int)C,
                     static_cast<float4 *>(packed));
  PG_LAUNCHED();
  return 0;
}

int pg_grid_fp32(const pg_frame *frames, const uint32_t *tile_ptr, int64_t F,
                 int64_t n_tiles, int64_t C, int64_t tile_cells,
                 int64_t n_blocks, pg_stream_t stream) {
  return launch_grid<0>(frames, tile_ptr, F, n_tiles, C, tile_cells, n_blocks,
                        stream);
}

int pg_grid_reduce_v(const pg_frame *frames, const uint32_t *tile_ptr,
                     int64_t F, int64_t n_tiles, int64_t C,
                     int64_t tile_cells, int64_t n_blocks,
                     pg_stream_t stream) {
  return launch_grid<1>(frames, tile_ptr, F, n_tiles, C, tile_cells, n_blocks,
                        stream);
}

int pg_grid_scatter(const pg_frame *frames, const uint32_t *tile_ptr,
                    int64_t F, int64_t n_tiles, int64_t C, int64_t tile_cells,
                    int64_t n_blocks, pg_stream_t stream) {
  return launch_grid<2>(frames, tile_ptr, F, n_tiles, C, tile_cells, n_blocks,
                        stream);
}

int pg_transpose_cn_c64(const void *y, int64_t C, int64_t N, void *y_t,
                        pg_stream_t stream) {
  PG_CHECK(C >= 0 && N >= 0 && C < 65535 * 32, "transpose: bad sizes");
  if (C == 0 || N == 0) return 0;
  const dim3 grid((unsigned)ceil_div(N, 32), (unsigned)ceil_div(C, 32));
  hipLaunchKernelGGL(k_transpose_cn, grid, dim3(32, 8), 0, as_stream(stream),
                     static_cast<const float2 *>(y), C, N,
                     static_cast<float2 *>(y_t));
  PG_LAUNCHED();
  return 0;
}

int pg_sparse_apply_c64(const void *y, const int64_t *interp_idx,
                        const void *table, void *out, int64_t B, int64_t C,
                        int64_t N, int64_t kw, int64_t K, pg_stream_t stream) {
  PG_CHECK(B >= 0 && C >= 0 && N >= 0 && kw >= 1 && K >= 1,
           "sparse_apply: bad sizes");
  const int64_t total = B * C * N * kw;
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_sparse_apply, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<const float2 *>(y), interp_idx,
                     static_cast<const float2 *>(table),
                     static_cast<float2 *>(out), B, C, N, kw, K);
  PG_LAUNCHED();
  return 0;
}

int pg_gather_perm_c64(const void *grid, int64_t G, const int64_t *indices,
                       const float *weights, const int64_t *inv_perm,
                       int64_t n, int64_t B, void *out, pg_stream_t stream) {
  PG_CHECK(G >= 0 && n >= 0 && B >= 0, "gather_perm: bad sizes");
  const int64_t total = B * n;
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_gather_perm, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<const float2 *>(grid), G, indices, weights,
                     inv_perm, n, B, static_cast<float2 *>(out));
  PG_LAUNCHED();
  return 0;
}

int pg_unit_count(const uint32_t *cell_ptr, int64_t S, int64_t G,
                  int32_t chunk, uint32_t *unit_count, pg_stream_t stream) {
  PG_CHECK(chunk >= 1, "unit_count: chunk must be positive");
  const int64_t total = S * G;
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_unit_count, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream), cell_ptr, S, G,
                     chunk, unit_count);
  PG_LAUNCHED();
  return 0;
}

}  // extern "C"
