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
// loop is persistent (grid-stride over tiles).  Cells with many replicas (the
// radial DC and its neighbours) are reduced by a two-kernel pre-pass -- hot
// units in parallel, then an in-order fold per hot cell -- whose sums the
// tile kernel copies, so no workgroup serialises on a hot cell.
#include "common.h"

namespace pg {

constexpr int kGridThreads = 256;
constexpr int kLdsAcc = 2048;   // GS * TC complex accumulators (16 KiB)
constexpr int kLdsPart = 1024;  // GS * NB complex unit partials (8 KiB)
constexpr int kLdsRows = 2048;  // GS * RB staged V rows, MODE 1 (16 KiB)

template <int GS> struct TileCfg {
  static constexpr int TC = kLdsAcc / GS;  // cells per tile
  static constexpr int NB = kLdsPart / GS; // units per batch
  static constexpr int RB = kLdsRows / GS; // staged rows per batch (MODE 1)
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

// MODE 0: v = (K[e] y_t[src, :]) * s * s     (plain FP32 GROG apply; y is
//         sample-major (n_src, C): one contiguous row per replica)
// MODE 1: v = vbuf[r]                          (grouped modes, second pass)
// MODE 2: v = y[:, src] * s                    (SparseFFT scatter, identity op,
//         y coil-major (C, n_src))
struct Rec {
  uint32_t src, ent;
  float s;
};

template <int MODE>
__device__ __forceinline__ Rec load_rec(const pg_frame &fr, uint32_t r) {
  Rec q{0u, 0u, 0.0f};
  if (MODE != 1) {
    q.src = fr.rep_src[r];
    q.s = fr.rep_s[r];
    if (MODE == 0) q.ent = fr.rep_ent[r];
  }
  return q;
}

template <int GS, int MODE>
__device__ __forceinline__ float2 replica_value(const pg_frame &fr, uint32_t r,
                                                const Rec &q, int lane, int C,
                                                int P, const float2 *rows,
                                                uint32_t rbase) {
  float vx = 0.0f, vy = 0.0f;
  if (MODE == 1) {
    if (lane < C) {
      if (rows != nullptr) return rows[(r - rbase) * GS + lane];
      return reinterpret_cast<const float2 *>(fr.vbuf)[(int64_t)r * C + lane];
    }
    return make_float2(0.0f, 0.0f);
  }
  if (MODE == 2) {
    if (lane < C) {
      const float2 a =
          reinterpret_cast<const float2 *>(fr.y)[(int64_t)lane * fr.n_src + q.src];
      vx = a.x * q.s;
      vy = a.y * q.s;
    }
    return make_float2(vx, vy);
  }
  const float4 *tb = reinterpret_cast<const float4 *>(fr.table) +
                     (int64_t)q.ent * P * C + lane;
  const float2 *yr = reinterpret_cast<const float2 *>(fr.y) + (int64_t)q.src * C;
  if (lane < C) {
    if ((C & 1) == 0) {
      // row of C complex = P float4 pairs (16-byte aligned for even C)
      const float4 *y4 = reinterpret_cast<const float4 *>(yr);
#pragma unroll
      for (int p = 0; p < (GS + 1) / 2; ++p) {
        if (p < P) {
          const float4 k = tb[p * C];
          const float4 yy = y4[p];
          vx += k.x * yy.x - k.y * yy.y;
          vy += k.x * yy.y + k.y * yy.x;
          vx += k.z * yy.z - k.w * yy.w;
          vy += k.z * yy.w + k.w * yy.z;
        }
      }
    } else {
#pragma unroll
      for (int p = 0; p < (GS + 1) / 2; ++p) {
        if (p < P) {
          const float4 k = tb[p * C];
          const float2 y0 = yr[2 * p];
          vx += k.x * y0.x - k.y * y0.y;
          vy += k.x * y0.y + k.y * y0.x;
          if (2 * p + 1 < C) {
            const float2 y1 = yr[2 * p + 1];
            vx += k.z * y1.x - k.w * y1.y;
            vy += k.z * y1.y + k.w * y1.x;
          }
        }
      }
    }
  }
  vx = vx * q.s;
  vy = vy * q.s;
  return make_float2(vx * q.s, vy * q.s);
}

// Canonical fold of one unit: ((0 + v_r0) + v_r0+1) + ... ; the next
// replica's record is fetched while the current one is applied.
template <int GS, int MODE>
__device__ __forceinline__ float2 unit_partial(const pg_frame &fr, uint32_t u,
                                               int lane, int C, int P,
                                               const float2 *rows = nullptr,
                                               uint32_t rbase = 0) {
  const uint32_t r0 = fr.unit_rep[u], r1 = fr.unit_rep[u + 1];
  float px = 0.0f, py = 0.0f;
  if (r0 >= r1) return make_float2(px, py);
  if (MODE == 1 && rows == nullptr) {
    // contiguous rows: issue 8 loads, then add them in order
    const float2 *vb = reinterpret_cast<const float2 *>(fr.vbuf) + lane;
    if (lane >= C) return make_float2(px, py);
    for (uint32_t rb = r0; rb < r1; rb += 8) {
      float2 v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        v[k] = (rb + k < r1) ? vb[(int64_t)(rb + k) * C] : make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (rb + k < r1) {
          px = px + v[k].x;
          py = py + v[k].y;
        }
      }
    }
    return make_float2(px, py);
  }
  Rec q = load_rec<MODE>(fr, r0);
  for (uint32_t r = r0; r < r1; ++r) {
    Rec nq = q;
    if (r + 1 < r1) nq = load_rec<MODE>(fr, r + 1);
    const float2 v = replica_value<GS, MODE>(fr, r, q, lane, C, P, rows, rbase);
    px = px + v.x;
    py = py + v.y;
    q = nq;
  }
  return make_float2(px, py);
}

template <int GS, int MODE>
__global__ __launch_bounds__(kGridThreads) void k_grid(
    const pg_frame *__restrict__ frames, const uint32_t *__restrict__ tile_ptr,
    int F, int C, uint32_t n_tiles) {
  constexpr int TC = TileCfg<GS>::TC;
  constexpr int NB = TileCfg<GS>::NB;
  constexpr int RB = TileCfg<GS>::RB;
  constexpr int NG = kGridThreads / GS;
  __shared__ float2 acc[GS * TC];
  __shared__ float2 part[NB * GS];
  // MODE 1 stages the batch's V rows (contiguous: units are cell-sorted)
  // with one coalesced cooperative load, for memory-level parallelism.
  __shared__ float2 rows_lds[MODE == 1 ? RB * GS : 1];
  const int lane = threadIdx.x % GS;
  const int grp = threadIdx.x / GS;
  const int P = (C + 1) / 2;

  // Round-robin tile order: neighbouring tiles (the dense k-space centre
  // spans many consecutive tiles) land on different workgroups and XCDs.
  for (uint32_t L = blockIdx.x; L < n_tiles; L += gridDim.x) {
    const int f = find_frame(tile_ptr, F, L);
    const pg_frame fr = frames[f];
    const int64_t cell0 = (int64_t)(L - tile_ptr[f]) * TC;
    const int64_t rem = fr.n_cell - cell0;
    const int ncell = rem < TC ? (int)rem : TC;
    const uint32_t u0 = fr.cell_unit[cell0];
    const uint32_t u1 = fr.cell_unit[cell0 + ncell];
    const bool has_hot = fr.n_hot > 0;

    for (int i = threadIdx.x; i < GS * TC; i += kGridThreads)
      acc[i] = make_float2(0.0f, 0.0f);
    __syncthreads();

    for (uint32_t ub = u0; ub < u1; ub += NB) {
      const int nbat = (u1 - ub) < (uint32_t)NB ? (int)(u1 - ub) : NB;
      const float2 *rows = nullptr;
      uint32_t rbase = 0;
      if (MODE == 1) {
        rbase = fr.unit_rep[ub];
        const uint32_t nrows = fr.unit_rep[ub + nbat] - rbase;
        if (nrows <= (uint32_t)RB) {
          const float2 *vb = reinterpret_cast<const float2 *>(fr.vbuf) +
                             (int64_t)rbase * C;
          for (int i = threadIdx.x; i < (int)nrows * C; i += kGridThreads) {
            const int row = i / C, co = i - row * C;
            rows_lds[row * GS + co] = vb[i];
          }
          rows = rows_lds;
          __syncthreads();
        }
      }
      for (int j = grp; j < nbat; j += NG) {
        const uint32_t u = ub + j;
        float2 pv = make_float2(0.0f, 0.0f);
        if (!has_hot || fr.unit_hot[u] < 0)
          pv = unit_partial<GS, MODE>(fr, u, lane, C, P, rows, rbase);
        part[j * GS + lane] = pv;
      }
      __syncthreads();
      // Fold unit partials into the cell accumulators, units of a cell in
      // order (one thread per (run head, coil)); a hot cell takes its sum
      // from the pre-pass once, at the batch holding its first unit.
      for (int i = threadIdx.x; i < nbat * GS; i += kGridThreads) {
        const int j = i / GS;
        const int co = i % GS;
        if (co >= C) continue;
        const uint32_t cell = fr.unit_cell[ub + j];
        if (j > 0 && fr.unit_cell[ub + j - 1] == cell) continue;
        const int t = (int)((int64_t)cell - cell0);
        const int h = has_hot ? fr.unit_hot[ub + j] : -1;
        if (h >= 0) {
          if (ub + j == fr.cell_unit[cell]) {
            const float2 pc =
                reinterpret_cast<const float2 *>(fr.p_cell)[(int64_t)h * C + co];
            float2 a = acc[co * TC + t];
            a.x = a.x + pc.x;
            a.y = a.y + pc.y;
            acc[co * TC + t] = a;
          }
          continue;
        }
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

// Hot pre-pass 1: partial of every unit of every hot cell, in parallel
// (grid.x over hot units, grid.y over frames).
template <int GS, int MODE>
__global__ __launch_bounds__(kGridThreads) void k_hot_units(
    const pg_frame *__restrict__ frames, int C) {
  constexpr int NG = kGridThreads / GS;
  const pg_frame fr = frames[blockIdx.y];
  const int lane = threadIdx.x % GS;
  const int64_t i = (int64_t)blockIdx.x * NG + threadIdx.x / GS;
  if (i >= fr.n_hot_units) return;
  const float2 pv = unit_partial<GS, MODE>(fr, fr.hot_units[i], lane, C, (C + 1) / 2);
  if (lane < C) reinterpret_cast<float2 *>(fr.p_unit)[i * C + lane] = pv;
}

// Hot pre-pass 2: in-order fold of each hot cell's unit partials.  The
// partials are fetched 16 at a time (independent loads in flight) and added
// in order, so the dependent chain is the adds, not the loads.
template <int GS>
__global__ __launch_bounds__(kGridThreads) void k_hot_cells(
    const pg_frame *__restrict__ frames, int C) {
  constexpr int NG = kGridThreads / GS;
  constexpr int KB = 16;
  const pg_frame fr = frames[blockIdx.y];
  const int lane = threadIdx.x % GS;
  const int64_t h = (int64_t)blockIdx.x * NG + threadIdx.x / GS;
  if (h >= fr.n_hot || lane >= C) return;
  const float2 *pu = reinterpret_cast<const float2 *>(fr.p_unit);
  float ax = 0.0f, ay = 0.0f;
  const uint32_t i0 = fr.hot_ptr[h], i1 = fr.hot_ptr[h + 1];
  for (uint32_t ib = i0; ib < i1; ib += KB) {
    float2 v[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k)
      v[k] = (ib + k < i1) ? pu[(int64_t)(ib + k) * C + lane] : make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      if (ib + k < i1) {
        ax = ax + v[k].x;
        ay = ay + v[k].y;
      }
    }
  }
  reinterpret_cast<float2 *>(fr.p_cell)[h * C + lane] = make_float2(ax, ay);
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
                       int64_t tile_cells, int64_t n_blocks,
                       int64_t max_hot_units, int64_t max_hot, pg_stream_t s) {
  PG_CHECK(C >= 1 && C <= 64, "grid: coil count must be in [1, 64], got %lld",
           (long long)C);
  PG_CHECK(F >= 1 && frames && tile_ptr, "grid: empty frame list");
  PG_CHECK(F < 65536, "grid: too many frames in one launch");
  PG_CHECK(n_tiles >= 0 && n_tiles < (int64_t)UINT32_MAX, "grid: bad tiles");
  PG_CHECK(max_hot_units >= 0 && max_hot >= 0, "grid: bad hot counts");
  if (n_tiles == 0) return 0;
  int64_t nb = n_blocks > 0 ? n_blocks : 2048;
  if (nb > n_tiles) nb = n_tiles;
  nb = ((nb + 7) / 8) * 8;
  const hipStream_t st = as_stream(s);
  const dim3 grid((unsigned)nb), block(kGridThreads);
#define PG_GRID_CASE(GSV)                                                    \
  if (C <= GSV) {                                                            \
    PG_CHECK(tile_cells == TileCfg<GSV>::TC,                                 \
             "grid: tile_cells must be %d for C=%lld", TileCfg<GSV>::TC,     \
             (long long)C);                                                  \
    constexpr int NG = kGridThreads / GSV;                                   \
    if (max_hot_units > 0) {                                                 \
      hipLaunchKernelGGL((k_hot_units<GSV, MODE>),                           \
                         dim3((unsigned)ceil_div(max_hot_units, NG),         \
                              (unsigned)F),                                  \
                         block, 0, st, frames, (int)C);                      \
      PG_LAUNCHED();                                                         \
      hipLaunchKernelGGL((k_hot_cells<GSV>),                                 \
                         dim3((unsigned)ceil_div(max_hot, NG), (unsigned)F), \
                         block, 0, st, frames, (int)C);                      \
      PG_LAUNCHED();                                                         \
    }                                                                        \
    hipLaunchKernelGGL((k_grid<GSV, MODE>), grid, block, 0, st, frames,      \
                       tile_ptr, (int)F, (int)C, (uint32_t)n_tiles);         \
    PG_LAUNCHED();                                                           \
    return 0;                                                                \
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
                 int64_t n_blocks, int64_t max_hot_units, int64_t max_hot,
                 pg_stream_t stream) {
  return launch_grid<0>(frames, tile_ptr, F, n_tiles, C, tile_cells, n_blocks,
                        max_hot_units, max_hot, stream);
}

int pg_grid_reduce_v(const pg_frame *frames, const uint32_t *tile_ptr,
                     int64_t F, int64_t n_tiles, int64_t C,
                     int64_t tile_cells, int64_t n_blocks,
                     int64_t max_hot_units, int64_t max_hot,
                     pg_stream_t stream) {
  return launch_grid<1>(frames, tile_ptr, F, n_tiles, C, tile_cells, n_blocks,
                        max_hot_units, max_hot, stream);
}

int pg_grid_scatter(const pg_frame *frames, const uint32_t *tile_ptr,
                    int64_t F, int64_t n_tiles, int64_t C, int64_t tile_cells,
                    int64_t n_blocks, int64_t max_hot_units, int64_t max_hot,
                    pg_stream_t stream) {
  return launch_grid<2>(frames, tile_ptr, F, n_tiles, C, tile_cells, n_blocks,
                        max_hot_units, max_hot, stream);
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
