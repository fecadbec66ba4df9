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
// Work decomposition (the launch shape picks one; none changes a bit):
//   tile_cells 256: wave tiles -- each wavefront owns 64 consecutive cells,
//     lane = cell for the bookkeeping, the fold and the coalesced write
//     (k_wtile: rows staged in LDS, no workgroup barrier; k_wrows, the V
//     reduction above 16 coils: rows read straight into registers);
//   tile_cells 128 / 64 / 32: workgroup tiles (k_tile) -- 256 threads own TC
//     consecutive cells, rows staged in LDS between workgroup barriers.
// Either way the grid is written coil row by coil row (coalesced, zeros
// included, no read-modify-write, no atomics) and the tile loop is
// persistent (grid-stride over tiles).  Cells with many replicas (the radial
// DC and its neighbours) are reduced by a two-kernel pre-pass -- hot units in
// parallel, then an in-order fold per hot cell -- written straight into the
// grid, so no tile serialises on a hot cell.
#include "common.h"

#include <mutex>

extern "C" int pg_grid_tile_options(int64_t C, int32_t *opts, int32_t max_opts);

namespace pg {

constexpr int kGridThreads = 256;
constexpr int kMaxDevices = 64;

// Default tile (cells per workgroup) per coil group GS: 256 selects the
// wave-tile kernels (one wavefront per 64 cells); 64 and 32 the workgroup
// tiles, whose threads own one cell and GS * TC / 256 coils.
__host__ __device__ constexpr int default_tile(int GS) {
  return GS <= 16 ? 256 : GS == 32 ? 64 : 32;
}

// Coil group of a coil count: the lanes that evaluate one replica (one per
// output coil).  The packed table holds GS / 2 coil pairs per (entry, coil),
// zero padded past C, so a staged (zero padded) sample row can be applied
// without a branch per pair.
__host__ __device__ constexpr int coil_group(int64_t C) {
  return C <= 4 ? 4 : C <= 8 ? 8 : C <= 16 ? 16 : C <= 32 ? 32 : 64;
}

// MODE 0: v = (K[e] y_t[src, :]) * s * s     (plain FP32 GROG apply; y is
//         sample-major (n_src, C): one contiguous row per replica)
// MODE 1: v = vbuf[r]                          (grouped modes, second pass)
// MODE 2: v = y[:, src] * s                    (SparseFFT scatter, identity op,
//         y coil-major (C, n_src))
// MODE 3: v = vbuf[r] decoded                  (MODE 1 with V in the
//         block-floating-point row format of common.h: 16 ceil(C/4) bytes a
//         row; wave-tile launches and the hot pre-pass only)
template <int MODE> constexpr bool kVMode = MODE == 1 || MODE == 3;

// coil co of V row r: MODE 1 complex64 rows, MODE 3 block-floating-point rows
template <int MODE>
__device__ __forceinline__ float2 v_coil(const void *vbuf, uint64_t r, int co, int C) {
  if constexpr (MODE == 3) {
    const uint4 g = gld(reinterpret_cast<const uint4 *>(static_cast<const unsigned char *>(vbuf) +
                                                       r * (uint64_t)bfp_row_bytes(C) + 16 * (co >> 2)));
    return bfp_decode_coil(g, co & 3);
  } else {
    return gld(reinterpret_cast<const float2 *>(vbuf) + r * (uint32_t)C + co);
  }
}
struct Rec {
  uint32_t src, ent;
  float s;
};

template <int MODE>
__device__ __forceinline__ Rec load_rec(const pg_frame &fr, uint32_t r) {
  Rec q{0u, 0u, 0.0f};
  if (!kVMode<MODE>) {
    q.src = fr.rep_src[r];
    q.s = fr.rep_s[r];
    if (MODE == 0) q.ent = fr.rep_ent[r];
  }
  return q;
}

// K[e][co][:] . y_row, then (v * s) * s.  tb points at the packed table
// pairs of (entry, coil) (stride C float4), yr at the replica's sample row
// in global memory (16-byte aligned when C is even); P = ceil(C / 2) pairs
// are used.  The products are spelled out as fmaf chains in ascending input
// coil, the same chain as apply_row_lds: a padding pair adds fmaf(0, 0, v),
// which leaves v unchanged but for the sign of a zero, and every fold starts
// from +0, so both give the same grid bits.  The table values of up to 16
// pairs are loaded together, every load from an in-bounds address.
template <int GS>
__device__ __forceinline__ float2 apply_row(const float4 *__restrict__ tb,
                                           const float2 *yr, int C, int P,
                                           float s) {
  constexpr int PM = (GS + 1) / 2;        // pairs at most
  constexpr int PB = PM < 16 ? PM : 16;   // pairs loaded together
  float vx = 0.0f, vy = 0.0f;
#pragma unroll
  for (int p0 = 0; p0 < PM; p0 += PB) {
    if (p0 >= P) break;
    float4 k[PB];
#pragma unroll
    for (int i = 0; i < PB; ++i) {
      const int p = p0 + i;
      k[i] = gld(tb + (p < P ? p : P - 1) * C);
    }
    if ((C & 1) == 0) {
      const float4 *y4 = reinterpret_cast<const float4 *>(yr);
#pragma unroll
      for (int i = 0; i < PB; ++i) {
        const int p = p0 + i;
        if (p < P) {
          const float4 yy = y4[p];
          vx = fmaf(k[i].x, yy.x, vx);
          vx = fmaf(-k[i].y, yy.y, vx);
          vy = fmaf(k[i].x, yy.y, vy);
          vy = fmaf(k[i].y, yy.x, vy);
          vx = fmaf(k[i].z, yy.z, vx);
          vx = fmaf(-k[i].w, yy.w, vx);
          vy = fmaf(k[i].z, yy.w, vy);
          vy = fmaf(k[i].w, yy.z, vy);
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < PB; ++i) {
        const int p = p0 + i;
        if (p < P) {
          const float2 y0 = yr[2 * p];
          vx = fmaf(k[i].x, y0.x, vx);
          vx = fmaf(-k[i].y, y0.y, vx);
          vy = fmaf(k[i].x, y0.y, vy);
          vy = fmaf(k[i].y, y0.x, vy);
          if (2 * p + 1 < C) {
            const float2 y1 = yr[2 * p + 1];
            vx = fmaf(k[i].z, y1.x, vx);
            vx = fmaf(-k[i].w, y1.y, vx);
            vy = fmaf(k[i].z, y1.y, vy);
            vy = fmaf(k[i].w, y1.x, vy);
          }
        }
      }
    }
  }
  vx = vx * s;
  vy = vy * s;
  return make_float2(vx * s, vy * s);
}

// The same product for a staged sample row: y4 is the row in LDS as GS / 2
// float4 pairs, zero padded past C, and the table is padded alike, so every
// pair is evaluated without a branch and the loads of a block of up to 16
// pairs (and of two rows, when the caller evaluates two) issue together.
template <int GS>
__device__ __forceinline__ float2 apply_row_lds(const float4 *__restrict__ tb,
                                               const float4 *y4, int C,
                                               float s) {
  constexpr int PM = GS / 2;
  constexpr int PB = PM < 16 ? PM : 16;
  float vx = 0.0f, vy = 0.0f;
#pragma unroll
  for (int p0 = 0; p0 < PM; p0 += PB) {
    float4 k[PB];
#pragma unroll
    for (int i = 0; i < PB; ++i) k[i] = gld(tb + (p0 + i) * C);
#pragma unroll
    for (int i = 0; i < PB; ++i) {
      const float4 yy = y4[p0 + i];
      vx = fmaf(k[i].x, yy.x, vx);
      vx = fmaf(-k[i].y, yy.y, vx);
      vy = fmaf(k[i].x, yy.y, vy);
      vy = fmaf(k[i].y, yy.x, vy);
      vx = fmaf(k[i].z, yy.z, vx);
      vx = fmaf(-k[i].w, yy.w, vx);
      vy = fmaf(k[i].z, yy.w, vy);
      vy = fmaf(k[i].w, yy.z, vy);
    }
  }
  vx = vx * s;
  vy = vy * s;
  return make_float2(vx * s, vy * s);
}

// One output coil of one replica's value (before the fold).  The products
// are spelled out as fmaf chains, so every kernel that evaluates a replica
// (tile kernel, hot pre-pass, multi-unit path) produces the same bits.
template <int GS, int MODE>
__device__ __forceinline__ float2 coil_value(const pg_frame &fr, uint32_t r,
                                             const Rec &q, int co, int C,
                                             int P) {
  if (kVMode<MODE>) return v_coil<MODE>(fr.vbuf, r, co, C);
  if (MODE == 2) {
    const float2 a =
        reinterpret_cast<const float2 *>(fr.y)[(int64_t)co * fr.n_src + q.src];
    return make_float2(a.x * q.s, a.y * q.s);
  }
  const float4 *tb = reinterpret_cast<const float4 *>(fr.table) +
                     (int64_t)q.ent * (GS / 2) * C + co;
  const float2 *yr = reinterpret_cast<const float2 *>(fr.y) + (int64_t)q.src * C;
  return apply_row<GS>(tb, yr, C, P, q.s);
}

// Canonical fold of one unit for output coil co: ((0 + v_r0) + v_r0+1) + ...;
// the next replica's record is fetched while the current one is applied.
template <int GS, int MODE>
__device__ __forceinline__ float2 unit_partial(const pg_frame &fr, uint32_t u,
                                               int co, int C, int P) {
#pragma clang fp contract(off)
  const uint32_t r0 = fr.unit_rep[u], r1 = fr.unit_rep[u + 1];
  float px = 0.0f, py = 0.0f;
  if (r0 >= r1 || co >= C) return make_float2(px, py);
  if (kVMode<MODE>) {
    // contiguous rows: issue 8 loads, then add them in order
    for (uint32_t rb = r0; rb < r1; rb += 8) {
      float2 v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        v[k] = (rb + k < r1) ? v_coil<MODE>(fr.vbuf, rb + k, co, C) : make_float2(0.f, 0.f);
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
    const float2 v = coil_value<GS, MODE>(fr, r, q, co, C, P);
    // the value's last product is not contracted into the fold (every
    // other path folds values that went through memory): see the pragma
    px = px + v.x;
    py = py + v.y;
    q = nq;
  }
  return make_float2(px, py);
}

// Tile kernel.  A workgroup owns TC consecutive cells of one frame.  The
// rows (replicas) of its ordinary cells -- not hot, one unit -- are numbered
// 0 .. NR-1 in cell order ("virtual rows": the rows of hot cells are left
// out, so a tile holding the k-space centre does no more work than any
// other) and processed in batches of RB rows:
//   phase A  cell offsets and tags into LDS (one coalesced load of
//            cell_unit, then unit_rep / unit_hot), then a block scan of the
//            ordinary cells' row counts;
//   B1       each batch row's replica index, and for MODE 0/2 its record
//            (source, entry, weight), into LDS: one round of loads;
//   B2       MODE 1: the V rows, MODE 2: the scaled source values, MODE 0:
//            the source sample rows, into LDS rows: one round of loads;
//   B3       MODE 0: groups of GS lanes (lane = output coil) apply the table
//            row of each replica to its staged sample row, in place;
//   C        each thread owns one cell and CPT coils and adds its cell's
//            rows of the batch in order, in registers.
// The tile is then written coil-major, consecutive threads on consecutive
// cells (zeros included, no read-modify-write, no atomics).  Hot cells take
// their sum from the pre-pass; cells with several units that are not hot
// (only when the pre-pass threshold is raised) fold their units in canonical
// order straight from memory.  The tile loop is persistent.
constexpr int kRowBytes = 16384;

// Cell with several units that is not reduced by the hot pre-pass: fold its
// unit partials for one coil in canonical order, straight from memory, with
// every loop rolled: one replica, one input coil at a time (the chain of
// apply_row: vx += Kr yr, vx -= Ki yi, vy += Kr yi, vy += Ki yr in ascending
// input coil, then (v * s) * s).  The tile kernels take this path only for
// plans whose pre-pass threshold leaves multi-unit cells out of the hot list;
// an unrolled copy would multiply their code size and scalar register demand.
template <int GS, int MODE>
__device__ __forceinline__ float2 multi_unit_coil(const pg_frame &fr, uint32_t ua,
                                                      uint32_t ub, int co, int C) {
#pragma clang fp contract(off)
  constexpr int PM = GS / 2;
  float cx = 0.0f, cy = 0.0f;
#pragma unroll 1
  for (uint32_t u = ua; u < ub; ++u) {
    const uint32_t r0 = gld(fr.unit_rep + u), r1 = gld(fr.unit_rep + u + 1);
    float px = 0.0f, py = 0.0f;
#pragma unroll 1
    for (uint32_t r = r0; r < r1; ++r) {
      float vx = 0.0f, vy = 0.0f;
      if (kVMode<MODE>) {
        const float2 v = v_coil<MODE>(fr.vbuf, r, co, C);
        vx = v.x;
        vy = v.y;
      } else if (MODE == 2) {
        const float s = gld(fr.rep_s + r);
        const float2 a = gld(reinterpret_cast<const float2 *>(fr.y) + (uint64_t)co * fr.n_src + gld(fr.rep_src + r));
        vx = a.x * s;
        vy = a.y * s;
      } else {
        const float s = gld(fr.rep_s + r);
        const float2 *yr = reinterpret_cast<const float2 *>(fr.y) + (uint64_t)gld(fr.rep_src + r) * (uint32_t)C;
        const float2 *tb = reinterpret_cast<const float2 *>(fr.table) +
                           2 * ((uint64_t)gld(fr.rep_ent + r) * PM * (uint32_t)C + co);
#pragma unroll 1
        for (int ci = 0; ci < C; ++ci) {
          const float2 k = gld(tb + 2 * (ci >> 1) * C + (ci & 1));
          const float2 yv = gld(yr + ci);
          vx = fmaf(k.x, yv.x, vx);
          vx = fmaf(-k.y, yv.y, vx);
          vy = fmaf(k.x, yv.y, vy);
          vy = fmaf(k.y, yv.x, vy);
        }
        vx = vx * s;
        vy = vy * s;
        vx = vx * s;
        vy = vy * s;
      }
      px = px + vx;
      py = py + vy;
    }
    cx = cx + px;
    cy = cy + py;
  }
  return make_float2(cx, cy);
}

// Largest t in [0, ncell) with vr[t] <= k (vr non-decreasing, k < vr[ncell]).
__device__ __forceinline__ int find_row_cell(const uint32_t *vr, int ncell,
                                             uint32_t k) {
  int lo = 0, hi = ncell - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (vr[mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int MODE, int GS> struct TileRows {
  static constexpr int RS = MODE == 0 ? GS + 2 : GS + 1;  // staged row stride (float2)
  static constexpr int OS = GS + 1;                       // MODE 0 output row stride
  static constexpr int REC = MODE == 1 ? 4 : 12;          // record bytes / row
  static constexpr int RB = kRowBytes / (RS * 8 + (MODE == 0 ? OS * 8 : 0) + REC);  // rows per batch
};

template <int GS, int MODE, int TC>
__global__ __launch_bounds__(kGridThreads) void k_tile(
    const pg_frame *__restrict__ frames, const uint32_t *__restrict__ tile_ptr,
    int F, int C, uint32_t n_tiles) {
  constexpr int RS = TileRows<MODE, GS>::RS;
  constexpr int OS = TileRows<MODE, GS>::OS;
  constexpr int RB = TileRows<MODE, GS>::RB;
  constexpr int FS = MODE == 0 ? OS : RS;  // stride of the rows phase C folds
  constexpr int TPC = kGridThreads / TC;  // threads per cell
  constexpr int CPT = GS / TPC;           // coils per thread
  constexpr int NG = kGridThreads / GS;   // apply groups
  static_assert(TPC * TC == kGridThreads && CPT >= 1 && CPT <= 16 && TC < kGridThreads,
                "tile shape");
  __shared__ __align__(16) float2 vrow[RB * RS];
  __shared__ float2 vout[MODE == 0 ? RB * OS : 1];  // MODE 0: applied rows
  __shared__ uint32_t rmap[RB];                    // MODE 1: replica of row
  __shared__ uint32_t rsrc[MODE == 1 ? 1 : RB];    // MODE 0/2: record
  __shared__ uint32_t rent[MODE == 0 ? RB : 1];
  __shared__ float rsw[MODE == 1 ? 1 : RB];
  __shared__ uint32_t crow[TC + 1];
  __shared__ uint32_t vr[TC + 1];
  __shared__ int tag[TC];
  __shared__ uint32_t cu_lo[TC], cu_hi[TC];

  const int tid = threadIdx.x;
  const int lane = tid % GS;
  const int grp = tid / GS;
  const int P = (C + 1) / 2;
  const int t_own = tid % TC;
  const int co0 = (tid / TC) * CPT;
  const uint32_t tp1 = tile_ptr[1];

  // frame of tile L: the first guess is exact when every frame has tp1 tiles
  auto frame_of = [&](uint32_t L) -> int {
    int f = tp1 > 0 ? (int)(L / tp1) : 0;
    if (f > F - 1) f = F - 1;
    while (f > 0 && tile_ptr[f] > L) --f;
    while (f + 1 < F && tile_ptr[f + 1] <= L) ++f;
    return f;
  };
  // the phase-A values this thread stages: entry tid (tid <= TC < 256)
  auto fetch_head = [&](uint32_t L, int f, uint32_t &c0v, uint32_t &v0v) {
    const pg_frame *q = frames + f;
    const int64_t cell0 = (int64_t)(L - tile_ptr[f]) * TC;
    const int64_t rem = q->n_cell - cell0;
    const int nc = rem < TC ? (int)rem : TC;
    if (tid <= nc) {
      c0v = q->cell_row[cell0 + tid];
      v0v = q->cell_vrow[cell0 + tid];
    }
  };

  uint32_t L = blockIdx.x;
  if (L >= n_tiles) return;
  int f = frame_of(L);
  uint32_t pc0 = 0, pv0 = 0;
  fetch_head(L, f, pc0, pv0);
  for (;;) {
    const pg_frame fr = frames[f];
    const int64_t cell0 = (int64_t)(L - tile_ptr[f]) * TC;
    const int64_t rem = fr.n_cell - cell0;
    const int ncell = rem < TC ? (int)rem : TC;

    // phase A: this tile's CSR row starts and ordinary-row prefix were
    // fetched during the previous tile; fetch the next tile's now
    if (tid <= ncell) {
      crow[tid] = pc0;
      vr[tid] = pv0;
    }
    const uint32_t Ln = L + gridDim.x;
    const bool more = Ln < n_tiles;
    int fn = f;
    if (more) {
      fn = frame_of(Ln);
      fetch_head(Ln, fn, pc0, pv0);
    }
    __syncthreads();
    // cells with rows but no ordinary rows are hot or multi-unit
    for (int t = tid; t < ncell; t += kGridThreads) {
      int tg = -1;
      if (vr[t + 1] == vr[t] && crow[t + 1] > crow[t]) {
        const uint32_t a = fr.cell_unit[cell0 + t];
        const uint32_t b = fr.cell_unit[cell0 + t + 1];
        const int h = fr.n_hot > 0 ? fr.unit_hot[a] : -1;
        tg = h >= 0 ? h : -2;
        cu_lo[t] = a;
        cu_hi[t] = b;
      }
      tag[t] = tg;
    }
    __syncthreads();
    const uint32_t V0 = vr[0], NR = vr[ncell] - V0;

    const bool own = t_own < ncell;
    const uint32_t my_lo = own ? vr[t_own] - V0 : 0u;
    const uint32_t my_hi = own ? vr[t_own + 1] - V0 : 0u;
    float2 acc[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) acc[k] = make_float2(0.0f, 0.0f);

    for (uint32_t k0 = 0; k0 < NR; k0 += RB) {
      const int nk = (NR - k0) < (uint32_t)RB ? (int)(NR - k0) : RB;
      // B1: row -> replica (and its record)
      for (int j = tid; j < nk; j += kGridThreads) {
        const uint32_t k = V0 + k0 + j;
        const int t = find_row_cell(vr, ncell, k);
        const uint32_t r = crow[t] + (k - vr[t]);
        if (MODE == 1) {
          rmap[j] = r;
        } else {
          rsrc[j] = fr.rep_src[r];
          rsw[j] = fr.rep_s[r];
          if (MODE == 0) rent[j] = fr.rep_ent[r];
        }
      }
      __syncthreads();
      // B2: stage the rows
      if (MODE != 2 && (C & 1) == 0) {
        // rows of C complex as C/2 float4 (16-byte aligned for even C)
        const int H = C >> 1;
        for (int i = tid; i < nk * H; i += kGridThreads) {
          const int j = i / H;
          const int c2 = i - j * H;
          const int64_t row = MODE == 1 ? (int64_t)rmap[j] : (int64_t)rsrc[j];
          const float4 v = reinterpret_cast<const float4 *>(MODE == 1 ? fr.vbuf : fr.y)[row * H + c2];
          float2 *d = &vrow[j * RS + 2 * c2];
          d[0] = make_float2(v.x, v.y);
          d[1] = make_float2(v.z, v.w);
        }
      } else {
        for (int i = tid; i < nk * C; i += kGridThreads) {
          const int j = i / C;
          const int c = i - j * C;
          float2 v;
          if (MODE == 1) {
            v = reinterpret_cast<const float2 *>(fr.vbuf)[(int64_t)rmap[j] * C + c];
          } else if (MODE == 2) {
            const float2 a = reinterpret_cast<const float2 *>(fr.y)[(int64_t)c * fr.n_src + rsrc[j]];
            v = make_float2(a.x * rsw[j], a.y * rsw[j]);
          } else {
            v = reinterpret_cast<const float2 *>(fr.y)[(int64_t)rsrc[j] * C + c];
          }
          vrow[j * RS + c] = v;
        }
      }
      __syncthreads();
      // B3 (MODE 0): apply each replica's table row to its staged sample
      // row, two rows per group in flight, into the output rows
      if (MODE == 0) {
        if (lane < C) {
          const float4 *tb = reinterpret_cast<const float4 *>(fr.table) + lane;
          for (int j1 = grp; j1 < nk; j1 += 2 * NG) {
            const int j2 = j1 + NG;
            const float2 v1 = apply_row<GS>(tb + (int64_t)rent[j1] * (GS / 2) * C, &vrow[j1 * RS], C, P, rsw[j1]);
            float2 v2 = make_float2(0.0f, 0.0f);
            if (j2 < nk) v2 = apply_row<GS>(tb + (int64_t)rent[j2] * (GS / 2) * C, &vrow[j2 * RS], C, P, rsw[j2]);
            vout[j1 * OS + lane] = v1;
            if (j2 < nk) vout[j2 * OS + lane] = v2;
          }
        }
        __syncthreads();
      }
      // C: fold this batch's rows of the owned cell
      {
        const uint32_t lo = my_lo > k0 ? my_lo : k0;
        const uint32_t hi = my_hi < k0 + nk ? my_hi : k0 + nk;
        for (uint32_t k = lo; k < hi; ++k) {
          const float2 *vp = MODE == 0 ? &vout[(k - k0) * FS + co0] : &vrow[(k - k0) * FS + co0];
#pragma unroll
          for (int c = 0; c < CPT; ++c) {
            const float2 v = vp[c];
            acc[c].x = acc[c].x + v.x;
            acc[c].y = acc[c].y + v.y;
          }
        }
      }
      __syncthreads();
    }

    if (own && tag[t_own] != -1 && tag[t_own] < 0) {
      // multi-unit cell that the pre-pass does not reduce
      float2 *g = reinterpret_cast<float2 *>(fr.grid) + cell0 + t_own;
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (co0 + k < C)
          g[(int64_t)(co0 + k) * fr.n_cell] =
              multi_unit_coil<GS, MODE>(fr, cu_lo[t_own], cu_hi[t_own], co0 + k, C);
    } else if (own && tag[t_own] == -1) {
      float2 *g = reinterpret_cast<float2 *>(fr.grid) + cell0 + t_own;
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (co0 + k < C)
          g[(int64_t)(co0 + k) * fr.n_cell] = make_float2(0.0f + acc[k].x, 0.0f + acc[k].y);
    }
    // hot cells (tag >= 0) are written by the pre-pass
    __syncthreads();
    if (!more) break;
    L = Ln;
    f = fn;
  }
}

// Wave-tile kernel (tile_cells = 256).  Same canonical arithmetic and the
// same cells as k_tile, but the unit of work is one wavefront owning 64
// consecutive cells of a frame: lane t is cell t for the row bookkeeping, the
// fold and the write, so no phase waits on another wavefront (wavefront-local
// LDS ordering only, no workgroup barrier).  A 256-thread workgroup runs the
// four wave tiles of a 256-cell tile; the tile loop is grid-stride.  The
// ordinary rows of the wave tile (virtual-row order, as in k_tile) go through
// batches of at most RB rows:
//   R  lane t writes the records of its cell's rows into LDS (replica for
//      MODE 1; source and weight for MODE 2; plus entry for MODE 0);
//   S  the wavefront stages the rows with 16-byte loads (MODE 0: sample rows
//      of all C coils; MODE 1: V rows; MODE 2: scaled coil-major values);
//   A  MODE 0: groups of CW lanes (lane = output coil) read their row's
//      staged sample values (LDS broadcast) and the table row from memory and
//      write the applied values over the staged row (every lane of the group
//      has read the row before any lane writes it: one instruction stream);
//   F  lane t adds its cell's rows in order, in registers.
// After the last batch each lane writes its cell's coils: consecutive lanes,
// consecutive cells (512-byte stores per coil).  More than 32 coils are done
// in passes of 32 output coils.
constexpr int kWaveCells = 64;
constexpr int kWaveTile = (kGridThreads / kWave) * kWaveCells;

// Streaming accesses of the wave-tile kernel (sample or V rows, records, cell
// ranges, grid stores) are non-temporal: the operator table it applies is
// read from L2 for every replica and is displaced less by data touched once
// (measured on MI355X, 200 frames 192^2 x 16 coils: plain kernel -3 %,
// V reduction -5 %).  The row-direct reduction keeps cached loads (each V row
// is read by several load instructions there, and non-temporal loads made it
// 1.7x slower) but stores the grid non-temporally (-5 %, kooshball 192^3 x 32
// coils).
#ifndef PG_GRID_NT
#define PG_GRID_NT 1
#endif
template <class T> __device__ __forceinline__ T sld(const T *p) {
  if constexpr (PG_GRID_NT != 0) return gld_nt(p);
  else return gld(p);
}
__device__ __forceinline__ void sst(float2 *p, float2 v) {
  if constexpr (PG_GRID_NT != 0) gst_nt(p, v);
  else gst(p, v);
}

template <int GS> struct WaveRows {
  static constexpr int CW = GS < 32 ? GS : 32;  // output coils per pass
  static constexpr int RS = GS + 2;              // staged row stride (float2): 16-byte rows
  // rows per batch: at least the rows of one ordinary cell (one unit of at
  // most 32 replicas), so a batch always holds whole cells
  static constexpr int RB = GS <= 8 ? 64 : GS == 16 ? 48 : 32;
};

// Table pairs of one (entry, coil), N at a time, and their products with a
// staged row (the chain of apply_row / apply_row_lds).
template <int N>
__device__ __forceinline__ void load_pairs(const float4 *__restrict__ tb, int C,
                                           float4 (&k)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) k[i] = gld(tb + i * C);
}
template <int N>
__device__ __forceinline__ void dot_pairs(const float4 (&k)[N], const float4 *y4,
                                          float &vx, float &vy) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const float4 yy = y4[i];
    vx = fmaf(k[i].x, yy.x, vx);
    vx = fmaf(-k[i].y, yy.y, vx);
    vy = fmaf(k[i].x, yy.y, vy);
    vy = fmaf(k[i].y, yy.x, vy);
    vx = fmaf(k[i].z, yy.z, vx);
    vx = fmaf(-k[i].w, yy.w, vx);
    vy = fmaf(k[i].z, yy.w, vy);
    vy = fmaf(k[i].w, yy.z, vy);
  }
}

template <int GS, int MODE>
__global__ __launch_bounds__(kGridThreads) void k_wtile(
    const pg_frame *__restrict__ frames, const uint32_t *__restrict__ tile_ptr,
    int F, int C, uint32_t n_tiles) {
  constexpr int CW = WaveRows<GS>::CW;
  constexpr int RS = WaveRows<GS>::RS;
  constexpr int RB = WaveRows<GS>::RB;
  constexpr int NW = kGridThreads / kWave;
  constexpr int PM = GS / 2;  // table pairs per (entry, coil)
  static_assert(GS % 2 == 0 && (RS % 2) == 0 && kWave % CW == 0, "wave tile shape");
  __shared__ __align__(16) float2 s_rows[NW][RB * RS];
  __shared__ uint32_t s_ra[NW][RB];                 // MODE 1/3: replica; else source
  __shared__ uint32_t s_rb[NW][MODE == 0 ? RB : 1];  // MODE 0: entry
  __shared__ float s_rw[NW][kVMode<MODE> ? 1 : RB];  // MODE 0/2: weight

  const int w = threadIdx.x / kWave;
  const int l = threadIdx.x % kWave;
  float2 *rows = s_rows[w];
  uint32_t *ra = s_ra[w];
  uint32_t *rb = s_rb[w];
  float *rw = s_rw[w];
  const uint32_t tp1 = tile_ptr[1];

  for (uint32_t L = blockIdx.x; L < n_tiles; L += gridDim.x) {
    int f = tp1 > 0 ? (int)(L / tp1) : 0;
    if (f > F - 1) f = F - 1;
    while (f > 0 && tile_ptr[f] > L) --f;
    while (f + 1 < F && tile_ptr[f + 1] <= L) ++f;
    const pg_frame fr = frames[f];
    const int64_t c0 = (int64_t)(L - tile_ptr[f]) * kWaveTile + (int64_t)w * kWaveCells;
    if (c0 >= fr.n_cell) continue;  // uniform over the wavefront
    const int64_t remc = fr.n_cell - c0;
    const int ncw = remc < kWaveCells ? (int)remc : kWaveCells;
    const bool own = l < ncw;
    // record slots hold valid indices (source / replica / entry 0) until the
    // batches overwrite them: the staging loads of lanes past a batch's rows
    // read them (the frame has at least one source and replica once it has a
    // row)
    for (int j = l; j < RB; j += kWave) {
      ra[j] = 0u;
      if (MODE == 0) rb[j] = 0u;
      if (!kVMode<MODE>) rw[j] = 0.0f;
    }

    // the cell's row ranges; cells with rows but no ordinary rows are hot
    // (written by the pre-pass) or multi-unit (folded from memory below)
    const int64_t cl = c0 + (own ? l : 0);
    const uint32_t vr0 = sld(fr.cell_vrow + cl), vr1 = sld(fr.cell_vrow + cl + 1);
    const uint32_t cr0 = sld(fr.cell_row + cl), cr1 = sld(fr.cell_row + cl + 1);
    // (a cell with more ordinary rows than a batch holds -- only with a
    // unit chunk above RB -- is also folded from memory; its rows still pass
    // through the batches, so every staged row slot holds a real record)
    const bool special = own && vr1 == vr0 && cr1 > cr0;
    const bool big = own && vr1 - vr0 > (uint32_t)RB;
    int tag = -1;
    uint32_t ua = 0, ub = 0;
    if (special || big) {
      ua = gld(fr.cell_unit + cl);
      ub = gld(fr.cell_unit + cl + 1);
      const int h = (special && fr.n_hot > 0) ? gld(fr.unit_hot + ua) : -1;
      tag = h >= 0 ? h : -2;
    }
    const uint32_t base = __builtin_amdgcn_readfirstlane(vr0);
    const uint32_t lo = own ? vr0 - base : 0u;
    const uint32_t hi = own ? vr1 - base : 0u;
    const uint32_t hf = big ? lo : hi;  // rows the batch must hold for this cell

    for (int cc = 0; cc < C; cc += CW) {
      // batches of whole cells [t0, t1) whose rows fit in RB
      for (int t0 = 0; t0 < ncw;) {
        // (wave-uniform values: read from one lane into scalar registers)
        const uint32_t k0 = __builtin_amdgcn_readfirstlane(__shfl(lo, t0));
        const uint64_t over = __ballot(own && l >= t0 && hf - k0 > (uint32_t)RB);
        const int t1 = over ? __ffsll((unsigned long long)over) - 1 : ncw;  // > t0
        const int nk = (int)(__builtin_amdgcn_readfirstlane(__shfl(hf, t1 - 1)) - k0);
        const bool in_b = own && l >= t0 && l < t1;
        // this lane's rows inside the batch rows [k0, k0 + nk)
        const uint32_t a = lo > k0 ? lo : k0;
        const uint32_t bb = hi < k0 + nk ? hi : k0 + nk;
        const uint32_t b = bb > a ? bb : a;
        // R: records of this lane's rows in the batch, four rows' loads at a
        // time (rows past the lane's last re-read its first)
        for (uint32_t k1 = a; k1 < b; k1 += 4) {
          uint32_t r[4], q0[4], q1[4];
          float q2[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) r[i] = cr0 + ((k1 + i < b ? k1 + i : k1) - lo);
          if (!kVMode<MODE>) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              q0[i] = sld(fr.rep_src + r[i]);
              q2[i] = sld(fr.rep_s + r[i]);
              if (MODE == 0) q1[i] = sld(fr.rep_ent + r[i]);
            }
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (k1 + i < b) {
              const int j = (int)(k1 + i - k0);
              if (kVMode<MODE>) {
                ra[j] = r[i];
              } else {
                ra[j] = q0[i];
                rw[j] = q2[i];
                if (MODE == 0) rb[j] = q1[i];
              }
            }
          }
        }
        wave_lds_sync();
        // S: stage the rows: lane = (row, coil) with SW coils per row (all GS
        // coils for MODE 0, zero padded past C; the pass's CW coils else),
        // NS loads per lane issued before their LDS writes.  Record
        // slots past the batch hold valid indices (reset per tile), so the
        // loads of lanes past the batch are in bounds; they store nothing.
        {
          constexpr int SW = MODE == 0 ? GS : CW;
          constexpr int RPI = kWave / SW;       // rows per load instruction
          // loads per lane and round: 8 for the plain apply at 16 coils
          // (fewer live registers: four waves per SIMD instead of three,
          // -5 % on MI355X, 200 frames 192^2 x 16 coils), 16 otherwise
          constexpr int NS = (MODE == 0 && GS == 16) ? 8 : 16;
          constexpr int NB = NS * RPI;          // rows per load round
          const int c = l % SW;
          const int j0 = l / SW;
          const int cg = MODE == 0 ? c : cc + c;  // coil of this lane's values
          const bool cv = cg < C;
          const int cs = cv ? cg : 0;
          // this lane's first value of every row (MODE 2: its coil's row)
          const float2 *ybase =
              kVMode<MODE> ? reinterpret_cast<const float2 *>(fr.vbuf) + cs
              : MODE == 0 ? reinterpret_cast<const float2 *>(fr.y) + cs
                          : reinterpret_cast<const float2 *>(fr.y) + (uint64_t)cs * fr.n_src;
          if constexpr (MODE == 3) {
            // block-floating-point V rows: lane = (row, 4-coil group of the
            // pass), one 16-byte load per row, 4 coils decoded into LDS
            constexpr int GW = CW / 4;
            constexpr int RPG = kWave / GW;  // rows per load instruction
            constexpr int NG = 8;            // loads per lane and round
            const int gl = l % GW, jg = l / GW;
            const int ga = cc / 4 + gl;
            const unsigned char *gb = static_cast<const unsigned char *>(fr.vbuf) + 16 * (4 * ga < C ? ga : 0);
            const uint32_t vrb = (uint32_t)bfp_row_bytes(C);
            for (int jb = 0; jb < nk; jb += NG * RPG) {
              uint32_t qg[NG];
              uint4 gq[NG];
#pragma unroll
              for (int i = 0; i < NG; ++i) {
                const int jj = jb + jg + i * RPG;
                qg[i] = ra[jj < RB ? jj : RB - 1];
              }
#pragma unroll
              for (int i = 0; i < NG; ++i)
                if (jb + i * RPG < nk) gq[i] = sld(reinterpret_cast<const uint4 *>(gb + (uint64_t)qg[i] * vrb));
#pragma unroll
              for (int i = 0; i < NG; ++i) {
                if (jb + i * RPG >= nk) break;
                const int j = jb + jg + i * RPG;
                if (j < nk) {
                  float o[8];
                  bfp_decode8(gq[i].x, gq[i].y, gq[i].z, gq[i].w, o);
#pragma unroll
                  for (int k = 0; k < 4; ++k)
                    if (cc + 4 * gl + k < C) rows[j * RS + 4 * gl + k] = make_float2(o[2 * k], o[2 * k + 1]);
                }
              }
            }
          } else
          for (int jb = 0; jb < nk; jb += NB) {
            // the round's record indices first (all LDS reads in flight
            // together), then its loads
            uint32_t qi[NS];
            float2 v[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              const int jj = jb + j0 + i * RPI;
              qi[i] = ra[jj < RB ? jj : RB - 1];
            }
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              if (jb + i * RPI < nk) {
                if (MODE == 2) v[i] = gld(ybase + qi[i]);
                else v[i] = sld(ybase + (uint64_t)qi[i] * (uint32_t)C);
              }
            }
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              if (jb + i * RPI >= nk) break;
              const int j = jb + j0 + i * RPI;
              if (j < nk) {
                float2 x = cv ? v[i] : make_float2(0.0f, 0.0f);
                if (MODE == 2) {
                  const float sj = rw[j];
                  x = make_float2(x.x * sj, x.y * sj);
                }
                if (MODE == 0 || cv) rows[j * RS + c] = x;
              }
            }
          }
        }
        wave_lds_sync();
        // A (MODE 0): apply each row's table row.  Up to 8 coils: two rows
        // per group, their 2 x GS/2 table loads issued together; 16 and 32
        // coils: one row; 64 coils: one row in two halves.
        if (MODE == 0) {
          constexpr int NG = kWave / CW;
          constexpr int RIF = GS <= 8 ? 2 : 1;
          const int co = l % CW;
          const int g = l / CW;
          const bool cv = cc + co < C;
          const float4 *tb = reinterpret_cast<const float4 *>(fr.table) + cc + (cv ? co : 0);
          for (int j1 = g; j1 < nk; j1 += RIF * NG) {
            const float4 *y1 = reinterpret_cast<const float4 *>(&rows[j1 * RS]);
            float ax = 0.0f, ay = 0.0f;
            if (RIF == 2) {
              // the second row's slot holds a valid entry even past the batch
              const int j2 = j1 + NG;
              const float4 *y2 = reinterpret_cast<const float4 *>(&rows[j2 * RS]);
              float4 k1[PM], k2[PM];
              load_pairs<PM>(tb + (uint64_t)rb[j1] * (uint32_t)(PM * C), C, k1);
              load_pairs<PM>(tb + (uint64_t)rb[j2] * (uint32_t)(PM * C), C, k2);
              float bx = 0.0f, by = 0.0f;
              dot_pairs<PM>(k1, y1, ax, ay);
              dot_pairs<PM>(k2, y2, bx, by);
              const float s2 = rw[j2];
              bx = bx * s2;
              by = by * s2;
              if (cv && j2 < nk) rows[j2 * RS + co] = make_float2(bx * s2, by * s2);
            } else {
              constexpr int PB = PM < 16 ? PM : 16;
              const float4 *t1 = tb + (uint64_t)rb[j1] * (uint32_t)(PM * C);
#pragma unroll
              for (int p0 = 0; p0 < PM; p0 += PB) {
                float4 k1[PB];
                load_pairs<PB>(t1 + p0 * C, C, k1);
                dot_pairs<PB>(k1, y1 + p0, ax, ay);
              }
            }
            const float s1 = rw[j1];
            ax = ax * s1;
            ay = ay * s1;
            if (cv) rows[j1 * RS + co] = make_float2(ax * s1, ay * s1);
          }
          wave_lds_sync();
        }
        // F: this lane's rows, in order, then its cell's coils of this pass
        float2 *g = reinterpret_cast<float2 *>(fr.grid) + c0 + l;
        if (in_b && tag == -1) {
          // eight coils at a time (few live registers; the rows stay in LDS)
          constexpr int FC = CW < 8 ? CW : 8;
#pragma unroll 1
          for (int c8 = 0; c8 < CW; c8 += FC) {
            float2 acc[FC];
#pragma unroll
            for (int c = 0; c < FC; ++c) acc[c] = make_float2(0.0f, 0.0f);
#pragma unroll 1
            for (uint32_t k = a; k < b; ++k) {
              const float2 *vp = &rows[(k - k0) * RS + c8];
#pragma unroll
              for (int c = 0; c < FC; ++c) {
                const float2 v = vp[c];
                acc[c].x = acc[c].x + v.x;
                acc[c].y = acc[c].y + v.y;
              }
            }
#pragma unroll
            for (int c = 0; c < FC; ++c)
              if (cc + c8 + c < C)
                sst(g + (int64_t)(cc + c8 + c) * fr.n_cell, make_float2(0.0f + acc[c].x, 0.0f + acc[c].y));
          }
        } else if (in_b && tag == -2) {
#pragma unroll 1
          for (int c = 0; c < CW; ++c)
            if (cc + c < C)
              gst(g + (int64_t)(cc + c) * fr.n_cell, multi_unit_coil<GS, MODE>(fr, ua, ub, cc + c, C));
        }
        wave_lds_sync();
        t0 = t1;
      }
    }
  }
}

// Row-direct reduction (MODE 1, tile_cells = 256): lane t of a wavefront owns
// cell t of 64 consecutive cells and reads its cell's V rows straight from
// memory (the rows of consecutive cells are consecutive, so the wavefront's
// loads cover one contiguous block) and adds them in order in registers; the
// grid is then written coil row by coil row, consecutive lanes on
// consecutive cells.  No LDS, no barrier.  The first two rows of every cell
// (all coils up to 32) are loaded in one round (addresses clamped in bounds,
// unused values discarded), so a wavefront waits on memory about once per
// tile; further rows follow one at a time.  The next tile's cell ranges are
// fetched while the current tile is processed.  Same canonical fold as
// k_tile: ((0 + v_r0) + v_r1) + ..., written as 0 + sum.
template <int W>
__device__ __forceinline__ void load_row(const float2 *vr, int C, int cc, bool even,
                                         float2 (&v)[W]) {
  if (even) {
#pragma unroll
    for (int i = 0; i < W / 2; ++i) {
      const int c = cc + 2 * i < C ? 2 * i : 0;
      const float4 q = gld(reinterpret_cast<const float4 *>(vr + c));
      v[2 * i] = make_float2(q.x, q.y);
      v[2 * i + 1] = make_float2(q.z, q.w);
    }
  } else {
#pragma unroll
    for (int i = 0; i < W; ++i) v[i] = gld(vr + (cc + i < C ? i : 0));
  }
}

// the same for block-floating-point V rows (MODE 3): W coils from coil cc
// (a multiple of 4) of the row at vr, W / 4 groups of 16 bytes
template <int W>
__device__ __forceinline__ void load_row_bfp(const unsigned char *vr, int C, int cc,
                                             float2 (&v)[W]) {
  static_assert(W % 4 == 0, "whole groups");
  uint4 g[W / 4];
#pragma unroll
  for (int i = 0; i < W / 4; ++i) {
    const int gi = cc / 4 + i;
    g[i] = gld(reinterpret_cast<const uint4 *>(vr + 16 * (4 * gi < C ? gi : 0)));
  }
#pragma unroll
  for (int i = 0; i < W / 4; ++i) {
    float o[8];
    bfp_decode8(g[i].x, g[i].y, g[i].z, g[i].w, o);
#pragma unroll
    for (int k = 0; k < 4; ++k) v[4 * i + k] = make_float2(o[2 * k], o[2 * k + 1]);
  }
}

// One block-floating-point V row added into acc, group by group as it is
// decoded (acc[c] + v_c, the fold's next addition; no decoded copy of the
// row is held, which keeps the row-direct reduction at four wavefronts per
// SIMD at 32 coils).
template <int W>
__device__ __forceinline__ void add_row_bfp(const unsigned char *vr, int C, int cc, float2 (&acc)[W]) {
  static_assert(W % 4 == 0, "whole groups");
  uint4 g[W / 4];
#pragma unroll
  for (int i = 0; i < W / 4; ++i) {
    const int gi = cc / 4 + i;
    g[i] = gld(reinterpret_cast<const uint4 *>(vr + 16 * (4 * gi < C ? gi : 0)));
  }
#pragma unroll
  for (int i = 0; i < W / 4; ++i) {
    float o[8];
    bfp_decode8(g[i].x, g[i].y, g[i].z, g[i].w, o);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[4 * i + k].x = acc[4 * i + k].x + o[2 * k];
      acc[4 * i + k].y = acc[4 * i + k].y + o[2 * k + 1];
    }
  }
}

#ifndef PG_WROWS_CW
#define PG_WROWS_CW 16
#endif
#ifndef PG_WROWS_TWO
#define PG_WROWS_TWO 0
#endif
#ifndef PG_WROWS_PREFETCH
#define PG_WROWS_PREFETCH 1
#endif
#ifndef PG_WROWS_MASKED
#define PG_WROWS_MASKED 1
#endif
#ifndef PG_WROWS_NTST
#define PG_WROWS_NTST 1
#endif
// smallest coil group whose block-floating-point V rows are reduced row-direct
#ifndef PG_WROWS_MIN_GS_BFP
#define PG_WROWS_MIN_GS_BFP 32
#endif
// block-floating-point V rows: a whole 32-coil row (8 groups) per pass
#ifndef PG_WROWS_CWB
#define PG_WROWS_CWB 32
#endif
template <int GS, int MODE>
__global__ __launch_bounds__(kGridThreads) void k_wrows(
    const pg_frame *__restrict__ frames, const uint32_t *__restrict__ tile_ptr,
    int F, int C, uint32_t n_tiles) {
  constexpr int CWM = MODE == 3 ? PG_WROWS_CWB : PG_WROWS_CW;
  constexpr int CW = GS < CWM ? GS : CWM;  // coils per pass
  constexpr bool TWO = PG_WROWS_TWO != 0;  // second row loaded with the first
  const int w = threadIdx.x / kWave;
  const int l = threadIdx.x % kWave;
  const uint32_t tp1 = tile_ptr[1];
  const bool even = (C & 1) == 0;
  auto frame_of = [&](uint32_t L) -> int {
    int f = tp1 > 0 ? (int)(L / tp1) : 0;
    if (f > F - 1) f = F - 1;
    while (f > 0 && tile_ptr[f] > L) --f;
    while (f + 1 < F && tile_ptr[f + 1] <= L) ++f;
    return f;
  };
  // this lane's cell of tile L and its row ranges
  struct Head {
    int f;
    int64_t cl;
    uint32_t vr0, vr1, cr0, cr1;
  };
  auto head = [&](uint32_t L) -> Head {
    Head h;
    h.f = frame_of(L);
    const pg_frame *q = frames + h.f;
    const int64_t nc = q->n_cell;
    h.cl = (int64_t)(L - tile_ptr[h.f]) * kWaveTile + (int64_t)w * kWaveCells + l;
    const int64_t c = h.cl < nc ? h.cl : nc - 1;  // in bounds; unused past the end
    h.vr0 = gld(q->cell_vrow + c);
    h.vr1 = gld(q->cell_vrow + c + 1);
    h.cr0 = gld(q->cell_row + c);
    h.cr1 = gld(q->cell_row + c + 1);
    return h;
  };
  uint32_t L = blockIdx.x;
  if (L >= n_tiles) return;
  Head h = head(L);
  for (;;) {
    const uint32_t Ln = L + gridDim.x;
    const bool more = Ln < n_tiles;
    Head hn = h;
    if (PG_WROWS_PREFETCH && more) hn = head(Ln);
    const pg_frame fr = frames[h.f];
    if (h.cl < fr.n_cell) {
      float2 *g = reinterpret_cast<float2 *>(fr.grid) + h.cl;
      if (h.vr1 == h.vr0 && h.cr1 > h.cr0) {
        // hot cell (the pre-pass writes it) or several units: fold from memory
        const uint32_t ua = gld(fr.cell_unit + h.cl), ub = gld(fr.cell_unit + h.cl + 1);
        if (!(fr.n_hot > 0 && gld(fr.unit_hot + ua) >= 0)) {
#pragma unroll 1
          for (int c = 0; c < C; ++c)
            gst(g + (int64_t)c * fr.n_cell, multi_unit_coil<GS, MODE>(fr, ua, ub, c, C));
        }
      } else {
        const uint32_t n = h.cr1 - h.cr0;
        // rows 0 and 1 of the cell: in bounds even for fewer rows (row 0 of
        // the frame's buffer always exists)
        const uint32_t r0 = n >= 1 ? h.cr0 : 0u;
        const uint32_t r1 = n >= 2 ? h.cr0 + 1 : r0;
        const uint32_t rn = TWO ? h.cr0 + 2 : h.cr0 + 1;  // first row of the loop
        const float2 *vb = reinterpret_cast<const float2 *>(fr.vbuf);
        const unsigned char *vq = static_cast<const unsigned char *>(fr.vbuf);
        const int64_t vrb = bfp_row_bytes(C);
        auto ldrow = [&](uint32_t r, int cc, float2 (&v)[CW]) {
          if constexpr (MODE == 3) load_row_bfp<CW>(vq + (int64_t)r * vrb, C, cc, v);
          else load_row<CW>(vb + (int64_t)r * C + cc, C, cc, even, v);
        };
#pragma unroll 1
        for (int cc = 0; cc < C; cc += CW) {
          if constexpr (MODE == 3 && !TWO) {
            // rows decoded straight into the sums: ((0 + v_r0) + v_r1) + ...
            float2 acc[CW];
#pragma unroll
            for (int c = 0; c < CW; ++c) acc[c] = make_float2(0.0f, 0.0f);
            if (n >= 1) add_row_bfp<CW>(vq + (int64_t)r0 * vrb, C, cc, acc);
#pragma unroll 1
            for (uint32_t r = rn; r < h.cr1; ++r) add_row_bfp<CW>(vq + (int64_t)r * vrb, C, cc, acc);
#pragma unroll
            for (int c = 0; c < CW; ++c)
              if (cc + c < C) {
                const float2 o = make_float2(0.0f + acc[c].x, 0.0f + acc[c].y);
                if (PG_WROWS_NTST) gst_nt(g + (int64_t)(cc + c) * fr.n_cell, o);
                else gst(g + (int64_t)(cc + c) * fr.n_cell, o);
              }
            continue;
          }
          float2 a[CW], b[CW];
#pragma unroll
          for (int c = 0; c < CW; ++c) a[c] = b[c] = make_float2(0.0f, 0.0f);
          // only lanes that own rows issue loads (empty cells -- half the
          // kooshball grid -- would otherwise spend load bandwidth)
          if (!PG_WROWS_MASKED || n >= 1) ldrow(r0, cc, a);
          if (TWO && (!PG_WROWS_MASKED || n >= 2)) ldrow(r1, cc, b);
          float2 acc[CW];
#pragma unroll
          for (int c = 0; c < CW; ++c) {
            acc[c] = make_float2(0.0f, 0.0f);
            if (n >= 1) {
              acc[c].x = acc[c].x + a[c].x;
              acc[c].y = acc[c].y + a[c].y;
            }
            if (TWO && n >= 2) {
              acc[c].x = acc[c].x + b[c].x;
              acc[c].y = acc[c].y + b[c].y;
            }
          }
#pragma unroll 1
          for (uint32_t r = rn; r < h.cr1; ++r) {
            ldrow(r, cc, a);
#pragma unroll
            for (int c = 0; c < CW; ++c) {
              acc[c].x = acc[c].x + a[c].x;
              acc[c].y = acc[c].y + a[c].y;
            }
          }
#pragma unroll
          for (int c = 0; c < CW; ++c)
            if (cc + c < C) {
              const float2 o = make_float2(0.0f + acc[c].x, 0.0f + acc[c].y);
              if (PG_WROWS_NTST) gst_nt(g + (int64_t)(cc + c) * fr.n_cell, o);
              else gst(g + (int64_t)(cc + c) * fr.n_cell, o);
            }
        }
      }
    }
    if (!more) break;
    L = Ln;
    h = PG_WROWS_PREFETCH ? hn : head(Ln);
  }
}

// Hot pre-pass 1: partial of every unit of every hot cell, in parallel
// (grid.x over hot units, grid.y over frames).  MODE 1 / 3 (values read from V):
// one group of GS lanes per unit folds its rows straight from memory.
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

// Hot pre-pass 1 for the modes that evaluate replicas (0: operator apply, 2:
// scaled source values): one workgroup per hot unit.  The unit's records go
// to LDS, then 256 / GS groups of GS lanes (lane = output coil) evaluate its
// replicas in parallel, each with coil_value (the chain every kernel uses),
// into LDS rows; lane co then folds the rows in replica order --
// ((0 + v_r0) + v_r1) + ... -- exactly as unit_partial does, so the partial
// has the same bits while the replicas' memory round trips overlap.  A unit
// longer than kUnitMax replicas (a unit chunk above it) falls back to
// unit_partial.
constexpr int kUnitMax = 32;

template <int GS, int MODE>
__global__ __launch_bounds__(kGridThreads) void k_hot_units_par(
    const pg_frame *__restrict__ frames, int C) {
  constexpr int NG = kGridThreads / GS;
  constexpr int VS = GS + 1;
  __shared__ float2 s_v[kUnitMax * VS];
  __shared__ uint32_t s_src[kUnitMax], s_ent[kUnitMax];
  __shared__ float s_s[kUnitMax];
  const pg_frame fr = frames[blockIdx.y];
  const int64_t i = blockIdx.x;
  if (i >= fr.n_hot_units) return;  // uniform over the workgroup
  const int tid = threadIdx.x;
  const int co = tid % GS, g = tid / GS;
  const int P = (C + 1) / 2;
  const uint32_t u = fr.hot_units[i];
  const uint32_t r0 = fr.unit_rep[u], r1 = fr.unit_rep[u + 1];
  const int n = (int)(r1 - r0);
  float2 *pu = reinterpret_cast<float2 *>(fr.p_unit) + i * C;
  if (n > kUnitMax) {
    if (g == 0 && co < C) pu[co] = unit_partial<GS, MODE>(fr, u, co, C, P);
    return;
  }
  if (tid < n) {
    s_src[tid] = gld(fr.rep_src + r0 + tid);
    s_s[tid] = gld(fr.rep_s + r0 + tid);
    s_ent[tid] = MODE == 0 ? gld(fr.rep_ent + r0 + tid) : 0u;
  }
  __syncthreads();
  if (co < C) {
    for (int j = g; j < n; j += NG) {
      const Rec q{s_src[j], s_ent[j], s_s[j]};
      s_v[j * VS + co] = coil_value<GS, MODE>(fr, r0 + j, q, co, C, P);
    }
  }
  __syncthreads();
  if (g == 0 && co < C) {
    float px = 0.0f, py = 0.0f;
    for (int j = 0; j < n; ++j) {
      const float2 v = s_v[j * VS + co];
      px = px + v.x;
      py = py + v.y;
    }
    pu[co] = make_float2(px, py);
  }
}

// Hot pre-pass 2: in-order fold of each hot cell's unit partials, written
// straight into the grid (the tile kernel leaves hot cells alone).  The
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
  // grid[c] = ((0 + p_u0) + p_u1) + ... : 0 + (the fold) as for every cell
  const uint32_t cell = fr.unit_cell[fr.hot_units[i0]];
  reinterpret_cast<float2 *>(fr.grid)[(int64_t)lane * fr.n_cell + cell] =
      make_float2(0.0f + ax, 0.0f + ay);
}

// packed[e][p][co] = (K[e][co][2p], K[e][co][2p+1]), p < coil_group(C) / 2,
// zero past C
__global__ void k_table_pack(const float2 *__restrict__ K, int64_t nK, int C,
                             float4 *__restrict__ packed) {
  const int P = coil_group(C) / 2;
  const int64_t total = nK * P * C;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / ((int64_t)P * C);
    const int rem = (int)(t - e * P * C);
    const int p = rem / C;
    const int co = rem - p * C;
    const float2 *row = K + (e * C + co) * C;
    const float2 a = (2 * p < C) ? row[2 * p] : make_float2(0.f, 0.f);
    const float2 b = (2 * p + 1 < C) ? row[2 * p + 1] : make_float2(0.f, 0.f);
    packed[t] = make_float4(a.x, a.y, b.x, b.y);
  }
}

// (C, N) coil-major -> (N, C) sample-major through an LDS tile of 64 samples
// x 32 coils per 256-thread block: each thread reads four 16-byte pairs of
// samples of one coil and writes four 16-byte pairs of coils of one sample,
// so both sides move whole 512-byte / 256-byte runs.  Odd C or N (pairs not
// 16-byte aligned) and ragged tiles take an element-wise path.
// The full-tile path reads and writes non-temporally: both sides are touched
// once, and the sample-major copy written that way is also read faster by the
// grouped apply that follows (measured on MI355X, kooshball 192^3 x 32 coils:
// transpose 0.70 -> 0.67 ms, apply 0.93 -> 0.91 ms).
constexpr int kTrS = 64, kTrC = 32;
#ifndef PG_TR_NTLD
#define PG_TR_NTLD 1
#endif
#ifndef PG_TR_NTST
#define PG_TR_NTST 1
#endif

template <int RF>
__device__ __forceinline__ void transpose_tile(const float2 *__restrict__ y,
                                               int64_t C, int64_t N,
                                               void *__restrict__ out,
                                               int64_t n0, int64_t c0,
                                               float2 (*tile)[kTrS + 1]) {
  float2 *yt = static_cast<float2 *>(out);
  unsigned char *ys = static_cast<unsigned char *>(out);
  const int64_t rb = RF == kRowS24 ? s24_row_bytes(C) : bfp_row_bytes(C);
  const int t = threadIdx.x;
  const bool full = n0 + kTrS <= N && ((C | N) & 1) == 0;
  if (full) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int cl = t / 32 + 8 * i, sp = t % 32;
      if (c0 + cl < C) {
        const float4 *src = reinterpret_cast<const float4 *>(y + (c0 + cl) * N + n0) + sp;
        const float4 v = PG_TR_NTLD ? gld_nt(src) : gld(src);
        tile[cl][2 * sp] = make_float2(v.x, v.y);
        tile[cl][2 * sp + 1] = make_float2(v.z, v.w);
      }
    }
  } else {
    for (int i = t; i < kTrS * kTrC; i += 256) {
      const int cl = i / kTrS, sl = i % kTrS;
      if (c0 + cl < C && n0 + sl < N) tile[cl][sl] = y[(c0 + cl) * N + n0 + sl];
    }
  }
  __syncthreads();
  const int64_t cw = C - c0 < kTrC ? C - c0 : kTrC;  // coils of this tile
  if constexpr (RF == kRowBfp) {
    // one (sample, 4-coil group) per thread and round: 8 consecutive
    // threads write one sample's 128 bytes (32 coils), rows back to back
    const int ng = (int)((cw + 3) / 4);
#pragma unroll
    for (int i = 0; i < kTrS * kTrC / 4 / 256; ++i) {
      const int idx = t + 256 * i;
      const int sl = idx / (kTrC / 4), gi = idx % (kTrC / 4);
      if (gi < ng && n0 + sl < N) {
        float x[8];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const bool v = 4 * gi + k < cw;
          const float2 a = v ? tile[4 * gi + k][sl] : make_float2(0.0f, 0.0f);
          x[2 * k] = a.x;
          x[2 * k + 1] = a.y;
        }
        const uint4 w = bfp_encode8(x);
        PG_GAS pg_u4n *dst = (PG_GAS pg_u4n *)(uintptr_t)(ys + (n0 + sl) * rb + 16 * (c0 / 4 + gi));
        pg_u4n n;
        n.x = w.x;
        n.y = w.y;
        n.z = w.z;
        n.w = w.w;
        if (PG_TR_NTST) __builtin_nontemporal_store(n, dst);
        else *dst = n;
      }
    }
    return;
  }
  if (full) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int sl = t / 16 + 16 * i, cp = t % 16;
      if (2 * cp < cw) {
        const float2 a = tile[2 * cp][sl], b = tile[2 * cp + 1][sl];
        if constexpr (RF == kRowS24) {
          // two coils = 12 bytes at byte 6 (c0 + 2 cp) of the row
          PG_GAS pg_u3n *dst = (PG_GAS pg_u3n *)(uintptr_t)(ys + (n0 + sl) * rb + 6 * (c0 + 2 * cp));
          __builtin_nontemporal_store(s24_pack4(a.x, a.y, b.x, b.y), dst);
        } else {
          float4 *dst = reinterpret_cast<float4 *>(yt + (n0 + sl) * C + c0) + cp;
          if (PG_TR_NTST) gst_nt(dst, make_float4(a.x, a.y, b.x, b.y));
          else gst(dst, make_float4(a.x, a.y, b.x, b.y));
        }
      }
    }
  } else {
    for (int i = t; i < kTrS * kTrC; i += 256) {
      const int sl = i / kTrC, cl = i % kTrC;
      if (cl < cw && n0 + sl < N) {
        if constexpr (RF == kRowS24) {
          // one coil = 6 bytes at an even byte offset: three 16-bit stores
          const uint32_t re = s24_round(tile[cl][sl].x), im = s24_round(tile[cl][sl].y);
          unsigned short *d = reinterpret_cast<unsigned short *>(ys + (n0 + sl) * rb + 6 * (c0 + cl));
          d[0] = (unsigned short)re;
          d[1] = (unsigned short)((re >> 16) | (im << 8));
          d[2] = (unsigned short)(im >> 8);
        } else {
          yt[(n0 + sl) * C + c0 + cl] = tile[cl][sl];
        }
      }
    }
  }
}

template <int RF>
__global__ __launch_bounds__(256) void k_transpose_cn(const float2 *__restrict__ y, int64_t C,
                                                      int64_t N, void *__restrict__ yt) {
  __shared__ float2 tile[kTrC][kTrS + 1];
  transpose_tile<RF>(y, C, N, yt, (int64_t)blockIdx.x * kTrS, (int64_t)blockIdx.y * kTrC, tile);
}

// Frames of uneven length back to back: frame f holds (C, N_f) coil-major
// at sample offset off[f] (element offset C * off[f]); its (N_f, C)
// sample-major rows go to the same offset of yt.  blockIdx.z = frame.
template <int RF>
__global__ __launch_bounds__(256) void k_transpose_frames(const float2 *__restrict__ y,
                                                          const int64_t *__restrict__ off,
                                                          int64_t C, void *__restrict__ yt) {
  __shared__ float2 tile[kTrC][kTrS + 1];
  const int64_t base = off[blockIdx.z];
  const int64_t N = off[blockIdx.z + 1] - base;
  const int64_t n0 = (int64_t)blockIdx.x * kTrS;
  if (n0 >= N) return;
  const int64_t rb = RF == kRowS24 ? s24_row_bytes(C) : RF == kRowBfp ? bfp_row_bytes(C) : 8 * C;
  transpose_tile<RF>(y + C * base, C, N, static_cast<unsigned char *>(yt) + rb * base, n0,
                     (int64_t)blockIdx.y * kTrC, tile);
}

// out[b, co, m] = K[idx[m]][co] . y[b, :, m / kw]
// rows (b, s) of B x S blocks: y (B, S, C, N), idx (S, N * kw) -- element
// s's entries at s * N * kw --, out (B, S, C, N * kw)
__global__ void k_sparse_apply(const float2 *__restrict__ y,
                               const int64_t *__restrict__ idx,
                               const float2 *__restrict__ K,
                               float2 *__restrict__ out, int64_t B, int64_t S,
                               int64_t C, int64_t N, int64_t kw, int64_t nK) {
  const int64_t M = N * kw;
  const int64_t total = B * S * C * M;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t bc = t / M;
    const int64_t m = t - bc * M;
    const int64_t bs = bc / C;
    const int64_t co = bc - bs * C;
    const int64_t n = m / kw;
    const int64_t e = idx[(bs % S) * M + m];
    float2 acc = make_float2(0.f, 0.f);
    if (e >= 0 && e < nK) {
      const float2 *row = K + (e * C + co) * C;
      const float2 *ys = y + bs * C * N + n;
      for (int64_t ci = 0; ci < C; ++ci) cmac<float>(acc, row[ci], ys[ci * N]);
    }
    out[t] = acc;
  }
}

// out[b, m] = w[inv[m]] * grid[b, idx[inv[m]]]  (gather + unsort fused)
// rows of n outputs; row r reads grid row r and the plan of stack element
// (r / C) % S (indices, weights and inv_perm of that element start at
// s * n; inv_perm holds element-local positions)
__global__ void k_gather_perm(const float2 *__restrict__ grid, int64_t G,
                              const int64_t *__restrict__ idx,
                              const float *__restrict__ w,
                              const int64_t *__restrict__ inv, int64_t n,
                              int64_t C, int64_t S, int64_t rows,
                              float2 *__restrict__ out) {
  const int64_t total = rows * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / n;
    const int64_t m = t - r * n;
    const int64_t base = ((r / C) % S) * n;
    const int64_t i = inv[base + m];
    const int64_t k = idx[base + i];
    float2 o = make_float2(0.f, 0.f);
    if (k >= 0 && k < G) {
      const float wi = w[base + i];
      const float2 g = grid[r * G + k];
      o = make_float2(wi * g.x, wi * g.y);
    }
    out[t] = o;
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

template <int GS, int MODE>
static int launch_tile(int64_t tile_cells, dim3 grid, hipStream_t st,
                       const pg_frame *frames, const uint32_t *tile_ptr,
                       int F, int C, uint32_t n_tiles) {
  const dim3 block(kGridThreads);
#define PG_TILE_LAUNCH(TCV)                                                   \
  if (tile_cells == TCV) {                                                    \
    hipLaunchKernelGGL((k_tile<GS, MODE, TCV>), grid, block, 0, st, frames,   \
                       tile_ptr, F, C, n_tiles);                              \
    PG_LAUNCHED();                                                            \
    return 0;                                                                 \
  }
  if (tile_cells == kWaveTile) {
    // the V reduction: row-direct above 16 coils (rows of 256 bytes and
    // more), LDS-staged wave tiles below (measured on MI355X)
    if constexpr (kVMode<MODE> && GS >= (MODE == 3 ? PG_WROWS_MIN_GS_BFP : 32))
      hipLaunchKernelGGL((k_wrows<GS, MODE>), grid, block, 0, st, frames, tile_ptr,
                         F, C, n_tiles);
    else
      hipLaunchKernelGGL((k_wtile<GS, MODE>), grid, block, 0, st, frames,
                         tile_ptr, F, C, n_tiles);
    PG_LAUNCHED();
    return 0;
  }
  if constexpr (MODE == 3) {
    PG_CHECK(false, "grid: block-floating-point V rows are reduced by wave tiles only (tile_cells %d), got %lld",
             kWaveTile, (long long)tile_cells);
  } else if constexpr (GS <= 32) {
    PG_TILE_LAUNCH(128) PG_TILE_LAUNCH(64)
  } else {
    PG_TILE_LAUNCH(64) PG_TILE_LAUNCH(32)
  }
#undef PG_TILE_LAUNCH
  PG_CHECK(false, "grid: tile_cells %lld is not compiled for C=%d",
           (long long)tile_cells, C);
}

template <int GS, int MODE>
static int launch_hot(hipStream_t st, const pg_frame *frames, int F, int C,
                      int64_t max_hot_units, int64_t max_hot) {
  constexpr int NG = kGridThreads / GS;
  const dim3 block(kGridThreads);
  if constexpr (kVMode<MODE>)
    hipLaunchKernelGGL((k_hot_units<GS, MODE>),
                       dim3((unsigned)ceil_div(max_hot_units, NG), (unsigned)F),
                       block, 0, st, frames, C);
  else
    hipLaunchKernelGGL((k_hot_units_par<GS, MODE>),
                       dim3((unsigned)max_hot_units, (unsigned)F), block, 0, st,
                       frames, C);
  PG_LAUNCHED();
  hipLaunchKernelGGL((k_hot_cells<GS>),
                     dim3((unsigned)ceil_div(max_hot, NG), (unsigned)F), block,
                     0, st, frames, C);
  PG_LAUNCHED();
  return 0;
}

// Helper stream per device for the hot pre-pass, forked from and joined back
// into the caller's stream with two events (created once, outside any graph
// capture; default priority).  The lock keeps one call's record/wait pairs
// together.  Launches of fewer tiles than kHotForkMinTiles run the pre-pass
// in line instead: the fork's four stream operations cost more host time
// than the pre-pass takes on the device there (a 256^2 frame: 8 us of
// kernels against about 30 us of record/wait calls), and the order of the
// writes does not change a bit (the pre-pass and the tile kernel write
// disjoint cells).
constexpr int64_t kHotForkMinTiles = 4096;
struct HotFork {
  std::mutex m;
  hipStream_t helper[kMaxDevices] = {};
  hipEvent_t fork[kMaxDevices] = {};
  hipEvent_t join[kMaxDevices] = {};
};
static HotFork &hot_fork() {
  static HotFork h;
  return h;
}

template <int GS, int MODE>
static int launch_gs(int64_t tile_cells, dim3 grid, hipStream_t st,
                     const pg_frame *frames, const uint32_t *tile_ptr, int F,
                     int C, uint32_t n_tiles, int64_t max_hot_units,
                     int64_t max_hot) {
  if (max_hot_units <= 0)
    return launch_tile<GS, MODE>(tile_cells, grid, st, frames, tile_ptr, F, C, n_tiles);
  if ((int64_t)n_tiles < kHotForkMinTiles) {
    if (launch_hot<GS, MODE>(st, frames, F, C, max_hot_units, max_hot)) return 1;
    return launch_tile<GS, MODE>(tile_cells, grid, st, frames, tile_ptr, F, C, n_tiles);
  }
  int dev = 0;
  PG_HIP(hipGetDevice(&dev));
  HotFork &hf = hot_fork();
  std::lock_guard<std::mutex> lk(hf.m);
  bool fork = dev >= 0 && dev < kMaxDevices;
  if (fork && hf.helper[dev] == nullptr) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    PG_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) {
      fork = false;  // no resource creation inside a capture: run in line
    } else {
      PG_HIP(hipStreamCreateWithFlags(&hf.helper[dev], hipStreamNonBlocking));
      PG_HIP(hipEventCreateWithFlags(&hf.fork[dev], hipEventDisableTiming));
      PG_HIP(hipEventCreateWithFlags(&hf.join[dev], hipEventDisableTiming));
    }
  }
  if (!fork) {
    if (launch_hot<GS, MODE>(st, frames, F, C, max_hot_units, max_hot)) return 1;
    return launch_tile<GS, MODE>(tile_cells, grid, st, frames, tile_ptr, F, C, n_tiles);
  }
  hipStream_t hs = hf.helper[dev];
  PG_HIP(hipEventRecord(hf.fork[dev], st));
  PG_HIP(hipStreamWaitEvent(hs, hf.fork[dev], 0));
  if (launch_hot<GS, MODE>(hs, frames, F, C, max_hot_units, max_hot)) return 1;
  PG_HIP(hipEventRecord(hf.join[dev], hs));
  const int rc = launch_tile<GS, MODE>(tile_cells, grid, st, frames, tile_ptr, F, C, n_tiles);
  PG_HIP(hipStreamWaitEvent(st, hf.join[dev], 0));
  return rc;
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
  PG_CHECK(max_hot_units >= 0 && max_hot >= 0 && max_hot_units < (int64_t)INT32_MAX,
           "grid: bad hot counts");
  {
    int32_t opts[3];
    const int n = pg_grid_tile_options(C, opts, 3);
    bool ok = false;
    for (int i = 0; i < n; ++i) ok |= (opts[i] == tile_cells);
    PG_CHECK(ok, "grid: tile_cells %lld is not compiled for C=%lld",
             (long long)tile_cells, (long long)C);
  }
  if (n_tiles == 0) return 0;
  int64_t nb = n_blocks > 0 ? n_blocks : 2048;
  if (nb > n_tiles) nb = n_tiles;
  const hipStream_t st = as_stream(s);
  const dim3 grid((unsigned)nb);
  const int c = (int)C, f = (int)F;
  const uint32_t nt = (uint32_t)n_tiles;
  if (C <= 4) return launch_gs<4, MODE>(tile_cells, grid, st, frames, tile_ptr, f, c, nt, max_hot_units, max_hot);
  if (C <= 8) return launch_gs<8, MODE>(tile_cells, grid, st, frames, tile_ptr, f, c, nt, max_hot_units, max_hot);
  if (C <= 16) return launch_gs<16, MODE>(tile_cells, grid, st, frames, tile_ptr, f, c, nt, max_hot_units, max_hot);
  if (C <= 32) return launch_gs<32, MODE>(tile_cells, grid, st, frames, tile_ptr, f, c, nt, max_hot_units, max_hot);
  return launch_gs<64, MODE>(tile_cells, grid, st, frames, tile_ptr, f, c, nt, max_hot_units, max_hot);
}

}  // namespace pg

using namespace pg;

extern "C" {

int pg_grid_tile_cells(int64_t C) {
  if (C <= 4) return default_tile(4);
  if (C <= 8) return default_tile(8);
  if (C <= 16) return default_tile(16);
  if (C <= 32) return default_tile(32);
  if (C <= 64) return default_tile(64);
  return 0;
}

int pg_grid_tile_options(int64_t C, int32_t *opts, int32_t max_opts) {
  int32_t o[3] = {0, 0, 0};
  int n = 0;
  if (C < 1 || C > 64) return 0;
  if (C <= 32) { o[0] = kWaveTile; o[1] = 128; o[2] = 64; n = 3; }
  else { o[0] = kWaveTile; o[1] = 64; o[2] = 32; n = 3; }
  for (int i = 0; i < n && i < max_opts; ++i) opts[i] = o[i];
  return n;
}

int64_t pg_table_packed_elems(int64_t K, int64_t C) {
  if (K < 1 || C < 1 || C > 64) return 0;
  return K * (coil_group(C) / 2) * C;
}

int pg_table_pack_c64(const void *table, int64_t K, int64_t C, void *packed,
                      pg_stream_t stream) {
  PG_CHECK(K >= 1 && C >= 1 && C <= 64, "table_pack: bad sizes");
  const int64_t total = pg_table_packed_elems(K, C);
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

int pg_grid_reduce_v_bfp(const pg_frame *frames, const uint32_t *tile_ptr,
                         int64_t F, int64_t n_tiles, int64_t C,
                         int64_t tile_cells, int64_t n_blocks,
                         int64_t max_hot_units, int64_t max_hot,
                         pg_stream_t stream) {
  PG_CHECK(tile_cells == kWaveTile,
           "grid_reduce_v_bfp: block-floating-point V rows are reduced by wave tiles only (tile_cells %d), got %lld",
           kWaveTile, (long long)tile_cells);
  return launch_grid<3>(frames, tile_ptr, F, n_tiles, C, tile_cells, n_blocks,
                        max_hot_units, max_hot, stream);
}

int pg_grid_scatter(const pg_frame *frames, const uint32_t *tile_ptr,
                    int64_t F, int64_t n_tiles, int64_t C, int64_t tile_cells,
                    int64_t n_blocks, int64_t max_hot_units, int64_t max_hot,
                    pg_stream_t stream) {
  return launch_grid<2>(frames, tile_ptr, F, n_tiles, C, tile_cells, n_blocks,
                        max_hot_units, max_hot, stream);
}

static int transpose_cn(const void *y, int64_t C, int64_t N, void *y_t,
                        int rf, pg_stream_t stream) {
  PG_CHECK(C >= 0 && N >= 0 && C < 65535 * 32, "transpose: bad sizes");
  PG_CHECK(rf == kRowC64 || ((uintptr_t)y_t & 15) == 0, "transpose: packed output not 16-byte aligned");
  if (C == 0 || N == 0) return 0;
  const dim3 grid((unsigned)ceil_div(N, kTrS), (unsigned)ceil_div(C, kTrC));
#define PG_TR_LAUNCH(RF)                                                   \
  hipLaunchKernelGGL(k_transpose_cn<RF>, grid, dim3(256), 0, as_stream(stream), \
                     static_cast<const float2 *>(y), C, N, y_t)
  if (rf == kRowS24) PG_TR_LAUNCH(kRowS24);
  else if (rf == kRowBfp) PG_TR_LAUNCH(kRowBfp);
  else PG_TR_LAUNCH(kRowC64);
#undef PG_TR_LAUNCH
  PG_LAUNCHED();
  return 0;
}

static int transpose_frames(const void *y, const int64_t *frame_off, int64_t F,
                            int64_t max_n, int64_t C, void *y_t, int rf,
                            pg_stream_t stream) {
  PG_CHECK(F >= 0 && F < 65536 && C >= 0 && C < 65535 * 32 && max_n >= 0,
           "transpose_frames: bad sizes (F=%lld, C=%lld)", (long long)F,
           (long long)C);
  PG_CHECK(F == 0 || frame_off, "transpose_frames: null offsets");
  PG_CHECK(rf == kRowC64 || ((uintptr_t)y_t & 15) == 0, "transpose_frames: packed output not 16-byte aligned");
  if (F == 0 || C == 0 || max_n == 0) return 0;
  const dim3 grid((unsigned)ceil_div(max_n, kTrS), (unsigned)ceil_div(C, kTrC),
                  (unsigned)F);
#define PG_TR_LAUNCH(RF)                                                  \
  hipLaunchKernelGGL(k_transpose_frames<RF>, grid, dim3(256), 0,         \
                     as_stream(stream), static_cast<const float2 *>(y),  \
                     frame_off, C, y_t)
  if (rf == kRowS24) PG_TR_LAUNCH(kRowS24);
  else if (rf == kRowBfp) PG_TR_LAUNCH(kRowBfp);
  else PG_TR_LAUNCH(kRowC64);
#undef PG_TR_LAUNCH
  PG_LAUNCHED();
  return 0;
}

int pg_transpose_cn_c64(const void *y, int64_t C, int64_t N, void *y_t,
                        pg_stream_t stream) {
  return transpose_cn(y, C, N, y_t, kRowC64, stream);
}

int pg_transpose_frames_c64(const void *y, const int64_t *frame_off,
                            int64_t F, int64_t max_n, int64_t C, void *y_t,
                            pg_stream_t stream) {
  return transpose_frames(y, frame_off, F, max_n, C, y_t, kRowC64, stream);
}

int64_t pg_s24_row_bytes(int64_t C) { return C >= 0 ? s24_row_bytes(C) : -1; }

int pg_transpose_cn_s24(const void *y, int64_t C, int64_t N, void *y_s,
                        pg_stream_t stream) {
  return transpose_cn(y, C, N, y_s, kRowS24, stream);
}

int pg_transpose_frames_s24(const void *y, const int64_t *frame_off,
                            int64_t F, int64_t max_n, int64_t C, void *y_s,
                            pg_stream_t stream) {
  return transpose_frames(y, frame_off, F, max_n, C, y_s, kRowS24, stream);
}

int64_t pg_bfp_row_bytes(int64_t C) { return C >= 0 ? bfp_row_bytes(C) : -1; }

int pg_transpose_cn_bfp(const void *y, int64_t C, int64_t N, void *y_b,
                        pg_stream_t stream) {
  return transpose_cn(y, C, N, y_b, kRowBfp, stream);
}

int pg_transpose_frames_bfp(const void *y, const int64_t *frame_off,
                            int64_t F, int64_t max_n, int64_t C, void *y_b,
                            pg_stream_t stream) {
  return transpose_frames(y, frame_off, F, max_n, C, y_b, kRowBfp, stream);
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
                     static_cast<float2 *>(out), B, (int64_t)1, C, N, kw, K);
  PG_LAUNCHED();
  return 0;
}

int pg_sparse_apply_stack_c64(const void *y, const int64_t *interp_idx,
                              const void *table, void *out, int64_t B, int64_t S,
                              int64_t C, int64_t N, int64_t kw, int64_t K,
                              pg_stream_t stream) {
  PG_CHECK(B >= 0 && S >= 1 && C >= 0 && N >= 0 && kw >= 1 && K >= 1,
           "sparse_apply_stack: bad sizes");
  const int64_t total = B * S * C * N * kw;
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_sparse_apply, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<const float2 *>(y), interp_idx,
                     static_cast<const float2 *>(table),
                     static_cast<float2 *>(out), B, S, C, N, kw, K);
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
                     inv_perm, n, (int64_t)1, (int64_t)1, B,
                     static_cast<float2 *>(out));
  PG_LAUNCHED();
  return 0;
}

int pg_gather_perm_stack_c64(const void *grid, int64_t G, const int64_t *indices,
                             const float *weights, const int64_t *inv_perm,
                             int64_t n, int64_t C, int64_t S, int64_t B,
                             void *out, pg_stream_t stream) {
  PG_CHECK(G >= 0 && n >= 0 && C >= 0 && S >= 1 && B >= 0, "gather_perm_stack: bad sizes");
  const int64_t total = B * S * C * n;
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_gather_perm, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<const float2 *>(grid), G, indices, weights,
                     inv_perm, n, C, S, B * S * C,
                     static_cast<float2 *>(out));
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
