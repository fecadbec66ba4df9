// The GROG plan compiler behind one call: every kernel, sort and scan of a
// plan (reference GrogInterpolator.__init__ -> _create_plan ->
// _create_plan_single -> _attach_fft_plan, src/pygrog/calib/_grog.py:302-331,
// 918-1234) enqueued in order on one stream, with all arrays carved from two
// caller-provided workspaces and one small host read at the end.  A Python
// orchestration of the same steps spends a millisecond of host time per
// plan in allocations and calls under the interpreter lock; here the host
// work is a few microseconds and the calling thread releases the lock for
// the whole call.
//
// Data-dependent sizes (reduction units, hot cells) are bounded above --
// units <= valid replicas <= S * nper, hot cells <= S * G -- so nothing
// needs to be read back before the end; the summary then gives the exact
// counts.
#include <cstring>
#include <utility>

#include "common.h"

namespace pg {

// 32-bit-key variants of pg_plan_targets / pg_plan_post (plan.hip)
int plan_targets_k32(const float *coords, const float *cmax, int64_t S, int64_t npts, pg_plan_geom geom,
                     int64_t *target_idx, float *distances, float *gauss, uint32_t *sort_key, uint32_t *sort_val,
                     int64_t *n_invalid, pg_stream_t stream);
int plan_post_k32(const uint32_t *keys_sorted, const uint32_t *vals_sorted, const float *gauss,
                  const int64_t *n_invalid, int64_t S, int64_t nper, int64_t G, int64_t *indices, int64_t *sort_perm,
                  int64_t *inv_perm, float *gauss_sorted, uint32_t *cell_ptr, pg_stream_t stream);

int plan_weights_listed(const uint32_t *vals_sorted, const float *gauss_sorted, const int64_t *n_invalid,
                        const uint32_t *cell_ptr, int64_t S, int64_t nper, int64_t G, int32_t kw, int32_t chunk,
                        float *weights, float *sqrt_weights, uint32_t *rep_src, float *rep_s, float *density,
                        uint32_t *unit_count, int64_t rep_stride, int64_t *big_list, uint32_t *big_count,
                        pg_stream_t stream);
int64_t plan_big_cells_max(int64_t S, int64_t nper);
int sort_pairs_u32_iota(void *workspace, size_t *ws_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                        uint32_t *vals_out, int64_t n, int32_t end_bit, pg_stream_t stream);

// Sort keys are S (G + 1) values at most: 32 bits wide when that fits.
inline bool keys32(int64_t S, int64_t G) { return S * (G + 1) - 1 <= (int64_t)UINT32_MAX; }

namespace {

// Field ids of the output workspace (offsets written by the layout query).
enum PlanField {
  kTargetIdx = 0,   // (S, nper) int64
  kWeights,         // (S, nper) f32
  kIndices,         // (S, nper) int64
  kSqrtWeights,     // (S, nper) f32
  kSortPerm,        // (S, nper) int64
  kInvPerm,         // (S, nper) int64
  kCellPtr,         // (S, G + 1) uint32
  kRepSrc,          // (S, Rmax) uint32
  kRepS,            // (S, Rmax) f32
  kDensity,         // (S, G) f32
  kCellUnit,        // (S G + 1) uint32
  kUnitCell,        // (Umax) uint32
  kUnitRep,         // (Umax + S) uint32
  kUnitHot,         // (Umax) int32
  kHotUnits,        // (Umax) uint32
  kHotPtr,          // (S G + S) uint32
  kCellVrow,        // (S, G + 1) uint32
  kNInvalid,        // (S) int64
  kSummary,         // (3 S + 3) int64, device copy of the summary
  kNumFields
};

struct Sizes {
  int64_t S, npts, ndim, kw, G, nper, rmax, umax, ncell;
};

Sizes sizes_of(int64_t S, int64_t npts, int64_t ndim, int64_t kw, int64_t G) {
  Sizes z;
  z.S = S;
  z.npts = npts;
  z.ndim = ndim;
  z.kw = kw;
  z.G = G;
  z.nper = npts * kw;
  z.rmax = z.nper > 0 ? z.nper : 1;
  z.umax = S * z.nper > 0 ? S * z.nper : 1;
  z.ncell = S * G;
  return z;
}

constexpr int64_t kAlign = 256;
int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Element bytes of each output field.
void field_bytes(const Sizes &z, int64_t *b) {
  const int64_t rep = z.S * z.nper;
  b[kTargetIdx] = 8 * rep;
  b[kWeights] = 4 * rep;
  b[kIndices] = 8 * rep;
  b[kSqrtWeights] = 4 * rep;
  b[kSortPerm] = 8 * rep;
  b[kInvPerm] = 8 * rep;
  b[kCellPtr] = 4 * z.S * (z.G + 1);
  b[kRepSrc] = 4 * z.S * z.rmax;
  b[kRepS] = 4 * z.S * z.rmax;
  b[kDensity] = 4 * z.ncell;
  b[kCellUnit] = 4 * (z.ncell + 1);
  b[kUnitCell] = 4 * z.umax;
  b[kUnitRep] = 4 * (z.umax + z.S);
  b[kUnitHot] = 4 * z.umax;
  b[kHotUnits] = 4 * z.umax;
  b[kHotPtr] = 4 * (z.ncell + z.S);
  b[kCellVrow] = 4 * z.S * (z.G + 1);
  b[kNInvalid] = 8 * z.S;
  b[kSummary] = 8 * (3 * z.S + 3);
}

// Scratch fields, in carve order.
enum TmpField {
  kCmax = 0, kGauss, kKeys, kVals, kKeysS, kValsS, kGaussSorted, kUnitCount,
  kA, kHF, kHN, kAs, kHFs, kHNs, kSortWs, kScanWs, kBigList, kBigCount, kNumTmp
};

int tmp_bytes(const Sizes &z, int64_t *b) {
  const int64_t rep = z.S * z.nper;
  b[kCmax] = 4 * z.S * z.ndim;
  b[kGauss] = 4 * rep;
  b[kKeys] = 8 * rep;
  b[kVals] = 4 * rep;
  b[kKeysS] = 8 * rep;
  b[kValsS] = 4 * rep;
  b[kGaussSorted] = 4 * rep;
  for (int f : {kUnitCount, kA, kHF, kHN, kAs, kHFs, kHNs}) b[f] = 4 * (z.ncell + 1);
  size_t sort_ws = 0, scan_ws = 0;
  if (rep > 0) {
    const int end_bit = 64 - __builtin_clzll((unsigned long long)(z.S * (z.G + 1) - 1) | 1ull);
    const int eb = end_bit > 0 ? end_bit : 1;
    if (keys32(z.S, z.G) ? sort_pairs_u32_iota(nullptr, &sort_ws, nullptr, nullptr, nullptr, rep, eb, nullptr)
                         : pg_sort_pairs_u64(nullptr, &sort_ws, nullptr, nullptr, nullptr, nullptr, rep, eb, nullptr))
      return 1;
  }
  if (pg_exclusive_scan_u32(nullptr, &scan_ws, nullptr, nullptr, z.ncell, nullptr)) return 1;
  b[kSortWs] = (int64_t)sort_ws;
  b[kScanWs] = (int64_t)scan_ws;
  b[kBigList] = 8 * plan_big_cells_max(z.S, z.nper);
  b[kBigCount] = 4;
  return 0;
}

// summary = [n_invalid[0..S), U, HF[s G] for s = 0..S, HN[s G] for s = 0..S]
__global__ void k_plan_summary(const int64_t *__restrict__ n_invalid,
                               const uint32_t *__restrict__ cell_unit,
                               const uint32_t *__restrict__ HF,
                               const uint32_t *__restrict__ HN, int64_t S,
                               int64_t G, int64_t *__restrict__ out) {
  for (int64_t i = threadIdx.x; i < 3 * S + 3; i += blockDim.x) {
    int64_t v;
    if (i < S) v = n_invalid[i];
    else if (i == S) v = cell_unit[S * G];
    else if (i <= 2 * S + 1) v = HF[(i - S - 1) * G];
    else v = HN[(i - 2 * S - 2) * G];
    out[i] = v;
  }
}

// chunks of `chunk` replicas of each entry; cnt[K] = 0 (the scan's last input)
__global__ void k_entry_chunks(const uint32_t *__restrict__ ent_ptr, int64_t K, int64_t chunk,
                               uint32_t *__restrict__ cnt) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e <= K; e += (int64_t)gridDim.x * blockDim.x)
    cnt[e] = e < K ? (uint32_t)(((int64_t)(ent_ptr[e + 1] - ent_ptr[e]) + chunk - 1) / chunk) : 0u;
}

// work item first[e] + c = (entry e, offset c * chunk)
__global__ void k_entry_items(const uint32_t *__restrict__ first, const uint32_t *__restrict__ cnt, int64_t K,
                              int64_t chunk, uint32_t *__restrict__ items) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < K; e += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t f = first[e], n = cnt[e];
    for (uint32_t c = 0; c < n; ++c) {
      items[2 * (int64_t)(f + c)] = (uint32_t)e;
      items[2 * (int64_t)(f + c) + 1] = (uint32_t)(c * chunk);
    }
  }
}

// Work-list slots past its length (first[K]) get entry 0xFFFFFFFF, which the
// grouped apply skips: a launch over the whole capacity then needs no host
// copy of the length.
__global__ void k_items_tail(const uint32_t *__restrict__ total, int64_t items_max, uint32_t *__restrict__ items) {
  const int64_t n = *total;
  for (int64_t i = n + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < items_max;
       i += (int64_t)gridDim.x * blockDim.x) {
    items[2 * i] = 0xFFFFFFFFu;
    items[2 * i + 1] = 0u;
  }
}

__global__ void k_gather_records(const uint32_t *__restrict__ ent_rep, const uint32_t *__restrict__ rep_src,
                                 const float *__restrict__ rep_s, int64_t R, uint32_t *__restrict__ ent_src,
                                 float *__restrict__ ent_s) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < R; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t r = ent_rep[i];
    ent_src[i] = rep_src[r];
    ent_s[i] = rep_s[r];
  }
}

// Up to kZeroRanges buffers of 32-bit words cleared in one launch
// (blockIdx.y = buffer; every buffer starts on a 16-byte boundary).
constexpr int kZeroRanges = 12;
struct ZeroRanges {
  uint32_t *p[kZeroRanges];
  int64_t n[kZeroRanges];
};
__global__ void k_zero_ranges(ZeroRanges z) {
  uint32_t *p = z.p[blockIdx.y];
  const int64_t n = z.n[blockIdx.y];
  const int64_t n4 = n / 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  uint4 *q = reinterpret_cast<uint4 *>(p);
  for (int64_t i = t; i < n4; i += stride) q[i] = make_uint4(0u, 0u, 0u, 0u);
  for (int64_t i = 4 * n4 + t; i < n; i += stride) p[i] = 0u;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

}  // namespace pg

using namespace pg;

extern "C" {

int pg_plan_compile_layout(int64_t S, int64_t npts, int32_t ndim, int32_t kw,
                           int64_t G, int64_t *offsets, int64_t *out_bytes,
                           int64_t *tmp_bytes_out) {
  PG_CHECK(S >= 1 && S <= 65535 && npts >= 0 && ndim >= 1 && ndim <= 3 &&
               kw >= 1 && kw <= 128 && G >= 1,
           "plan_compile: bad sizes");
  PG_CHECK(S * npts * kw < (int64_t)UINT32_MAX && S * (G + 1) < (int64_t)UINT32_MAX,
           "plan_compile: too many candidate replicas or cells");
  PG_CHECK(offsets && out_bytes && tmp_bytes_out, "plan_compile: null output");
  const Sizes z = sizes_of(S, npts, ndim, kw, G);
  int64_t b[kNumFields];
  field_bytes(z, b);
  int64_t off = 0;
  for (int f = 0; f < kNumFields; ++f) {
    offsets[f] = off;
    off += align_up(b[f] > 0 ? b[f] : 1);
  }
  *out_bytes = off;
  int64_t t[kNumTmp];
  if (tmp_bytes(z, t)) return 1;
  int64_t toff = 0;
  for (int f = 0; f < kNumTmp; ++f) toff += align_up(t[f] > 0 ? t[f] : 1);
  *tmp_bytes_out = toff;
  return 0;
}

static int plan_compile(const float *coords, int64_t S, int64_t npts,
                        pg_plan_geom geom, int64_t G, int32_t chunk,
                        int64_t min_units, void *out_ws, int64_t out_ws_bytes,
                        void *tmp_ws, int64_t tmp_ws_bytes, float *distances,
                        int64_t *summary, bool sync, pg_stream_t stream) {
  PG_CHECK(chunk >= 1, "plan_compile: unit chunk must be positive");
  PG_CHECK(out_ws && tmp_ws && summary, "plan_compile: null workspace");
  int64_t offs[kNumFields], ob = 0, tb = 0;
  if (int rc = pg_plan_compile_layout(S, npts, geom.ndim, geom.kw, G, offs, &ob, &tb)) return rc;
  PG_CHECK(out_ws_bytes >= ob && tmp_ws_bytes >= tb,
           "plan_compile: workspaces of %lld / %lld bytes, %lld / %lld needed",
           (long long)out_ws_bytes, (long long)tmp_ws_bytes, (long long)ob, (long long)tb);
  const Sizes z = sizes_of(S, npts, geom.ndim, geom.kw, G);
  PG_CHECK(z.nper == 0 || (coords && distances), "plan_compile: null coordinates");
  const hipStream_t st = as_stream(stream);
  char *o = static_cast<char *>(out_ws);
  auto O = [&](int f) { return static_cast<void *>(o + offs[f]); };
  int64_t tbytes[kNumTmp];
  if (tmp_bytes(z, tbytes)) return 1;
  void *T[kNumTmp];
  {
    char *t = static_cast<char *>(tmp_ws);
    int64_t toff = 0;
    for (int f = 0; f < kNumTmp; ++f) {
      T[f] = t + toff;
      toff += align_up(tbytes[f] > 0 ? tbytes[f] : 1);
    }
  }
  const int64_t rep = S * z.nper;
  auto *target_idx = static_cast<int64_t *>(O(kTargetIdx));
  auto *weights = static_cast<float *>(O(kWeights));
  auto *indices = static_cast<int64_t *>(O(kIndices));
  auto *sqrt_w = static_cast<float *>(O(kSqrtWeights));
  auto *sort_perm = static_cast<int64_t *>(O(kSortPerm));
  auto *inv_perm = static_cast<int64_t *>(O(kInvPerm));
  auto *cell_ptr = static_cast<uint32_t *>(O(kCellPtr));
  auto *rep_src = static_cast<uint32_t *>(O(kRepSrc));
  auto *rep_s = static_cast<float *>(O(kRepS));
  auto *density = static_cast<float *>(O(kDensity));
  auto *cell_unit = static_cast<uint32_t *>(O(kCellUnit));
  auto *unit_cell = static_cast<uint32_t *>(O(kUnitCell));
  auto *unit_rep = static_cast<uint32_t *>(O(kUnitRep));
  auto *unit_hot = static_cast<int32_t *>(O(kUnitHot));
  auto *hot_units = static_cast<uint32_t *>(O(kHotUnits));
  auto *hot_ptr = static_cast<uint32_t *>(O(kHotPtr));
  auto *cell_vrow = static_cast<uint32_t *>(O(kCellVrow));
  auto *n_invalid = static_cast<int64_t *>(O(kNInvalid));
  auto *summary_dev = static_cast<int64_t *>(O(kSummary));
  auto *cmax = static_cast<float *>(T[kCmax]);
  auto *gauss = static_cast<float *>(T[kGauss]);
  auto *keys = static_cast<uint64_t *>(T[kKeys]);
  auto *vals = static_cast<uint32_t *>(T[kVals]);
  auto *keys_s = static_cast<uint64_t *>(T[kKeysS]);
  auto *vals_s = static_cast<uint32_t *>(T[kValsS]);
  auto *gauss_sorted = static_cast<float *>(T[kGaussSorted]);
  auto *unit_count = static_cast<uint32_t *>(T[kUnitCount]);
  const int64_t nc1 = z.ncell + 1;

  // zero what the kernels accumulate into or only partly write (one launch
  // for all of them; every buffer is a 256-byte aligned slice of a workspace)
  {
    ZeroRanges zr{};
    int nz = 0;
    auto add = [&](void *ptr, int64_t bytes) {
      zr.p[nz] = static_cast<uint32_t *>(ptr);
      zr.n[nz] = bytes / 4;
      ++nz;
    };
    add(cmax, tbytes[kCmax]);
    add(n_invalid, 8 * S);
    add(weights, 4 * rep);
    add(sqrt_w, 4 * rep);
    add(rep_src, 4 * S * z.rmax);
    add(rep_s, 4 * S * z.rmax);
    add(unit_count, 4 * nc1);
    add(T[kBigCount], 4);
    if (z.nper == 0) add(cell_ptr, 4 * S * (G + 1));
    int64_t most = 1;
    for (int i = 0; i < nz; ++i) most = zr.n[i] > most ? zr.n[i] : most;
    hipLaunchKernelGGL(k_zero_ranges, dim3(grid_for((most + 3) / 4, 256, 2048), nz), dim3(256), 0, st, zr);
    PG_LAUNCHED();
  }

  if (int rc = pg_plan_absmax(coords, S, npts, geom.ndim, cmax, stream)) return rc;
  const bool k32 = keys32(S, G);
  auto *keys32_in = reinterpret_cast<uint32_t *>(keys);
  auto *keys32_s = reinterpret_cast<uint32_t *>(keys_s);
  // (32-bit keys: the sort's values are the candidate indices themselves,
  // from a counting iterator, so the targets kernel writes none)
  if (int rc = k32 ? plan_targets_k32(coords, cmax, S, npts, geom, target_idx, distances, gauss, keys32_in, nullptr,
                                      n_invalid, stream)
                   : pg_plan_targets(coords, cmax, S, npts, geom, target_idx, distances, gauss, keys, vals,
                                     n_invalid, stream))
    return rc;
  if (rep > 0) {
    const int end_bit = 64 - __builtin_clzll((unsigned long long)(S * (G + 1) - 1) | 1ull);
    const int eb = end_bit > 0 ? end_bit : 1;
    size_t ws = (size_t)tbytes[kSortWs];
    if (int rc = k32 ? sort_pairs_u32_iota(T[kSortWs], &ws, keys32_in, keys32_s, vals_s, rep, eb, stream)
                     : pg_sort_pairs_u64(T[kSortWs], &ws, keys, keys_s, vals, vals_s, rep, eb, stream))
      return rc;
  }
  if (int rc = k32 ? plan_post_k32(keys32_s, vals_s, gauss, n_invalid, S, z.nper, G, indices, sort_perm, inv_perm,
                                   gauss_sorted, cell_ptr, stream)
                   : pg_plan_post(keys_s, vals_s, gauss, n_invalid, S, z.nper, G, indices, sort_perm, inv_perm,
                                  gauss_sorted, cell_ptr, stream))
    return rc;
  if (int rc = plan_weights_listed(vals_s, gauss_sorted, n_invalid, cell_ptr, S, z.nper, G, geom.kw, chunk, weights,
                                   sqrt_w, rep_src, rep_s, density, unit_count, z.rmax,
                                   static_cast<int64_t *>(T[kBigList]), static_cast<uint32_t *>(T[kBigCount]),
                                   stream))
    return rc;
  size_t scan_ws = (size_t)tbytes[kScanWs];
  if (int rc = pg_exclusive_scan_u32(T[kScanWs], &scan_ws, unit_count, cell_unit, z.ncell, stream)) return rc;
  if (int rc = pg_plan_units(cell_ptr, cell_unit, S, G, chunk, unit_cell, unit_rep, stream)) return rc;
  auto *a = static_cast<uint32_t *>(T[kA]);
  auto *hf = static_cast<uint32_t *>(T[kHF]);
  auto *hn = static_cast<uint32_t *>(T[kHN]);
  auto *A = static_cast<uint32_t *>(T[kAs]);
  auto *HF = static_cast<uint32_t *>(T[kHFs]);
  auto *HN = static_cast<uint32_t *>(T[kHNs]);
  if (int rc = pg_plan_hot_counts(cell_ptr, cell_unit, S, G, min_units, a, hf, hn, stream)) return rc;
  for (auto pr : {std::make_pair(a, A), std::make_pair(hf, HF), std::make_pair(hn, HN)}) {
    scan_ws = (size_t)tbytes[kScanWs];
    if (int rc = pg_exclusive_scan_u32(T[kScanWs], &scan_ws, pr.first, pr.second, z.ncell, stream)) return rc;
  }
  // units of no hot cell stay -1 (pg_plan_hot_rows writes every unit of the
  // bound when U < Umax only up to U; the rest is never read)
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(z.umax, 256)), dim3(256), 0, st,
                     reinterpret_cast<uint32_t *>(unit_hot), z.umax, 0xFFFFFFFFu);
  PG_LAUNCHED();
  if (int rc = pg_plan_hot_rows(cell_unit, S, G, min_units, A, HF, HN, cell_vrow, unit_hot, hot_units, hot_ptr,
                                stream))
    return rc;
  hipLaunchKernelGGL(k_plan_summary, dim3(1), dim3(256), 0, st, n_invalid, cell_unit, HF, HN, S, G, summary_dev);
  PG_LAUNCHED();
  PG_HIP(hipMemcpyAsync(summary, summary_dev, 8 * (3 * S + 3), hipMemcpyDeviceToHost, st));
  if (sync) PG_HIP(hipStreamSynchronize(st));
  return 0;
}

int pg_plan_compile(const float *coords, int64_t S, int64_t npts,
                    pg_plan_geom geom, int64_t G, int32_t chunk,
                    int64_t min_units, void *out_ws, int64_t out_ws_bytes,
                    void *tmp_ws, int64_t tmp_ws_bytes, float *distances,
                    int64_t *summary, pg_stream_t stream) {
  return plan_compile(coords, S, npts, geom, G, chunk, min_units, out_ws, out_ws_bytes, tmp_ws, tmp_ws_bytes,
                      distances, summary, true, stream);
}

int pg_plan_compile_async(const float *coords, int64_t S, int64_t npts,
                          pg_plan_geom geom, int64_t G, int32_t chunk,
                          int64_t min_units, void *out_ws, int64_t out_ws_bytes,
                          void *tmp_ws, int64_t tmp_ws_bytes, float *distances,
                          int64_t *summary_pinned, pg_stream_t stream) {
  return plan_compile(coords, S, npts, geom, G, chunk, min_units, out_ws, out_ws_bytes, tmp_ws, tmp_ws_bytes,
                      distances, summary_pinned, false, stream);
}

// bits of the entries 0 .. K-1 (at least one)
static int entry_bits(int64_t K) { return 64 - __builtin_clzll((unsigned long long)(K - 1) | 1ull); }

// Replicas grouped by table entry for the grouped apply, in one call: a
// stable radix sort of the entries (values = replica ids), each entry's range
// of replicas (CSR, K + 1), the work list of (entry, first offset) chunks of
// `chunk` replicas in entry order, and each replica's source and weight
// gathered into entry order.  Temporary arrays come from one caller
// workspace (pg_group_by_entry_layout); the work-list length is copied to
// the host at the end (the stream is synchronised).
int pg_group_by_entry_layout(int64_t R, int64_t K, int64_t chunk, int64_t *items_max,
                             int64_t *tmp_bytes) {
  PG_CHECK(R >= 1 && R < (int64_t)INT32_MAX && K >= 1 && K < (int64_t)INT32_MAX - 1 && chunk >= 1,
           "group_by_entry: bad sizes R=%lld K=%lld chunk=%lld", (long long)R, (long long)K,
           (long long)chunk);
  PG_CHECK(items_max && tmp_bytes, "group_by_entry: null output");
  size_t sort_ws = 0, scan_ws = 0;
  if (int rc = sort_pairs_u32_iota(nullptr, &sort_ws, nullptr, nullptr, nullptr, R, entry_bits(K), nullptr))
    return rc;
  if (int rc = pg_exclusive_scan_u32(nullptr, &scan_ws, nullptr, nullptr, K, nullptr)) return rc;
  auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
  *items_max = R / chunk + K + 1;
  *tmp_bytes = up((int64_t)sort_ws) + up((int64_t)scan_ws) + 2 * up(4 * R) + 2 * up(4 * (K + 1));
  return 0;
}

static int group_by_entry(const uint32_t *rep_ent, const uint32_t *rep_src, const float *rep_s, int64_t R,
                          int64_t K, int64_t chunk, uint32_t *ent_ptr, uint32_t *ent_rep, uint32_t *ent_src,
                          float *ent_s, uint32_t *items, int64_t items_max, void *tmp, int64_t tmp_bytes,
                          int64_t *n_items, bool sync, pg_stream_t stream) {
  int64_t im = 0, tb = 0;
  if (int rc = pg_group_by_entry_layout(R, K, chunk, &im, &tb)) return rc;
  PG_CHECK(items_max >= im && tmp_bytes >= tb, "group_by_entry: work list or workspace too small");
  PG_CHECK(rep_ent && rep_src && rep_s && ent_ptr && ent_rep && ent_src && ent_s && items && tmp && (n_items || !sync),
           "group_by_entry: null pointer");
  const hipStream_t st = as_stream(stream);
  size_t sort_ws = 0, scan_ws = 0;
  const int end_bit = entry_bits(K);
  if (int rc = sort_pairs_u32_iota(nullptr, &sort_ws, nullptr, nullptr, nullptr, R, end_bit, nullptr))
    return rc;
  if (int rc = pg_exclusive_scan_u32(nullptr, &scan_ws, nullptr, nullptr, K, nullptr)) return rc;
  auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
  auto *base = static_cast<unsigned char *>(tmp);
  void *sort_t = base;
  base += up((int64_t)sort_ws);
  void *scan_t = base;
  base += up((int64_t)scan_ws);
  auto *iota = reinterpret_cast<uint32_t *>(base);
  base += up(4 * R);
  auto *keys_o = reinterpret_cast<uint32_t *>(base);
  base += up(4 * R);
  auto *cnt = reinterpret_cast<uint32_t *>(base);
  base += up(4 * (K + 1));
  auto *first = reinterpret_cast<uint32_t *>(base);

  (void)iota;
  if (int rc = sort_pairs_u32_iota(sort_t, &sort_ws, rep_ent, keys_o, ent_rep, R, end_bit, stream)) return rc;
  if (int rc = pg_csr_from_sorted_u32(keys_o, R, K, ent_ptr, stream)) return rc;
  hipLaunchKernelGGL(k_entry_chunks, dim3(grid_for(K + 1, 256)), dim3(256), 0, st, ent_ptr, K, chunk, cnt);
  PG_LAUNCHED();
  if (int rc = pg_exclusive_scan_u32(scan_t, &scan_ws, cnt, first, K, stream)) return rc;
  hipLaunchKernelGGL(k_entry_items, dim3(grid_for(K, 256)), dim3(256), 0, st, first, cnt, K, chunk, items);
  PG_LAUNCHED();
  hipLaunchKernelGGL(k_gather_records, dim3(grid_for(R, 256)), dim3(256), 0, st, ent_rep, rep_src, rep_s, R,
                     ent_src, ent_s);
  PG_LAUNCHED();
  if (sync) {
    uint32_t total = 0;
    PG_HIP(hipMemcpyAsync(&total, first + K, 4, hipMemcpyDeviceToHost, st));
    PG_HIP(hipStreamSynchronize(st));
    *n_items = total;
    return 0;
  }
  hipLaunchKernelGGL(k_items_tail, dim3(grid_for(items_max, 256, 1024)), dim3(256), 0, st, first + K, items_max,
                     items);
  PG_LAUNCHED();
  // the length, for host readers once the stream has passed this point
  if (n_items != nullptr) {
    *n_items = 0;
    PG_HIP(hipMemcpyAsync(n_items, first + K, 4, hipMemcpyDeviceToHost, st));
  }
  return 0;
}

int pg_group_by_entry(const uint32_t *rep_ent, const uint32_t *rep_src, const float *rep_s, int64_t R,
                      int64_t K, int64_t chunk, uint32_t *ent_ptr, uint32_t *ent_rep, uint32_t *ent_src,
                      float *ent_s, uint32_t *items, int64_t items_max, void *tmp, int64_t tmp_bytes,
                      int64_t *n_items, pg_stream_t stream) {
  return group_by_entry(rep_ent, rep_src, rep_s, R, K, chunk, ent_ptr, ent_rep, ent_src, ent_s, items, items_max,
                        tmp, tmp_bytes, n_items, true, stream);
}

int pg_group_by_entry_async(const uint32_t *rep_ent, const uint32_t *rep_src, const float *rep_s, int64_t R,
                            int64_t K, int64_t chunk, uint32_t *ent_ptr, uint32_t *ent_rep, uint32_t *ent_src,
                            float *ent_s, uint32_t *items, int64_t items_max, void *tmp, int64_t tmp_bytes,
                            int64_t *n_items_pinned, pg_stream_t stream) {
  return group_by_entry(rep_ent, rep_src, rep_s, R, K, chunk, ent_ptr, ent_rep, ent_src, ent_s, items, items_max,
                        tmp, tmp_bytes, n_items_pinned, false, stream);
}

}  // extern "C"
