// Device-resident GROG plan compiler.
//
// This translation unit is compiled with -ffp-contract=off: every float
// operation below is a separately rounded IEEE op (division and rint are the
// correctly rounded gfx950 sequences), in the operation order of the
// reference's float32 torch code, so target indices, shifts and quantised
// table entries are bit-exact with it:
//   _rescale_coords            src/pygrog/calib/_grog.py:1318-1322
//   _create_plan_single        src/pygrog/calib/_grog.py:1081-1234
//   _attach_fft_plan           src/pygrog/calib/_grog.py:918-998
//   calc_interp_table (quant.) src/pygrog/calib/_grog.py:406-422
//
// Pipeline (driven from pygrog_amd/calib/_grog.py):
//   absmax -> targets -> radix sort (key = element, cell) -> post (perms,
//   CSR) -> weights (EGROG per cell, f64 sums in reference order) -> scan ->
//   units; later quantize (table entries) once the table is known.
#include <hipcub/hipcub.hpp>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "common.h"

namespace pg {

// cmax[s, d] = max |coords[s, :, d]| via unsigned-bit atomicMax (values are
// non-negative, so the IEEE bit order is the numeric order).  cmax zeroed by
// the caller.  Each thread walks whole points; the workgroup folds its
// wavefronts' maxima in LDS and adds one atomic per axis (per-wavefront
// atomics on the same three words had serialised at the L2).
constexpr int kAbsmaxThreads = 256;
__global__ __launch_bounds__(kAbsmaxThreads) void k_absmax(
    const float *__restrict__ coords, int64_t S, int64_t npts, int ndim,
    unsigned *__restrict__ cmax) {
  const int s = blockIdx.y;
  const float *c = coords + s * npts * ndim;
  unsigned m[3] = {0u, 0u, 0u};
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npts;
       p += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d < ndim) {
        const unsigned b = __float_as_uint(fabsf(c[p * ndim + d]));
        m[d] = b > m[d] ? b : m[d];
      }
    }
  }
  __shared__ unsigned s_m[3][kAbsmaxThreads / kWave];
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    unsigned v = m[d];
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned o = __shfl_down(v, off, kWave);
      v = o > v ? o : v;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) s_m[d][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < ndim) {
    unsigned v = 0u;
    for (int i = 0; i < kAbsmaxThreads / kWave; ++i) v = s_m[threadIdx.x][i] > v ? s_m[threadIdx.x][i] : v;
    atomicMax(cmax + s * ndim + threadIdx.x, v);
  }
}

struct GeomDev {
  int ndim, kw;
  int64_t gs[3], st[3];
  float ha[3], org[3], stp[3];
  int off[128][3];
};

// KT: the sort key's type -- 32 bits whenever S (G + 1) fits (every
// single-trajectory plan up to 2^32 cells), halving the key bytes the
// radix sort moves; the keys' values are the same either way.
template <class KT>
__global__ void k_plan_targets(const float *__restrict__ coords,
                               const float *__restrict__ cmax, int64_t S,
                               int64_t npts, GeomDev g,
                               int64_t *__restrict__ target_idx,
                               float *__restrict__ distances,
                               float *__restrict__ gauss,
                               KT *__restrict__ sort_key,
                               uint32_t *__restrict__ sort_val,
                               unsigned long long *__restrict__ n_invalid) {
  const int64_t nper = npts * g.kw;
  const int64_t total = S * nper;
  const int64_t G = g.gs[0] * (g.ndim > 1 ? g.gs[1] : 1) *
                    (g.ndim > 2 ? g.gs[2] : 1);
  const int lane = threadIdx.x & (kWave - 1);
  // invalid replicas of the element the wave is in, added to n_invalid when
  // the wave moves on to another element and at the end (one atomic per
  // wave and element, not per iteration: same-address atomics serialise)
  unsigned long long run = 0;
  int64_t run_s = -1;
  for (int64_t gi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
       gi - threadIdx.x < total; gi += (int64_t)gridDim.x * blockDim.x) {
    const bool active = gi < total;
    int64_t s = 0;
    bool valid = false;
    if (active) {
      s = gi / nper;
      const int64_t loc = gi - s * nper;
      const int64_t n = loc / g.kw;
      const int k = (int)(loc - n * g.kw);
      const float *c = coords + (s * npts + n) * g.ndim;
      bool inb = true;
      int64_t flat = 0;
      float dsq = 0.0f;
      float dist[3];
      for (int d = 0; d < g.ndim; ++d) {
        // _rescale_coords: (0.5 * amp) * coords / cmax
        const float cs = (g.ha[d] * c[d]) / cmax[s * g.ndim + d];
        // frac_idx = (coords - origins) / grid_steps; nearest = round()
        const float frac = (cs - g.org[d]) / g.stp[d];
        const int nearest = (int)__builtin_rintf(frac);
        const int t = nearest + g.off[k][d];
        // target_coords_space = origins + target_nd.float() * grid_steps
        const float tc = g.org[d] + (float)t * g.stp[d];
        const float dd = tc - cs;
        dist[d] = dd;
        if (t < 0 || (int64_t)t >= g.gs[d]) inb = false;
        if (!(fabsf(dd) <= 0.5f)) inb = false;
        flat += (int64_t)t * g.st[d];
        // dist_grid = distances / grid_steps; dist_sq = sum(dist_grid**2)
        const float dg = dd / g.stp[d];
        dsq = (d == 0) ? dg * dg : dsq + dg * dg;
      }
      valid = inb;
      for (int d = 0; d < g.ndim; ++d) distances[gi * g.ndim + d] = dist[d];
      target_idx[gi] = valid ? flat : (int64_t)-1;
      gauss[gi] = valid ? expf(-0.5f * dsq) : 0.0f;
      sort_key[gi] = (KT)((uint64_t)s * (uint64_t)(G + 1) +
                          (valid ? (uint64_t)(flat + 1) : 0ull));
      if (sort_val != nullptr) sort_val[gi] = (uint32_t)gi;
    }
    // lanes in lane 0's element count into the wave's run; the others (a
    // wave straddling two elements) add directly
    const bool inv = active && !valid;
    const int64_t s0 = __shfl(s, 0, kWave);
    const unsigned long long same = __ballot(inv && s == s0);
    if (s0 != run_s) {
      if (lane == 0 && run) atomicAdd(n_invalid + run_s, run);
      run = 0;
      run_s = s0;
    }
    run += (unsigned long long)__popcll(same);
    if (inv && s != s0) atomicAdd(n_invalid + s, 1ull);
  }
  if (lane == 0 && run) atomicAdd(n_invalid + run_s, run);
}

// cell_ptr[lo..hi] = v for every lane's range, by the whole wavefront: a
// lane fills a range of up to kShortGap cells itself; longer ranges (empty
// stretches of the grid -- the corners of a 3D grid, everything past the
// last replica -- can be most of G) are filled one after another by all 64
// lanes with consecutive stores.  Every lane of the wavefront must call it.
constexpr int64_t kShortGap = 8;
__device__ __forceinline__ void wave_fill(uint32_t *cp, int64_t lo, int64_t hi, uint32_t v) {
  const bool lng = hi - lo + 1 > kShortGap;
  if (!lng)
    for (int64_t c = lo; c <= hi; ++c) cp[c] = v;
  unsigned long long m = __ballot(lng);
  const int lane = threadIdx.x & (kWave - 1);
  while (m) {
    const int src = __ffsll((long long)m) - 1;
    m &= m - 1;
    uint32_t *p = reinterpret_cast<uint32_t *>(__shfl((long long)(uintptr_t)cp, src));
    const int64_t l = __shfl(lo, src), h = __shfl(hi, src);
    const uint32_t w = __shfl(v, src);
    for (int64_t c = l + lane; c <= h; c += kWave) p[c] = w;
  }
}

template <class KT>
__global__ void k_plan_post(const KT *__restrict__ keys,
                            const uint32_t *__restrict__ vals,
                            const float *__restrict__ gauss,
                            const int64_t *__restrict__ n_invalid, int64_t S,
                            int64_t nper, int64_t G,
                            int64_t *__restrict__ indices,
                            int64_t *__restrict__ sort_perm,
                            int64_t *__restrict__ inv_perm,
                            float *__restrict__ gauss_sorted,
                            uint32_t *__restrict__ cell_ptr) {
  const int64_t total = S * nper;
  // wave-uniform trip count (every lane takes part in wave_fill)
  for (int64_t gi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
       gi - (threadIdx.x & (kWave - 1)) < total; gi += (int64_t)gridDim.x * blockDim.x) {
    uint32_t *cp = cell_ptr;
    int64_t lo = 0, hi = -1, lo2 = 0, hi2 = -1;  // cp[lo..hi] = v, cp[lo2..hi2] = v2
    uint32_t v = 0, v2 = 0;
    if (gi < total) {
      const uint64_t key = (uint64_t)keys[gi];
      const int64_t s = gi / nper;
      const int64_t iloc = gi - s * nper;
      const int64_t kloc = (int64_t)(key - (uint64_t)s * (uint64_t)(G + 1));
      const uint32_t val = vals[gi];
      const int64_t vloc = (int64_t)val - s * nper;
      indices[gi] = kloc > 0 ? kloc - 1 : 0;
      sort_perm[gi] = vloc;
      inv_perm[s * nper + vloc] = iloc;
      gauss_sorted[gi] = gauss[val];
      cp = cell_ptr + s * (G + 1);
      const int64_t ninv = n_invalid[s];
      if (iloc >= ninv) {
        // cells after the previous replica's cell up to this one start here
        const int64_t r = iloc - ninv;
        const int64_t cell = kloc - 1;
        int64_t prev = -1;
        if (r > 0) {
          const int64_t pk = (int64_t)((uint64_t)keys[gi - 1] - (uint64_t)s * (uint64_t)(G + 1));
          prev = pk - 1;
        }
        lo = prev + 1;
        hi = cell;
        v = (uint32_t)r;
        if (iloc == nper - 1) {
          lo2 = cell + 1;
          hi2 = G;
          v2 = (uint32_t)(r + 1);
        }
      } else if (iloc == nper - 1) {
        // no valid replica in this element
        lo = 0;
        hi = G;
        v = 0u;
      }
    }
    wave_fill(cp, lo, hi, v);
    wave_fill(cp, lo2, hi2, v2);
  }
}

// EGROG weights per cell, with the reference's float64 per-cell sum of the
// Gaussian weights in natural replica order (gauss_sum.scatter_add_, the
// stable sort keeps that order inside a cell) and the float32 density sum
// in the same order.  A 256-thread workgroup takes 256 consecutive cells per
// step: a thread handles its own cell when it has at most kBigCell replicas;
// larger cells (the k-space centre: tens of thousands of replicas) are done
// by the whole workgroup one after the other -- chunks of the cell's values
// staged in LDS by all threads, the two ordered sums run by one thread from
// LDS (the same additions in the same order, so the same bits), and the
// order-free per-replica writes by all threads.
constexpr int kBigCell = 256;
constexpr int kPwThreads = 256;
constexpr int kPwChunk = 2048;

// One big cell (more than kBigCell replicas) by the whole workgroup: the
// float64 Gaussian sum and the float32 density sum in replica order, the
// per-replica writes in parallel.  Every thread of the workgroup calls it.
struct PwShared {
  float val[2][kPwChunk];
  double rd[kPwThreads];
  int ri[kPwThreads];
  float denom;
};
__device__ void big_cell_weights(int64_t tb, PwShared &sh, const uint32_t *__restrict__ vals,
                                 const float *__restrict__ gauss_sorted, const int64_t *__restrict__ n_invalid,
                                 const uint32_t *__restrict__ cell_ptr, int64_t nper, int64_t G, int kw,
                                 float *__restrict__ weights, float *__restrict__ sqrt_weights,
                                 uint32_t *__restrict__ rep_src, float *__restrict__ rep_s,
                                 float *__restrict__ density, int64_t rep_stride) {
  constexpr int KB = 16;
  const int tid = threadIdx.x;
  float *s_val = sh.val[0];
  double *s_rd = sh.rd;
  int *s_ri = sh.ri;
  const int64_t s = tb / G;
  const int64_t c = tb - s * G;
  const uint32_t *cp = cell_ptr + s * (G + 1);
  const int64_t r0 = cp[c], r1 = cp[c + 1];
  const int64_t base = s * nper + n_invalid[s];
  const float *gs = gauss_sorted + base;
  // The float64 sum of the cell's values in replica order.  Every value
  // is a non-negative float32, an integer multiple of 2^L (L: the
  // smallest exponent of a least significant bit among them); while a
  // sum stays below 2^(53+L) every partial sum, in any order, is exact.
  // So the sum is first formed in parallel; if it is below that bound it
  // is the ordered sum's exact value (a rounded partial sum would have
  // been at or above the bound, and so the total too), otherwise the
  // ordered sum is formed one value after another as before.
  double part = 0.0;
  int lsb = INT_MAX;
  for (int64_t r = r0 + tid; r < r1; r += kPwThreads) {
    const float g = gs[r];
    part += (double)g;
    const uint32_t e = (__float_as_uint(g) >> 23) & 0xffu;
    if (g > 0.0f) lsb = min(lsb, (int)(e == 0u ? 1u : e) - 150);
  }
  s_rd[tid] = part;
  s_ri[tid] = lsb;
  __syncthreads();
  for (int h = kPwThreads / 2; h > 0; h >>= 1) {
    if (tid < h) {
      s_rd[tid] += s_rd[tid + h];
      s_ri[tid] = min(s_ri[tid], s_ri[tid + h]);
    }
    __syncthreads();
  }
  const double psum = s_rd[0];
  const int plsb = s_ri[0];
  const bool exact = plsb == INT_MAX || psum < ldexp(1.0, 53 + plsb);
  __syncthreads();
  double gsum = exact ? psum : 0.0;
  for (int64_t cb = r0; !exact && cb < r1; cb += kPwChunk) {
    const int n = (int)(r1 - cb < kPwChunk ? r1 - cb : kPwChunk);
    for (int k = tid; k < n; k += kPwThreads) s_val[k] = gs[cb + k];
    __syncthreads();
    if (tid == 0) {
      for (int k0 = 0; k0 < n; k0 += KB) {
        float g[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) g[k] = (k0 + k < n) ? s_val[k0 + k] : 0.0f;
#pragma unroll
        for (int k = 0; k < KB; ++k)
          if (k0 + k < n) gsum += (double)g[k];
      }
    }
    __syncthreads();
  }
  if (tid == 0) sh.denom = (float)(gsum > 1e-12 ? gsum : 1e-12);
  __syncthreads();
  const float denom = sh.denom;
  // per-replica writes in parallel, sw^2 staged in LDS for the ordered
  // density sum, double-buffered: while thread 0 adds chunk i in order, the
  // other three wavefronts write chunk i + 1 (all four write chunk 0)
  float dens = 0.0f;
  auto write_chunk = [&](int64_t cb, float *buf, int t0, int nt) {
    const int n = (int)(r1 - cb < kPwChunk ? r1 - cb : kPwChunk);
    for (int k = tid - t0; k < n; k += nt) {
      const int64_t r = cb + k;
      const uint32_t v = vals[base + r];
      const float w = gs[r] / denom;
      const float sw = sqrtf(w > 0.0f ? w : 0.0f);
      weights[v] = w;
      sqrt_weights[base + r] = sw;
      rep_s[s * rep_stride + r] = sw;
      rep_src[s * rep_stride + r] = (uint32_t)(((int64_t)v - s * nper) / kw);
      buf[k] = sw * sw;
    }
  };
  write_chunk(r0, s_val, 0, kPwThreads);
  __syncthreads();
  int ib = 0;
  for (int64_t cb = r0; cb < r1; cb += kPwChunk, ib ^= 1) {
    const int n = (int)(r1 - cb < kPwChunk ? r1 - cb : kPwChunk);
    if (tid < kWave) {
      if (tid == 0) {
        // the ordered float32 sum: 64 values per round of 16-byte reads
        const float *q4 = sh.val[ib];
        int k0 = 0;
        for (; k0 + 64 <= n; k0 += 64) {
          float4 q[16];
#pragma unroll
          for (int k = 0; k < 16; ++k) q[k] = reinterpret_cast<const float4 *>(q4 + k0)[k];
#pragma unroll
          for (int k = 0; k < 16; ++k) {
            dens = dens + q[k].x;
            dens = dens + q[k].y;
            dens = dens + q[k].z;
            dens = dens + q[k].w;
          }
        }
        for (; k0 < n; ++k0) dens = dens + q4[k0];
      }
    } else if (cb + kPwChunk < r1) {
      write_chunk(cb + kPwChunk, sh.val[ib ^ 1], kWave, kPwThreads - kWave);
    }
    __syncthreads();
  }
  if (tid == 0) density[tb] = dens;
}

__global__ __launch_bounds__(kPwThreads) void k_plan_weights(
    const uint32_t *__restrict__ vals, const float *__restrict__ gauss_sorted,
    const int64_t *__restrict__ n_invalid, const uint32_t *__restrict__ cell_ptr,
    int64_t S, int64_t nper, int64_t G, int kw, int chunk,
    float *__restrict__ weights, float *__restrict__ sqrt_weights,
    uint32_t *__restrict__ rep_src, float *__restrict__ rep_s,
    float *__restrict__ density, uint32_t *__restrict__ unit_count,
    int64_t rep_stride, int64_t big_min, int64_t *__restrict__ big_list, uint32_t *__restrict__ big_count) {
  __shared__ int64_t s_big[kPwThreads];
  __shared__ int s_nbig;
  __shared__ __align__(16) PwShared sh;
  const int tid = threadIdx.x;
  const int64_t total = S * G;
  constexpr int KB = 16;
  for (int64_t t0 = (int64_t)blockIdx.x * kPwThreads; t0 < total;
       t0 += (int64_t)gridDim.x * kPwThreads) {
    if (tid == 0) s_nbig = 0;
    __syncthreads();
    const int64_t t = t0 + tid;
    if (t < total) {
      const int64_t s = t / G;
      const int64_t c = t - s * G;
      const uint32_t *cp = cell_ptr + s * (G + 1);
      const int64_t r0 = cp[c], r1 = cp[c + 1];
      const int64_t base = s * nper + n_invalid[s];
      unit_count[t] = (uint32_t)((r1 - r0 + chunk - 1) / chunk);
      if (r1 - r0 > big_min) {
        // big cells: listed for k_plan_weights_big when a list is given,
        // else done by this workgroup below
        if (big_list != nullptr) big_list[atomicAdd(big_count, 1u)] = t;
        else s_big[atomicAdd(&s_nbig, 1)] = t;
      } else {
        // loads go 16 at a time (independent, in flight together) and are
        // added in order
        const float *gs = gauss_sorted + base;
        double gsum = 0.0;
        for (int64_t rb = r0; rb < r1; rb += KB) {
          float g[KB];
#pragma unroll
          for (int k = 0; k < KB; ++k) g[k] = (rb + k < r1) ? gs[rb + k] : 0.0f;
#pragma unroll
          for (int k = 0; k < KB; ++k)
            if (rb + k < r1) gsum += (double)g[k];
        }
        // gauss_sum[...].clamp(min=1e-12).float()
        const float denom = (float)(gsum > 1e-12 ? gsum : 1e-12);
        float dens = 0.0f;
        for (int64_t rb = r0; rb < r1; rb += KB) {
          float g[KB];
          uint32_t v[KB];
#pragma unroll
          for (int k = 0; k < KB; ++k) {
            g[k] = (rb + k < r1) ? gs[rb + k] : 0.0f;
            v[k] = (rb + k < r1) ? vals[base + rb + k] : 0u;
          }
#pragma unroll
          for (int k = 0; k < KB; ++k) {
            if (rb + k < r1) {
              const int64_t r = rb + k;
              const float w = g[k] / denom;
              const float sw = sqrtf(w > 0.0f ? w : 0.0f);
              weights[v[k]] = w;
              sqrt_weights[base + r] = sw;
              rep_s[s * rep_stride + r] = sw;
              rep_src[s * rep_stride + r] = (uint32_t)(((int64_t)v[k] - s * nper) / kw);
              dens = dens + sw * sw;
            }
          }
        }
        density[t] = dens;
      }
    }
    __syncthreads();
    const int nbig = s_nbig;
    for (int b = 0; b < nbig; ++b)
      big_cell_weights(s_big[b], sh, vals, gauss_sorted, n_invalid, cell_ptr, nper, G, kw, weights, sqrt_weights,
                       rep_src, rep_s, density, rep_stride);
  }
}

// The big cells a k_plan_weights launch listed (big_list), one workgroup at a
// time each, in parallel across the grid.
__global__ __launch_bounds__(kPwThreads) void k_plan_weights_big(
    const int64_t *__restrict__ big_list, const uint32_t *__restrict__ big_count,
    const uint32_t *__restrict__ vals, const float *__restrict__ gauss_sorted,
    const int64_t *__restrict__ n_invalid, const uint32_t *__restrict__ cell_ptr,
    int64_t nper, int64_t G, int kw, float *__restrict__ weights, float *__restrict__ sqrt_weights,
    uint32_t *__restrict__ rep_src, float *__restrict__ rep_s, float *__restrict__ density,
    int64_t rep_stride) {
  __shared__ __align__(16) PwShared sh;
  const uint32_t n = *big_count;
  for (uint32_t b = blockIdx.x; b < n; b += gridDim.x)
    big_cell_weights(big_list[b], sh, vals, gauss_sorted, n_invalid, cell_ptr, nper, G, kw, weights, sqrt_weights,
                     rep_src, rep_s, density, rep_stride);
}

__global__ void k_plan_units(const uint32_t *__restrict__ cell_ptr,
                             const uint32_t *__restrict__ cell_unit, int64_t S,
                             int64_t G, int chunk,
                             uint32_t *__restrict__ unit_cell,
                             uint32_t *__restrict__ unit_rep) {
  const int64_t total = S * G;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / G;
    const int64_t c = t - s * G;
    const uint32_t *cp = cell_ptr + s * (G + 1);
    const uint32_t u0 = cell_unit[t];
    const uint32_t nu = cell_unit[t + 1] - u0;
    for (uint32_t k = 0; k < nu; ++k) {
      unit_cell[u0 + k] = (uint32_t)c;
      unit_rep[u0 + k + s] = cp[c] + k * (uint32_t)chunk;
    }
    if (c == G - 1) unit_rep[cell_unit[t + 1] + s] = cp[G];
  }
}

// Hot cells and ordinary rows from the units.  A cell with nu units is
// hot when nu >= min_units (its units are reduced by the pre-pass); it is
// special (left out of the tile kernels' row numbering) when it is hot or
// has more than one unit.  Pass 1 writes, per cell t of all S*G cells (and a
// zero at t = S*G for the exclusive scans that follow): a[t] = the cell's
// replicas if it is ordinary, else 0; hf[t] = 1 if hot; hn[t] = nu if hot.
__global__ void k_plan_hot_counts(const uint32_t *__restrict__ cell_ptr,
                                  const uint32_t *__restrict__ cell_unit,
                                  int64_t S, int64_t G, int64_t min_units,
                                  uint32_t *__restrict__ a, uint32_t *__restrict__ hf,
                                  uint32_t *__restrict__ hn) {
  const int64_t total = S * G;
  const int64_t lim = min_units > 1 ? min_units : 1;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t <= total;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t == total) {
      a[t] = hf[t] = hn[t] = 0u;
      continue;
    }
    const int64_t s = t / G;
    const int64_t c = t - s * G;
    const uint32_t *cp = cell_ptr + s * (G + 1);
    const int64_t nu = (int64_t)(cell_unit[t + 1] - cell_unit[t]);
    const bool hot = nu >= min_units;
    const bool special = nu >= lim || nu > 1;
    a[t] = special ? 0u : cp[c + 1] - cp[c];
    hf[t] = hot ? 1u : 0u;
    hn[t] = hot ? (uint32_t)nu : 0u;
  }
}

// Pass 2, from the exclusive scans A, HF, HN of pass 1 (S*G + 1 entries):
// cell_vrow (S, G+1) = A minus the element's base; for element s with hot
// cells h = 0..H_s-1 (cell order) the element's hot_ptr is the H_s + 1
// entries at hot_ptr + HF[s*G] + s (unit offsets from the element's first
// hot unit), its hot units the HN[(s+1)*G] - HN[s*G] entries at hot_units +
// HN[s*G] (global unit ids in cell, then unit order); unit_hot[u] = the
// element-local hot index of unit u's cell, -1 if the cell is not hot.
__global__ void k_plan_hot_rows(const uint32_t *__restrict__ cell_unit,
                                int64_t S, int64_t G, int64_t min_units,
                                const uint32_t *__restrict__ A,
                                const uint32_t *__restrict__ HF,
                                const uint32_t *__restrict__ HN,
                                uint32_t *__restrict__ cell_vrow,
                                int32_t *__restrict__ unit_hot,
                                uint32_t *__restrict__ hot_units,
                                uint32_t *__restrict__ hot_ptr) {
  const int64_t total = S * G;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / G;
    const int64_t c = t - s * G;
    const uint32_t a0 = A[s * G], h0 = HF[s * G], u0 = HN[s * G];
    uint32_t *vr = cell_vrow + s * (G + 1);
    vr[c] = A[t] - a0;
    const uint32_t cu = cell_unit[t];
    const int64_t nu = (int64_t)(cell_unit[t + 1] - cu);
    if (nu >= min_units) {
      const uint32_t h = HF[t] - h0;
      hot_ptr[h0 + s + h] = HN[t] - u0;
      for (int64_t k = 0; k < nu; ++k) {
        hot_units[HN[t] + k] = cu + (uint32_t)k;
        unit_hot[cu + k] = (int32_t)h;
      }
    } else {
      for (int64_t k = 0; k < nu; ++k) unit_hot[cu + k] = -1;
    }
    if (c == G - 1) {
      const int64_t e = (s + 1) * G;
      vr[G] = A[e] - a0;
      hot_ptr[h0 + s + (HF[e] - h0)] = HN[e] - u0;
    }
  }
}

// Operator table from the per-axis fractional powers P (ndim, n, C, C)
// complex128 (P[0] = Gx, P[1] = Gy, P[2] = Gz), one workgroup per entry and
// one thread per output element (r, c):
//   2D: T[i n + j]         = Gx[j] Gy[i]
//   3D: T[i n^2 + j n + k] = (F[k] Gy[j]) L[i],  (F, L) = (Gx, Gz), or (Gz, Gx)
//       with `corrected`
// in FP64 (every complex product an fma chain over the inner index in
// ascending order, the 3D intermediate rounded to FP64 as a separate
// product would be), written as complex64 (round to nearest) and, when
// t64 is given, complex128.  The operands of an entry are staged in LDS.
constexpr int kComposeMaxC = 32;
__device__ __forceinline__ double2 cmac(double2 acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}
__global__ __launch_bounds__(kComposeMaxC * kComposeMaxC) void k_table_compose(
    const double2 *__restrict__ P, int ndim, int64_t n, int C, int corrected,
    float2 *__restrict__ t32, double2 *__restrict__ t64) {
  __shared__ double2 s_a[kComposeMaxC * kComposeMaxC];
  __shared__ double2 s_b[kComposeMaxC * kComposeMaxC];
  __shared__ double2 s_l[kComposeMaxC * kComposeMaxC];
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x;
  const int CC = C * C;
  const int r = t / C, c = t - (t / C) * C;
  const int64_t mat = (int64_t)CC;
  const double2 *Gx = P, *Gy = P + n * mat, *Gz = P + 2 * n * mat;
  double2 acc = make_double2(0.0, 0.0);
  if (ndim == 2) {
    const int64_t i = e / n, j = e - (e / n) * n;
    if (t < CC) {
      s_a[t] = Gx[j * mat + t];
      s_b[t] = Gy[i * mat + t];
    }
    __syncthreads();
    if (t < CC)
      for (int l = 0; l < C; ++l) acc = cmac(acc, s_a[r * C + l], s_b[l * C + c]);
  } else {
    const int64_t i = e / (n * n), rem = e - (e / (n * n)) * (n * n);
    const int64_t j = rem / n, k = rem - (rem / n) * n;
    const double2 *F = corrected ? Gz : Gx, *L = corrected ? Gx : Gz;
    if (t < CC) {
      s_a[t] = F[k * mat + t];
      s_b[t] = Gy[j * mat + t];
      s_l[t] = L[i * mat + t];
    }
    __syncthreads();
    double2 m = make_double2(0.0, 0.0);
    if (t < CC)
      for (int l = 0; l < C; ++l) m = cmac(m, s_a[r * C + l], s_b[l * C + c]);
    __syncthreads();
    if (t < CC) s_a[t] = m;
    __syncthreads();
    if (t < CC)
      for (int l = 0; l < C; ++l) acc = cmac(acc, s_a[r * C + l], s_l[l * C + c]);
  }
  if (t < CC) {
    t32[e * mat + t] = make_float2((float)acc.x, (float)acc.y);
    if (t64 != nullptr) t64[e * mat + t] = acc;
  }
}

// Normal-equation products of the Cartesian GRAPPA fit, in FP64, from the
// coil-major calibration x (C, P) complex128 over P = n0 n1 n2 positions
// (C-order, unused dims 1).  Block (k, m): for m < 2 ndim, axis d = m / 2,
// output element (i, j) = k of
//   m even: G_d[i][j] = sum_p conj(x_i[p]) x_j[p]
//   m odd:  X_d[i][j] = sum_p conj(x_i[p]) x_j[p + stride_d]
// over the positions p not on the last slice along d (the rows of the
// reference's source / target matrices, _grappa.py:290-368); m = 2 ndim,
// k = i < C: sum_p |x_i[p]|^2.  Each thread sums its strided positions,
// then a fixed tree: the same bits on every run.
constexpr int kGramThreads = 256;
__global__ __launch_bounds__(kGramThreads) void k_grappa_gram(
    const double2 *__restrict__ x, int64_t C, int ndim, int64_t n0, int64_t n1,
    int64_t n2, double2 *__restrict__ out) {
  __shared__ double s_re[kGramThreads], s_im[kGramThreads];
  const int m = blockIdx.y;
  const int64_t k = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t P = n0 * n1 * n2;
  double re = 0.0, im = 0.0;
  if (m == 2 * ndim) {
    if (k >= C) return;
    const double2 *xi = x + k * P;
    for (int64_t p = tid; p < P; p += kGramThreads) {
      const double2 a = xi[p];
      re += a.x * a.x + a.y * a.y;
    }
  } else {
    if (k >= C * C) return;
    const int64_t i = k / C, j = k - (k / C) * C;
    const int d = m / 2;
    const bool cross = (m & 1) != 0;
    const int64_t nd = d == 0 ? n0 : d == 1 ? n1 : n2;
    const int64_t st = d == 0 ? n1 * n2 : d == 1 ? n2 : 1;  // stride of axis d
    const double2 *xi = x + i * P;
    const double2 *xj = x + j * P + (cross ? st : 0);
    for (int64_t p = tid; p < P; p += kGramThreads) {
      if ((p / st) % nd == nd - 1) continue;
      const double2 a = xi[p], b = xj[p];
      re += a.x * b.x + a.y * b.y;
      im += a.x * b.y - a.y * b.x;
    }
  }
  s_re[tid] = re;
  s_im[tid] = im;
  __syncthreads();
  for (int h = kGramThreads / 2; h > 0; h >>= 1) {
    if (tid < h) {
      s_re[tid] += s_re[tid + h];
      s_im[tid] += s_im[tid + h];
    }
    __syncthreads();
  }
  if (tid == 0) {
    double2 r;
    r.x = s_re[0];
    r.y = s_im[0];
    out[m == 2 * ndim ? 2 * ndim * C * C + k : (int64_t)m * C * C + k] = r;
  }
}

__global__ void k_plan_quantize(const float *__restrict__ distances,
                                const int64_t *__restrict__ inv_perm,
                                const int64_t *__restrict__ n_invalid,
                                const int64_t *__restrict__ target_idx,
                                int64_t S, int64_t nper, int ndim, float pfac,
                                float stepsize, float radius, int64_t nsteps,
                                int64_t n_entries,
                                int64_t *__restrict__ interp_idx,
                                uint32_t *__restrict__ rep_ent,
                                int64_t rep_stride) {
  const int64_t total = S * nper;
  for (int64_t gi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; gi < total;
       gi += (int64_t)gridDim.x * blockDim.x) {
    // (radius + round(d * pfac) / pfac) / stepsize -> round -> * nsteps^d ->
    // round -> int32 -> sum
    int64_t e = 0;
    float stride_f = 1.0f;
    for (int d = 0; d < ndim; ++d) {
      const float dd = distances[gi * ndim + d];
      const float a = __builtin_rintf(dd * pfac) / pfac;
      const float q = __builtin_rintf((radius + a) / stepsize);
      e += (int64_t)(int32_t)__builtin_rintf(q * stride_f);
      stride_f = (float)(stride_f * (float)nsteps);
    }
    if (e < 0) e = 0;
    if (e > n_entries - 1) e = n_entries - 1;
    interp_idx[gi] = e;
    if (target_idx[gi] >= 0) {
      const int64_t s = gi / nper;
      const int64_t r = inv_perm[gi] - n_invalid[s];
      rep_ent[s * rep_stride + r] = (uint32_t)e;
    }
  }
}

// cell_ptr from sorted indices (all inside [0, G)).
template <class I>
__global__ void k_csr_from_sorted(const I *__restrict__ idx, int64_t n,
                                  int64_t G, uint32_t *__restrict__ cp) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t cur = (i < n) ? (int64_t)idx[i] : G;
    const int64_t prev = (i > 0) ? (int64_t)idx[i - 1] : -1;
    for (int64_t c = prev + 1; c <= cur; ++c) cp[c] = (uint32_t)i;
  }
}

}  // namespace pg

using namespace pg;

namespace pg {

template <class KT>
int plan_targets_keyed(const float *coords, const float *cmax, int64_t S, int64_t npts, pg_plan_geom geom,
                       int64_t *target_idx, float *distances, float *gauss, KT *sort_key, uint32_t *sort_val,
                       int64_t *n_invalid, pg_stream_t stream) {
  PG_CHECK(geom.ndim >= 1 && geom.ndim <= 3, "plan: ndim must be 1..3");
  PG_CHECK(geom.kw >= 1 && geom.kw <= 128, "plan: kernel size must be 1..128");
  const int64_t total = S * npts * geom.kw;
  PG_CHECK(total < (int64_t)UINT32_MAX, "plan: too many candidate replicas");
  if (total == 0) return 0;
  GeomDev g;
  g.ndim = geom.ndim;
  g.kw = geom.kw;
  for (int d = 0; d < 3; ++d) {
    g.gs[d] = geom.grid_shape[d];
    g.st[d] = geom.strides[d];
    g.ha[d] = geom.half_amp[d];
    g.org[d] = geom.origins[d];
    g.stp[d] = geom.steps[d];
  }
  for (int k = 0; k < 128; ++k)
    for (int d = 0; d < 3; ++d) g.off[k][d] = geom.offsets[k][d];
  const int threads = 256;
  // a persistent grid of at most 2048 workgroups (8 per CU): each wavefront
  // walks many replicas and adds its invalid count once
  hipLaunchKernelGGL(k_plan_targets<KT>, dim3(grid_for(total, threads, 2048)),
                     dim3(threads), 0, as_stream(stream), coords, cmax, S,
                     npts, g, target_idx, distances, gauss, sort_key, sort_val,
                     reinterpret_cast<unsigned long long *>(n_invalid));
  PG_LAUNCHED();
  return 0;
}

template <class KT>
int plan_post_keyed(const KT *keys_sorted, const uint32_t *vals_sorted, const float *gauss, const int64_t *n_invalid,
                    int64_t S, int64_t nper, int64_t G, int64_t *indices, int64_t *sort_perm, int64_t *inv_perm,
                    float *gauss_sorted, uint32_t *cell_ptr, pg_stream_t stream) {
  const int64_t total = S * nper;
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_plan_post<KT>, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream), keys_sorted,
                     vals_sorted, gauss, n_invalid, S, nper, G, indices,
                     sort_perm, inv_perm, gauss_sorted, cell_ptr);
  PG_LAUNCHED();
  return 0;
}

// Cells above this many replicas go to k_plan_weights_big when the caller
// lists them: one workgroup per cell, in parallel, so a thread walks at most
// this many replicas in order.
constexpr int64_t kBigCellListed = 64;
int64_t plan_big_cells_max(int64_t S, int64_t nper) { return S * nper / (kBigCellListed + 1) + 1; }

// pg_plan_weights with the big cells listed (big_list: room for
// plan_big_cells_max cells; *big_count zero on entry) and done by
// k_plan_weights_big, one workgroup per cell across the grid, instead of
// one after another by the workgroup that holds them (the cells round the
// k-space centre of a 3D kooshball sit in a few workgroups)
int plan_weights_listed(const uint32_t *vals_sorted, const float *gauss_sorted, const int64_t *n_invalid,
                        const uint32_t *cell_ptr, int64_t S, int64_t nper, int64_t G, int32_t kw, int32_t chunk,
                        float *weights, float *sqrt_weights, uint32_t *rep_src, float *rep_s, float *density,
                        uint32_t *unit_count, int64_t rep_stride, int64_t *big_list, uint32_t *big_count,
                        pg_stream_t stream) {
  PG_CHECK(chunk >= 1, "plan: chunk must be positive");
  const int64_t total = S * G;
  if (total == 0) return 0;
  const hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_plan_weights, dim3(grid_for(total, kPwThreads)), dim3(kPwThreads), 0, st, vals_sorted,
                     gauss_sorted, n_invalid, cell_ptr, S, nper, G, kw, chunk, weights, sqrt_weights, rep_src,
                     rep_s, density, unit_count, rep_stride, (int64_t)kBigCellListed, big_list, big_count);
  PG_LAUNCHED();
  const int64_t most = plan_big_cells_max(S, nper);
  hipLaunchKernelGGL(k_plan_weights_big, dim3((unsigned)(most < 2048 ? most : 2048)), dim3(kPwThreads), 0, st,
                     big_list, big_count, vals_sorted, gauss_sorted, n_invalid, cell_ptr, nper, G, kw, weights,
                     sqrt_weights, rep_src, rep_s, density, rep_stride);
  PG_LAUNCHED();
  return 0;
}

// 32-bit-key plan steps for the one-call compiler (plan_compile.hip)
int plan_targets_k32(const float *coords, const float *cmax, int64_t S, int64_t npts, pg_plan_geom geom,
                     int64_t *target_idx, float *distances, float *gauss, uint32_t *sort_key, uint32_t *sort_val,
                     int64_t *n_invalid, pg_stream_t stream) {
  return plan_targets_keyed(coords, cmax, S, npts, geom, target_idx, distances, gauss, sort_key, sort_val,
                            n_invalid, stream);
}
int plan_post_k32(const uint32_t *keys_sorted, const uint32_t *vals_sorted, const float *gauss,
                  const int64_t *n_invalid, int64_t S, int64_t nper, int64_t G, int64_t *indices, int64_t *sort_perm,
                  int64_t *inv_perm, float *gauss_sorted, uint32_t *cell_ptr, pg_stream_t stream) {
  return plan_post_keyed(keys_sorted, vals_sorted, gauss, n_invalid, S, nper, G, indices, sort_perm, inv_perm,
                         gauss_sorted, cell_ptr, stream);
}

}  // namespace pg

extern "C" {

int pg_plan_absmax(const float *coords, int64_t S, int64_t npts, int32_t ndim,
                   float *cmax, pg_stream_t stream) {
  PG_CHECK(ndim >= 1 && ndim <= 3, "plan: ndim must be 1..3");
  PG_CHECK(S >= 1 && npts >= 0 && S <= 65535, "plan: bad stack/points");
  if (npts == 0) return 0;
  dim3 grid(grid_for(npts, kAbsmaxThreads, 512), (unsigned)S);
  hipLaunchKernelGGL(k_absmax, grid, dim3(kAbsmaxThreads), 0, as_stream(stream),
                     coords, S, npts, ndim, reinterpret_cast<unsigned *>(cmax));
  PG_LAUNCHED();
  return 0;
}

int pg_plan_targets(const float *coords, const float *cmax, int64_t S,
                    int64_t npts, pg_plan_geom geom, int64_t *target_idx,
                    float *distances, float *gauss, uint64_t *sort_key,
                    uint32_t *sort_val, int64_t *n_invalid,
                    pg_stream_t stream) {
  return pg::plan_targets_keyed(coords, cmax, S, npts, geom, target_idx, distances, gauss, sort_key, sort_val,
                                n_invalid, stream);
}

}  // extern "C"

// The library's key-value sorts: rocPRIM's onesweep radix sort for every
// size above one block (its default switches to a merge sort below 2^20
// items, about 18 launches for a 2^17-item plan against 8; both are stable,
// so the permutation is the same).
namespace {
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;
}

extern "C" {

int pg_sort_pairs_u64(void *workspace, size_t *ws_bytes,
                      const uint64_t *keys_in, uint64_t *keys_out,
                      const uint32_t *vals_in, uint32_t *vals_out, int64_t n,
                      int32_t end_bit, pg_stream_t stream) {
  PG_CHECK(ws_bytes != nullptr, "sort: ws_bytes is null");
  PG_CHECK(n >= 0 && n < (int64_t)INT32_MAX, "sort: bad n");
  PG_CHECK(end_bit >= 1 && end_bit <= 64, "sort: bad end_bit");
  PG_HIP(rocprim::radix_sort_pairs<SortConfig>(workspace, *ws_bytes, keys_in, keys_out, vals_in, vals_out,
                                               (unsigned int)n, 0u, (unsigned int)end_bit, as_stream(stream)));
  return 0;
}

int pg_sort_pairs_u32(void *workspace, size_t *ws_bytes,
                      const uint32_t *keys_in, uint32_t *keys_out,
                      const uint32_t *vals_in, uint32_t *vals_out, int64_t n,
                      int32_t end_bit, pg_stream_t stream) {
  PG_CHECK(ws_bytes != nullptr, "sort: ws_bytes is null");
  PG_CHECK(n >= 0 && n < (int64_t)INT32_MAX, "sort: bad n");
  PG_CHECK(end_bit >= 1 && end_bit <= 32, "sort: bad end_bit");
  PG_HIP(rocprim::radix_sort_pairs<SortConfig>(workspace, *ws_bytes, keys_in, keys_out, vals_in, vals_out,
                                               (unsigned int)n, 0u, (unsigned int)end_bit, as_stream(stream)));
  return 0;
}

}  // extern "C"

namespace pg {
// Key-value sort of (keys, 0 .. n-1): the values come from a counting
// iterator instead of a materialised index array (same sort, one launch and
// 8 n bytes less).
int sort_pairs_u32_iota(void *workspace, size_t *ws_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                        uint32_t *vals_out, int64_t n, int32_t end_bit, pg_stream_t stream) {
  PG_CHECK(ws_bytes != nullptr, "sort: ws_bytes is null");
  PG_CHECK(n >= 0 && n < (int64_t)INT32_MAX, "sort: bad n");
  PG_CHECK(end_bit >= 1 && end_bit <= 32, "sort: bad end_bit");
  PG_HIP(rocprim::radix_sort_pairs<SortConfig>(workspace, *ws_bytes, keys_in, keys_out,
                                               rocprim::counting_iterator<uint32_t>(0u), vals_out, (unsigned int)n,
                                               0u, (unsigned int)end_bit, as_stream(stream)));
  return 0;
}
}  // namespace pg

extern "C" {

int pg_exclusive_scan_u32(void *workspace, size_t *ws_bytes,
                          const uint32_t *in, uint32_t *out, int64_t n,
                          pg_stream_t stream) {
  PG_CHECK(ws_bytes != nullptr, "scan: ws_bytes is null");
  PG_CHECK(n >= 0 && n < (int64_t)INT32_MAX, "scan: bad n");
  // Scan n + 1 items where the last input is read as 0: the caller passes an
  // input array of n + 1 entries with in[n] == 0.
  PG_HIP(hipcub::DeviceScan::ExclusiveSum(workspace, *ws_bytes, in, out,
                                          (int)(n + 1), as_stream(stream)));
  return 0;
}

int pg_plan_post(const uint64_t *keys_sorted, const uint32_t *vals_sorted,
                 const float *gauss, const int64_t *n_invalid, int64_t S,
                 int64_t nper, int64_t G, int64_t *indices, int64_t *sort_perm,
                 int64_t *inv_perm, float *gauss_sorted, uint32_t *cell_ptr,
                 pg_stream_t stream) {
  return pg::plan_post_keyed(keys_sorted, vals_sorted, gauss, n_invalid, S, nper, G, indices, sort_perm, inv_perm,
                             gauss_sorted, cell_ptr, stream);
}

int pg_plan_weights(const uint32_t *vals_sorted, const float *gauss_sorted,
                    const int64_t *n_invalid, const uint32_t *cell_ptr,
                    int64_t S, int64_t nper, int64_t G, int32_t kw,
                    int32_t chunk, float *weights, float *sqrt_weights,
                    uint32_t *rep_src, float *rep_s, float *density,
                    uint32_t *unit_count, int64_t rep_stride,
                    pg_stream_t stream) {
  PG_CHECK(chunk >= 1, "plan: chunk must be positive");
  const int64_t total = S * G;
  if (total == 0) return 0;
  hipLaunchKernelGGL(k_plan_weights, dim3(grid_for(total, kPwThreads)),
                     dim3(kPwThreads), 0, as_stream(stream), vals_sorted,
                     gauss_sorted, n_invalid, cell_ptr, S, nper, G, kw, chunk,
                     weights, sqrt_weights, rep_src, rep_s, density,
                     unit_count, rep_stride, (int64_t)kBigCell, (int64_t *)nullptr, (uint32_t *)nullptr);
  PG_LAUNCHED();
  return 0;
}

int pg_plan_units(const uint32_t *cell_ptr, const uint32_t *cell_unit,
                  int64_t S, int64_t G, int32_t chunk, uint32_t *unit_cell,
                  uint32_t *unit_rep, pg_stream_t stream) {
  const int64_t total = S * G;
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_plan_units, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream), cell_ptr, cell_unit,
                     S, G, chunk, unit_cell, unit_rep);
  PG_LAUNCHED();
  return 0;
}

int pg_plan_hot_counts(const uint32_t *cell_ptr, const uint32_t *cell_unit,
                       int64_t S, int64_t G, int64_t min_units, uint32_t *a,
                       uint32_t *hf, uint32_t *hn, pg_stream_t stream) {
  PG_CHECK(S >= 1 && G >= 1, "plan_hot_counts: bad sizes");
  const int threads = 256;
  hipLaunchKernelGGL(k_plan_hot_counts, dim3(grid_for(S * G + 1, threads)),
                     dim3(threads), 0, as_stream(stream), cell_ptr, cell_unit,
                     S, G, min_units, a, hf, hn);
  PG_LAUNCHED();
  return 0;
}

int pg_plan_hot_rows(const uint32_t *cell_unit, int64_t S, int64_t G,
                     int64_t min_units, const uint32_t *A, const uint32_t *HF,
                     const uint32_t *HN, uint32_t *cell_vrow, int32_t *unit_hot,
                     uint32_t *hot_units, uint32_t *hot_ptr, pg_stream_t stream) {
  PG_CHECK(S >= 1 && G >= 1, "plan_hot_rows: bad sizes");
  const int threads = 256;
  hipLaunchKernelGGL(k_plan_hot_rows, dim3(grid_for(S * G, threads)),
                     dim3(threads), 0, as_stream(stream), cell_unit, S, G,
                     min_units, A, HF, HN, cell_vrow, unit_hot, hot_units,
                     hot_ptr);
  PG_LAUNCHED();
  return 0;
}

int pg_grappa_gram_f64(const void *calib, int64_t C, int32_t ndim,
                       const int64_t *dims, void *out, pg_stream_t stream) {
  PG_CHECK(C >= 1 && C <= 256 && (ndim == 2 || ndim == 3) && dims != nullptr,
           "grappa_gram: bad sizes");
  int64_t n[3] = {1, 1, 1};
  for (int d = 0; d < ndim; ++d) {
    PG_CHECK(dims[d] >= 2, "grappa_gram: every calibration axis needs 2 or more points");
    n[d] = dims[d];
  }
  const dim3 grid((unsigned)(C * C), (unsigned)(2 * ndim + 1));
  hipLaunchKernelGGL(k_grappa_gram, grid, dim3(kGramThreads), 0, as_stream(stream),
                     static_cast<const double2 *>(calib), C, (int)ndim, n[0], n[1],
                     n[2], static_cast<double2 *>(out));
  PG_LAUNCHED();
  return 0;
}

int pg_table_compose_f64(const void *powers, int32_t ndim, int64_t nsteps, int64_t C,
                         int32_t corrected, void *table_c64, void *table_c128,
                         pg_stream_t stream) {
  PG_CHECK((ndim == 2 || ndim == 3) && nsteps >= 1 && C >= 1 && C <= kComposeMaxC && powers && table_c64,
           "table_compose: bad arguments (ndim %d, nsteps %lld, C %lld; at most %d coils)", (int)ndim,
           (long long)nsteps, (long long)C, kComposeMaxC);
  const int64_t K = ndim == 2 ? nsteps * nsteps : nsteps * nsteps * nsteps;
  PG_CHECK(K < (int64_t)INT32_MAX, "table_compose: too many entries");
  const int threads = (int)((C * C + 63) / 64 * 64);
  hipLaunchKernelGGL(k_table_compose, dim3((unsigned)K), dim3(threads), 0, as_stream(stream),
                     static_cast<const double2 *>(powers), (int)ndim, nsteps, (int)C, (int)corrected,
                     static_cast<float2 *>(table_c64), static_cast<double2 *>(table_c128));
  PG_LAUNCHED();
  return 0;
}

int pg_plan_quantize(const float *distances, const int64_t *inv_perm,
                     const int64_t *n_invalid, const int64_t *target_idx,
                     int64_t S, int64_t nper, int32_t ndim, float pfac,
                     float stepsize, float radius, int64_t nsteps,
                     int64_t n_entries, int64_t *interp_idx, uint32_t *rep_ent,
                     int64_t rep_stride, pg_stream_t stream) {
  PG_CHECK(ndim >= 1 && ndim <= 3, "quantize: ndim must be 1..3");
  PG_CHECK(n_entries >= 1, "quantize: empty table");
  const int64_t total = S * nper;
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_plan_quantize, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream), distances, inv_perm,
                     n_invalid, target_idx, S, nper, ndim, pfac, stepsize,
                     radius, nsteps, n_entries, interp_idx, rep_ent,
                     rep_stride);
  PG_LAUNCHED();
  return 0;
}

int pg_csr_from_sorted(const int64_t *indices, int64_t n, int64_t G,
                       uint32_t *cell_ptr, pg_stream_t stream) {
  PG_CHECK(n >= 0 && G >= 0 && n < (int64_t)UINT32_MAX, "csr: bad sizes");
  const int threads = 256;
  hipLaunchKernelGGL(k_csr_from_sorted<int64_t>,
                     dim3(grid_for(n + 1, threads)), dim3(threads), 0,
                     as_stream(stream), indices, n, G, cell_ptr);
  PG_LAUNCHED();
  return 0;
}

int pg_csr_from_sorted_u32(const uint32_t *keys, int64_t n, int64_t G,
                           uint32_t *cell_ptr, pg_stream_t stream) {
  PG_CHECK(n >= 0 && G >= 0 && n < (int64_t)UINT32_MAX, "csr: bad sizes");
  const int threads = 256;
  hipLaunchKernelGGL(k_csr_from_sorted<uint32_t>,
                     dim3(grid_for(n + 1, threads)), dim3(threads), 0,
                     as_stream(stream), keys, n, G, cell_ptr);
  PG_LAUNCHED();
  return 0;
}

}  // extern "C"
