// The seven ops of the reference's native module (csrc/torch/module.cpp:54-87)
// as gfx950 kernels behind the C ABI of include/pygrog_hip.h.
//
//   grog_interpolate   one thread per (row, output coil); the input row is
//                      shared by the C threads of a row (one transaction per
//                      coil), the operator row streams per thread.
//   scatter_add(_binned), psf_scatter_scalar
//                      one element per lane; runs of equal indices inside a
//                      wavefront are reduced with a segmented shuffle scan
//                      and written with one hardware atomic per run, so the
//                      radial DC run costs one atomic per 64 samples instead
//                      of one per sample (the reference CUDA path serialises
//                      on it, sparse_ops_cuda.cu:27-63).
//   gather             one element per thread.
//   psf_scatter_outer(_basis)
//                      one sample per thread, M*M complex atomics.
#include <cstring>

#include "common.h"

namespace pg {

static thread_local char g_err[1024] = {0};

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------------------
// grog_interpolate: out[nb, co] = sum_ci K[idx[n], co, ci] * data[nb, ci]
// (reference grog_interp.cpp:39-73: same summation order over ci).
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_grog_interp(const v2_t<T> *__restrict__ data,
                              const int64_t *__restrict__ idx,
                              const v2_t<T> *__restrict__ kern,
                              v2_t<T> *__restrict__ out, int64_t N, int64_t B,
                              int64_t C, int64_t K) {
  const int64_t total = N * B * C;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t nb = t / C;
    const int64_t co = t - nb * C;
    const int64_t n = nb / B;
    const int64_t k = idx[n];
    v2_t<T> acc = cmake<T>(T(0), T(0));
    if (k >= 0 && k < K) {
      const v2_t<T> *krow = kern + (k * C + co) * C;
      const v2_t<T> *src = data + nb * C;
      for (int64_t ci = 0; ci < C; ++ci) cmac<T>(acc, krow[ci], src[ci]);
    }
    out[t] = acc;
  }
}

// ---------------------------------------------------------------------------
// Wavefront segmented reduction + one atomic per run.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void atomic_add_val(float *p, float v) {
  unsafeAtomicAdd(p, v);
}
__device__ __forceinline__ void atomic_add_val(double *p, double v) {
  unsafeAtomicAdd(p, v);
}

// NV = number of scalar components per element (2 complex, 1 real).
template <class R, int NV>
struct Vals {
  R v[NV];
};

template <class R, int NV>
__device__ __forceinline__ void run_reduce_store(int64_t key, Vals<R, NV> val,
                                                 R *grid) {
  const int lane = threadIdx.x & (kWave - 1);
  // Head flags -> segment ids (inclusive prefix count of heads).
  int64_t prev = __shfl_up(key, 1, kWave);
  int head = (lane == 0 || prev != key) ? 1 : 0;
  int seg = head;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    int o = __shfl_up(seg, off, kWave);
    if (lane >= off) seg += o;
  }
  // Hillis-Steele inclusive segmented scan of the values.
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    int oseg = __shfl_up(seg, off, kWave);
    Vals<R, NV> o;
#pragma unroll
    for (int j = 0; j < NV; ++j) o.v[j] = __shfl_up(val.v[j], off, kWave);
    if (lane >= off && oseg == seg) {
#pragma unroll
      for (int j = 0; j < NV; ++j) val.v[j] = o.v[j] + val.v[j];
    }
  }
  int nseg = __shfl_down(seg, 1, kWave);
  const bool tail = (lane == kWave - 1) || (nseg != seg);
  if (tail && key >= 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) atomic_add_val(grid + key * NV + j, val.v[j]);
  }
}

// grid[idx[i]] += w[i] * data[i]; complex (NV = 2) with real weights.
// blockIdx.y selects a row: grid (rows, G) and data (rows, n) share idx, w.
template <class R>
__global__ void k_scatter_add(R *__restrict__ grid, int64_t G,
                              const R *__restrict__ data,
                              const int64_t *__restrict__ idx,
                              const R *__restrict__ w, int64_t n) {
  grid += (int64_t)blockIdx.y * 2 * G;
  data += (int64_t)blockIdx.y * 2 * n;
  const int64_t nwaves_total = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t wave0 = blockIdx.x * (int64_t)(blockDim.x / kWave) +
                        threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  for (int64_t base = wave0 * kWave; base < n; base += nwaves_total * kWave) {
    const int64_t i = base + lane;
    int64_t key = -1;
    Vals<R, 2> v;
    v.v[0] = R(0);
    v.v[1] = R(0);
    if (i < n) {
      const int64_t k = idx[i];
      if (k >= 0 && k < G) {
        key = k;
        const R wi = w[i];
        v.v[0] = wi * data[2 * i];
        v.v[1] = wi * data[2 * i + 1];
      }
    }
    run_reduce_store<R, 2>(key, v, grid);
  }
}

// psf[idx[i]] += w_sq[i] (real).
template <class R>
__global__ void k_psf_scalar(R *__restrict__ psf, int64_t G,
                             const int64_t *__restrict__ idx,
                             const R *__restrict__ w, int64_t n) {
  const int64_t nwaves_total = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t wave0 = blockIdx.x * (int64_t)(blockDim.x / kWave) +
                        threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  for (int64_t base = wave0 * kWave; base < n; base += nwaves_total * kWave) {
    const int64_t i = base + lane;
    int64_t key = -1;
    Vals<R, 1> v;
    v.v[0] = R(0);
    if (i < n) {
      const int64_t k = idx[i];
      if (k >= 0 && k < G) {
        key = k;
        v.v[0] = w[i];
      }
    }
    run_reduce_store<R, 1>(key, v, psf);
  }
}

// out[i] = w[i] * grid[idx[i]].
template <class R>
__global__ void k_gather(const R *__restrict__ grid, int64_t G,
                         const int64_t *__restrict__ idx,
                         const R *__restrict__ w, int64_t n,
                         R *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = idx[i];
    R re = R(0), im = R(0);
    if (k >= 0 && k < G) {
      const R wi = w[i];
      re = wi * grid[2 * k];
      im = wi * grid[2 * k + 1];
    }
    out[2 * i] = re;
    out[2 * i + 1] = im;
  }
}

// psf[idx, l, l'] += w * conj(b_l) * b_l' (reference toep_psf.cpp:124-175).
template <class R>
__global__ void k_psf_outer(R *__restrict__ psf, int64_t G, int64_t M,
                            const int64_t *__restrict__ idx,
                            const R *__restrict__ w,
                            const R *__restrict__ basis, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = idx[i];
    if (k < 0 || k >= G) continue;
    const R wi = w[i];
    const R *b = basis + 2 * i * M;
    R *p = psf + 2 * k * M * M;
    for (int64_t l = 0; l < M; ++l) {
      // cb = conj(b_l) * w
      const R cr = b[2 * l] * wi, ci = -b[2 * l + 1] * wi;
      for (int64_t lp = 0; lp < M; ++lp) {
        const R br = b[2 * lp], bi = b[2 * lp + 1];
        atomic_add_val(p + 2 * (l * M + lp), cr * br - ci * bi);
        atomic_add_val(p + 2 * (l * M + lp) + 1, cr * bi + ci * br);
      }
    }
  }
}

// psf[idx, k, k'] += w * conj(B[k,t]) * B[k',t] (toep_psf.cpp:181-237).
template <class R>
__global__ void k_psf_outer_basis(R *__restrict__ psf, int64_t G, int64_t K,
                                  const int64_t *__restrict__ idx,
                                  const R *__restrict__ w,
                                  const R *__restrict__ basis, int64_t T,
                                  const int64_t *__restrict__ tidx,
                                  int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = idx[i];
    if (g < 0 || g >= G) continue;
    const int64_t t = tidx[i];
    if (t < 0 || t >= T) continue;
    const R wi = w[i];
    R *p = psf + 2 * g * K * K;
    for (int64_t k = 0; k < K; ++k) {
      const R cr = basis[2 * (k * T + t)] * wi;
      const R ci = -basis[2 * (k * T + t) + 1] * wi;
      for (int64_t kp = 0; kp < K; ++kp) {
        const R br = basis[2 * (kp * T + t)], bi = basis[2 * (kp * T + t) + 1];
        atomic_add_val(p + 2 * (k * K + kp), cr * br - ci * bi);
        atomic_add_val(p + 2 * (k * K + kp) + 1, cr * bi + ci * br);
      }
    }
  }
}

template <class T>
static int grog_interpolate_impl(const void *data, const int64_t *idx,
                                 const void *kern, void *out, int64_t N,
                                 int64_t B, int64_t C, int64_t K,
                                 pg_stream_t s) {
  PG_CHECK(N >= 0 && B >= 0 && C >= 0 && K >= 0,
           "grog_interpolate: negative size");
  const int64_t total = N * B * C;
  if (total == 0) return 0;
  PG_CHECK(data && idx && kern && out, "grog_interpolate: null pointer");
  const int threads = 256;
  hipLaunchKernelGGL(k_grog_interp<T>, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(s),
                     reinterpret_cast<const v2_t<T> *>(data), idx,
                     reinterpret_cast<const v2_t<T> *>(kern),
                     reinterpret_cast<v2_t<T> *>(out), N, B, C, K);
  PG_LAUNCHED();
  return 0;
}

template <class R>
static int scatter_impl(void *grid, int64_t G, const void *data,
                        const int64_t *idx, const R *w, int64_t n,
                        pg_stream_t s) {
  PG_CHECK(G >= 0 && n >= 0, "scatter_add: negative size");
  if (n == 0) return 0;
  PG_CHECK(grid && data && idx && w, "scatter_add: null pointer");
  const int threads = 256;
  hipLaunchKernelGGL(k_scatter_add<R>, dim3(grid_for(n, threads, 16384)),
                     dim3(threads), 0, as_stream(s), static_cast<R *>(grid), G,
                     static_cast<const R *>(data), idx, w, n);
  PG_LAUNCHED();
  return 0;
}

template <class R>
static int gather_impl(const void *grid, int64_t G, const int64_t *idx,
                       const R *w, int64_t n, void *out, pg_stream_t s) {
  PG_CHECK(G >= 0 && n >= 0, "gather: negative size");
  if (n == 0) return 0;
  PG_CHECK(grid && idx && w && out, "gather: null pointer");
  const int threads = 256;
  hipLaunchKernelGGL(k_gather<R>, dim3(grid_for(n, threads)), dim3(threads), 0,
                     as_stream(s), static_cast<const R *>(grid), G, idx, w, n,
                     static_cast<R *>(out));
  PG_LAUNCHED();
  return 0;
}

template <class R>
static int psf_scalar_impl(R *psf, int64_t G, const int64_t *idx, const R *w,
                           int64_t n, pg_stream_t s) {
  PG_CHECK(G >= 0 && n >= 0, "psf_scatter_scalar: negative size");
  if (n == 0) return 0;
  PG_CHECK(psf && idx && w, "psf_scatter_scalar: null pointer");
  const int threads = 256;
  hipLaunchKernelGGL(k_psf_scalar<R>, dim3(grid_for(n, threads, 16384)),
                     dim3(threads), 0, as_stream(s), psf, G, idx, w, n);
  PG_LAUNCHED();
  return 0;
}

template <class R>
static int psf_outer_impl(void *psf, int64_t G, int64_t M, const int64_t *idx,
                          const R *w, const void *basis, int64_t n,
                          pg_stream_t s) {
  PG_CHECK(G >= 0 && M >= 0 && n >= 0, "psf_scatter_outer: negative size");
  if (n == 0 || M == 0) return 0;
  PG_CHECK(psf && idx && w && basis, "psf_scatter_outer: null pointer");
  const int threads = 128;
  hipLaunchKernelGGL(k_psf_outer<R>, dim3(grid_for(n, threads)),
                     dim3(threads), 0, as_stream(s), static_cast<R *>(psf), G,
                     M, idx, w, static_cast<const R *>(basis), n);
  PG_LAUNCHED();
  return 0;
}

template <class R>
static int psf_outer_basis_impl(void *psf, int64_t G, int64_t K,
                                const int64_t *idx, const R *w,
                                const void *basis, int64_t T,
                                const int64_t *tidx, int64_t n,
                                pg_stream_t s) {
  PG_CHECK(G >= 0 && K >= 0 && T >= 0 && n >= 0,
           "psf_scatter_outer_basis: negative size");
  if (n == 0 || K == 0) return 0;
  PG_CHECK(psf && idx && w && basis && tidx,
           "psf_scatter_outer_basis: null pointer");
  const int threads = 128;
  hipLaunchKernelGGL(k_psf_outer_basis<R>, dim3(grid_for(n, threads)),
                     dim3(threads), 0, as_stream(s), static_cast<R *>(psf), G,
                     K, idx, w, static_cast<const R *>(basis), T, tidx, n);
  PG_LAUNCHED();
  return 0;
}

static int check_bins(const int64_t *bin_starts, int64_t nb, int64_t bin_size) {
  PG_CHECK(bin_size > 0, "bin_size must be positive");
  PG_CHECK(nb >= 0 && (nb == 0 || bin_starts != nullptr),
           "bin_starts must be a valid 1-D array");
  return 0;
}

}  // namespace pg

using namespace pg;

extern "C" {

const char *pg_last_error(void) { return g_err; }
int pg_abi_version(void) { return 1; }

int pg_grog_interpolate_c64(const void *data, const int64_t *indexes,
                            const void *kernel, void *out, int64_t N,
                            int64_t B, int64_t C, int64_t K,
                            pg_stream_t stream) {
  return grog_interpolate_impl<float>(data, indexes, kernel, out, N, B, C, K,
                                      stream);
}
int pg_grog_interpolate_c128(const void *data, const int64_t *indexes,
                             const void *kernel, void *out, int64_t N,
                             int64_t B, int64_t C, int64_t K,
                             pg_stream_t stream) {
  return grog_interpolate_impl<double>(data, indexes, kernel, out, N, B, C, K,
                                       stream);
}

int pg_scatter_add_c64(void *grid, int64_t grid_size, const void *data,
                       const int64_t *indices, const float *weights, int64_t n,
                       pg_stream_t stream) {
  return scatter_impl<float>(grid, grid_size, data, indices, weights, n,
                             stream);
}
int pg_scatter_add_rows_c64(void *grid, int64_t grid_size, const void *data,
                            const int64_t *indices, const float *weights,
                            int64_t n, int64_t rows, pg_stream_t stream) {
  PG_CHECK(grid_size >= 0 && n >= 0 && rows >= 0 && rows <= 65535,
           "scatter_add_rows: bad sizes");
  if (n == 0 || rows == 0) return 0;
  PG_CHECK(grid && data && indices && weights, "scatter_add_rows: null pointer");
  const int threads = 256;
  hipLaunchKernelGGL(k_scatter_add<float>,
                     dim3(grid_for(n, threads, 16384), (unsigned)rows),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<float *>(grid), grid_size,
                     static_cast<const float *>(data), indices, weights, n);
  PG_LAUNCHED();
  return 0;
}
int pg_scatter_add_c128(void *grid, int64_t grid_size, const void *data,
                        const int64_t *indices, const double *weights,
                        int64_t n, pg_stream_t stream) {
  return scatter_impl<double>(grid, grid_size, data, indices, weights, n,
                              stream);
}

int pg_scatter_add_binned_c64(void *grid, int64_t grid_size, const void *data,
                              const int64_t *indices, const float *weights,
                              int64_t n, const int64_t *bin_starts,
                              int64_t n_bin_starts, int64_t bin_size,
                              pg_stream_t stream) {
  if (int rc = check_bins(bin_starts, n_bin_starts, bin_size)) return rc;
  return scatter_impl<float>(grid, grid_size, data, indices, weights, n,
                             stream);
}
int pg_scatter_add_binned_c128(void *grid, int64_t grid_size,
                               const void *data, const int64_t *indices,
                               const double *weights, int64_t n,
                               const int64_t *bin_starts, int64_t n_bin_starts,
                               int64_t bin_size, pg_stream_t stream) {
  if (int rc = check_bins(bin_starts, n_bin_starts, bin_size)) return rc;
  return scatter_impl<double>(grid, grid_size, data, indices, weights, n,
                              stream);
}

int pg_gather_c64(const void *grid, int64_t grid_size, const int64_t *indices,
                  const float *weights, int64_t n, void *out,
                  pg_stream_t stream) {
  return gather_impl<float>(grid, grid_size, indices, weights, n, out, stream);
}
int pg_gather_c128(const void *grid, int64_t grid_size,
                   const int64_t *indices, const double *weights, int64_t n,
                   void *out, pg_stream_t stream) {
  return gather_impl<double>(grid, grid_size, indices, weights, n, out,
                             stream);
}

int pg_psf_scatter_scalar_f32(float *psf, int64_t grid_size,
                              const int64_t *indices, const float *w_sq,
                              int64_t n, pg_stream_t stream) {
  return psf_scalar_impl<float>(psf, grid_size, indices, w_sq, n, stream);
}
int pg_psf_scatter_scalar_f64(double *psf, int64_t grid_size,
                              const int64_t *indices, const double *w_sq,
                              int64_t n, pg_stream_t stream) {
  return psf_scalar_impl<double>(psf, grid_size, indices, w_sq, n, stream);
}

int pg_psf_scatter_outer_c64(void *psf, int64_t grid_size, int64_t M,
                             const int64_t *indices, const float *w_sq,
                             const void *basis_per_sample, int64_t n,
                             pg_stream_t stream) {
  return psf_outer_impl<float>(psf, grid_size, M, indices, w_sq,
                               basis_per_sample, n, stream);
}
int pg_psf_scatter_outer_c128(void *psf, int64_t grid_size, int64_t M,
                              const int64_t *indices, const double *w_sq,
                              const void *basis_per_sample, int64_t n,
                              pg_stream_t stream) {
  return psf_outer_impl<double>(psf, grid_size, M, indices, w_sq,
                                basis_per_sample, n, stream);
}

int pg_psf_scatter_outer_basis_c64(void *psf, int64_t grid_size, int64_t K,
                                   const int64_t *indices, const float *w_sq,
                                   const void *basis, int64_t T,
                                   const int64_t *time_index, int64_t n,
                                   pg_stream_t stream) {
  return psf_outer_basis_impl<float>(psf, grid_size, K, indices, w_sq, basis,
                                     T, time_index, n, stream);
}
int pg_psf_scatter_outer_basis_c128(void *psf, int64_t grid_size, int64_t K,
                                    const int64_t *indices,
                                    const double *w_sq, const void *basis,
                                    int64_t T, const int64_t *time_index,
                                    int64_t n, pg_stream_t stream) {
  return psf_outer_basis_impl<double>(psf, grid_size, K, indices, w_sq, basis,
                                      T, time_index, n, stream);
}

}  // extern "C"
