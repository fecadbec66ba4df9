// The seven ops of the reference's native module (csrc/torch/module.cpp:54-87)
// as gfx950 kernels behind the C ABI of include/pygrog_hip.h.
//
//   grog_interpolate   one thread per (row, output coil); the input row is
//                      shared by the C threads of a row (one transaction per
//                      coil), the operator row streams per thread.
//   scatter_add(_binned), psf_scatter_scalar, psf_scatter_outer(_basis)
//                      order-fixed: the first element of each run of equal
//                      sorted indices folds the run in input order from the
//                      cell's value, as the reference's CPU loops do, and
//                      commits with a compare-and-swap (see below); one
//                      thread per sample (per sample and PSF element for the
//                      outer products).  Bitwise the reference's CPU result
//                      for sorted indices, run to run.
//   gather             one element per thread.
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
// Order-fixed accumulation for the in-place ops (scatter_add, the PSFs).
//
// The reference's CPU ops take sorted indices and add each sample into its
// cell in input order, starting from the cell's current value
// (sparse_ops_cpu_impl.inl:77-120, toep_psf.cpp:58-118: a clean thread
// partition keeps every cell on one thread).  Here the first element of
// every run of equal consecutive indices folds the whole run the same way,
// from the value it reads, and commits the result with a compare-and-swap.
// With sorted indices a cell is one run, so the swap always succeeds and the
// cell holds exactly the reference's sum, whatever the launch shape.  If the
// swap fails, another run of the same cell wrote it (unsorted input): the
// run's partial sum, folded from zero, is then added atomically -- the sum
// is right, its rounding order is not fixed (as with the reference's CUDA
// atomics).  The file is built with -ffp-contract=off: every product and
// sum rounds separately, as in the reference's loops.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void atomic_add_val(float *p, float v) {
  unsafeAtomicAdd(p, v);
}
__device__ __forceinline__ void atomic_add_val(double *p, double v) {
  unsafeAtomicAdd(p, v);
}

// the cell's current value, read at device scope (not a stale cached copy)
template <class R>
__device__ __forceinline__ R coherent_load(R *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool cas_commit(float *p, float seen, float v) {
  const unsigned int e = __float_as_uint(seen);
  return atomicCAS(reinterpret_cast<unsigned int *>(p), e, __float_as_uint(v)) == e;
}
__device__ __forceinline__ bool cas_commit(double *p, double seen, double v) {
  const unsigned long long e = (unsigned long long)__double_as_longlong(seen);
  return atomicCAS(reinterpret_cast<unsigned long long *>(p), e,
                   (unsigned long long)__double_as_longlong(v)) == e;
}

// Run heads: the first element of a run of equal in-range indices.
__device__ __forceinline__ bool run_head(const int64_t *__restrict__ idx,
                                         int64_t i, int64_t G, int64_t &k) {
  k = idx[i];
  if (k < 0 || k >= G) return false;
  return i == 0 || idx[i - 1] != k;
}
__device__ __forceinline__ int64_t run_end(const int64_t *__restrict__ idx,
                                           int64_t i, int64_t n, int64_t k) {
  int64_t e = i + 1;
  while (e < n && idx[e] == k) ++e;
  return e;
}

// Fold NV components over the run [i, e): acc[v] = acc[v] + term(j, v) in j
// order from the cell's values at p[0..NV), commit each with a swap, and on a
// lost swap add the partial instead.
template <class R, int NV, class Term>
__device__ __forceinline__ void fold_commit(R *p, int64_t i, int64_t e, Term term) {
  R seen[NV], acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = seen[v] = coherent_load(p + v);
  for (int64_t j = i; j < e; ++j) {
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = acc[v] + term(j, v);
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (!cas_commit(p + v, seen[v], acc[v])) {
      R part = R(0);
      for (int64_t j = i; j < e; ++j) part = part + term(j, v);
      atomic_add_val(p + v, part);
    }
  }
}

// grid[idx[i]] += w[i] * data[i]; complex with real weights.  blockIdx.y
// selects a row: grid (rows, G) and data (rows, n) share idx and w.
template <class R>
__global__ void k_scatter_add(R *__restrict__ grid, int64_t G,
                              const R *__restrict__ data,
                              const int64_t *__restrict__ idx,
                              const R *__restrict__ w, int64_t n) {
  grid += (int64_t)blockIdx.y * 2 * G;
  data += (int64_t)blockIdx.y * 2 * n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t k;
    if (!run_head(idx, i, G, k)) continue;
    const int64_t e = run_end(idx, i, n, k);
    fold_commit<R, 2>(grid + 2 * k, i, e,
                      [&](int64_t j, int v) { return w[j] * data[2 * j + v]; });
  }
}

// psf[idx[i]] += w_sq[i] (real).
template <class R>
__global__ void k_psf_scalar(R *__restrict__ psf, int64_t G,
                             const int64_t *__restrict__ idx,
                             const R *__restrict__ w, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t k;
    if (!run_head(idx, i, G, k)) continue;
    const int64_t e = run_end(idx, i, n, k);
    fold_commit<R, 1>(psf + k, i, e, [&](int64_t j, int) { return w[j]; });
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

// psf[idx, l, l'] += (conj(b_l) * w) * b_l' (reference toep_psf.cpp:124-175):
// one thread per (sample, l, l'); the run head's threads fold the run.
template <class R>
__global__ void k_psf_outer(R *__restrict__ psf, int64_t G, int64_t M,
                            const int64_t *__restrict__ idx,
                            const R *__restrict__ w,
                            const R *__restrict__ basis, int64_t n) {
  const int64_t MM = M * M;
  const int64_t total = n * MM;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / MM;
    const int64_t el = t - i * MM;
    int64_t k;
    if (!run_head(idx, i, G, k)) continue;
    const int64_t e = run_end(idx, i, n, k);
    const int64_t l = el / M, lp = el - l * M;
    fold_commit<R, 2>(psf + 2 * (k * MM + el), i, e, [&](int64_t j, int v) {
      const R *b = basis + 2 * j * M;
      const R cr = b[2 * l] * w[j], ci = -b[2 * l + 1] * w[j];
      const R br = b[2 * lp], bi = b[2 * lp + 1];
      return v == 0 ? cr * br - ci * bi : cr * bi + ci * br;
    });
  }
}

// psf[idx, k, k'] += (conj(B[k,t]) * w) * B[k',t] (toep_psf.cpp:181-237);
// samples whose time index is out of range add nothing.
template <class R>
__global__ void k_psf_outer_basis(R *__restrict__ psf, int64_t G, int64_t K,
                                  const int64_t *__restrict__ idx,
                                  const R *__restrict__ w,
                                  const R *__restrict__ basis, int64_t T,
                                  const int64_t *__restrict__ tidx,
                                  int64_t n) {
  const int64_t KK = K * K;
  const int64_t total = n * KK;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / KK;
    const int64_t el = t - i * KK;
    int64_t g;
    if (!run_head(idx, i, G, g)) continue;
    const int64_t e = run_end(idx, i, n, g);
    const int64_t a = el / K, ap = el - a * K;
    R *p = psf + 2 * (g * KK + el);
    // a sample with its time index out of range is skipped, not added as 0
    // (0 + x and x differ for x = -0)
    R seen[2], acc[2];
    acc[0] = seen[0] = coherent_load(p);
    acc[1] = seen[1] = coherent_load(p + 1);
    auto term = [&](int64_t j, R &re, R &im) {
      const int64_t tt = tidx[j];
      if (tt < 0 || tt >= T) return false;
      const R cr = basis[2 * (a * T + tt)] * w[j];
      const R ci = -basis[2 * (a * T + tt) + 1] * w[j];
      const R br = basis[2 * (ap * T + tt)], bi = basis[2 * (ap * T + tt) + 1];
      re = cr * br - ci * bi;
      im = cr * bi + ci * br;
      return true;
    };
    for (int64_t j = i; j < e; ++j) {
      R re, im;
      if (term(j, re, im)) {
        acc[0] = acc[0] + re;
        acc[1] = acc[1] + im;
      }
    }
    for (int v = 0; v < 2; ++v) {
      if (!cas_commit(p + v, seen[v], acc[v])) {
        R part = R(0);
        for (int64_t j = i; j < e; ++j) {
          R re, im;
          if (term(j, re, im)) part = part + (v == 0 ? re : im);
        }
        atomic_add_val(p + v, part);
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
int pg_scatter_add_rows_c128(void *grid, int64_t grid_size, const void *data,
                             const int64_t *indices, const double *weights,
                             int64_t n, int64_t rows, pg_stream_t stream) {
  PG_CHECK(grid_size >= 0 && n >= 0 && rows >= 0 && rows <= 65535,
           "scatter_add_rows: bad sizes");
  if (n == 0 || rows == 0) return 0;
  PG_CHECK(grid && data && indices && weights, "scatter_add_rows: null pointer");
  const int threads = 256;
  hipLaunchKernelGGL(k_scatter_add<double>,
                     dim3(grid_for(n, threads, 16384), (unsigned)rows),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<double *>(grid), grid_size,
                     static_cast<const double *>(data), indices, weights, n);
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
