/*
 * grog_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain single-threaded C restatement of the reference's CPU kernels, used
 * as the checker of the HIP library (only tests/, the smoke entry and the
 * benchmark's CPU-baseline leg may load it).  Complex values are (re, im)
 * pairs multiplied with the c10::complex formula
 *   (a + ib)(c + id) = (ac - bd) + i(ad + bc).
 *
 *   or_interp_*        grog_interpolate_cpu_impl  csrc/torch/grog_interp.cpp:39-73
 *   or_scatter_add_*   scatter_add_impl            csrc/torch/sparse_ops_cpu_impl.inl:38-120
 *   or_gather_*        gather_impl                 csrc/torch/sparse_ops_cpu_impl.inl:125-168
 *   or_psf_scalar_*    psf_scatter_scalar_cpu      csrc/torch/toep_psf.cpp:80-118
 *   or_psf_outer_*     psf_scatter_outer_cpu       csrc/torch/toep_psf.cpp:124-175
 *   or_psf_outer_basis_*  psf_scatter_outer_basis_cpu  csrc/torch/toep_psf.cpp:181-237
 *
 * The reference splits the loops over OpenMP threads at index boundaries
 * (each cell is still summed in input order by one thread), so the
 * sequential loops below give the same per-cell summation order.
 */
#include <stdint.h>

#define DEFINE_OPS(T, SFX)                                                     \
  /* out[nb, co] = sum_ci K[idx[n], co, ci] * in[nb, ci] (ci ascending) */     \
  void or_interp_##SFX(T *out, const T *in, const int64_t *idx, const T *K,    \
                       int64_t N, int64_t B, int64_t C, int64_t nK) {          \
    for (int64_t nb = 0; nb < N * B; ++nb) {                                   \
      const int64_t k = idx[nb / B];                                           \
      for (int64_t co = 0; co < C; ++co) {                                     \
        T re = 0, im = 0;                                                      \
        if (k >= 0 && k < nK) {                                                \
          const T *row = K + 2 * (k * C * C + co * C);                         \
          const T *src = in + 2 * nb * C;                                      \
          for (int64_t ci = 0; ci < C; ++ci) {                                 \
            const T a = row[2 * ci], b = row[2 * ci + 1];                      \
            const T c = src[2 * ci], d = src[2 * ci + 1];                      \
            re += a * c - b * d;                                               \
            im += a * d + b * c;                                               \
          }                                                                    \
        }                                                                      \
        out[2 * (nb * C + co)] = re;                                           \
        out[2 * (nb * C + co) + 1] = im;                                       \
      }                                                                        \
    }                                                                          \
  }                                                                            \
  /* grid[idx[n]] += w[n] * data[n] (indices outside [0, G) skipped) */        \
  void or_scatter_add_##SFX(T *grid, int64_t G, const T *data,                 \
                            const int64_t *idx, const T *w, int64_t n) {       \
    for (int64_t i = 0; i < n; ++i) {                                          \
      const int64_t k = idx[i];                                                \
      if (k >= 0 && k < G) {                                                   \
        grid[2 * k] += w[i] * data[2 * i];                                     \
        grid[2 * k + 1] += w[i] * data[2 * i + 1];                             \
      }                                                                        \
    }                                                                          \
  }                                                                            \
  /* out[n] = w[n] * grid[idx[n]] (0 outside [0, G)) */                        \
  void or_gather_##SFX(T *out, const T *grid, int64_t G, const int64_t *idx,   \
                       const T *w, int64_t n) {                                \
    for (int64_t i = 0; i < n; ++i) {                                          \
      const int64_t k = idx[i];                                                \
      T re = 0, im = 0;                                                        \
      if (k >= 0 && k < G) {                                                   \
        re = w[i] * grid[2 * k];                                               \
        im = w[i] * grid[2 * k + 1];                                           \
      }                                                                        \
      out[2 * i] = re;                                                         \
      out[2 * i + 1] = im;                                                     \
    }                                                                          \
  }                                                                            \
  /* psf[idx[n]] += w_sq[n] (real) */                                          \
  void or_psf_scalar_##SFX(T *psf, int64_t G, const int64_t *idx,              \
                           const T *w, int64_t n) {                            \
    for (int64_t i = 0; i < n; ++i) {                                          \
      const int64_t k = idx[i];                                                \
      if (k >= 0 && k < G) psf[k] += w[i];                                     \
    }                                                                          \
  }                                                                            \
  /* psf[idx, l, l'] += (conj(b_l) * w) * b_l' */                              \
  void or_psf_outer_##SFX(T *psf, int64_t G, int64_t M, const int64_t *idx,    \
                          const T *w, const T *b, int64_t n) {                 \
    for (int64_t i = 0; i < n; ++i) {                                          \
      const int64_t k = idx[i];                                                \
      if (k < 0 || k >= G) continue;                                           \
      const T *bn = b + 2 * i * M;                                             \
      T *pn = psf + 2 * k * M * M;                                             \
      for (int64_t l = 0; l < M; ++l) {                                        \
        const T cr = bn[2 * l] * w[i], ci = -bn[2 * l + 1] * w[i];             \
        for (int64_t lp = 0; lp < M; ++lp) {                                   \
          const T br = bn[2 * lp], bi = bn[2 * lp + 1];                        \
          pn[2 * (l * M + lp)] += cr * br - ci * bi;                           \
          pn[2 * (l * M + lp) + 1] += cr * bi + ci * br;                       \
        }                                                                      \
      }                                                                        \
    }                                                                          \
  }                                                                            \
  /* psf[idx, k, k'] += (conj(B[k,t]) * w) * B[k',t]; basis (K, T) */          \
  void or_psf_outer_basis_##SFX(T *psf, int64_t G, int64_t K,                  \
                                const int64_t *idx, const T *w, const T *bs,   \
                                int64_t Tn, const int64_t *tidx, int64_t n) {  \
    for (int64_t i = 0; i < n; ++i) {                                          \
      const int64_t g = idx[i];                                                \
      if (g < 0 || g >= G) continue;                                           \
      const int64_t t = tidx[i];                                               \
      if (t < 0 || t >= Tn) continue;                                          \
      T *pn = psf + 2 * g * K * K;                                             \
      for (int64_t k = 0; k < K; ++k) {                                        \
        const T cr = bs[2 * (k * Tn + t)] * w[i];                              \
        const T ci = -bs[2 * (k * Tn + t) + 1] * w[i];                         \
        for (int64_t kp = 0; kp < K; ++kp) {                                   \
          const T br = bs[2 * (kp * Tn + t)], bi = bs[2 * (kp * Tn + t) + 1];  \
          pn[2 * (k * K + kp)] += cr * br - ci * bi;                           \
          pn[2 * (k * K + kp) + 1] += cr * bi + ci * br;                       \
        }                                                                      \
      }                                                                        \
    }                                                                          \
  }

DEFINE_OPS(float, f32)
DEFINE_OPS(double, f64)
