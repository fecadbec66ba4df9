// SENSE expand / combine around torch.fft with the centring folded in.
//
// For an even length N (M = N/2) the centred transforms of the reference
// (src/pygrog/_base/_fftc.py:114-139: ifftshift -> fftn/ifftn -> fftshift)
// are plain transforms between two sign modulations:
//   fftc(x)[k]  = (-1)^(k-M) * fftn((-1)^n x)[k]
//   ifftc(x)[n] = (-1)^(n-M) * ifftn((-1)^k x)[n]
// (shift theorem, per axis; multiplying by +-1 is exact).  These kernels do
// the image-side half of that in the same pass as the coil expansion + zero
// pad (forward direction) or the crop + conj(S) coil sum (adjoint
// direction); the k-space-side signs are folded into the scatter / gather
// weights by the callers.  Rolls, pad buffers and the separate multiply and
// reduction passes of the unfused path disappear.
//
//   expand:  out[b, c, g] = s(g) * S[c, g - off] * x[b, g - off]   (g inside
//            the image window, else 0); without maps x is (B, C, *image)
//   combine: out[b, p]    = sign * sum_c conj(S[c, p]) * s(p + off) * full[b, c, p + off]
//            (coils in order); without maps out[b, c, p] = sign * s(p + off) * full[b, c, p + off]
// with s(g) = (-1)^(g0 + g1 + g2) when modulate, else 1, and off = (grid -
// image) // 2 per axis (the reference's centred crop / pad).
#include <mutex>

#include "common.h"

namespace pg {

struct Dims {
  int nd;
  int64_t img[3], grd[3], off[3];
  int64_t P, G;  // image and grid sizes
};

__device__ __forceinline__ float parity_sign(int64_t a, int64_t b, int64_t c) {
  return ((a + b + c) & 1) ? -1.0f : 1.0f;
}

__global__ void k_sense_expand(const float2 *__restrict__ x,
                               const float2 *__restrict__ smaps, int64_t B,
                               int64_t C, Dims d, int modulate,
                               float2 *__restrict__ out) {
  const int64_t total = B * C * d.G;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t bc = t / d.G;
    int64_t g = t - bc * d.G;
    const int64_t b = bc / C, c = bc - b * C;
    int64_t gi[3] = {0, 0, 0};
    for (int a = d.nd - 1; a >= 0; --a) {
      gi[a] = g % d.grd[a];
      g /= d.grd[a];
    }
    bool inside = true;
    int64_t p = 0;
    for (int a = 0; a < d.nd; ++a) {
      const int64_t q = gi[a] - d.off[a];
      inside = inside && q >= 0 && q < d.img[a];
      p = p * d.img[a] + q;
    }
    float2 v = make_float2(0.0f, 0.0f);
    if (inside) {
      if (smaps) {
        const float2 s = smaps[c * d.P + p];
        const float2 xv = x[b * d.P + p];
        v = make_float2(s.x * xv.x - s.y * xv.y, s.x * xv.y + s.y * xv.x);
      } else {
        v = x[(b * C + c) * d.P + p];
      }
      if (modulate) {
        const float sg = parity_sign(gi[0], gi[1], gi[2]);
        v.x *= sg;
        v.y *= sg;
      }
    }
    out[t] = v;
  }
}

__global__ void k_sense_combine(const float2 *__restrict__ full,
                                const float2 *__restrict__ smaps, int64_t B,
                                int64_t C, Dims d, int modulate, float sign,
                                float2 *__restrict__ out) {
  const int64_t per_b = smaps ? d.P : C * d.P;
  const int64_t total = B * per_b;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / per_b;
    const int64_t rem = t - b * per_b;
    const int64_t c1 = smaps ? 0 : rem / d.P;
    int64_t p = rem - c1 * d.P;
    const int64_t p_flat = p;
    int64_t gi[3] = {0, 0, 0};
    int64_t g = 0;
    {
      int64_t pi[3] = {0, 0, 0};
      for (int a = d.nd - 1; a >= 0; --a) {
        pi[a] = p % d.img[a];
        p /= d.img[a];
      }
      for (int a = 0; a < d.nd; ++a) {
        gi[a] = pi[a] + d.off[a];
        g = g * d.grd[a] + gi[a];
      }
    }
    const float sg = sign * (modulate ? parity_sign(gi[0], gi[1], gi[2]) : 1.0f);
    float2 acc;
    if (smaps) {
      float ax = 0.0f, ay = 0.0f;
      for (int64_t c = 0; c < C; ++c) {
        const float2 s = smaps[c * d.P + p_flat];
        const float2 f = full[(b * C + c) * d.G + g];
        // conj(s) * f
        ax = ax + (s.x * f.x + s.y * f.y);
        ay = ay + (s.x * f.y - s.y * f.x);
      }
      acc = make_float2(ax * sg, ay * sg);
    } else {
      const float2 f = full[(b * C + c1) * d.G + g];
      acc = make_float2(f.x * sg, f.y * sg);
    }
    out[t] = acc;
  }
}

static int make_dims(int32_t ndim, const int64_t *image_shape,
                     const int64_t *grid_shape, Dims &d) {
  PG_CHECK(ndim >= 1 && ndim <= 3, "sense: ndim must be 1..3, got %d", ndim);
  PG_CHECK(image_shape && grid_shape, "sense: null shapes");
  d.nd = ndim;
  d.P = 1;
  d.G = 1;
  for (int a = 0; a < 3; ++a) {
    d.img[a] = a < ndim ? image_shape[a] : 1;
    d.grd[a] = a < ndim ? grid_shape[a] : 1;
    PG_CHECK(d.img[a] >= 1 && d.grd[a] >= d.img[a],
             "sense: image axis %d (%lld) must be in [1, grid %lld]", a,
             (long long)d.img[a], (long long)d.grd[a]);
    d.off[a] = (d.grd[a] - d.img[a]) / 2;
    d.P *= d.img[a];
    d.G *= d.grd[a];
  }
  return 0;
}

}  // namespace pg

using namespace pg;

extern "C" {

int pg_sense_expand_c64(const void *x, const void *smaps, int64_t B, int64_t C,
                        int32_t ndim, const int64_t *image_shape,
                        const int64_t *grid_shape, int32_t modulate, void *out,
                        pg_stream_t stream) {
  Dims d;
  if (make_dims(ndim, image_shape, grid_shape, d)) return 1;
  PG_CHECK(B >= 0 && C >= 1 && x && out, "sense_expand: bad arguments");
  const int64_t total = B * C * d.G;
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_sense_expand, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<const float2 *>(x),
                     static_cast<const float2 *>(smaps), B, C, d, (int)modulate,
                     static_cast<float2 *>(out));
  PG_LAUNCHED();
  return 0;
}

int pg_sense_combine_c64(const void *full, const void *smaps, int64_t B,
                         int64_t C, int32_t ndim, const int64_t *image_shape,
                         const int64_t *grid_shape, int32_t modulate,
                         float sign, void *out, pg_stream_t stream) {
  Dims d;
  if (make_dims(ndim, image_shape, grid_shape, d)) return 1;
  PG_CHECK(B >= 0 && C >= 1 && full && out, "sense_combine: bad arguments");
  const int64_t total = B * (smaps ? d.P : C * d.P);
  if (total == 0) return 0;
  const int threads = 256;
  hipLaunchKernelGGL(k_sense_combine, dim3(grid_for(total, threads)),
                     dim3(threads), 0, as_stream(stream),
                     static_cast<const float2 *>(full),
                     static_cast<const float2 *>(smaps), B, C, d, (int)modulate,
                     sign, static_cast<float2 *>(out));
  PG_LAUNCHED();
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// FP64-accumulated inner product of complex64 vectors (the CG / LSMR scalars):
// out[0] + i out[1] = sum_i conj(a_i) b_i, summed in double, deterministic
// (fixed per-block ranges, fixed tree; the block partials by a fixed tree).
namespace pg {

constexpr int kDotThreads = 256;
constexpr int kDotBlocks = 1024;

// conj(x) y accumulated in double, each product and sum rounded on its own
// (no contraction), so every kernel that forms these sums over the same
// elements in the same order gets the same bits
__device__ __forceinline__ void dot_acc(double &re, double &im, float2 x, float2 y) {
#pragma clang fp contract(off)
  re = re + ((double)x.x * (double)y.x + (double)x.y * (double)y.y);
  im = im + ((double)x.x * (double)y.y - (double)x.y * (double)y.x);
}

// Block b of nb owns elements [b per, min((b + 1) per, n)), per = ceil(n / nb);
// thread t adds elements t, t + 256, ... of its block's range in order; the
// block's 256 sums then go through a fixed tree.
__host__ __device__ __forceinline__ int dot_blocks(int64_t n) {
  int64_t nb = (n + kDotThreads - 1) / kDotThreads;
  if (nb > kDotBlocks) nb = kDotBlocks;
  return nb < 1 ? 1 : (int)nb;
}

__device__ __forceinline__ void block_tree(double *sr, double *si) {
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kDotThreads) void k_vdot_partial(
    const float2 *__restrict__ a, const float2 *__restrict__ b, int64_t n,
    double2 *__restrict__ part) {
  __shared__ double sr[kDotThreads], si[kDotThreads];
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  double re = 0.0, im = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kDotThreads) dot_acc(re, im, a[i], b[i]);
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  block_tree(sr, si);
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sr[0], si[0]);
}

// The block partials summed by one workgroup: thread t adds partials t,
// t + 256, ... in order, then the same fixed tree as k_vdot_partial (every
// load in flight at once; one thread walking the partials took a memory
// round trip per partial: 34 us for 400 of them).
__global__ __launch_bounds__(kDotThreads) void k_vdot_final(const double2 *__restrict__ part, int nb,
                                                            double *__restrict__ out) {
  __shared__ double sr[kDotThreads], si[kDotThreads];
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < nb; i += kDotThreads) {
    const double2 v = part[i];
    re += v.x;
    im += v.y;
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  block_tree(sr, si);
  if (threadIdx.x == 0) {
    out[0] = sr[0];
    out[1] = si[0];
  }
}

// ---------------------------------------------------------------------------
// One CG iteration's vector and scalar work, fused (cg with FP64 scalars and
// no preconditioner): the same arithmetic as the host loop -- pAp = Re<p, Ap>,
// alpha = rz / (pAp + 1e-30), a = f32(alpha), x' = x + a p, r = r - a Ap,
// rr = Re<r, r>, rel = sqrt(rr) / r0, beta = rr / (rz + 1e-30), rz = rr,
// p = r + f32(beta) p -- with every product and sum rounded on its own as the
// separate torch operations round them, and the sums in the FP64 dot's block
// decomposition and trees, so the iterates are the host loop's bit for bit.
// sc (FP64, device): [0] rz, [1] alpha, [2] beta, [3] r0, [4] rel, [5] rr.
enum CgScalar { kRz = 0, kAlpha, kBeta, kR0, kRel, kRr };

// block partials -> Re of the sum, by the same tree as k_vdot_final
__device__ __forceinline__ double reduce_parts_re(const double2 *__restrict__ part, int nb, double *sr, double *si) {
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < nb; i += kDotThreads) {
    const double2 v = part[i];
    re += v.x;
    im += v.y;
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  block_tree(sr, si);
  return sr[0];
}

__global__ __launch_bounds__(kDotThreads) void k_cg_alpha(const double2 *__restrict__ part, int nb,
                                                          double *__restrict__ sc) {
#pragma clang fp contract(off)
  __shared__ double sr[kDotThreads], si[kDotThreads];
  const double pap = reduce_parts_re(part, nb, sr, si);
  if (threadIdx.x == 0) sc[kAlpha] = sc[kRz] / (pap + 1e-30);
}

// x_out = x_in + a p, r = r - a Ap, and the block partials of <r, r>
__global__ __launch_bounds__(kDotThreads) void k_cg_xr(const float2 *__restrict__ p, const float2 *__restrict__ ap,
                                                       const float2 *__restrict__ x_in, float2 *__restrict__ x_out,
                                                       float2 *__restrict__ r, int64_t n,
                                                       const double *__restrict__ sc, double2 *__restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double sr[kDotThreads], si[kDotThreads];
  const float a = __double2float_rn(sc[kAlpha]);
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  double re = 0.0, im = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kDotThreads) {
    const float2 pv = p[i], av = ap[i], xv = x_in[i];
    float2 rv = r[i];
    // plain operators under this kernel's contract(off): the products and
    // sums are rounded separately, as the separate torch operations round
    // them (an inlined helper's operators would not be covered)
    x_out[i] = make_float2(xv.x + a * pv.x, xv.y + a * pv.y);
    rv = make_float2(rv.x - a * av.x, rv.y - a * av.y);
    r[i] = rv;
    dot_acc(re, im, rv, rv);
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  block_tree(sr, si);
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sr[0], si[0]);
}

__global__ __launch_bounds__(kDotThreads) void k_cg_beta(const double2 *__restrict__ part, int nb,
                                                         double *__restrict__ sc) {
#pragma clang fp contract(off)
  __shared__ double sr[kDotThreads], si[kDotThreads];
  const double rr = reduce_parts_re(part, nb, sr, si);
  if (threadIdx.x == 0) {
    const double rz = sc[kRz];
    sc[kRr] = rr;
    sc[kRel] = sqrt(rr) / sc[kR0];
    sc[kBeta] = rr / (rz + 1e-30);
    sc[kRz] = rr;
  }
}

// p = r + f32(beta) p
__global__ void k_cg_p(const float2 *__restrict__ r, float2 *__restrict__ p, int64_t n,
                       const double *__restrict__ sc) {
#pragma clang fp contract(off)
  const float b = __double2float_rn(sc[kBeta]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 pv = p[i], rv = r[i];
    p[i] = make_float2(rv.x + b * pv.x, rv.y + b * pv.y);
  }
}

}  // namespace pg

namespace pg {

// Block partials of pg_vdot_c64_f64_noalloc: the library's own device memory
// (one copy per GPU, held until the process ends) instead of a caller
// workspace.  Host code never uses the symbol itself as an address: the
// device address of each GPU's copy comes from hipGetSymbolAddress, taken
// once per device with that device current.
static __device__ __align__(256) double2 vdot_scratch[kDotBlocks];

constexpr int kMaxDevices = 64;
static std::mutex g_vdot_scratch_mu;
static void *g_vdot_scratch_ptr[kMaxDevices] = {nullptr};

static int vdot_scratch_address(void **out) {
  int dev = -1;
  PG_HIP(hipGetDevice(&dev));
  PG_CHECK(dev >= 0 && dev < kMaxDevices, "vdot_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_vdot_scratch_mu);
  if (g_vdot_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    PG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(vdot_scratch)));
    PG_CHECK(p != nullptr, "vdot_noalloc: no device address for the scratch array");
    g_vdot_scratch_ptr[dev] = p;
  }
  *out = g_vdot_scratch_ptr[dev];
  return 0;
}

static int vdot_launch(const void *a, const void *b, int64_t n, void *part,
                       double *out, hipStream_t st) {
  const int nb = dot_blocks(n);
  hipLaunchKernelGGL(k_vdot_partial, dim3(nb), dim3(kDotThreads), 0, st,
                     static_cast<const float2 *>(a),
                     static_cast<const float2 *>(b), n,
                     static_cast<double2 *>(part));
  PG_LAUNCHED();
  hipLaunchKernelGGL(k_vdot_final, dim3(1), dim3(kDotThreads), 0, st,
                     static_cast<const double2 *>(part), nb, out);
  PG_LAUNCHED();
  return 0;
}

}  // namespace pg

extern "C" {

int64_t pg_vdot_workspace_bytes(void) { return (int64_t)kDotBlocks * 16; }

int pg_cg_step_xr(const void *p, const void *ap, const void *x_in, void *x_out, void *r, int64_t n,
                  double *sc, void *workspace, pg_stream_t stream) {
  PG_CHECK(n >= 1 && p && ap && x_in && x_out && r && sc && workspace, "cg_step_xr: bad arguments");
  PG_CHECK(x_in != x_out, "cg_step_xr: x_in and x_out must differ");
  const hipStream_t st = as_stream(stream);
  const int nb = dot_blocks(n);
  auto *part = static_cast<double2 *>(workspace);
  hipLaunchKernelGGL(k_vdot_partial, dim3(nb), dim3(kDotThreads), 0, st, static_cast<const float2 *>(p),
                     static_cast<const float2 *>(ap), n, part);
  PG_LAUNCHED();
  hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kDotThreads), 0, st, part, nb, sc);
  PG_LAUNCHED();
  hipLaunchKernelGGL(k_cg_xr, dim3(nb), dim3(kDotThreads), 0, st, static_cast<const float2 *>(p),
                     static_cast<const float2 *>(ap), static_cast<const float2 *>(x_in),
                     static_cast<float2 *>(x_out), static_cast<float2 *>(r), n, sc, part);
  PG_LAUNCHED();
  hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(kDotThreads), 0, st, part, nb, sc);
  PG_LAUNCHED();
  return 0;
}

int pg_cg_step_p(const void *r, void *p, int64_t n, const double *sc, pg_stream_t stream) {
  PG_CHECK(n >= 1 && r && p && sc, "cg_step_p: bad arguments");
  hipLaunchKernelGGL(k_cg_p, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream),
                     static_cast<const float2 *>(r), static_cast<float2 *>(p), n, sc);
  PG_LAUNCHED();
  return 0;
}

int pg_vdot_c64_f64_noalloc(const void *a, const void *b, int64_t n,
                            double *out, pg_stream_t stream) {
  PG_CHECK(n >= 0 && out, "vdot_noalloc: bad arguments");
  PG_CHECK(n <= PG_VDOT_NOALLOC_MAX_N,
           "vdot_noalloc: n = %lld exceeds the maximum %lld of the allocation-free entry",
           (long long)n, (long long)PG_VDOT_NOALLOC_MAX_N);
  PG_CHECK(n == 0 || (a && b), "vdot_noalloc: null vectors");
  void *part = nullptr;
  if (int rc = vdot_scratch_address(&part)) return rc;
  return vdot_launch(a, b, n, part, out, as_stream(stream));
}

int pg_vdot_c64_f64(const void *a, const void *b, int64_t n, void *workspace,
                    double *out, pg_stream_t stream) {
  PG_CHECK(n >= 0 && out && workspace, "vdot: bad arguments");
  PG_CHECK(n == 0 || (a && b), "vdot: null vectors");
  return vdot_launch(a, b, n, workspace, out, as_stream(stream));
}

}  // extern "C"
