// Generic wavefront kernels: PDE solves from increments already in HBM (any static kernel),
// pairwise / Gram / symmetric Gram, scaled and unscaled, float32-reference and float64 storage,
// plus the fused robust-normalisation root-find. C ABI in include/sigpde_hip.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "capi_util.h"
#include "wave_solver.h"

namespace sigpde {

// Rows per lane H and band count for a refined grid of M rows: H = ceil(M / (64 * nb)) with at
// most kMaxH rows per lane, adjusted so that a lane's rows never straddle an original row
// unevenly (H a multiple of R = 2^order, or a divisor of R).
constexpr int kMaxH = 16;

static int choose_rows_per_lane(int M, int order) {
  const int R = 1 << order;
  int nb = (M + kWave * kMaxH - 1) / (kWave * kMaxH);
  if (nb < 1) nb = 1;
  for (;; ++nb) {
    int H = (M + kWave * nb - 1) / (kWave * nb);
    if (H < 1) H = 1;
    if (H >= R) {
      H = (H + R - 1) / R * R;
    } else {
      int p = 1;
      while (p < H) p <<= 1;
      H = p;  // power of two below R divides R
    }
    if (H <= kMaxH) return H;
  }
}

// ------------------------------------------------------------------ pair descriptors
template <typename T>
struct PairwiseArgs {
  const T *inc;
  int64_t B, stride_b;
  int m, n, order;
  const T *sx, *sy;
  T *out;
};

template <typename T>
struct GramArgs {
  const T *inc;
  int64_t bx, by, stride_x, stride_y;
  int m, n, order;
  const T *sx, *sy;
  int symmetric;
  int64_t off_x, off_y, ldo;
  T *out;
};

constexpr int kWavesPerBlock = 4;

template <typename T, int H>
__global__ __launch_bounds__(kWave *kWavesPerBlock) void k_pde_pairwise(PairwiseArgs<T> a,
                                                                        int bndStride) {
  extern __shared__ unsigned char smem_raw[];
  const int wid = threadIdx.x / kWave;
  const int64_t b = (int64_t)blockIdx.x * kWavesPerBlock + wid;
  if (b >= a.B) return;
  T *bnd = reinterpret_cast<T *>(smem_raw) + (int64_t)wid * bndStride;
  T s = (T)1;
  if (a.sx || a.sy) s = (a.sy ? a.sy[b] : (T)1) * (a.sx ? a.sx[b] : (T)1);
  T r = wave_solve<T, H, false>(a.inc + b * a.stride_b, a.n, a.m, a.n, a.order, s, bnd);
  if ((threadIdx.x & (kWave - 1)) == 0) a.out[b] = r;
}

template <typename T, int H>
__global__ __launch_bounds__(kWave *kWavesPerBlock) void k_pde_gram(GramArgs<T> a, int bndStride) {
  extern __shared__ unsigned char smem_raw[];
  const int wid = threadIdx.x / kWave;
  const int64_t p = (int64_t)blockIdx.x * kWavesPerBlock + wid;
  if (p >= a.bx * a.by) return;
  const int64_t i = p / a.by, j = p - (p / a.by) * a.by;
  const int64_t gi = a.off_x + i, gj = a.off_y + j;
  if (a.symmetric && gj > gi) return;
  T *bnd = reinterpret_cast<T *>(smem_raw) + (int64_t)wid * bndStride;
  T s = (T)1;
  if (a.sx || a.sy) s = (a.sx ? a.sx[i] : (T)1) * (a.sy ? a.sy[j] : (T)1);
  T r = wave_solve<T, H, false>(a.inc + i * a.stride_x + j * a.stride_y, a.n, a.m, a.n, a.order,
                                s, bnd);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    a.out[gi * a.ldo + gj] = r;
    if (a.symmetric && gi != gj) a.out[gj * a.ldo + gi] = r;
  }
}

// ------------------------------------------------------------------ robust normalisation
// One wave per path. Chandrupatla's bracketing iteration as in
// sigpde/cuda_pairwise_kernels.py:88-166: every lane carries the same double-precision bracket
// state; the wavefront result is broadcast with readlane before it is used, so all lanes take the
// same branches and no shared scratch is written between iterations.
template <typename T>
struct RobustArgs {
  const T *inc;
  int64_t B, stride_b;
  int m, order, normalizer, maxit;
  double tol;
  const T *target;
  T *knorm;
  T *lambda;
  int32_t *iters;
};

template <typename T>
__device__ __forceinline__ T log_normalizer(T x);
template <>
__device__ __forceinline__ float log_normalizer(float x) {
  return 2.0f - 1.0f / (1.0f + logf(x));
}
template <>
__device__ __forceinline__ double log_normalizer(double x) {
  return 2.0 - 1.0 / (1.0 + log(x));
}

template <typename T, int H>
__global__ __launch_bounds__(kWave *kWavesPerBlock) void k_robust_norm(RobustArgs<T> a,
                                                                       int bndStride) {
  extern __shared__ unsigned char smem_raw[];
  const int wid = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t b = (int64_t)blockIdx.x * kWavesPerBlock + wid;
  if (b >= a.B) return;
  T *bnd = reinterpret_cast<T *>(smem_raw) + (int64_t)wid * bndStride;
  const T *inc = a.inc + b * a.stride_b;

  T K1, Tt;
  if (a.normalizer == 0) {
    // _norm_factors (sigpde/torch/kernel.py:181-184): K(x,x), NaN -> inf, target
    K1 = wave_solve<T, H, false>(inc, a.m, a.m, a.m, a.order, (T)1, bnd);
    if (isnan(K1)) K1 = (T)INFINITY;
    Tt = log_normalizer<T>(K1);
  } else {
    K1 = a.knorm[b];
    Tt = a.target[b];
  }

  double av = 0.0, cv = 0.0, bv = 1.0;
  double fa = 1.0 - (double)Tt;
  double fb = (double)(T)(K1 - Tt);
  double fc = fa, t = 0.5;
  double outv = 0.0;
  int it = 0;
  while (it < a.maxit) {
    double xt = av + t * (bv - av);
    ++it;
    T sol = wave_solve<T, H, true>(inc, a.m, a.m, a.m, a.order, xt * xt, bnd);
    double ft = isnan(sol) ? (double)INFINITY : (double)(T)(sol - Tt);
    if (ft * fa >= 0) {
      cv = av;
      fc = fa;
    } else {
      cv = bv;
      bv = av;
      fc = fb;
      fb = fa;
    }
    av = xt;
    fa = ft;
    if (fabs(fa) < fabs(fb)) {
      outv = av;
      if (fabs(fa) < a.tol) break;
    } else {
      outv = bv;
      if (fabs(fb) < a.tol) break;
    }
    double xi = (av - bv) / (cv - bv);
    double phi = (fa - fb) / (fc - fb);
    if (phi * phi < xi && (1 - phi) * (1 - phi) < 1 - xi)
      t = fa / (fb - fa) * fc / (fb - fc) + (cv - av) / (bv - av) * fa / (fc - fa) * fb / (fc - fb);
    else
      t = 0.5;
    if (!(0 < t && t < 1)) t = 0.5;
  }
  if (lane == 0) {
    // scales[norms < 1] = 0 (sigpde/torch/kernel.py:195)
    a.lambda[b] = (K1 < (T)1) ? (T)0 : (T)outv;
    if (a.normalizer == 0) a.knorm[b] = K1;
    if (a.iters) a.iters[b] = it;
  }
}

// ------------------------------------------------------------------ dispatch
template <template <typename, int> class K, typename T, typename Args>
struct Dispatch;

#define SIGPDE_H_CASES(X) \
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

template <typename T>
static hipError_t launch_pairwise(const PairwiseArgs<T> &a, hipStream_t st) {
  const int M = a.m << a.order, N = a.n << a.order;
  const int H = choose_rows_per_lane(M, a.order);
  const bool multiBand = M > kWave * H;
  const int bndStride = multiBand ? N : 0;
  const size_t shmem = (size_t)kWavesPerBlock * bndStride * sizeof(T);
  const dim3 grid((unsigned)((a.B + kWavesPerBlock - 1) / kWavesPerBlock));
  const dim3 block(kWave * kWavesPerBlock);
  switch (H) {
#define X(h)                                                                   \
  case h:                                                                      \
    hipLaunchKernelGGL((k_pde_pairwise<T, h>), grid, block, shmem, st, a, bndStride); \
    break;
    SIGPDE_H_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_gram(const GramArgs<T> &a, hipStream_t st) {
  const int M = a.m << a.order, N = a.n << a.order;
  const int H = choose_rows_per_lane(M, a.order);
  const bool multiBand = M > kWave * H;
  const int bndStride = multiBand ? N : 0;
  const size_t shmem = (size_t)kWavesPerBlock * bndStride * sizeof(T);
  const int64_t np = a.bx * a.by;
  const dim3 grid((unsigned)((np + kWavesPerBlock - 1) / kWavesPerBlock));
  const dim3 block(kWave * kWavesPerBlock);
  switch (H) {
#define X(h)                                                                 \
  case h:                                                                    \
    hipLaunchKernelGGL((k_pde_gram<T, h>), grid, block, shmem, st, a, bndStride); \
    break;
    SIGPDE_H_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_robust(const RobustArgs<T> &a, hipStream_t st) {
  const int M = a.m << a.order, N = M;
  const int H = choose_rows_per_lane(M, a.order);
  const bool multiBand = M > kWave * H;
  const int bndStride = multiBand ? N : 0;
  const size_t shmem = (size_t)kWavesPerBlock * bndStride * sizeof(T);
  const dim3 grid((unsigned)((a.B + kWavesPerBlock - 1) / kWavesPerBlock));
  const dim3 block(kWave * kWavesPerBlock);
  switch (H) {
#define X(h)                                                                    \
  case h:                                                                       \
    hipLaunchKernelGGL((k_robust_norm<T, h>), grid, block, shmem, st, a, bndStride); \
    break;
    SIGPDE_H_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static int check_grid(int64_t m, int64_t n, int order) {
  if (m < 1 || n < 1) return set_error(SIGPDE_EINVAL, "paths need length >= 2 (m, n >= 1 increments)");
  if (order < 0 || order > 12) return set_error(SIGPDE_EINVAL, "dyadic order must be in [0, 12]");
  if ((m << order) > kMaxRefinedRows || (n << order) > kMaxRefinedRows)
    return set_error(SIGPDE_EINVAL, "refined grid side (length-1) << order exceeds %d", kMaxRefinedRows);
  return SIGPDE_OK;
}

}  // namespace sigpde

using namespace sigpde;

extern "C" int sigpde_pde_pairwise(int dtype, const void *inc, int64_t B, int64_t m, int64_t n,
                                   int64_t inc_stride_b, int order, const void *sx, const void *sy,
                                   void *out, void *stream) {
  clear_error();
  if (B == 0) return SIGPDE_OK;
  if (B < 0 || !inc || !out) return set_error(SIGPDE_EINVAL, "pde_pairwise: bad batch or null pointer");
  if (int rc = check_grid(m, n, order)) return rc;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (dtype == SIGPDE_F32) {
    PairwiseArgs<float> a{(const float *)inc, B, inc_stride_b, (int)m, (int)n, order,
                          (const float *)sx, (const float *)sy, (float *)out};
    e = launch_pairwise(a, st);
  } else if (dtype == SIGPDE_F64) {
    PairwiseArgs<double> a{(const double *)inc, B, inc_stride_b, (int)m, (int)n, order,
                           (const double *)sx, (const double *)sy, (double *)out};
    e = launch_pairwise(a, st);
  } else {
    return set_error(SIGPDE_EINVAL, "pde_pairwise: unknown dtype %d", dtype);
  }
  return hip_status(e, "pde_pairwise launch");
}

extern "C" int sigpde_pde_gram(int dtype, const void *inc, int64_t bx, int64_t by, int64_t m,
                               int64_t n, int64_t stride_x, int64_t stride_y, int order,
                               const void *sx, const void *sy, int symmetric, int64_t off_x,
                               int64_t off_y, void *out, int64_t ldo, void *stream) {
  clear_error();
  if (bx == 0 || by == 0) return SIGPDE_OK;
  if (bx < 0 || by < 0 || !inc || !out) return set_error(SIGPDE_EINVAL, "pde_gram: bad tile or null pointer");
  if (int rc = check_grid(m, n, order)) return rc;
  if (symmetric && m != n) return set_error(SIGPDE_EINVAL, "pde_gram: symmetric Gram needs m == n");
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (dtype == SIGPDE_F32) {
    GramArgs<float> a{(const float *)inc, bx, by, stride_x, stride_y, (int)m, (int)n, order,
                      (const float *)sx, (const float *)sy, symmetric, off_x, off_y, ldo, (float *)out};
    e = launch_gram(a, st);
  } else if (dtype == SIGPDE_F64) {
    GramArgs<double> a{(const double *)inc, bx, by, stride_x, stride_y, (int)m, (int)n, order,
                       (const double *)sx, (const double *)sy, symmetric, off_x, off_y, ldo,
                       (double *)out};
    e = launch_gram(a, st);
  } else {
    return set_error(SIGPDE_EINVAL, "pde_gram: unknown dtype %d", dtype);
  }
  return hip_status(e, "pde_gram launch");
}

extern "C" int sigpde_robust_norm(int dtype, const void *inc, int64_t B, int64_t m,
                                  int64_t inc_stride_b, int order, int normalizer,
                                  const void *target, void *knorm, int maxit, double tol,
                                  void *lambda_out, int32_t *iters_out, void *stream) {
  clear_error();
  if (B == 0) return SIGPDE_OK;
  if (B < 0 || !inc || !lambda_out || !knorm)
    return set_error(SIGPDE_EINVAL, "robust_norm: bad batch or null pointer");
  if (normalizer != SIGPDE_NORM_LOG && normalizer != SIGPDE_NORM_GIVEN)
    return set_error(SIGPDE_EINVAL, "robust_norm: unknown normalizer id %d", normalizer);
  if (normalizer == SIGPDE_NORM_GIVEN && !target)
    return set_error(SIGPDE_EINVAL, "robust_norm: SIGPDE_NORM_GIVEN needs target");
  if (maxit < 0) return set_error(SIGPDE_EINVAL, "robust_norm: maxit must be >= 0");
  if (int rc = check_grid(m, m, order)) return rc;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (dtype == SIGPDE_F32) {
    RobustArgs<float> a{(const float *)inc, B, inc_stride_b, (int)m, order, normalizer, maxit, tol,
                        (const float *)target, (float *)knorm, (float *)lambda_out, iters_out};
    e = launch_robust(a, st);
  } else if (dtype == SIGPDE_F64) {
    RobustArgs<double> a{(const double *)inc, B, inc_stride_b, (int)m, order, normalizer, maxit, tol,
                         (const double *)target, (double *)knorm, (double *)lambda_out, iters_out};
    e = launch_robust(a, st);
  } else {
    return set_error(SIGPDE_EINVAL, "robust_norm: unknown dtype %d", dtype);
  }
  return hip_status(e, "robust_norm launch");
}
