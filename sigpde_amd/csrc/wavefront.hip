// Generic wavefront kernels: PDE solves from increments already in HBM (any static kernel),
// pairwise / Gram / symmetric Gram, scaled and unscaled, float32-reference and float64 storage,
// plus the fused robust-normalisation root-find. C ABI in include/sigpde_hip.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "capi_util.h"
#include <type_traits>

#include "ring16.h"
#include "ring_solver.h"

namespace sigpde {

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

template <typename T>
struct RefArith;
template <>
struct RefArith<float> {
  static constexpr Arith kPlain = Arith::Ref32, kScaled = Arith::Scaled32;
};
template <>
struct RefArith<double> {
  static constexpr Arith kPlain = Arith::Ref64, kScaled = Arith::Scaled64;
};

// Groups of 16 lanes run the skewed-ring solver (ring16.h), groups of 64 the generic one.
template <typename T, int G>
using ProducerFor = typename std::conditional<G == 16, Global16Producer<T>, GlobalProducer<T, G>>::type;

template <Arith A, int G, int R, class Producer, class Out>
__device__ __forceinline__ void solve_wave(Producer &P, const PairGeom<G> &g, double s, void *lds,
                                           Out &&out) {
  if constexpr (G == 16)
    ring16_solve<A, R>(P, g, s, lds, out);
  else
    ring_solve<A, G>(P, g, s, lds, out);
}

// One wave per block; a wave solves Layout<G>::kPairs pairs (one per lane group of G lanes).
// Increments stream from HBM through the wave's LDS ring.
// R: refined rows per lane of the 16-lane solver (1 for 64-lane groups).
template <typename T, int G, int R>
__global__ __launch_bounds__(kWave) void k_pde_pairwise(PairwiseArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using L = Layout<G>;
  const int lane = threadIdx.x & (kWave - 1);
  PairGeom<G> g(a.m, a.n, a.order, 1, R);
  ProducerFor<T, G> P;
  P.ld = a.n;
  P.scaled = a.sx || a.sy;
#pragma unroll
  for (int q = 0; q < L::kPairs; ++q) {
    const int64_t b = (int64_t)blockIdx.x * L::kPairs + q;
    const bool ok = b < a.B;
    P.inc[q] = ok ? a.inc + b * a.stride_b : nullptr;
    P.scale[q] = (P.scaled && ok) ? (a.sy ? a.sy[b] : (T)1) * (a.sx ? a.sx[b] : (T)1) : (T)1;
  }
  const int64_t b = (int64_t)blockIdx.x * L::kPairs + lane / G;
  solve_wave<RefArith<T>::kPlain, G, R>(P, g, 1.0, smem_raw, [&](int, double r) {
    if (b < a.B) a.out[b] = (T)r;
  });
}

// Pair p of a Gram tile -> (i, j). Symmetric diagonal tiles (off_x == off_y, bx == by) list only
// the lower triangle j <= i; other tiles list all bx*by pairs (symmetric ones skip j above i).
__device__ __forceinline__ void tile_pair(int64_t p, int64_t by, bool tri, int64_t &i, int64_t &j) {
  if (tri) {
    int64_t r = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > p) --r;
    while ((r + 1) * (r + 2) / 2 <= p) ++r;
    i = r;
    j = p - r * (r + 1) / 2;
  } else {
    i = p / by;
    j = p - i * by;
  }
}

template <typename T, int G, int R>
__global__ __launch_bounds__(kWave) void k_pde_gram(GramArgs<T> a, int64_t npairs, int tri) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using L = Layout<G>;
  const int lane = threadIdx.x & (kWave - 1);
  PairGeom<G> g(a.m, a.n, a.order, 1, R);
  ProducerFor<T, G> P;
  P.ld = a.n;
  P.scaled = a.sx || a.sy;
  int64_t gi[L::kPairs], gj[L::kPairs];
#pragma unroll
  for (int q = 0; q < L::kPairs; ++q) {
    const int64_t p = (int64_t)blockIdx.x * L::kPairs + q;
    int64_t i = 0, j = 0;
    bool ok = p < npairs;
    if (ok) tile_pair(p, a.by, tri, i, j);
    gi[q] = a.off_x + i;
    gj[q] = a.off_y + j;
    if (a.symmetric && gj[q] > gi[q]) ok = false;
    P.inc[q] = ok ? a.inc + i * a.stride_x + j * a.stride_y : nullptr;
    P.scale[q] = (P.scaled && ok) ? (a.sx ? a.sx[i] : (T)1) * (a.sy ? a.sy[j] : (T)1) : (T)1;
    if (!ok) gi[q] = -1;
  }
  const int q = lane / G;
  int64_t xi = gi[0], yj = gj[0];
#pragma unroll
  for (int w = 1; w < L::kPairs; ++w)
    if (q == w) {
      xi = gi[w];
      yj = gj[w];
    }
  solve_wave<RefArith<T>::kPlain, G, R>(P, g, 1.0, smem_raw, [&](int, double r) {
    if (xi >= 0) {
      a.out[xi * a.ldo + yj] = (T)r;
      if (a.symmetric && xi != yj) a.out[yj * a.ldo + xi] = (T)r;
    }
  });
}

// ------------------------------------------------------------------ robust normalisation
// Chandrupatla's bracketing iteration as in sigpde/cuda_pairwise_kernels.py:88-166, one path per
// lane group. Every lane of a group carries the same double-precision bracket state; the wavefront
// result is broadcast within the group before it is used, so a group's lanes take the same
// branches, and no shared scratch is written between iterations. The wave iterates until all its
// groups have stopped; a stopped group keeps its answer.
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

// Broadcasts the value held by lane (group base + src) to every lane of its group of G lanes.
template <int G>
__device__ __forceinline__ double group_bcast(double v, int src) {
  const int lane = threadIdx.x & (kWave - 1);
  const int from = ((lane & ~(G - 1)) + src) << 2;
  int2 a = *reinterpret_cast<int2 *>(&v);
  int2 r;
  r.x = __builtin_amdgcn_ds_bpermute(from, a.x);
  r.y = __builtin_amdgcn_ds_bpermute(from, a.y);
  return *reinterpret_cast<double *>(&r);
}

template <typename T, int G, int R>
__global__ __launch_bounds__(kWave) void k_robust_norm(RobustArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using L = Layout<G>;
  const int lane = threadIdx.x & (kWave - 1);
  const int q = lane / G;
  const int64_t b = (int64_t)blockIdx.x * L::kPairs + q;
  const bool live = b < a.B;
  PairGeom<G> g(a.m, a.m, a.order, 1, R);
  const int rM = G == 16 ? ((g.M - 1) % (16 * R)) / R : (g.M - 1) & (G - 1);
  ProducerFor<T, G> P;
  P.ld = a.m;
  P.scaled = false;
#pragma unroll
  for (int w = 0; w < L::kPairs; ++w) {
    const int64_t bw = (int64_t)blockIdx.x * L::kPairs + w;
    P.inc[w] = bw < a.B ? a.inc + bw * a.stride_b : nullptr;
    P.scale[w] = (T)1;
  }

  T K1 = (T)1, Tt = (T)1;
  if (a.normalizer == 0) {
    // _norm_factors (sigpde/torch/kernel.py:181-184): K(x,x), NaN -> inf, target
    double r = 0.0;
    solve_wave<RefArith<T>::kPlain, G, R>(P, g, 1.0, smem_raw, [&](int, double v) { r = v; });
    K1 = (T)group_bcast<G>(r, rM);
    if (isnan(K1)) K1 = (T)INFINITY;
    Tt = log_normalizer<T>(K1);
  } else if (live) {
    K1 = a.knorm[b];
    Tt = a.target[b];
  }

  double av = 0.0, cv = 0.0, bv = 1.0;
  double fa = 1.0 - (double)Tt;
  double fb = (double)(T)(K1 - Tt);
  double fc = fa, t = 0.5;
  double outv = 0.0;
  int it = 0;
  bool active = live && a.maxit > 0;
  while (__any(active)) {
    double xt = av + t * (bv - av);
    double r = 0.0;
    solve_wave<RefArith<T>::kScaled, G, R>(P, g, xt * xt, smem_raw, [&](int, double v) { r = v; });
    double sol = group_bcast<G>(r, rM);
    if (!active) continue;
    ++it;
    T solT = (T)sol;
    double ft = isnan(solT) ? (double)INFINITY : (double)(T)(solT - Tt);
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
    bool stop;
    if (fabs(fa) < fabs(fb)) {
      outv = av;
      stop = fabs(fa) < a.tol;
    } else {
      outv = bv;
      stop = fabs(fb) < a.tol;
    }
    double xi = (av - bv) / (cv - bv);
    double phi = (fa - fb) / (fc - fb);
    if (phi * phi < xi && (1 - phi) * (1 - phi) < 1 - xi)
      t = fa / (fb - fa) * fc / (fb - fc) + (cv - av) / (bv - av) * fa / (fc - fa) * fb / (fc - fb);
    else
      t = 0.5;
    if (!(0 < t && t < 1)) t = 0.5;
    if (stop || it >= a.maxit) active = false;
  }
  if (live && (lane & (G - 1)) == 0) {
    // scales[norms < 1] = 0 (sigpde/torch/kernel.py:195)
    a.lambda[b] = (K1 < (T)1) ? (T)0 : (T)outv;
    if (a.normalizer == 0) a.knorm[b] = K1;
    if (a.iters) a.iters[b] = it;
  }
}

// ------------------------------------------------------------------ launchers
// Lane groups of 16 (four pairs per wave) need half the LDS per pair and fill the CUs better when
// there are many pairs; groups of 64 finish a single pair in fewer sequential steps.
constexpr int64_t kManyPairs = 2048;

template <typename T, int G>
static size_t lds_for(int N) {
  if (G == 16) return r16_lds_bytes<RefArith<T>::kPlain>(N);
  return wave_lds_bytes<RefArith<T>::kPlain, G>(N);
}

// Lane groups of 16 keep four boundary lines per wave, groups of 64 one: past the refined length
// whose four lines fill the LDS, every batch runs on groups of 64.
template <typename T>
static bool g16_fits(int N) { return lds_for<T, 16>(N) <= kMaxLdsBytes; }

// Longest refined column count (n << order) that groups of 64 hold in LDS.
template <typename T>
static int64_t max_columns() {
  int64_t lo = 1, hi = 1 << 20;  // lds_for<T, 64>(lo) fits, lds_for<T, 64>(hi) does not
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    (lds_for<T, 64>((int)mid) <= kMaxLdsBytes ? lo : hi) = mid;
  }
  return lo;
}

template <typename T>
static hipError_t launch_pairwise(const PairwiseArgs<T> &a, hipStream_t st) {
  const int N = a.n << a.order;
  if (a.B >= kManyPairs && g16_fits<T>(N)) {
    const dim3 grid((unsigned)((a.B + 3) / 4));
    switch (r16_rows(a.order)) {
      case 1: k_pde_pairwise<T, 16, 1><<<grid, dim3(kWave), lds_for<T, 16>(N), st>>>(a); break;
      case 2: k_pde_pairwise<T, 16, 2><<<grid, dim3(kWave), lds_for<T, 16>(N), st>>>(a); break;
      case 4: k_pde_pairwise<T, 16, 4><<<grid, dim3(kWave), lds_for<T, 16>(N), st>>>(a); break;
      default: k_pde_pairwise<T, 16, 8><<<grid, dim3(kWave), lds_for<T, 16>(N), st>>>(a); break;
    }
  } else {
    k_pde_pairwise<T, 64, 1><<<dim3((unsigned)a.B), dim3(kWave), lds_for<T, 64>(N), st>>>(a);
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_gram(const GramArgs<T> &a, hipStream_t st) {
  const int N = a.n << a.order;
  const bool tri = a.symmetric && a.off_x == a.off_y && a.bx == a.by;
  const int64_t np = tri ? a.bx * (a.bx + 1) / 2 : a.bx * a.by;
  if (np >= kManyPairs && g16_fits<T>(N)) {
    const dim3 grid((unsigned)((np + 3) / 4));
    const size_t lds = lds_for<T, 16>(N);
    switch (r16_rows(a.order)) {
      case 1: k_pde_gram<T, 16, 1><<<grid, dim3(kWave), lds, st>>>(a, np, (int)tri); break;
      case 2: k_pde_gram<T, 16, 2><<<grid, dim3(kWave), lds, st>>>(a, np, (int)tri); break;
      case 4: k_pde_gram<T, 16, 4><<<grid, dim3(kWave), lds, st>>>(a, np, (int)tri); break;
      default: k_pde_gram<T, 16, 8><<<grid, dim3(kWave), lds, st>>>(a, np, (int)tri); break;
    }
  } else {
    k_pde_gram<T, 64, 1><<<dim3((unsigned)np), dim3(kWave), lds_for<T, 64>(N), st>>>(a, np, (int)tri);
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_robust(const RobustArgs<T> &a, hipStream_t st) {
  const int N = a.m << a.order;
  if (a.B >= kManyPairs && g16_fits<T>(N)) {
    const dim3 grid((unsigned)((a.B + 3) / 4));
    switch (r16_rows(a.order)) {
      case 1: k_robust_norm<T, 16, 1><<<grid, dim3(kWave), lds_for<T, 16>(N), st>>>(a); break;
      case 2: k_robust_norm<T, 16, 2><<<grid, dim3(kWave), lds_for<T, 16>(N), st>>>(a); break;
      case 4: k_robust_norm<T, 16, 4><<<grid, dim3(kWave), lds_for<T, 16>(N), st>>>(a); break;
      default: k_robust_norm<T, 16, 8><<<grid, dim3(kWave), lds_for<T, 16>(N), st>>>(a); break;
    }
  } else {
    k_robust_norm<T, 64, 1><<<dim3((unsigned)a.B), dim3(kWave), lds_for<T, 64>(N), st>>>(a);
  }
  return hipGetLastError();
}

// m: increments along the rows (bands), n: along the columns (boundary lines).
static int check_grid(int dtype, int64_t m, int64_t n, int order) {
  if (m < 1 || n < 1) return set_error(SIGPDE_EINVAL, "paths need length >= 2 (m, n >= 1 increments)");
  if (order < 0 || order > 12) return set_error(SIGPDE_EINVAL, "dyadic order must be in [0, 12]");
  if ((m << order) > kMaxRefinedRows)
    return set_error(SIGPDE_EINVAL, "refined rows (length-1) << order = %lld exceed %lld",
                     (long long)(m << order), (long long)kMaxRefinedRows);
  const int64_t maxc = dtype == SIGPDE_F32 ? max_columns<float>() : max_columns<double>();
  if ((n << order) > maxc)
    return set_error(SIGPDE_EINVAL,
                     "refined columns (length-1) << order = %lld exceed %lld, the longest "
                     "boundary line one wave's LDS holds for this dtype",
                     (long long)(n << order), (long long)maxc);
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
  if (int rc = check_grid(dtype, m, n, order)) return rc;
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
  if (int rc = check_grid(dtype, m, n, order)) return rc;
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
  if (int rc = check_grid(dtype, m, m, order)) return rc;
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
