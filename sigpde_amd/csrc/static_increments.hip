// Increments of the static kernels in one pass (the reference's point kernels,
// sigpde/torch/static_kernels.py:3-54, followed by the 4-term second difference of
// sigpde/torch/kernel.py:13-21):
//   inc[i][j] = (((k[i+1][j+1] + k[i][j]) - k[i+1][j]) - k[i][j+1]) * 4^-order
// The reference materialises the point Gram and then about ten full-size temporaries (the
// distance, its scaling, exp, three slice differences, the division); here a tile of increments
// (32 x 255 from inner products: a thread per point column, rows of 256 contiguous reads and 255
// contiguous writes; 16 x 16
// for L1-RBF, whose cost is the distance sums) reads its point values once, keeps k in LDS and
// writes the increments once.
//   linear: k = kscale * <x,y> with <x,y> the caller's GEMM output (the reference scheme's
//           increments for precision "ref", bit-equal to the torch formulation);
//   RBF:    k = exp(-((-2 <x,y>) + (|x|^2 + |y|^2)) * (1/sigma)) with <x,y> the caller's GEMM
//           output and |x|^2, |y|^2 the caller's row sums: the same float operations in the order
//           as the torch formulation (division by the Python scalar sigma is a multiplication by
//           its reciprocal in the operand type), so the increments equal the torch path's.
//   L1-RBF: k = exp(-(sum_k |x_k - y_k|) * (1/sigma)), the distance summed here over k in order
//           (no (A, B, M, N, D) temporary); the torch reduction sums in another order, so these
//           agree with it to float rounding.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "capi_util.h"

#pragma clang fp contract(off)

namespace sigpde {

constexpr int kIpRows = 32, kIpCols = 255;  // inner-product route: increments per tile (a thread
                                             // per point column: 256 = the block)
constexpr int kL1Tile = 16;                  // L1-RBF: increments per tile side
constexpr int kDimChunk = 32;                // L1: dimensions staged in LDS at a time

// Tile decomposition of a 1-D grid: blocks enumerate (pair, tile row, tile column).
struct IncGeom {
  int64_t pairs;  // A * B (Gram) or A (pairwise)
  int64_t B;      // y paths (1 for pairwise)
  int m, n;       // increments per x / y path (M - 1, N - 1)
  int tr, tc;     // increments per tile (rows, columns)
  int ti, tj;     // tiles per side
  __device__ void tile(int64_t blk, int64_t &a, int64_t &b, int &i0, int &j0) const {
    const int64_t per = (int64_t)ti * tj;
    const int64_t p = blk / per;
    const int r = (int)(blk - p * per);
    a = p / B;
    b = p - a * B;
    i0 = (r / tj) * tr;
    j0 = (r % tj) * tc;
  }
};

// The TR x TC increments of a tile from the (TR + 1) x (TC + 1) point values in `kt` (row
// stride KS), each thread taking columns t % TC of rows t / TC, t / TC + 256 / TC, ...
template <typename T, int TR, int TC, int KS>
__device__ __forceinline__ void write_tile(const T *kt, const IncGeom &g, int64_t a, int64_t b,
                                           int i0, int j0, T inv4, T *out) {
  if (threadIdx.x >= TC * (256 / TC)) return;
  const int lj = threadIdx.x % TC;
  const int j = j0 + lj;
  if (j >= g.n) return;
  for (int li = threadIdx.x / TC; li < TR; li += 256 / TC) {
    const int i = i0 + li;
    if (i >= g.m) break;
    const T *k0 = kt + li * KS + lj;
    const T d = ((k0[KS + 1] + k0[0]) - k0[KS]) - k0[1];
    out[((a * g.B + b) * g.m + i) * (int64_t)g.n + j] = d * inv4;
  }
}

// From point inner products. ip element (a, b, i, j) at a*sa + b*sb + i*si + j*sj (points
// i < m+1, j < n+1). RBF (RBF = true): xx (a, i) at a*xa + i, yy (a, b, j) at a*ya + b*yb + j,
// c = 1/sigma; linear: c = kscale (xx, yy unused).
template <typename T, bool RBF>
__global__ __launch_bounds__(256) void k_ip_increments(const T *__restrict__ ip, int64_t sa,
                                                       int64_t sb, int64_t si, int64_t sj,
                                                       const T *__restrict__ xx, int64_t xa,
                                                       const T *__restrict__ yy, int64_t ya,
                                                       int64_t yb, IncGeom g, T c, T inv4,
                                                       T *__restrict__ out) {
  constexpr int KR = kIpRows + 1, KC = kIpCols + 1;
  static_assert(KC == 256, "a thread per point column");
  __shared__ T kt[KR * KC];
  int64_t a, b;
  int i0, j0;
  g.tile(blockIdx.x, a, b, i0, j0);
  const int lj = threadIdx.x;
  const int j = min(j0 + lj, g.n);  // clamped: unused past the edge
  const T *col = ip + a * sa + b * sb + j * sj;
  T yv = (T)0;
  if constexpr (RBF) yv = yy[a * ya + b * yb + j];
  for (int li = 0; li < KR; ++li) {
    const int i = min(i0 + li, g.m);
    const T v = col[i * si];
    if constexpr (RBF) {
      T dist = (T)-2 * v;
      dist += xx[a * xa + i] + yv;
      kt[li * KC + lj] = exp((-dist) * c);
    } else {
      kt[li * KC + lj] = c * v;
    }
  }
  __syncthreads();
  write_tile<T, kIpRows, kIpCols, KC>(kt, g, a, b, i0, j0, inv4, out);
}

// L1-RBF. X (A, m+1, D), Y (By, n+1, D) contiguous; pairwise: y path = x path (ypair = 1).
template <typename T>
__global__ __launch_bounds__(256) void k_l1rbf_increments(const T *__restrict__ X,
                                                          const T *__restrict__ Y, int D,
                                                          int ypair, IncGeom g, T inv_sigma,
                                                          T inv4, T *__restrict__ out) {
  constexpr int kPt = kL1Tile + 1;  // point values per tile side
  __shared__ T kt[kPt * kPt];
  __shared__ T xs[kPt][kDimChunk + 1], ys[kPt][kDimChunk + 1];
  int64_t a, b;
  int i0, j0;
  g.tile(blockIdx.x, a, b, i0, j0);
  const int64_t yp = ypair ? a : b;
  const T *xr = X + a * (int64_t)(g.m + 1) * D;
  const T *yr = Y + yp * (int64_t)(g.n + 1) * D;
  // each thread owns point values p = threadIdx.x and threadIdx.x + 256 (289 per tile)
  T acc0 = 0, acc1 = 0;
  const int p0 = threadIdx.x, p1 = threadIdx.x + 256;
  for (int k0 = 0; k0 < D; k0 += kDimChunk) {
    const int kc = min(kDimChunk, D - k0);
    for (int e = threadIdx.x; e < kPt * kDimChunk; e += blockDim.x) {
      const int r = e / kDimChunk, k = e % kDimChunk;
      const int i = min(i0 + r, g.m), j = min(j0 + r, g.n);
      xs[r][k] = k < kc ? xr[(int64_t)i * D + k0 + k] : (T)0;
      ys[r][k] = k < kc ? yr[(int64_t)j * D + k0 + k] : (T)0;
    }
    __syncthreads();
    {
      const int li = p0 / kPt, lj = p0 % kPt;
      for (int k = 0; k < kc; ++k) acc0 += fabs(xs[li][k] - ys[lj][k]);
    }
    if (p1 < kPt * kPt) {
      const int li = p1 / kPt, lj = p1 % kPt;
      for (int k = 0; k < kc; ++k) acc1 += fabs(xs[li][k] - ys[lj][k]);
    }
    __syncthreads();
  }
  kt[p0] = exp((-acc0) * inv_sigma);
  if (p1 < kPt * kPt) kt[p1] = exp((-acc1) * inv_sigma);
  __syncthreads();
  write_tile<T, kL1Tile, kL1Tile, kPt>(kt, g, a, b, i0, j0, inv4, out);
}

static IncGeom inc_geom(int64_t A, int64_t B, int64_t m, int64_t n, int tr, int tc) {
  IncGeom g;
  g.pairs = A * B;
  g.B = B;
  g.m = (int)m;
  g.n = (int)n;
  g.tr = tr;
  g.tc = tc;
  g.ti = (int)((m + tr - 1) / tr);
  g.tj = (int)((n + tc - 1) / tc);
  return g;
}

static int check_inc(int dtype, int64_t A, int64_t B, int64_t m, int64_t n, int order,
                     double sigma) {
  if (dtype != SIGPDE_F32 && dtype != SIGPDE_F64)
    return set_error(SIGPDE_EINVAL, "increments: unknown dtype %d", dtype);
  if (A < 0 || B < 0 || m < 1 || n < 1)
    return set_error(SIGPDE_EINVAL, "increments: bad sizes (paths need length >= 2)");
  if (m > (1 << 24) || n > (1 << 24))
    return set_error(SIGPDE_EINVAL, "increments: paths longer than 2^24 points");
  if (order < 0 || order > 12) return set_error(SIGPDE_EINVAL, "increments: dyadic order must be in [0, 12]");
  if (!(sigma != 0.0)) return set_error(SIGPDE_EINVAL, "increments: sigma must be nonzero");
  return SIGPDE_OK;
}

// Reciprocals in the operand type, as torch forms them for a division by a Python scalar.
template <typename T>
static T recip(double v) {
  return (T)1 / (T)v;
}

}  // namespace sigpde

using namespace sigpde;

extern "C" int sigpde_rbf_increments(int dtype, const void *ip, int64_t A, int64_t B, int64_t m,
                                     int64_t n, const int64_t *ip_strides, const void *xx,
                                     int64_t xx_stride_a, const void *yy, int64_t yy_stride_a,
                                     int64_t yy_stride_b, double sigma, int order, void *out,
                                     void *stream) {
  clear_error();
  if (int rc = check_inc(dtype, A, B, m, n, order, sigma)) return rc;
  if (A == 0 || B == 0) return SIGPDE_OK;
  if (!ip || !ip_strides || !xx || !yy || !out)
    return set_error(SIGPDE_EINVAL, "rbf_increments: null pointer");
  const IncGeom g = inc_geom(A, B, m, n, kIpRows, kIpCols);
  const int64_t blocks = g.pairs * g.ti * g.tj;
  if (blocks > INT32_MAX) return set_error(SIGPDE_EINVAL, "rbf_increments: too many tiles");
  hipStream_t st = (hipStream_t)stream;
  const int64_t *s = ip_strides;
  if (dtype == SIGPDE_F32) {
    k_ip_increments<float, true><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
        (const float *)ip, s[0], s[1], s[2], s[3], (const float *)xx, xx_stride_a,
        (const float *)yy, yy_stride_a, yy_stride_b, g, recip<float>(sigma),
        recip<float>((double)(1ll << (2 * order))), (float *)out);
  } else {
    k_ip_increments<double, true><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
        (const double *)ip, s[0], s[1], s[2], s[3], (const double *)xx, xx_stride_a,
        (const double *)yy, yy_stride_a, yy_stride_b, g, recip<double>(sigma),
        recip<double>((double)(1ll << (2 * order))), (double *)out);
  }
  return hip_status(hipGetLastError(), "rbf_increments launch");
}

extern "C" int sigpde_linear_increments(int dtype, const void *ip, int64_t A, int64_t B,
                                        int64_t m, int64_t n, const int64_t *ip_strides,
                                        double kscale, int order, void *out, void *stream) {
  clear_error();
  if (int rc = check_inc(dtype, A, B, m, n, order, 1.0)) return rc;
  if (A == 0 || B == 0) return SIGPDE_OK;
  if (!ip || !ip_strides || !out) return set_error(SIGPDE_EINVAL, "linear_increments: null pointer");
  const IncGeom g = inc_geom(A, B, m, n, kIpRows, kIpCols);
  const int64_t blocks = g.pairs * g.ti * g.tj;
  if (blocks > INT32_MAX) return set_error(SIGPDE_EINVAL, "linear_increments: too many tiles");
  hipStream_t st = (hipStream_t)stream;
  const int64_t *s = ip_strides;
  if (dtype == SIGPDE_F32) {
    k_ip_increments<float, false><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
        (const float *)ip, s[0], s[1], s[2], s[3], nullptr, 0, nullptr, 0, 0, g, (float)kscale,
        recip<float>((double)(1ll << (2 * order))), (float *)out);
  } else {
    k_ip_increments<double, false><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
        (const double *)ip, s[0], s[1], s[2], s[3], nullptr, 0, nullptr, 0, 0, g, kscale,
        recip<double>((double)(1ll << (2 * order))), (double *)out);
  }
  return hip_status(hipGetLastError(), "linear_increments launch");
}

extern "C" int sigpde_l1rbf_increments(int dtype, const void *X, int64_t A, const void *Y,
                                       int64_t B, int64_t m, int64_t n, int64_t d, int pairwise,
                                       double sigma, int order, void *out, void *stream) {
  clear_error();
  if (int rc = check_inc(dtype, A, pairwise ? 1 : B, m, n, order, sigma)) return rc;
  if (A == 0 || (!pairwise && B == 0)) return SIGPDE_OK;
  if (!X || !Y || !out) return set_error(SIGPDE_EINVAL, "l1rbf_increments: null pointer");
  if (d < 1 || d > (1 << 20)) return set_error(SIGPDE_EINVAL, "l1rbf_increments: bad dim");
  const IncGeom g = inc_geom(A, pairwise ? 1 : B, m, n, kL1Tile, kL1Tile);
  const int64_t blocks = g.pairs * g.ti * g.tj;
  if (blocks > INT32_MAX) return set_error(SIGPDE_EINVAL, "l1rbf_increments: too many tiles");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == SIGPDE_F32) {
    k_l1rbf_increments<float><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
        (const float *)X, (const float *)Y, (int)d, pairwise, g, recip<float>(sigma),
        recip<float>((double)(1ll << (2 * order))), (float *)out);
  } else {
    k_l1rbf_increments<double><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
        (const double *)X, (const double *)Y, (int)d, pairwise, g, recip<double>(sigma),
        recip<double>((double)(1ll << (2 * order))), (double *)out);
  }
  return hip_status(hipGetLastError(), "l1rbf_increments launch");
}
