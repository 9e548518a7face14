// C ABI of the fused reference-scheme static-kernel Gram (fused_point.h): a prepass pads every
// path's points to the producer's K width (and, for RBF, forms |point|^2 as the reference's
// torch.sum(X**2, dim=2) does: squares rounded, summed in order), then one launch solves the
// whole Gram.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include "capi_util.h"
#include "fused_point.h"

#pragma clang fp contract(off)

namespace sigpde {

// Points of one path per block: rows of dpad elements (zero padded) and, when xx is given,
// |point|^2 = sum_k fl(x_k * x_k) in order.
template <typename T>
__global__ __launch_bounds__(256) void k_point_prep(const T *__restrict__ X, int L, int d,
                                                    int dpad, T *__restrict__ xp,
                                                    T *__restrict__ xx) {
  const int64_t b = blockIdx.x;
  const T *x = X + b * (int64_t)L * d;
  T *o = xp + b * (int64_t)L * dpad;
  for (int i = threadIdx.x; i < L * dpad; i += blockDim.x) {
    const int s = i / dpad, k = i - s * dpad;
    o[i] = k < d ? x[s * d + k] : (T)0;
  }
  if (xx) {
    for (int s = threadIdx.x; s < L; s += blockDim.x) {
      T acc = (T)0;
      for (int k = 0; k < d; ++k) {
        const T v = x[s * d + k];
        acc += v * v;
      }
      xx[b * L + s] = acc;
    }
  }
}

template <typename T>
__global__ void k_scales_to_double(const T *src, int64_t n, double *dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}

constexpr int kMaxPointOrder = 3;

static bool point_kind_ok(int kind) {
  return kind == kPointLinear || kind == kPointRBF || kind == kPointL1;
}

// Elements of K per lane (MFMA kinds, dpad = 4 NE) or per column (L1-RBF, dpad = NE) for dim d, or
// 0 past the route's limit (MFMA: float32 64, float64 32; L1-RBF: float32 16, float64 8, where
// the whole column sits in registers).
static int point_ne(int dtype, int kind, int64_t d) {
  if (kind == kPointL1) return d <= 8 ? 8 : (d <= 16 && dtype == SIGPDE_F32 ? 16 : 0);
  const int64_t cap = dtype == SIGPDE_F64 ? 8 : 16;
  for (int ne = 2; ne <= cap; ne *= 2)
    if (4 * ne >= d) return ne;
  return 0;
}
static int point_dpad(int kind, int ne) { return kind == kPointL1 ? ne : 4 * ne; }

static size_t align_up256(size_t v) { return (v + 255) & ~(size_t)255; }

struct PointWorkspace {
  void *xp, *yp, *xx, *yy;
  double *sxd, *syd;
};

static size_t point_workspace_layout(int dtype, int kind, int64_t Bx, int64_t Lx, int64_t By,
                                     int64_t Ly, int64_t d, int same, char *base,
                                     PointWorkspace *w) {
  const int ne = point_ne(dtype, kind, d);
  const size_t dp = (size_t)point_dpad(kind, ne);
  const size_t es = dtype == SIGPDE_F64 ? 8 : 4;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += align_up256(bytes);
    return p;
  };
  PointWorkspace t{};
  t.xp = take((size_t)Bx * Lx * dp * es);
  t.yp = same ? t.xp : take((size_t)By * Ly * dp * es);
  if (kind == kPointRBF) {
    t.xx = take((size_t)Bx * Lx * es);
    t.yy = same ? t.xx : take((size_t)By * Ly * es);
  }
  t.sxd = (double *)take((size_t)Bx * 8);
  t.syd = (double *)take((size_t)By * 8);
  if (w) *w = t;
  return off;
}

// 0 when the fused route takes this problem, else SIGPDE_EINVAL with the reason.
static int point_check(int dtype, int kind, int64_t Bx, int64_t By, int64_t Lx, int64_t Ly,
                       int64_t d, int order) {
  if (dtype != SIGPDE_F32 && dtype != SIGPDE_F64)
    return set_error(SIGPDE_EINVAL, "static_gram: dtype must be SIGPDE_F32 or SIGPDE_F64");
  if (!point_kind_ok(kind))
    return set_error(SIGPDE_EINVAL, "static_gram: unknown static kernel %d", kind);
  if (Bx < 0 || By < 0 || Lx < 2 || Ly < 2 || d < 1)
    return set_error(SIGPDE_EINVAL, "static_gram: bad shape (Bx %lld, By %lld, Lx %lld, Ly %lld, "
                     "d %lld)", (long long)Bx, (long long)By, (long long)Lx, (long long)Ly,
                     (long long)d);
  if (order < 0 || order > kMaxPointOrder)
    return set_error(SIGPDE_EINVAL, "static_gram: dyadic order %d, the fused route takes 0..%d",
                     order, kMaxPointOrder);
  if (point_ne(dtype, kind, d) == 0)
    return set_error(SIGPDE_EINVAL, "static_gram: dim %lld exceeds the fused route's %d",
                     (long long)d,
                     kind == kPointL1 ? (dtype == SIGPDE_F64 ? 8 : 16) : (dtype == SIGPDE_F64 ? 32 : 64));
  if (((Lx - 1) << order) > kMaxRefinedRows || Lx > INT32_MAX / 64)
    return set_error(SIGPDE_EINVAL, "static_gram: refined rows (Lx-1) << order = %lld exceed "
                     "%lld", (long long)((Lx - 1) << order), (long long)kMaxRefinedRows);
  const int64_t N = (Ly - 1) << order;
  const int dp = point_dpad(kind, point_ne(dtype, kind, d));
  const size_t lds = dtype == SIGPDE_F64 ? point_lds_bytes<double>(kind, (int)(N < (1 << 22) ? N : (1 << 22)), dp)
                                         : point_lds_bytes<float>(kind, (int)(N < (1 << 22) ? N : (1 << 22)), dp);
  if (N > kMaxRefinedCols || lds > kMaxLdsBytes)
    return set_error(SIGPDE_EINVAL, "static_gram: %lld refined columns need %zu bytes of LDS per "
                     "wave, more than the fused route holds (%zu)", (long long)N, lds,
                     kMaxLdsBytes);
  return SIGPDE_OK;
}

template <typename T>
static hipError_t point_launch(int kind, int ne, PointGramArgs<T> a, hipStream_t st) {
#define SIGPDE_NE_CASES(KIND)                                          \
  switch (ne) {                                                        \
    case 2: return launch_point_gram<T, KIND, 2>(a, st);               \
    case 4: return launch_point_gram<T, KIND, 4>(a, st);               \
    case 8: return launch_point_gram<T, KIND, 8>(a, st);               \
    default: break;                                                    \
  }
  if (kind == kPointL1) {
    if (ne == 8) return launch_point_gram<T, kPointL1, 8>(a, st);
    if constexpr (std::is_same<T, float>::value) return launch_point_gram<T, kPointL1, 16>(a, st);
    return hipErrorInvalidValue;
  }
  if (kind == kPointRBF) {
    SIGPDE_NE_CASES(kPointRBF)
    if constexpr (std::is_same<T, float>::value) return launch_point_gram<T, kPointRBF, 16>(a, st);
  } else {
    SIGPDE_NE_CASES(kPointLinear)
    if constexpr (std::is_same<T, float>::value)
      return launch_point_gram<T, kPointLinear, 16>(a, st);
  }
#undef SIGPDE_NE_CASES
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t point_run(int kind, const void *X, int64_t Bx, int64_t Lx, const void *Y,
                            int64_t By, int64_t Ly, int64_t d, double param, int order,
                            const void *sx, const void *sy, int symmetric, void *out,
                            int64_t ldo, const PointWorkspace &w, hipStream_t st) {
  const int dtype = std::is_same<T, double>::value ? SIGPDE_F64 : SIGPDE_F32;
  const int ne = point_ne(dtype, kind, d);
  const int dp = point_dpad(kind, ne);
  const bool rbf = kind == kPointRBF;
  k_point_prep<T><<<dim3((unsigned)Bx), dim3(256), 0, st>>>((const T *)X, (int)Lx, (int)d, dp,
                                                            (T *)w.xp, rbf ? (T *)w.xx : nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (w.yp != w.xp) {
    k_point_prep<T><<<dim3((unsigned)By), dim3(256), 0, st>>>((const T *)Y, (int)Ly, (int)d, dp,
                                                              (T *)w.yp,
                                                              rbf ? (T *)w.yy : nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (sx) {
    k_scales_to_double<T><<<dim3((unsigned)((Bx + 255) / 256)), dim3(256), 0, st>>>(
        (const T *)sx, Bx, w.sxd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (sy) {
    k_scales_to_double<T><<<dim3((unsigned)((By + 255) / 256)), dim3(256), 0, st>>>(
        (const T *)sy, By, w.syd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  PointGramArgs<T> a{};
  a.op.xp = (const T *)w.xp;
  a.op.yp = (const T *)w.yp;
  a.op.xx = (const T *)w.xx;
  a.op.yy = (const T *)w.yy;
  a.sx = sx ? w.sxd : nullptr;
  a.sy = sy ? w.syd : nullptr;
  a.Bx = Bx;
  a.By = By;
  a.Lx = (int)Lx;
  a.Ly = (int)Ly;
  a.order = order;
  a.dpad = dp;
  // linear: the scale as a tensor-dtype scalar; RBF, L1-RBF: 1 / sigma in the tensor's dtype
  a.prm = kind == kPointLinear ? (T)param : (T)1 / (T)param;
  a.inv4 = (T)1 / (T)(1ll << (2 * order));
  a.symmetric = symmetric;
  a.out = (T *)out;
  a.ldo = ldo;
  return point_launch<T>(kind, ne, a, st);
}

}  // namespace sigpde

using namespace sigpde;

extern "C" int sigpde_static_gram_supported(int dtype, int kernel, int64_t Lx, int64_t Ly,
                                            int64_t d, int order) {
  clear_error();
  return point_check(dtype, kernel, 1, 1, Lx, Ly, d, order) == SIGPDE_OK;
}

extern "C" size_t sigpde_static_workspace_bytes(int dtype, int kernel, int64_t Bx, int64_t Lx,
                                                int64_t By, int64_t Ly, int64_t d,
                                                int same_paths) {
  if ((dtype != SIGPDE_F32 && dtype != SIGPDE_F64) || !point_kind_ok(kernel)) return 0;
  if (Bx < 0 || By < 0 || Lx < 2 || Ly < 2 || d < 1 || point_ne(dtype, kernel, d) == 0) return 0;
  return point_workspace_layout(dtype, kernel, Bx, Lx, By, Ly, d, same_paths, nullptr, nullptr);
}

extern "C" int sigpde_static_gram(int dtype, int kernel, const void *X, int64_t Bx, int64_t Lx,
                                  const void *Y, int64_t By, int64_t Ly, int64_t d, double param,
                                  int order, const void *sx, const void *sy, int symmetric,
                                  void *out, int64_t ldo, void *workspace, size_t ws_bytes,
                                  void *stream) {
  clear_error();
  if (int rc = point_check(dtype, kernel, Bx, By, Lx, Ly, d, order)) return rc;
  if (Bx == 0 || By == 0) return SIGPDE_OK;
  if (!X || !Y || !out || !workspace)
    return set_error(SIGPDE_EINVAL, "static_gram: null pointer");
  if (!(param == param) || (kernel != kPointLinear && param == 0.0))
    return set_error(SIGPDE_EINVAL, "static_gram: sigma must be a non-zero number");
  if (ldo < By) return set_error(SIGPDE_EINVAL, "static_gram: ldo %lld < By %lld",
                                 (long long)ldo, (long long)By);
  if (symmetric && (X != Y || Bx != By || Lx != Ly))
    return set_error(SIGPDE_EINVAL, "static_gram: symmetric needs Y == X");
  if (Bx > INT32_MAX / 2 || By > INT32_MAX / 2)
    return set_error(SIGPDE_EINVAL, "static_gram: more than 2^30 paths");
  const int same = X == Y && Bx == By && Lx == Ly;
  PointWorkspace w;
  const size_t need = point_workspace_layout(dtype, kernel, Bx, Lx, By, Ly, d, same,
                                             (char *)workspace, &w);
  if (ws_bytes < need)
    return set_error(SIGPDE_EINVAL, "static_gram: workspace of %zu bytes, need %zu", ws_bytes,
                     need);
  hipStream_t st = (hipStream_t)stream;
  const hipError_t e =
      dtype == SIGPDE_F32
          ? point_run<float>(kernel, X, Bx, Lx, Y, By, Ly, d, param, order, sx, sy, symmetric,
                             out, ldo, w, st)
          : point_run<double>(kernel, X, Bx, Lx, Y, By, Ly, d, param, order, sx, sy, symmetric,
                              out, ldo, w, st);
  return hip_status(e, "static_gram launch");
}
