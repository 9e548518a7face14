// Fused linear-kernel signature Gram / pairwise on gfx950: the increment Gram
//   c[i][j] = kscale * <x_{i+1} - x_i, y_{j+1} - y_j> / 4^order
// (equal to the reference's 4-term second difference of the point Gram, sigpde/torch/kernel.py:
// 13-21 with sigpde/torch/static_kernels.py:7-11) is computed on MFMA straight into the wavefront's
// LDS ring; no increment tensor is ever written to HBM.
//
// Precision "fp16x2": each path's increments are formed exactly in double, scaled by a per-path
// power of two 2^e (so max |dx| * 2^e lies in [2^13, 2^14)) and split into fp16 hi + lo. Products
// lo*hi + hi*lo and hi*hi (the lo*lo term is below the fp16x2 resolution) accumulate in fp32 on
// v_mfma_f32_16x16x32_f16, in one accumulator per magnitude joined by a float add, and go to the
// ring unscaled; the wavefront runs in double and applies
// the pair's scale m = kscale 4^-order 2^-(ex+ey) (x_scale y_scale) through its coefficients
// (Arith::Mixed). Relative error against the float64 reference: ~3e-7 on Brownian paths.
//
// Precision "bf16x3": the same scaled increments split into three bf16 parts (hi, mid, lo: each
// the bf16 rounding of what the previous parts leave), the six products of magnitude >= 2^-16 of
// hi*hi on v_mfma_f32_16x16x32_bf16 in fp32, the same double wavefront (Arith::Mixed).
//
// Precision "fp32" (the plain FP32 fused path): increments rounded to float, products on
// v_mfma_f32_16x16x4f32 in fp32, scaled in fp32, and the wavefront in the reference's float32
// typing (Arith::Ref32: c and c^2 in float, the rest in double, cells stored as float).
//
// Precision "fp64": increments in double (exact for float inputs), products on
// v_mfma_f64_16x16x4f64, scaled in double, and the wavefront all in double (Arith::Ref64): the
// reference scheme in float64 without the increment tensor (dim <= 64).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <climits>

#include "capi_util.h"
#include "gram16f.h"
#include "mfma_producer.h"

namespace sigpde {

// ------------------------------------------------------------------ split prepass
// One block per path: dx = x[s+1] - x[s] in double (exact for float inputs), per-path scale 2^e
// (max |dx| 2^e in [2^13, 2^14): fp16 parts stay in range), NPART 16-bit parts of E per element,
// each the rounding of what the previous parts leave (computed exactly in double), rows of dpad
// elements (zero padded), and inv = 2^-e.
template <typename T, typename E, int NPART>
__global__ __launch_bounds__(256) void k_split16(const T *__restrict__ X, int64_t B, int L, int d,
                                                 int dpad, E *__restrict__ p0, E *__restrict__ p1,
                                                 E *__restrict__ p2, float *__restrict__ inv) {
  __shared__ double red[256];
  const int64_t b = blockIdx.x;
  const T *x = X + b * (int64_t)L * d;
  const int m = L - 1;
  double mx = 0.0;
  for (int i = threadIdx.x; i < m * d; i += blockDim.x) {
    const int s = i / d, k = i - s * d;
    double v = (double)x[(s + 1) * d + k] - (double)x[s * d + k];
    mx = fmax(mx, fabs(v));
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  mx = red[0];
  const int e = (mx > 0.0 && isfinite(mx)) ? 13 - ilogb(mx) : 0;
  if (threadIdx.x == 0) inv[b] = ldexpf(1.0f, -e);
  const int64_t row0 = b * (int64_t)m * dpad;
  for (int i = threadIdx.x; i < m * dpad; i += blockDim.x) {
    const int s = i / dpad, k = i - s * dpad;
    double v = 0.0;
    if (k < d) v = ldexp((double)x[(s + 1) * d + k] - (double)x[s * d + k], e);
    const E a = (E)(float)v;
    v -= (double)(float)a;
    const E c = (E)(float)v;
    p0[row0 + i] = a;
    if (NPART == 2) {
      p1[row0 + i] = c;
    } else {
      v -= (double)(float)c;
      p1[row0 + i] = c;
      p2[row0 + i] = (E)(float)v;
    }
  }
}

// fp32 operands: one block per path, dx = x[s+1] - x[s] in the input type, rounded to float,
// rows of dpad floats (zero padded).
template <typename T>
__global__ __launch_bounds__(256) void k_split_f32(const T *__restrict__ X, int L, int d, int dpad,
                                                   float *__restrict__ out) {
  const int64_t b = blockIdx.x;
  const T *x = X + b * (int64_t)L * d;
  const int m = L - 1;
  float *o = out + b * (int64_t)m * dpad;
  for (int i = threadIdx.x; i < m * dpad; i += blockDim.x) {
    const int s = i / dpad, k = i - s * dpad;
    o[i] = k < d ? (float)(x[(s + 1) * d + k] - x[s * d + k]) : 0.0f;
  }
}

// fp64 operands: one block per path, dx = x[s+1] - x[s] in double, rows of dpad doubles (zero
// padded).
template <typename T>
__global__ __launch_bounds__(256) void k_split_f64(const T *__restrict__ X, int L, int d, int dpad,
                                                   double *__restrict__ out) {
  const int64_t b = blockIdx.x;
  const T *x = X + b * (int64_t)L * d;
  const int m = L - 1;
  double *o = out + b * (int64_t)m * dpad;
  for (int i = threadIdx.x; i < m * dpad; i += blockDim.x) {
    const int s = i / dpad, k = i - s * dpad;
    o[i] = k < d ? (double)x[(s + 1) * d + k] - (double)x[s * d + k] : 0.0;
  }
}

// Arithmetic of the wavefront for each producer precision.
// ------------------------------------------------------------------ kernels
// Gram: a wave takes four x paths (a quad) and streams `ys` y paths through them.
// Symmetric: only y <= x pairs are written, with their mirror.

// One work unit of a fused Gram: x paths xb + 4 xq + g (g < 4, below xe) against the ys y paths
// from yb + yr ys (below ye); the result of x path i and y path j goes to
// out[(i - xb) * ldo + (j - yb)]. symmetric: the unit's x and y paths are the same paths (X == Y
// over the unit's request), only y <= x pairs are solved and each value is also written to its
// mirror.
template <typename TO, int KB, int PREC, int R, bool GL>
__device__ __forceinline__ void gram_unit(const GramFusedArgs<TO> &a, const double *sxd,
                                          const double *syd, void *gl, int64_t xb, int64_t xe,
                                          int64_t xq, int64_t yb, int64_t ye, int64_t yr,
                                          bool symmetric, TO *out, int64_t ldo) {
  using Prod = MfmaProducer<KB, 4, 1, PREC>;
  using E = typename Prod::elt;
  constexpr Arith A = FusedArith<PREC>::A;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int G = 16;
  const int lane = threadIdx.x & 63;
  const int64_t bx = xe - xb, by = ye - yb;
  const int64_t y0 = yr * a.ys;  // local
  int ys = (int)((by - y0) < a.ys ? (by - y0) : a.ys);
  if (symmetric) {
    // only y <= max x of the quad
    const int64_t xmax = xq * 4 + 3 < bx - 1 ? xq * 4 + 3 : bx - 1;
    if (y0 > xmax) return;
    if (y0 + ys - 1 > xmax) ys = (int)(xmax - y0 + 1);
  }
  PairGeom<G> geo(a.m, a.n, a.order, ys, R);
  const int64_t xme = xq * 4 + lane / G;  // local
  auto write = [&](int sp, double r) {
    const int64_t y = y0 + sp;
    if (xme >= bx || y >= by) return;
    if (symmetric) {
      if (y > xme) return;
      out[xme * ldo + y] = (TO)r;
      out[y * ldo + xme] = (TO)r;
    } else {
      out[xme * ldo + y] = (TO)r;
    }
  };
  Prod P;
  P.xh = (const E *)a.xh;
  P.xl = (const E *)a.xl;
  P.yh = (const E *)a.yh;
  P.yl = (const E *)a.yl;
  P.xm = (const E *)a.xm;
  P.ym = (const E *)a.ym;
  P.dpad = a.dpad;
  P.m = a.m;
  P.n = a.n;
  P.ystep = 1;
  P.invy = a.invy;
  P.sx = sxd;
  P.sy = syd;
  P.kmult = a.kmult;
  P.ny = ye;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int64_t x = xq * 4 + g;
    P.xa[g] = x < bx ? xb + x : -1;
  }
  P.ybase[0] = yb + y0;
  P.ylane = yb + y0;
  {
    const int64_t xc = xb + (xme < bx ? xme : bx - 1);
    P.xsl = Prod::kSplit ? a.kmult * (double)a.invx[xc] * (sxd ? sxd[xc] : 1.0) : 1.0;
    P.pmY = 1.0;
  }
  ring16_solve<A, R, GL>(P, geo, 1.0, smem_raw, write, fused_wave_lines<A, GL>(gl, geo.N));
}

// Occupancy: at dim <= 32 (one K block) the fp16x2 kernel fits three waves per SIMD (measured 9%
// faster at dim 16); the other instantiations would spill there and keep the compiler's choice.
// GL: boundary lines in global memory at gl (one block per wave of the launch), else in LDS.
template <typename TO, int KB, int PREC, int R, bool GL>
__global__ __launch_bounds__(PREC == kPrecFp16x2F ? 128 : 64)
__attribute__((amdgpu_waves_per_eu(KB == 1 && PREC == kPrecFp16x2 && R == 1 ? 3 : 1))) void
k_linear_gram_fused(GramFusedArgs<TO> a, const double *sxd, const double *syd, void *gl) {
  const int64_t nq = (a.Bx + 3) / 4;
  const int64_t nyr = (a.By + a.ys - 1) / a.ys;
  // XCD-aware order: blocks are dealt round-robin to the 8 XCDs; each XCD takes a contiguous run
  // of work units, ordered in super-tiles of tx x-quads by ty y-ranges, so that the waves resident
  // on one XCD at a time share their x and y operands in that XCD's L2.
  const int64_t bid = blockIdx.x;             // grid = 8 * per_xcd blocks
  const int64_t per_xcd = gridDim.x / 8;
  const int64_t lin = (bid % 8) * per_xcd + bid / 8;
  const int64_t tsz = (int64_t)a.tx * a.ty;
  const int64_t tile = lin / tsz, w = lin - tile * tsz;
  const int64_t ntx = (nq + a.tx - 1) / a.tx;
  const int64_t xq = (tile % ntx) * a.tx + w % a.tx;
  const int64_t yr = (tile / ntx) * a.ty + w / a.tx;
  if (xq >= nq || yr >= nyr) return;
  gram_unit<TO, KB, PREC, R, GL>(a, sxd, syd, gl, 0, a.Bx, xq, 0, a.By, yr, a.symmetric != 0,
                                 a.out, a.ldo);
}

// Several Grams of one shape class in one launch: request r takes the units prefix[r] ..
// prefix[r+1] - 1 (x-quads fastest), its x paths x0 .. x0 + bx - 1 and y paths y0 .. y0 + by - 1
// of the launch's split operands, and writes its (bx, by) result at out + out_off, row-major.
// The request table travels by value in the kernel arguments (no device table, nothing to
// allocate; a captured launch holds it).
constexpr int kMaxGramBatch = 64;
struct GramBatchReq {
  int32_t x0, bx, y0, by, prefix, symmetric;
  int64_t out_off;
};
struct GramBatch {
  int32_t n, nunits;
  GramBatchReq r[kMaxGramBatch];
};

template <typename TO, int KB, int PREC, int R, bool GL>
__global__ __launch_bounds__(PREC == kPrecFp16x2F ? 128 : 64)
__attribute__((amdgpu_waves_per_eu(KB == 1 && PREC == kPrecFp16x2 && R == 1 ? 3 : 1))) void
k_linear_gram_batch(GramFusedArgs<TO> a, const double *sxd, const double *syd, void *gl,
                    GramBatch b) {
  // blocks dealt round-robin to the 8 XCDs, each XCD a contiguous run of units (of few requests)
  const int64_t bid = blockIdx.x;
  const int64_t per_xcd = gridDim.x / 8;
  const int64_t lin = (bid % 8) * per_xcd + bid / 8;
  if (lin >= b.nunits) return;
  int lo = 0, hi = b.n - 1;  // the last request whose first unit is <= lin
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.r[mid].prefix <= lin) lo = mid;
    else hi = mid - 1;
  }
  const GramBatchReq q = b.r[lo];
  const int64_t u = lin - q.prefix;
  const int64_t nq = (q.bx + 3) / 4;
  gram_unit<TO, KB, PREC, R, GL>(a, sxd, syd, gl, q.x0, (int64_t)q.x0 + q.bx, u % nq, q.y0,
                                 (int64_t)q.y0 + q.by, u / nq, q.symmetric != 0,
                                 a.out + q.out_off, q.by);
}

// ------------------------------------------------------------------ fp16x2f Gram (gram16f.h)
// One work unit of the fp16x2f Gram on a workgroup of NC consumer / producer wave pairs: x quads
// NC xg .. NC xg + NC - 1 (consumer and producer p take quad NC xg + p) against the ys y paths
// from yb + yr ys, results as in gram_unit. symmetric: the stream is cut at the group's largest x
// path (each consumer writes only its y <= x pairs and their mirrors).
template <typename TO, int KB, int R, int NC>
__device__ __forceinline__ void gram_unit_f16f(const GramFusedArgs<TO> &a, const double *sxd,
                                               const double *syd, int64_t xb, int64_t xe,
                                               int64_t xg, int64_t yb, int64_t ye, int64_t yr,
                                               bool symmetric, TO *out, int64_t ldo) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & 63;
  const int pi = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) % NC);
  const int64_t xq = xg * NC + pi;
  const int64_t bx = xe - xb, by = ye - yb;
  const int64_t y0 = yr * a.ys;  // local
  int ys = (int)((by - y0) < a.ys ? (by - y0) : a.ys);
  if (symmetric) {
    // only y <= the group's largest x path
    const int64_t xmax = xg * NC * 4 + 4 * NC - 1 < bx - 1 ? xg * NC * 4 + 4 * NC - 1 : bx - 1;
    if (y0 > xmax) return;
    if (y0 + ys - 1 > xmax) ys = (int)(xmax - y0 + 1);
  }
  PairGeom<16> geo(a.m, a.n, a.order, ys, R);
  const int64_t xme = xq * 4 + lane / 16;  // local
  G16FProducer<KB, NC> Q;
  Q.xh = (const _Float16 *)a.xh;
  Q.xl = (const _Float16 *)a.xl;
  Q.yh = (const _Float16 *)a.yh;
  Q.yl = (const _Float16 *)a.yl;
  Q.dpad = a.dpad;
  Q.m = a.m;
  Q.n = a.n;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int64_t x = xq * 4 + g;
    Q.xa[g] = x < bx ? xb + x : -1;
  }
  Q.ybase = yb + y0;
  Q.ny = ye;
  const int64_t xc = xb + (xme < bx ? xme : bx - 1);
  const double xsl = a.kmult * (double)a.invx[xc] * (sxd ? sxd[xc] : 1.0);
  // scale m = kscale 4^-order 2^-(ex+ey) (x_scale y_scale) of stream pair sp of this lane's group
  auto pair_scale = [&](int sp) -> double {
    int64_t yp = yb + y0 + sp;
    yp = yp < ye ? yp : ye - 1;
    return xsl * (syd ? (double)a.invy[yp] * syd[yp] : (double)a.invy[yp]);
  };
  auto write = [&](int sp, double r) {
    const int64_t y = y0 + sp;
    if (xme >= bx || y >= by) return;
    if (symmetric) {
      if (y > xme) return;
      out[xme * ldo + y] = (TO)r;
      out[y * ldo + xme] = (TO)r;
    } else {
      out[xme * ldo + y] = (TO)r;
    }
  };
  g16f_solve<R, KB, NC>(Q, geo, smem_raw, pair_scale, write);
}

// The fp16x2f Gram: one x group (NC x quads) by one y range per workgroup, in the XCD-aware
// super-tile order of k_linear_gram_fused (tx x groups by ty y ranges).
template <typename TO, int KB, int R, int NC>
__global__ __launch_bounds__(128 * NC) void k_linear_gram_f16f(GramFusedArgs<TO> a,
                                                               const double *sxd,
                                                               const double *syd) {
  const int64_t nq = (a.Bx + 3) / 4;
  const int64_t nxg = (nq + NC - 1) / NC;
  const int64_t nyr = (a.By + a.ys - 1) / a.ys;
  const int64_t bid = blockIdx.x;
  const int64_t per_xcd = gridDim.x / 8;
  const int64_t lin = (bid % 8) * per_xcd + bid / 8;
  const int64_t tsz = (int64_t)a.tx * a.ty;
  const int64_t tile = lin / tsz, w = lin - tile * tsz;
  const int64_t ntx = (nxg + a.tx - 1) / a.tx;
  const int64_t xg = (tile % ntx) * a.tx + w % a.tx;
  const int64_t yr = (tile / ntx) * a.ty + w / a.tx;
  if (xg >= nxg || yr >= nyr) return;
  gram_unit_f16f<TO, KB, R, NC>(a, sxd, syd, 0, a.Bx, xg, 0, a.By, yr, a.symmetric != 0, a.out,
                                a.ldo);
}

// Several fp16x2f Grams in one launch (k_linear_gram_batch's request table; units are x groups by
// y ranges, x groups fastest).
template <typename TO, int KB, int R, int NC>
__global__ __launch_bounds__(128 * NC) void k_linear_gram_batch_f16f(GramFusedArgs<TO> a,
                                                                     const double *sxd,
                                                                     const double *syd,
                                                                     GramBatch b) {
  const int64_t bid = blockIdx.x;
  const int64_t per_xcd = gridDim.x / 8;
  const int64_t lin = (bid % 8) * per_xcd + bid / 8;
  if (lin >= b.nunits) return;
  int lo = 0, hi = b.n - 1;  // the last request whose first unit is <= lin
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.r[mid].prefix <= lin) lo = mid;
    else hi = mid - 1;
  }
  const GramBatchReq q = b.r[lo];
  const int64_t u = lin - q.prefix;
  const int64_t nxg = ((q.bx + 3) / 4 + NC - 1) / NC;
  gram_unit_f16f<TO, KB, R, NC>(a, sxd, syd, q.x0, (int64_t)q.x0 + q.bx, u % nxg, q.y0,
                                (int64_t)q.y0 + q.by, u / nxg, q.symmetric != 0,
                                a.out + q.out_off, q.by);
}

// double copies of optional scales
template <typename T>
__global__ void k_to_double(const T *src, int64_t n, double *dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}

// ------------------------------------------------------------------ host side
constexpr int kTileX = 8;   // x-quads per super-tile
constexpr int kTileY = 32;  // y-ranges per super-tile

struct Workspace {
  void *xh, *xl, *yh, *yl;  // fp16x2: hi / lo halves; fp32: floats in xh / yh (xl = yl = null)
  void *xm, *ym;            // bf16x3: middle parts (hi / lo in xh / xl)
  float *invx, *invy;       // fp16x2, bf16x3
  double *sxd, *syd;
};

static inline size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

static int dpad_for(int64_t d) { return (int)((d + kKBlock - 1) / kKBlock * kKBlock); }

static size_t workspace_layout(int precision, int64_t Bx, int64_t Lx, int64_t By, int64_t Ly,
                               int64_t d, int same, char *base, Workspace *w) {
  const size_t dp = (size_t)dpad_for(d);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += align_up(bytes);
    return p;
  };
  const size_t nx = (size_t)Bx * (Lx - 1) * dp, ny = (size_t)By * (Ly - 1) * dp;
  Workspace t{};
  if (precision == SIGPDE_PREC_FP32 || precision == SIGPDE_PREC_FP64) {
    const size_t es = precision == SIGPDE_PREC_FP64 ? 8 : 4;
    t.xh = take(nx * es);
    t.yh = same ? t.xh : take(ny * es);
  } else {
    t.xh = take(nx * 2);
    t.xl = take(nx * 2);
    t.yh = same ? t.xh : take(ny * 2);
    t.yl = same ? t.xl : take(ny * 2);
    if (precision == SIGPDE_PREC_BF16X3) {
      t.xm = take(nx * 2);
      t.ym = same ? t.xm : take(ny * 2);
    }
    t.invx = (float *)take(Bx * 4);
    t.invy = same ? t.invx : (float *)take(By * 4);
  }
  t.sxd = (double *)take(Bx * 8);
  t.syd = (double *)take(By * 8);
  if (w) *w = t;
  return off;
}

template <typename T, int PREC>
static hipError_t run_split(const T *X, int64_t B, int64_t L, int64_t d, int dp, void *h, void *mid,
                            void *l, float *inv, hipStream_t st) {
  if (B == 0) return hipSuccess;
  const dim3 grid((unsigned)B), block(256);
  if (PREC == kPrecFp32)
    k_split_f32<T><<<grid, block, 0, st>>>(X, (int)L, (int)d, dp, (float *)h);
  else if (PREC == kPrecFp64)
    k_split_f64<T><<<grid, block, 0, st>>>(X, (int)L, (int)d, dp, (double *)h);
  else if (PREC == kPrecBf16x3)
    k_split16<T, __bf16, 3><<<grid, block, 0, st>>>(X, B, (int)L, (int)d, dp, (__bf16 *)h,
                                                    (__bf16 *)mid, (__bf16 *)l, inv);
  else
    k_split16<T, _Float16, 2><<<grid, block, 0, st>>>(X, B, (int)L, (int)d, dp, (_Float16 *)h,
                                                      (_Float16 *)l, nullptr, inv);
  return hipGetLastError();
}

template <typename T>
static hipError_t run_to_double(const void *src, int64_t n, double *dst, hipStream_t st) {
  if (!src || n == 0) return hipSuccess;
  k_to_double<T><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>((const T *)src, n, dst);
  return hipGetLastError();
}

// fp16x2f Gram launch with NC consumer / producer pairs per workgroup (a.ys set; batch as in
// launch_fused).
template <typename T, int KB, int R, int NC>
static hipError_t launch_gram_f16f_nc(GramFusedArgs<T> a, const double *sxd, const double *syd,
                                      const GramBatch *batch, int N, hipStream_t st) {
  const int64_t nxg = ((a.Bx + 3) / 4 + NC - 1) / NC;
  const int64_t nyr = (a.By + a.ys - 1) / a.ys;
  const int64_t txm = kTileX / NC;  // super-tiles of about kTileX x quads by kTileY y ranges
  a.tx = (int)(nxg < txm ? nxg : txm);
  a.ty = (int)(nyr < kTileY ? nyr : kTileY);
  int64_t nblk = (nxg + a.tx - 1) / a.tx * a.tx * ((nyr + a.ty - 1) / a.ty * a.ty);
  GramBatch gb{};
  if (batch) {
    gb = *batch;
    int64_t pre = 0;
    for (int r = 0; r < gb.n; ++r) {
      gb.r[r].prefix = (int32_t)pre;
      pre += (((int64_t)gb.r[r].bx + 3) / 4 + NC - 1) / NC * ((gb.r[r].by + a.ys - 1) / a.ys);
    }
    gb.nunits = (int32_t)pre;
    nblk = pre;
  }
  // a multiple of 8 blocks so that the XCD-major block order covers every work unit
  const int64_t grid = (nblk + 7) / 8 * 8;
  const size_t shmem = g16f_wg_lds_bytes<KB, NC>(N);
  if (batch)
    k_linear_gram_batch_f16f<T, KB, R, NC><<<dim3((unsigned)grid), dim3(128 * NC), shmem, st>>>(
        a, sxd, syd, gb);
  else
    k_linear_gram_f16f<T, KB, R, NC><<<dim3((unsigned)grid), dim3(128 * NC), shmem, st>>>(a, sxd,
                                                                                         syd);
  return hipGetLastError();
}

// Consumer waves one CU holds for NC pairs per workgroup (0 when the query fails or the LDS of
// a workgroup exceeds the CU's).
template <typename T, int KB, int R, int NC>
static int f16f_consumers_per_cu(int N) {
  const size_t lds = g16f_wg_lds_bytes<KB, NC>(N);
  if (lds > kMaxLdsBytes) return 0;
  return NC * blocks_per_cu(reinterpret_cast<const void *>(&k_linear_gram_f16f<T, KB, R, NC>),
                            128 * NC, lds);
}

// The fp16x2f Gram: NC = 4, 2 or 1 pairs of waves per workgroup, the one that keeps the most
// consumer waves resident per CU (the larger NC on a tie: a chunk's y operands are fetched once
// per workgroup); every choice gives the same results.
template <typename T, int KB, int R>
static hipError_t launch_gram_f16f(const GramFusedArgs<T> &a, const double *sxd, const double *syd,
                                   const GramBatch *batch, int N, hipStream_t st) {
  const int c4 = f16f_consumers_per_cu<T, KB, R, 4>(N);
  const int c2 = f16f_consumers_per_cu<T, KB, R, 2>(N);
  const int c1 = f16f_consumers_per_cu<T, KB, R, 1>(N);
  if (c4 >= c2 && c4 >= c1 && c4 > 0) return launch_gram_f16f_nc<T, KB, R, 4>(a, sxd, syd, batch, N, st);
  if (c2 >= c1 && c2 > 0) return launch_gram_f16f_nc<T, KB, R, 2>(a, sxd, syd, batch, N, st);
  return launch_gram_f16f_nc<T, KB, R, 1>(a, sxd, syd, batch, N, st);
}

// batch (Gram only, may be null): several Grams over the Bx x paths and By y paths (see
// k_linear_gram_batch); its units and prefixes are filled in here, once the stream length is set.
template <typename T, int KB, int PREC>
static hipError_t launch_fused(bool pairwise, const void *X, const void *Y, int64_t Bx, int64_t By,
                               int64_t Lx, int64_t Ly, int64_t d, double kscale, int order,
                               const void *sx, const void *sy, int symmetric, void *out,
                               int64_t ldo, const Workspace &w, const GramBatch *batch,
                               hipStream_t st) {
  const int dp = dpad_for(d);
  hipError_t e = run_split<T, PREC>((const T *)X, Bx, Lx, d, dp, w.xh, w.xm, w.xl, w.invx, st);
  if (e != hipSuccess) return e;
  if (w.yh != w.xh) {
    e = run_split<T, PREC>((const T *)Y, By, Ly, d, dp, w.yh, w.ym, w.yl, w.invy, st);
    if (e != hipSuccess) return e;
  }
  if ((e = run_to_double<T>(sx, Bx, w.sxd, st)) != hipSuccess) return e;
  if ((e = run_to_double<T>(sy, By, w.syd, st)) != hipSuccess) return e;
  GramFusedArgs<T> a{};
  a.xh = w.xh;
  a.xl = w.xl;
  a.yh = w.yh;
  a.yl = w.yl;
  a.xm = w.xm;
  a.ym = w.ym;
  a.invx = w.invx;
  a.invy = w.invy;
  a.Bx = Bx;
  a.By = By;
  a.m = (int)(Lx - 1);
  a.n = (int)(Ly - 1);
  a.order = order;
  a.dpad = dp;
  a.kmult = kscale / (double)(1ll << (2 * order));
  a.symmetric = symmetric;
  a.out = (T *)out;
  a.ldo = ldo;
  const double *sxd = sx ? w.sxd : nullptr;
  const double *syd = sy ? w.syd : nullptr;
  if (pairwise) {
    return launch_linear_pairwise_fused<T, KB, PREC>(a, sxd, syd, st);
  } else {
    // y paths streamed per wave: enough waves to fill the GPU, long enough streams to amortise
    // the 15-step pipeline fill of each pair
    const int64_t nq = (Bx + 3) / 4;
    auto units = [&](int ysv) -> int64_t {
      if (!batch) return nq * ((By + ysv - 1) / ysv);
      int64_t u = 0;
      for (int r = 0; r < batch->n; ++r)
        u += (int64_t)((batch->r[r].bx + 3) / 4) * ((batch->r[r].by + ysv - 1) / ysv);
      return u;
    };
    int ys = 16;
    while (ys > 1 && units(ys) < 4096) ys >>= 1;
    // a wave's virtual-column count (ys pairs x bands x band length) must stay within int
    {
      const PairGeom<16> g1(a.m, a.n, order, 1, r16_rows(order));
      while (ys > 1 && (int64_t)ys * g1.pair_columns() > (int64_t)1 << 30) ys >>= 1;
    }
    a.ys = ys;
    if constexpr (PREC == kPrecFp16x2F) {
      const int N = (int)((Ly - 1) << order);
      switch (r16_rows(order)) {  // dyadic order <= 2 (check_fused)
        case 1: return launch_gram_f16f<T, KB, 1>(a, sxd, syd, batch, N, st);
        case 2: return launch_gram_f16f<T, KB, 2>(a, sxd, syd, batch, N, st);
        default: return launch_gram_f16f<T, KB, 4>(a, sxd, syd, batch, N, st);
      }
    } else {
    const int64_t nyr = (By + ys - 1) / ys;
    // super-tiles of about one XCD's resident waves
    a.tx = (int)(nq < kTileX ? nq : kTileX);
    a.ty = (int)(nyr < kTileY ? nyr : kTileY);
    GramBatch gb{};
    int64_t nblk = (nq + a.tx - 1) / a.tx * a.tx * ((nyr + a.ty - 1) / a.ty * a.ty);
    if (batch) {
      gb = *batch;
      int64_t pre = 0;
      for (int r = 0; r < gb.n; ++r) {
        gb.r[r].prefix = (int32_t)pre;
        pre += (int64_t)((gb.r[r].bx + 3) / 4) * ((gb.r[r].by + ys - 1) / ys);
      }
      gb.nunits = (int32_t)pre;
      nblk = pre;
    }
    // a multiple of 8 blocks so that the XCD-major block order covers every work unit
    const int64_t grid = (nblk + 7) / 8 * 8;
    const dim3 g3((unsigned)grid), b3(64);
    constexpr Arith A = FusedArith<PREC>::A;
    const int N = (int)((Ly - 1) << order);
    auto go = [&](auto rc) -> hipError_t {
      constexpr int R = decltype(rc)::value;
      const bool gl = fused_lines_in_global<A>(
          reinterpret_cast<const void *>(&k_linear_gram_fused<T, KB, PREC, R, false>),
          reinterpret_cast<const void *>(&k_linear_gram_fused<T, KB, PREC, R, true>), N, grid, st);
      return with_fused_lines<A>(gl, N, grid, st, [&](void *lines, size_t shmem) {
        if (batch) {
          if (lines)
            k_linear_gram_batch<T, KB, PREC, R, true><<<g3, b3, shmem, st>>>(a, sxd, syd, lines,
                                                                               gb);
          else
            k_linear_gram_batch<T, KB, PREC, R, false><<<g3, b3, shmem, st>>>(a, sxd, syd,
                                                                                nullptr, gb);
        } else if (lines) {
          k_linear_gram_fused<T, KB, PREC, R, true><<<g3, b3, shmem, st>>>(a, sxd, syd, lines);
        } else {
          k_linear_gram_fused<T, KB, PREC, R, false><<<g3, b3, shmem, st>>>(a, sxd, syd, nullptr);
        }
      });
    };
    switch (r16_rows(order)) {
      case 1: return go(std::integral_constant<int, 1>());
      case 2: return go(std::integral_constant<int, 2>());
      case 4: return go(std::integral_constant<int, 4>());
      default: return go(std::integral_constant<int, 8>());
    }
    }
  }
}

template <typename T, int PREC>
static hipError_t dispatch_kb(bool pairwise, const void *X, const void *Y, int64_t Bx, int64_t By,
                              int64_t Lx, int64_t Ly, int64_t d, double kscale, int order,
                              const void *sx, const void *sy, int symmetric, void *out,
                              int64_t ldo, const Workspace &w, const GramBatch *batch,
                              hipStream_t st) {
  const int kb = dpad_for(d) / kKBlock;
#define SIGPDE_KB_CASE(K)                                                                          \
  case K:                                                                                          \
    return launch_fused<T, K, PREC>(pairwise, X, Y, Bx, By, Lx, Ly, d, kscale, order, sx, sy,       \
                                    symmetric, out, ldo, w, batch, st);
  if constexpr (PREC == kPrecFp64) {  // dim <= 64
    switch (kb) {
      SIGPDE_KB_CASE(1)
      SIGPDE_KB_CASE(2)
      default:
        return hipErrorInvalidValue;
    }
  } else {
    switch (kb) {
      SIGPDE_KB_CASE(1)
      SIGPDE_KB_CASE(2)
      SIGPDE_KB_CASE(3)
      SIGPDE_KB_CASE(4)
      default:
        return hipErrorInvalidValue;
    }
  }
#undef SIGPDE_KB_CASE
}

template <typename T>
static hipError_t dispatch_prec(int precision, bool pairwise, const void *X, const void *Y,
                                int64_t Bx, int64_t By, int64_t Lx, int64_t Ly, int64_t d,
                                double kscale, int order, const void *sx, const void *sy,
                                int symmetric, void *out, int64_t ldo, const Workspace &w,
                                const GramBatch *batch, hipStream_t st) {
  if (precision == SIGPDE_PREC_FP32)
    return dispatch_kb<T, kPrecFp32>(pairwise, X, Y, Bx, By, Lx, Ly, d, kscale, order, sx, sy,
                                     symmetric, out, ldo, w, batch, st);
  if (precision == SIGPDE_PREC_BF16X3)
    return dispatch_kb<T, kPrecBf16x3>(pairwise, X, Y, Bx, By, Lx, Ly, d, kscale, order, sx, sy,
                                       symmetric, out, ldo, w, batch, st);
  if (precision == SIGPDE_PREC_FP64)
    return dispatch_kb<T, kPrecFp64>(pairwise, X, Y, Bx, By, Lx, Ly, d, kscale, order, sx, sy,
                                     symmetric, out, ldo, w, batch, st);
  if (precision == SIGPDE_PREC_FP16X2F)
    return dispatch_kb<T, kPrecFp16x2F>(pairwise, X, Y, Bx, By, Lx, Ly, d, kscale, order, sx, sy,
                                        symmetric, out, ldo, w, batch, st);
  return dispatch_kb<T, kPrecFp16x2>(pairwise, X, Y, Bx, By, Lx, Ly, d, kscale, order, sx, sy,
                                     symmetric, out, ldo, w, batch, st);
}

constexpr int64_t kMaxDim = 4 * kKBlock;
constexpr int64_t kMaxDimFp64 = 2 * kKBlock;  // fp64: x operands of four pairs in registers

// Longest refined column count whose four boundary lines and ring fit one wave's LDS (float line
// and ring; double for fp64).
static int64_t fused_max_columns(int precision) {
  int64_t lo = 1, hi = 1 << 20;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    const size_t need = precision == SIGPDE_PREC_FP64 ? r16_lds_bytes<Arith::Ref64>((int)mid)
                                                      : r16_lds_bytes<Arith::Mixed>((int)mid);
    (need <= kMaxLdsBytes ? lo : hi) = mid;
  }
  return lo;
}

static bool known_precision(int precision) {
  return precision == SIGPDE_PREC_FP16X2 || precision == SIGPDE_PREC_FP32 ||
         precision == SIGPDE_PREC_BF16X3 || precision == SIGPDE_PREC_FP64 ||
         precision == SIGPDE_PREC_FP16X2F;
}

// The float32 wavefront (SIGPDE_PREC_FP16X2F) holds 1e-5 against float64 up to 512 refined rows
// and columns at dyadic order <= 2 (measured on the CPU: 4e-6 at 255 x 255 cells, 5.5e-6 at
// 396 x 396 at dyadic order 2, 8.6e-6 at 511 x 511; at order 3 the increments c/64 leave c^2/12
// near float resolution: 4.5e-5 at 1016 x 1016, 1e-4 on single entries at 256 x 256, measured on
// the GPU); other grids are refused.
constexpr int64_t kMaxRefinedF32Wave = 512;
constexpr int kMaxOrderF32Wave = 2;

static int check_fused(int dtype, int precision, const void *X, const void *Y, int64_t Bx,
                       int64_t By, int64_t Lx, int64_t Ly, int64_t d, int order, void *out,
                       void *ws) {
  if (dtype != SIGPDE_F32 && dtype != SIGPDE_F64)
    return set_error(SIGPDE_EINVAL, "linear: unknown dtype %d", dtype);
  if (!known_precision(precision))
    return set_error(SIGPDE_EINVAL, "linear: unsupported precision mode %d", precision);
  if (Bx < 0 || By < 0) return set_error(SIGPDE_EINVAL, "linear: negative batch");
  // an empty result reads and writes nothing: its pointers may be null (empty device tensors)
  if (Bx > 0 && By > 0 && (!X || !Y || !out || !ws))
    return set_error(SIGPDE_EINVAL, "linear: null pointer");
  if (Lx < 2 || Ly < 2) return set_error(SIGPDE_EINVAL, "linear: paths need length >= 2");
  if (d < 1 || d > kMaxDim) return set_error(SIGPDE_EINVAL, "linear: dim must be in [1, %d]", (int)kMaxDim);
  if (precision == SIGPDE_PREC_FP64 && d > kMaxDimFp64)
    return set_error(SIGPDE_EINVAL, "linear: precision fp64 takes dim <= %d", (int)kMaxDimFp64);
  if (order < 0 || order > 12) return set_error(SIGPDE_EINVAL, "linear: dyadic order must be in [0, 12]");
  if (precision == SIGPDE_PREC_FP16X2F && order > kMaxOrderF32Wave)
    return set_error(SIGPDE_EINVAL,
                     "linear: precision fp16x2f (float32 wavefront) takes dyadic order <= %d, got "
                     "%d; use fp16x2 (double wavefront)", kMaxOrderF32Wave, order);
  if (precision == SIGPDE_PREC_FP16X2F &&
      (((Lx - 1) << order) > kMaxRefinedF32Wave || ((Ly - 1) << order) > kMaxRefinedF32Wave))
    return set_error(SIGPDE_EINVAL,
                     "linear: precision fp16x2f (float32 wavefront) takes at most %lld refined "
                     "rows and columns, got %lld x %lld; use fp16x2 (double wavefront)",
                     (long long)kMaxRefinedF32Wave, (long long)((Lx - 1) << order),
                     (long long)((Ly - 1) << order));
  if (((Lx - 1) << order) > kMaxRefinedRows)
    return set_error(SIGPDE_EINVAL, "linear: refined rows (Lx-1) << order = %lld exceed %lld",
                     (long long)((Lx - 1) << order), (long long)kMaxRefinedRows);
  const int64_t maxc = fused_max_columns(precision);
  if (((Ly - 1) << order) > maxc)
    return set_error(SIGPDE_EINVAL,
                     "linear: refined columns (Ly-1) << order = %lld exceed %lld, the longest "
                     "boundary lines one wave's LDS holds",
                     (long long)((Ly - 1) << order), (long long)maxc);
  return SIGPDE_OK;
}

}  // namespace sigpde

using namespace sigpde;

extern "C" size_t sigpde_linear_workspace_bytes(int precision, int64_t Bx, int64_t Lx, int64_t By,
                                               int64_t Ly, int64_t d, int same_paths) {
  if (Bx < 0 || By < 0 || Lx < 2 || Ly < 2 || d < 1) return 0;
  if (!known_precision(precision)) return 0;
  return workspace_layout(precision, Bx, Lx, By, Ly, d, same_paths, nullptr, nullptr);
}

extern "C" int sigpde_linear_gram(int dtype, int precision, const void *X, int64_t Bx, int64_t Lx,
                                  const void *Y, int64_t By, int64_t Ly, int64_t d, double kscale,
                                  int order, const void *sx, const void *sy, int symmetric,
                                  void *out, int64_t ldo, void *workspace, size_t ws_bytes,
                                  void *stream) {
  clear_error();
  if (int rc = check_fused(dtype, precision, X, Y, Bx, By, Lx, Ly, d, order, out, workspace))
    return rc;
  if (Bx == 0 || By == 0) return SIGPDE_OK;
  if (symmetric && (X != Y || Bx != By || Lx != Ly))
    return set_error(SIGPDE_EINVAL, "linear_gram: symmetric needs Y == X");
  const int same = X == Y && Bx == By && Lx == Ly;
  Workspace w;
  const size_t need = workspace_layout(precision, Bx, Lx, By, Ly, d, same, (char *)workspace, &w);
  if (ws_bytes < need)
    return set_error(SIGPDE_EINVAL, "linear_gram: workspace of %zu bytes, need %zu", ws_bytes, need);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = dtype == SIGPDE_F32
                     ? dispatch_prec<float>(precision, false, X, Y, Bx, By, Lx, Ly, d, kscale, order, sx, sy,
                                          symmetric, out, ldo, w, nullptr, st)
                     : dispatch_prec<double>(precision, false, X, Y, Bx, By, Lx, Ly, d, kscale, order, sx, sy,
                                           symmetric, out, ldo, w, nullptr, st);
  return hip_status(e, "linear_gram launch");
}

extern "C" int sigpde_linear_gram_batch(int dtype, int precision, const void *X, int64_t nx,
                                        int64_t Lx, const void *Y, int64_t ny, int64_t Ly,
                                        int64_t d, double kscale, int order,
                                        const int64_t *requests, int64_t nreq, const void *sx,
                                        const void *sy, void *out, void *workspace,
                                        size_t ws_bytes, void *stream) {
  clear_error();
  if (nreq < 0 || nreq > kMaxGramBatch)
    return set_error(SIGPDE_EINVAL, "linear_gram_batch: %lld requests, at most %d per launch",
                     (long long)nreq, kMaxGramBatch);
  if (nreq > 0 && !requests) return set_error(SIGPDE_EINVAL, "linear_gram_batch: null request table");
  GramBatch gb{};
  int64_t out_elems = 0;
  for (int64_t r = 0; r < nreq; ++r) {
    const int64_t *q = requests + 5 * r;
    const int64_t x0 = q[0], bx = q[1], y0 = q[2], by = q[3], sym = q[4];
    if (bx < 0 || by < 0 || x0 < 0 || y0 < 0 || x0 + bx > nx || y0 + by > ny)
      return set_error(SIGPDE_EINVAL, "linear_gram_batch: request %lld out of the %lld x / %lld y "
                       "paths", (long long)r, (long long)nx, (long long)ny);
    if (sym && (bx != by || Lx != Ly))
      return set_error(SIGPDE_EINVAL, "linear_gram_batch: symmetric request %lld needs bx == by "
                       "and Lx == Ly", (long long)r);
    if (bx == 0 || by == 0) continue;  // empty result: no units
    GramBatchReq &g = gb.r[gb.n++];
    g.x0 = (int32_t)x0;
    g.bx = (int32_t)bx;
    g.y0 = (int32_t)y0;
    g.by = (int32_t)by;
    g.symmetric = sym ? 1 : 0;
    g.out_off = out_elems;
    out_elems += bx * by;
  }
  if (nx > INT32_MAX || ny > INT32_MAX)
    return set_error(SIGPDE_EINVAL, "linear_gram_batch: more than 2^31 paths");
  {
    int64_t units = 0;  // at the longest y stream (one y path a wave): an upper bound
    for (int r = 0; r < gb.n; ++r) units += (int64_t)((gb.r[r].bx + 3) / 4) * gb.r[r].by;
    if (units > INT32_MAX)
      return set_error(SIGPDE_EINVAL, "linear_gram_batch: more than 2^31 work units");
  }
  if (int rc = check_fused(dtype, precision, X, Y, nx, ny, Lx, Ly, d, order, out, workspace))
    return rc;
  if (gb.n == 0) return SIGPDE_OK;
  Workspace w;
  const size_t need = workspace_layout(precision, nx, Lx, ny, Ly, d, 0, (char *)workspace, &w);
  if (ws_bytes < need)
    return set_error(SIGPDE_EINVAL, "linear_gram_batch: workspace of %zu bytes, need %zu",
                     ws_bytes, need);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = dtype == SIGPDE_F32
                     ? dispatch_prec<float>(precision, false, X, Y, nx, ny, Lx, Ly, d, kscale,
                                            order, sx, sy, 0, out, 0, w, &gb, st)
                     : dispatch_prec<double>(precision, false, X, Y, nx, ny, Lx, Ly, d, kscale,
                                             order, sx, sy, 0, out, 0, w, &gb, st);
  return hip_status(e, "linear_gram_batch launch");
}

extern "C" int sigpde_linear_pairwise(int dtype, int precision, const void *X, const void *Y,
                                      int64_t B, int64_t Lx, int64_t Ly, int64_t d, double kscale,
                                      int order, const void *sx, const void *sy, void *out,
                                      void *workspace, size_t ws_bytes, void *stream) {
  clear_error();
  if (int rc = check_fused(dtype, precision, X, Y, B, B, Lx, Ly, d, order, out, workspace))
    return rc;
  if (B == 0) return SIGPDE_OK;
  const int same = X == Y && Lx == Ly;
  Workspace w;
  const size_t need = workspace_layout(precision, B, Lx, B, Ly, d, same, (char *)workspace, &w);
  if (ws_bytes < need)
    return set_error(SIGPDE_EINVAL, "linear_pairwise: workspace of %zu bytes, need %zu", ws_bytes, need);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = dtype == SIGPDE_F32
                     ? dispatch_prec<float>(precision, true, X, Y, B, B, Lx, Ly, d, kscale, order, sx, sy, 0,
                                          out, 1, w, nullptr, st)
                     : dispatch_prec<double>(precision, true, X, Y, B, B, Lx, Ly, d, kscale, order, sx, sy, 0,
                                           out, 1, w, nullptr, st);
  return hip_status(e, "linear_pairwise launch");
}
