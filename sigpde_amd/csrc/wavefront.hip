// Generic wavefront kernels: PDE solves from increments already in HBM (any static kernel),
// pairwise / Gram / symmetric Gram, scaled and unscaled, float32-reference and float64 storage,
// plus the fused robust-normalisation root-find. C ABI in include/sigpde_hip.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "capi_util.h"
#include <mutex>
#include <type_traits>

#include "mw_solver.h"
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

// GL: boundary lines in global memory (16-lane groups only), `lines` = this wave's block of
// r16_line_bytes<A>(N) bytes.
template <Arith A, int G, int R, bool GL, class Producer, class Out>
__device__ __forceinline__ void solve_wave(Producer &P, const PairGeom<G> &g, double s, void *lds,
                                           void *lines, Out &&out) {
  static_assert(!GL || G == 16, "global boundary lines need 16-lane groups");
  if constexpr (G == 16)
    ring16_solve<A, R, GL>(P, g, s, lds, out, reinterpret_cast<typename R16Bnd<A>::T *>(lines));
  else
    ring_solve<A, G>(P, g, s, lds, out);
}

// This wave's block of global boundary lines (GL launches: `gl` holds one block per wave of the
// launch; other launches pass null).
template <Arith A, bool GL>
__device__ __forceinline__ void *wave_lines(void *gl, int N) {
  if constexpr (GL)
    return static_cast<char *>(gl) + (size_t)blockIdx.x * r16_line_bytes<A>(N);
  else
    return nullptr;
}

// One wave per block; a wave solves Layout<G>::kPairs pairs (one per lane group of G lanes).
// Increments stream from HBM through the wave's LDS ring.
// R: refined rows per lane of the 16-lane solver (1 for 64-lane groups).
// w0: index of the launch's first wave (launches of global-line chunks), gl: global lines (GL).
template <typename T, int G, int R, bool GL>
__global__ __launch_bounds__(kWave) void k_pde_pairwise(PairwiseArgs<T> a, int64_t w0, void *gl) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using L = Layout<G>;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wv = w0 + blockIdx.x;
  PairGeom<G> g(a.m, a.n, a.order, 1, R);
  ProducerFor<T, G> P;
  P.ld = a.n;
  P.scaled = a.sx || a.sy;
#pragma unroll
  for (int q = 0; q < L::kPairs; ++q) {
    const int64_t b = wv * L::kPairs + q;
    const bool ok = b < a.B;
    P.inc[q] = ok ? a.inc + b * a.stride_b : nullptr;
    P.scale[q] = (P.scaled && ok) ? (a.sy ? a.sy[b] : (T)1) * (a.sx ? a.sx[b] : (T)1) : (T)1;
  }
  const int64_t b = wv * L::kPairs + lane / G;
  constexpr Arith A = RefArith<T>::kPlain;
  solve_wave<A, G, R, GL>(P, g, 1.0, smem_raw, wave_lines<A, GL>(gl, g.N), [&](int, double r) {
    if (b < a.B) a.out[b] = (T)r;
  });
}

// One pair per workgroup of W waves (mw_solver.h), for few long pairs; R refined rows per lane.
template <typename T, int W, int R>
__global__ __launch_bounds__(64 * W) void k_pde_pairwise_mw(PairwiseArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t b = blockIdx.x;
  const MwGeom g(a.m, a.n, a.order, W, R);
  const bool scaled = a.sx || a.sy;
  const T scale = scaled ? (a.sy ? a.sy[b] : (T)1) * (a.sx ? a.sx[b] : (T)1) : (T)1;
  const double r = mw_solve<RefArith<T>::kPlain, W, R>(a.inc + b * a.stride_b, a.n, scale,
                                                       scaled, g, 1.0, smem_raw);
  if (threadIdx.x == 0) a.out[b] = (T)r;
}

// Pairs spread over G workgroups each (mw_solve XW). Per pair, ws holds a control block of
// kXwgCtrlBytes (G progress counters from offset 0, the root-find's result sequence at
// kXwgSeqOff and result at kXwgResOff) and then the pair's windows; all zero at launch.
constexpr size_t kXwgCtrlBytes = 256;
constexpr size_t kXwgSeqOff = 128, kXwgResOff = 136;
constexpr int kXwgMaxGroups = 32;
constexpr int kXwgDmin = 7;  // band offset of a spread pair: two periods of slack on the reads

struct XwgLaunch {
  char *ws;
  int *err;          // the launch's failure flag
  size_t pairBytes;  // control block + windows of one pair
  long long limit;   // wall-clock ticks one wait may take
  int G, wing, hook;
};

__device__ inline Xwg xwg_state(const XwgLaunch &x, int64_t pair, int gi) {
  char *blk = x.ws + pair * x.pairBytes;
  Xwg s;
  s.win = blk + kXwgCtrlBytes;
  s.done = reinterpret_cast<int *>(blk);
  s.err = x.err;
  s.limit = x.limit;
  s.G = x.G;
  s.gi = gi;
  s.wing = x.wing;
  s.hook = x.hook;
  s.base = 0;
  s.seenIn = s.seenOut = 0;
  s.dead = false;
  return s;
}

// One pair per G workgroups of W waves: the pair's bands over all G W waves (few long pairs on a
// GPU with more CUs than pairs). Workgroup gi of pair b is block b G + gi.
template <typename T, int W, int R>
__global__ __launch_bounds__(64 * W) void k_pde_pairwise_xw(PairwiseArgs<T> a, XwgLaunch x) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t b = blockIdx.x / x.G;
  const int gi = (int)(blockIdx.x - b * x.G);
  const MwGeom g(a.m, a.n, a.order, W * x.G, R, kXwgDmin);
  const bool scaled = a.sx || a.sy;
  const T scale = scaled ? (a.sy ? a.sy[b] : (T)1) * (a.sx ? a.sx[b] : (T)1) : (T)1;
  Xwg xs = xwg_state(x, b, gi);
  const double r = mw_solve<RefArith<T>::kPlain, W, R, true>(a.inc + b * a.stride_b, a.n, scale,
                                                             scaled, g, 1.0, smem_raw, &xs);
  if (gi == mw_corner_group(g, W, x.G) && threadIdx.x == 0) a.out[b] = (T)r;
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

template <typename T, int G, int R, bool GL>
__global__ __launch_bounds__(kWave) void k_pde_gram(GramArgs<T> a, int64_t npairs, int tri,
                                                    int64_t w0, void *gl) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using L = Layout<G>;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wv = w0 + blockIdx.x;
  PairGeom<G> g(a.m, a.n, a.order, 1, R);
  ProducerFor<T, G> P;
  P.ld = a.n;
  P.scaled = a.sx || a.sy;
  int64_t gi[L::kPairs], gj[L::kPairs];
#pragma unroll
  for (int q = 0; q < L::kPairs; ++q) {
    const int64_t p = wv * L::kPairs + q;
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
  constexpr Arith A = RefArith<T>::kPlain;
  solve_wave<A, G, R, GL>(P, g, 1.0, smem_raw, wave_lines<A, GL>(gl, g.N), [&](int, double r) {
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
  double np0, np1;  // built-in normaliser parameters
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

// Target of a built-in normaliser for K(x,x) = x (NaN already mapped to +inf), in T:
//   SIGPDE_NORM_LOG:    2 - 1/(1 + log x), the reference's default (sigpde/torch/kernel.py:23-24);
//   SIGPDE_NORM_TENSOR: x for x <= C, else C + C^(1+a) (C^-a - x^-a) / a (C = p0, a = p1 > 0):
//                       increasing, at most x, bounded by C (1 + 1/a);
//   SIGPDE_NORM_CLIP:   min(x, C) (C = p0).
template <typename T>
__device__ __forceinline__ T builtin_target(int id, T x, double p0, double p1) {
  if (id == SIGPDE_NORM_TENSOR) {
    const T c = (T)p0, a = (T)p1;
    if (x <= c) return x;
    return c + pow(c, (T)1 + a) * (pow(c, -a) - pow(x, -a)) / a;
  }
  if (id == SIGPDE_NORM_CLIP) {
    const T c = (T)p0;
    return x < c ? x : c;
  }
  return log_normalizer<T>(x);
}

// Chandrupatla's bracketing iteration for one path (sigpde/cuda_pairwise_kernels.py:93-166):
// F(x) = PDE(x^2 inc) - target on [0, 1], F(0) = 1 - target, F(1) = K(x,x) - target, the
// bookkeeping and the interpolation test operation for operation, in double; function values
// rounded through the path dtype T as the reference's scratch stores them. Stops when the smaller
// bracket end's |F| < tol (function value only) or after maxit evaluations. The bracket
// arithmetic is not contracted into FMAs (every product and quotient rounded as written), so that
// with the same function values the iterates, and lambda, equal the oracle's bit for bit.
// It also stops at a fixed point: when an iteration leaves the whole state (a, b, c, fa, fb, fc,
// t) unchanged, every later iteration repeats it (the wavefront is deterministic), so the result
// is the one maxit evaluations would give, bit for bit. Float32 paths whose function values
// cannot reach tol (most of them at tol 1e-8: F is quantised to float32) reach such a point after
// about 40-50 evaluations instead of running all maxit.
template <typename T>
struct RootFind {
  double av, cv, bv, fa, fb, fc, t, outv;
  int it;
  T Tt;
  __device__ void init(T K1, T target) {
    Tt = target;
    av = 0.0;
    cv = 0.0;
    bv = 1.0;
    fa = 1.0 - (double)Tt;
    fb = (double)(T)(K1 - Tt);
    fc = fa;
    t = 0.5;
    outv = 0.0;
    it = 0;
  }
  __device__ double next() const {
#pragma clang fp contract(off)
    return av + t * (bv - av);
  }
  // Takes F's evaluation at xt = next() (the wavefront result sol); true when the search stops.
  __device__ bool update(double xt, double sol, double tol, int maxit) {
#pragma clang fp contract(off)
    const double a0 = av, b0 = bv, c0 = cv, fa0 = fa, fb0 = fb, fc0 = fc, t0 = t;
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
      stop = fabs(fa) < tol;
    } else {
      outv = bv;
      stop = fabs(fb) < tol;
    }
    double xi = (av - bv) / (cv - bv);
    double phi = (fa - fb) / (fc - fb);
    if (phi * phi < xi && (1 - phi) * (1 - phi) < 1 - xi)
      t = fa / (fb - fa) * fc / (fb - fc) + (cv - av) / (bv - av) * fa / (fc - fa) * fb / (fc - fb);
    else
      t = 0.5;
    if (!(0 < t && t < 1)) t = 0.5;
    const bool fixed = av == a0 && bv == b0 && cv == c0 && fa == fa0 && fb == fb0 && fc == fc0 &&
                       t == t0;
    return stop || fixed || it >= maxit;
  }
};

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

template <typename T, int G, int R, bool GL>
__global__ __launch_bounds__(kWave) void k_robust_norm(RobustArgs<T> a, int64_t w0, void *gl) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using L = Layout<G>;
  const int lane = threadIdx.x & (kWave - 1);
  const int q = lane / G;
  const int64_t wv = w0 + blockIdx.x;
  const int64_t b = wv * L::kPairs + q;
  const bool live = b < a.B;
  PairGeom<G> g(a.m, a.m, a.order, 1, R);
  const int rM = G == 16 ? ((g.M - 1) % (16 * R)) / R : (g.M - 1) & (G - 1);
  ProducerFor<T, G> P;
  P.ld = a.m;
  P.scaled = false;
#pragma unroll
  for (int w = 0; w < L::kPairs; ++w) {
    const int64_t bw = wv * L::kPairs + w;
    P.inc[w] = bw < a.B ? a.inc + bw * a.stride_b : nullptr;
    P.scale[w] = (T)1;
  }

  T K1 = (T)1, Tt = (T)1;
  if (a.normalizer != SIGPDE_NORM_GIVEN) {
    // _norm_factors (sigpde/torch/kernel.py:181-184): K(x,x), NaN -> inf, target
    double r = 0.0;
    solve_wave<RefArith<T>::kPlain, G, R, GL>(
        P, g, 1.0, smem_raw, wave_lines<RefArith<T>::kPlain, GL>(gl, g.N), [&](int, double v) { r = v; });
    K1 = (T)group_bcast<G>(r, rM);
    if (isnan(K1)) K1 = (T)INFINITY;
    Tt = builtin_target<T>(a.normalizer, K1, a.np0, a.np1);
  } else if (live) {
    K1 = a.knorm[b];
    Tt = a.target[b];
  }

  RootFind<T> rf;
  rf.init(K1, Tt);
  bool active = live && a.maxit > 0;
  while (__any(active)) {
    const double xt = rf.next();
    double r = 0.0;
    solve_wave<RefArith<T>::kScaled, G, R, GL>(
        P, g, xt * xt, smem_raw, wave_lines<RefArith<T>::kScaled, GL>(gl, g.N), [&](int, double v) { r = v; });
    const double sol = group_bcast<G>(r, rM);
    if (!active) continue;
    if (rf.update(xt, sol, a.tol, a.maxit)) active = false;
  }
  if (live && (lane & (G - 1)) == 0) {
    // scales[norms < 1] = 0 (sigpde/torch/kernel.py:195)
    a.lambda[b] = (K1 < (T)1) ? (T)0 : (T)rf.outv;
    if (a.normalizer != SIGPDE_NORM_GIVEN) a.knorm[b] = K1;
    if (a.iters) a.iters[b] = rf.it;
  }
}

// The same root-find with one path per workgroup of W waves (mw_solver.h: the bands of each
// wavefront evaluation run concurrently), for few long paths; R refined rows per lane. The
// float32 kernel of 8 waves and two rows per lane is held to four waves per SIMD (two workgroups
// per CU: up to 512 paths at once; the compiler's own choice, 130 registers, held one).
template <typename T, int W, int R>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(
    sizeof(T) == 4 && W == 8 && R == 2 ? 4 : 1))) void k_robust_norm_mw(RobustArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t b = blockIdx.x;
  const MwGeom g(a.m, a.m, a.order, W, R);
  const T *inc = a.inc + b * a.stride_b;
  T K1 = (T)1, Tt = (T)1;
  if (a.normalizer != SIGPDE_NORM_GIVEN) {
    K1 = (T)mw_solve<RefArith<T>::kPlain, W, R>(inc, a.m, (T)1, false, g, 1.0, smem_raw);
    if (isnan(K1)) K1 = (T)INFINITY;
    Tt = builtin_target<T>(a.normalizer, K1, a.np0, a.np1);
  } else {
    K1 = a.knorm[b];
    Tt = a.target[b];
  }
  RootFind<T> rf;
  rf.init(K1, Tt);
  bool active = a.maxit > 0;  // uniform over the workgroup: one path
  while (active) {
    const double xt = rf.next();
    const double sol = mw_solve<RefArith<T>::kScaled, W, R>(inc, a.m, (T)1, false, g, xt * xt,
                                                           smem_raw);
    if (rf.update(xt, sol, a.tol, a.maxit)) active = false;
  }
  if (threadIdx.x == 0) {
    a.lambda[b] = (K1 < (T)1) ? (T)0 : (T)rf.outv;
    if (a.normalizer != SIGPDE_NORM_GIVEN) a.knorm[b] = K1;
    if (a.iters) a.iters[b] = rf.it;
  }
}

// The root-find with each path spread over G workgroups of W waves. Every workgroup of a path
// runs the same root-find: the corner workgroup publishes each evaluation's K[M][N] (release) and
// the others take it from there (acquire, bounded wait), so all update the same state the same
// way and run the same evaluations.
template <typename T, int W, int R>
__global__ __launch_bounds__(64 * W) void k_robust_norm_xw(RobustArgs<T> a, XwgLaunch x) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double shared_result;
  const int64_t b = blockIdx.x / x.G;
  const int gi = (int)(blockIdx.x - b * x.G);
  const MwGeom g(a.m, a.m, a.order, W * x.G, R, kXwgDmin);
  const T *inc = a.inc + b * a.stride_b;
  Xwg xs = xwg_state(x, b, gi);
  char *blk = x.ws + b * x.pairBytes;
  int *seqp = reinterpret_cast<int *>(blk + kXwgSeqOff);
  double *resp = reinterpret_cast<double *>(blk + kXwgResOff);
  const bool corner = gi == mw_corner_group(g, W, x.G);
  int seq = 0, seen = 0;
  auto share = [&](double r) -> double {
    ++seq;
    if (corner) {
      if (threadIdx.x == 0) {
        *resp = r;
        __hip_atomic_store(seqp, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
      return r;
    }
    if (threadIdx.x < 64) {
      xwg_wait(seqp, seq, seen, xs);
      if (threadIdx.x == 0) shared_result = *resp;
    }
    __syncthreads();
    const double v = shared_result;
    __syncthreads();
    return v;
  };
  T K1 = (T)1, Tt = (T)1;
  if (a.normalizer != SIGPDE_NORM_GIVEN) {
    K1 = (T)share(
        mw_solve<RefArith<T>::kPlain, W, R, true>(inc, a.m, (T)1, false, g, 1.0, smem_raw, &xs));
    if (isnan(K1)) K1 = (T)INFINITY;
    Tt = builtin_target<T>(a.normalizer, K1, a.np0, a.np1);
  } else {
    K1 = a.knorm[b];
    Tt = a.target[b];
  }
  RootFind<T> rf;
  rf.init(K1, Tt);
  bool active = a.maxit > 0;  // uniform over the path's workgroups
  while (active) {
    const double xt = rf.next();
    const double sol = share(mw_solve<RefArith<T>::kScaled, W, R, true>(
        inc, a.m, (T)1, false, g, xt * xt, smem_raw, &xs));
    if (rf.update(xt, sol, a.tol, a.maxit)) active = false;
  }
  if (corner && threadIdx.x == 0) {
    a.lambda[b] = (K1 < (T)1) ? (T)0 : (T)rf.outv;
    if (a.normalizer != SIGPDE_NORM_GIVEN) a.knorm[b] = K1;
    if (a.iters) a.iters[b] = rf.it;
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

// Where the band-boundary lines live. Lane groups of 16 keep four lines per wave, groups of 64
// one: past the refined length whose four lines fill the LDS, batches run on groups of 64; past
// the length whose one line fills it, on groups of 16 with their lines in global memory.
template <typename T>
static bool g16_fits(int N) { return lds_for<T, 16>(N) <= kMaxLdsBytes; }
template <typename T>
static bool g64_fits(int N) { return lds_for<T, 64>(N) <= kMaxLdsBytes; }

// Global boundary lines: at most this much per launch (a launch covers that many waves; longer
// batches run as several launches on the stream, reusing the block).
constexpr size_t kGlobalLineBytes = (size_t)4 << 30;

// Launches a 16-lane kernel: launch(integral_constant R, bool_constant GL, w0, waves, lines) with
// R from the dyadic order. global: lines in global memory, allocated stream-ordered for the call,
// in launches of as many waves as kGlobalLineBytes holds (at least one: the block never exceeds
// that cap by more than one wave's lines: r16_line_bytes is at most 134 MB, double lines at 2^22
// refined columns).
template <Arith A, class F>
static hipError_t launch16(int order, int N, int64_t waves, bool global, hipStream_t st,
                           F &&launch) {
  auto by_rows = [&](auto glc, int64_t w0, int64_t nw, void *gl) {
    switch (r16_rows(order)) {
      case 1: launch(std::integral_constant<int, 1>(), glc, w0, nw, gl); break;
      case 2: launch(std::integral_constant<int, 2>(), glc, w0, nw, gl); break;
      case 4: launch(std::integral_constant<int, 4>(), glc, w0, nw, gl); break;
      default: launch(std::integral_constant<int, 8>(), glc, w0, nw, gl); break;
    }
  };
  if (!global) {
    by_rows(std::false_type(), 0, waves, nullptr);
    return hipGetLastError();
  }
  const size_t per = r16_line_bytes<A>(N);
  int64_t chunk = (int64_t)(kGlobalLineBytes / per);
  if (chunk < 1) chunk = 1;
  if (chunk > waves) chunk = waves;
  void *gl = nullptr;
  hipError_t e = hipMallocAsync(&gl, (size_t)chunk * per, st);
  if (e != hipSuccess) return e;
  for (int64_t w0 = 0; w0 < waves && e == hipSuccess; w0 += chunk) {
    by_rows(std::true_type(), w0, waves - w0 < chunk ? waves - w0 : chunk, gl);
    e = hipGetLastError();
  }
  const hipError_t ef = hipFreeAsync(gl, st);
  return e != hipSuccess ? e : ef;
}

// Workgroups of `kernel` (64 W threads, lds bytes of LDS) the current GPU holds at once.
static int64_t resident_groups(const void *kernel, int W, size_t lds) {
  return (int64_t)cu_count() * blocks_per_cu(kernel, 64 * W, lds);
}

// Pairs of several bands of 64 refined rows run one pair per workgroup of W waves (mw_solver.h)
// when the GPU holds all of their workgroups at once and the windows fit LDS: W = 16 for at least
// 8 bands, else W = 8 for at least 4, and bands longer than the D periods between band starts.
// Pipelining the bands over the waves shortens each pair's sequential chain; with more pairs than
// resident workgroups the one-wave routes, which fill the SIMDs with pairs instead, are faster. kern16 / kern8: the kernels of the two widths (kern16 null
// for float64, whose registers would spill at 16 waves per CU). Sets W and returns the
// workgroup's LDS bytes, or returns 0 when the pairs take the other routes.
template <typename T>
static size_t mw_route(int64_t pairs, int m, int n, int order, int R, const void *kern16,
                       const void *kern8, int &W) {
  for (const int w : {16, 8}) {
    const void *kern = w == 16 ? kern16 : kern8;
    if (!kern) continue;
    const MwGeom g(m, n, order, w, R);
    if (g.nb < w / 2 || g.D >= g.P) continue;  // D >= P: bands would not overlap
    const size_t lds = mw_lds_bytes<RefArith<T>::kPlain>(g, w);
    if (lds > kMaxLdsBytes || pairs > resident_groups(kern, w, lds)) continue;
    W = w;
    return lds;
  }
  return 0;
}

// The multi-wave route with R refined rows per lane. From dyadic order 1, R = 2 is tried first:
// two rows share each row of increments, so the coefficients and the shift from the lane above
// serve two cells (root-finds 20-35% faster); its extra registers can hold fewer workgroups per
// CU, and pairs past that count try R = 1 before the one-wave routes. Launches and returns true
// when mw_route takes the pairs (e: the launch status).
template <typename T, int R>
static bool mw_pairwise(const PairwiseArgs<T> &a, hipStream_t st, hipError_t &e) {
  int W = 0;
  const void *k16 = nullptr;
  if constexpr (sizeof(T) == 4) k16 = reinterpret_cast<const void *>(&k_pde_pairwise_mw<T, 16, R>);
  const void *k8 = reinterpret_cast<const void *>(&k_pde_pairwise_mw<T, 8, R>);
  const size_t lds = mw_route<T>(a.B, a.m, a.n, a.order, R, k16, k8, W);
  if (!lds) return false;
  if constexpr (sizeof(T) == 4) {
    if (W == 16) {
      k_pde_pairwise_mw<T, 16, R><<<dim3((unsigned)a.B), dim3(64 * 16), lds, st>>>(a);
      e = hipGetLastError();
      return true;
    }
  }
  k_pde_pairwise_mw<T, 8, R><<<dim3((unsigned)a.B), dim3(64 * 8), lds, st>>>(a);
  e = hipGetLastError();
  return true;
}

template <typename T, int R>
static bool mw_robust(const RobustArgs<T> &a, hipStream_t st, hipError_t &e) {
  int W = 0;
  const void *k16 = nullptr;
  if constexpr (sizeof(T) == 4) k16 = reinterpret_cast<const void *>(&k_robust_norm_mw<T, 16, R>);
  const void *k8 = reinterpret_cast<const void *>(&k_robust_norm_mw<T, 8, R>);
  const size_t lds = mw_route<T>(a.B, a.m, a.m, a.order, R, k16, k8, W);
  if (!lds) return false;
  if constexpr (sizeof(T) == 4) {
    if (W == 16) {
      k_robust_norm_mw<T, 16, R><<<dim3((unsigned)a.B), dim3(64 * 16), lds, st>>>(a);
      e = hipGetLastError();
      return true;
    }
  }
  k_robust_norm_mw<T, 8, R><<<dim3((unsigned)a.B), dim3(64 * 8), lds, st>>>(a);
  e = hipGetLastError();
  return true;
}

// ------------------------------------------------------------------ spread route
// Settings of the spread route (sigpde_set_spread): 0 = automatic, 1 = never, G >= 2 = G
// workgroups per pair whenever they are all resident; the wall-clock limit of one wait; the
// failure-path test hook. Spread launches run one at a time in the process: each needs all of its
// workgroups resident at once, which two such launches in flight could deny each other.
static std::mutex g_xwg_mutex;
static int g_xwg_groups = 0;
static double g_xwg_wait_s = 2.0;
static int g_xwg_hook = 0;

// Workgroups per pair for the spread route over W waves each (0: not taken). The wave chain of a
// band step is latency-bound, so the spread runs one wave per SIMD (W = 4, one workgroup per CU)
// and takes about ((P + 1) / kXwgDmin + 1) waves per pair plus one workgroup: enough that D
// falls to kXwgDmin (the pair's sequential periods, (nb - 1) D + P, at their least). All the
// launch's workgroups must be resident at once (the waits between them rely on it), at most one
// per CU, and the sequential periods must fall to at most 4/5 of the single-workgroup route's
// (whose workgroups hold up to 16 waves). Sets the LDS bytes of a workgroup.
template <typename T>
static int xwg_plan(int64_t pairs, int m, int n, int order, int R, int W, const void *kern,
                    hipStream_t st, size_t &lds_out) {
  constexpr Arith A = RefArith<T>::kPlain;
  if (g_xwg_groups == 1 || pairs < 1) return 0;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
    (void)hipGetLastError();  // the launch that follows reports its own status
    return 0;
  }
  if (cs != hipStreamCaptureStatusNone) return 0;
  const int forced = g_xwg_groups;
  int G = forced;
  const MwGeom g16(m, n, order, 16, R);
  if (forced < 2) {
    const int waves = (g16.P + 1 + kXwgDmin - 1) / kXwgDmin + 1;
    G = std::min((waves + W - 1) / W + 1, kXwgMaxGroups);
    G = (int)std::min<int64_t>(G, cu_count() / pairs);
    if (G < 2) return 0;
  }
  const MwGeom g(m, n, order, W * G, R, kXwgDmin);
  const size_t lds = mw_lds_bytes<A>(g, W);
  if (lds > kMaxLdsBytes) return 0;
  if (pairs * G > (int64_t)cu_count() * blocks_per_cu(kern, 64 * W, lds)) return 0;
  if (forced < 2 && (int64_t)g.T * 5 > (int64_t)g16.T * 4) return 0;
  lds_out = lds;
  return G;
}

// Runs launch(XwgLaunch) with the pairs' workspace (zeroed), waits for it, checks its failure flag
// and frees the workspace. Returns a SIGPDE status.
template <typename BT, class F>
static int xwg_run(int64_t pairs, int G, const MwGeom &g, hipStream_t st, const char *what,
                   F &&launch) {
  std::lock_guard<std::mutex> lock(g_xwg_mutex);
  XwgLaunch x{};
  x.G = G;
  x.wing = xwg_window_entries(g.D);
  x.hook = g_xwg_hook;
  x.pairBytes = kXwgCtrlBytes + ((size_t)G * 2 * x.wing * sizeof(BT) + 255) / 256 * 256;
  int dev = 0, khz = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  if (e != hipSuccess) return hip_status(e, what);
  if (khz <= 0) khz = 100000;  // the constant 100 MHz counter of CDNA GPUs
  x.limit = (long long)(g_xwg_wait_s * 1e3 * (double)khz);
  const size_t bytes = 256 + (size_t)pairs * x.pairBytes;
  void *ws = nullptr;
  e = hipMallocAsync(&ws, bytes, st);
  if (e != hipSuccess) return hip_status(e, what);
  x.err = static_cast<int *>(ws);
  x.ws = static_cast<char *>(ws) + 256;
  e = hipMemsetAsync(ws, 0, bytes, st);
  if (e == hipSuccess) {
    launch(x);
    e = hipGetLastError();
  }
  int failed = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&failed, ws, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  const hipError_t ef = hipFreeAsync(ws, st);
  if (e == hipSuccess) e = ef;
  if (e != hipSuccess) return hip_status(e, what);
  if (failed)
    return set_error(SIGPDE_ETIMEOUT,
                     "%s: a wait between the %d workgroups of a pair gave up after %g s (were "
                     "they all resident?); the outputs are invalid",
                     what, G, g_xwg_wait_s);
  return SIGPDE_OK;
}

// The spread route for pairwise solves (float32: 16 waves a workgroup, float64: 8); R as for the
// multi-wave route. Returns true when taken, with its status in rc.
constexpr int kXwgWaves = 4;  // waves per workgroup of the spread route

template <typename T, int R>
static bool xwg_pairwise(const PairwiseArgs<T> &a, hipStream_t st, int &rc) {
  constexpr int W = kXwgWaves;
  size_t lds = 0;
  const int G = xwg_plan<T>(a.B, a.m, a.n, a.order, R, W,
                            reinterpret_cast<const void *>(&k_pde_pairwise_xw<T, W, R>), st, lds);
  if (G < 2) return false;
  const MwGeom g(a.m, a.n, a.order, W * G, R, kXwgDmin);
  rc = xwg_run<typename R16Bnd<RefArith<T>::kPlain>::T>(
      a.B, G, g, st, "pde_pairwise (spread)", [&](const XwgLaunch &x) {
        k_pde_pairwise_xw<T, W, R><<<dim3((unsigned)(a.B * G)), dim3(64 * W), lds, st>>>(a, x);
      });
  return true;
}

template <typename T, int R>
static bool xwg_robust(const RobustArgs<T> &a, hipStream_t st, int &rc) {
  constexpr int W = kXwgWaves;
  size_t lds = 0;
  const int G = xwg_plan<T>(a.B, a.m, a.m, a.order, R, W,
                            reinterpret_cast<const void *>(&k_robust_norm_xw<T, W, R>), st, lds);
  if (G < 2) return false;
  const MwGeom g(a.m, a.m, a.order, W * G, R, kXwgDmin);
  rc = xwg_run<typename R16Bnd<RefArith<T>::kPlain>::T>(
      a.B, G, g, st, "robust_norm (spread)", [&](const XwgLaunch &x) {
        k_robust_norm_xw<T, W, R><<<dim3((unsigned)(a.B * G)), dim3(64 * W), lds, st>>>(a, x);
      });
  return true;
}

template <typename T>
static hipError_t launch_pairwise_local(const PairwiseArgs<T> &a, hipStream_t st) {
  constexpr Arith A = RefArith<T>::kPlain;
  hipError_t e = hipSuccess;
  if ((a.order >= 1 && mw_pairwise<T, 2>(a, st, e)) || mw_pairwise<T, 1>(a, st, e)) return e;
  const int N = a.n << a.order;
  const bool f16 = g16_fits<T>(N), f64 = g64_fits<T>(N);
  if ((a.B >= kManyPairs && f16) || !f64) {
    const size_t lds = f16 ? lds_for<T, 16>(N) : r16_ring_bytes<A>();
    return launch16<A>(a.order, N, (a.B + 3) / 4, !f16, st,
                       [&](auto rc, auto glc, int64_t w0, int64_t nw, void *gl) {
                         k_pde_pairwise<T, 16, decltype(rc)::value, decltype(glc)::value>
                             <<<dim3((unsigned)nw), dim3(kWave), lds, st>>>(a, w0, gl);
                       });
  }
  k_pde_pairwise<T, 64, 1, false><<<dim3((unsigned)a.B), dim3(kWave), lds_for<T, 64>(N), st>>>(
      a, 0, nullptr);
  return hipGetLastError();
}

// Pairs spread over workgroups when that shortens their sequential chain, else the routes of
// one workgroup per pair or per wave. Returns a SIGPDE status.
template <typename T>
static int launch_pairwise(const PairwiseArgs<T> &a, hipStream_t st) {
  int rc = SIGPDE_OK;
  if ((a.order >= 1 && xwg_pairwise<T, 2>(a, st, rc)) || xwg_pairwise<T, 1>(a, st, rc)) return rc;
  return hip_status(launch_pairwise_local(a, st), "pde_pairwise launch");
}

template <typename T>
static hipError_t launch_gram(const GramArgs<T> &a, hipStream_t st) {
  constexpr Arith A = RefArith<T>::kPlain;
  const int N = a.n << a.order;
  const bool tri = a.symmetric && a.off_x == a.off_y && a.bx == a.by;
  const int64_t np = tri ? a.bx * (a.bx + 1) / 2 : a.bx * a.by;
  const bool f16 = g16_fits<T>(N), f64 = g64_fits<T>(N);
  if ((np >= kManyPairs && f16) || !f64) {
    const size_t lds = f16 ? lds_for<T, 16>(N) : r16_ring_bytes<A>();
    return launch16<A>(a.order, N, (np + 3) / 4, !f16, st,
                       [&](auto rc, auto glc, int64_t w0, int64_t nw, void *gl) {
                         k_pde_gram<T, 16, decltype(rc)::value, decltype(glc)::value>
                             <<<dim3((unsigned)nw), dim3(kWave), lds, st>>>(a, np, (int)tri, w0, gl);
                       });
  }
  k_pde_gram<T, 64, 1, false><<<dim3((unsigned)np), dim3(kWave), lds_for<T, 64>(N), st>>>(
      a, np, (int)tri, 0, nullptr);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_robust_local(const RobustArgs<T> &a, hipStream_t st) {
  constexpr Arith A = RefArith<T>::kPlain;
  hipError_t e = hipSuccess;
  if ((a.order >= 1 && mw_robust<T, 2>(a, st, e)) || mw_robust<T, 1>(a, st, e)) return e;
  const int N = a.m << a.order;
  const bool f16 = g16_fits<T>(N), f64 = g64_fits<T>(N);
  if ((a.B >= kManyPairs && f16) || !f64) {
    const size_t lds = f16 ? lds_for<T, 16>(N) : r16_ring_bytes<A>();
    return launch16<A>(a.order, N, (a.B + 3) / 4, !f16, st,
                       [&](auto rc, auto glc, int64_t w0, int64_t nw, void *gl) {
                         k_robust_norm<T, 16, decltype(rc)::value, decltype(glc)::value>
                             <<<dim3((unsigned)nw), dim3(kWave), lds, st>>>(a, w0, gl);
                       });
  }
  k_robust_norm<T, 64, 1, false><<<dim3((unsigned)a.B), dim3(kWave), lds_for<T, 64>(N), st>>>(
      a, 0, nullptr);
  return hipGetLastError();
}

template <typename T>
static int launch_robust(const RobustArgs<T> &a, hipStream_t st) {
  int rc = SIGPDE_OK;
  if ((a.order >= 1 && xwg_robust<T, 2>(a, st, rc)) || xwg_robust<T, 1>(a, st, rc)) return rc;
  return hip_status(launch_robust_local(a, st), "robust_norm launch");
}

// m: increments along the rows (bands), n: along the columns (boundary lines).
static int check_grid(int dtype, int64_t m, int64_t n, int order) {
  if (m < 1 || n < 1) return set_error(SIGPDE_EINVAL, "paths need length >= 2 (m, n >= 1 increments)");
  if (order < 0 || order > 12) return set_error(SIGPDE_EINVAL, "dyadic order must be in [0, 12]");
  const int64_t M = m << order, N = n << order;
  if (M > kMaxRefinedRows)
    return set_error(SIGPDE_EINVAL, "refined rows (length-1) << order = %lld exceed %lld",
                     (long long)M, (long long)kMaxRefinedRows);
  if (N > kMaxRefinedCols)
    return set_error(SIGPDE_EINVAL, "refined columns (length-1) << order = %lld exceed %lld",
                     (long long)N, (long long)kMaxRefinedCols);
  // virtual columns of a pair on 16-lane groups (bands x band length) must stay within int
  if ((M / 16 + 1) * (N + 48) > kMaxVirtualColumns)
    return set_error(SIGPDE_EINVAL,
                     "refined grid %lld x %lld too large (bands x band length above %lld)",
                     (long long)M, (long long)N, (long long)kMaxVirtualColumns);
  (void)dtype;
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
  if (dtype == SIGPDE_F32) {
    PairwiseArgs<float> a{(const float *)inc, B, inc_stride_b, (int)m, (int)n, order,
                          (const float *)sx, (const float *)sy, (float *)out};
    return launch_pairwise(a, st);
  }
  if (dtype == SIGPDE_F64) {
    PairwiseArgs<double> a{(const double *)inc, B, inc_stride_b, (int)m, (int)n, order,
                           (const double *)sx, (const double *)sy, (double *)out};
    return launch_pairwise(a, st);
  }
  return set_error(SIGPDE_EINVAL, "pde_pairwise: unknown dtype %d", dtype);
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

extern "C" int sigpde_robust_norm_builtin(int dtype, const void *inc, int64_t B, int64_t m,
                                          int64_t inc_stride_b, int order, int normalizer,
                                          double p0, double p1, void *knorm, int maxit,
                                          double tol, void *lambda_out, int32_t *iters_out,
                                          void *stream) {
  clear_error();
  if (B == 0) return SIGPDE_OK;
  if (B < 0 || !inc || !lambda_out || !knorm)
    return set_error(SIGPDE_EINVAL, "robust_norm: bad batch or null pointer");
  if (normalizer != SIGPDE_NORM_LOG && normalizer != SIGPDE_NORM_TENSOR &&
      normalizer != SIGPDE_NORM_CLIP)
    return set_error(SIGPDE_EINVAL, "robust_norm_builtin: unknown built-in normalizer id %d",
                     normalizer);
  if (normalizer == SIGPDE_NORM_TENSOR && !(p0 >= 1.0 && p1 > 0.0 && p0 < INFINITY && p1 < INFINITY))
    return set_error(SIGPDE_EINVAL, "robust_norm_builtin: SIGPDE_NORM_TENSOR needs C >= 1 and "
                     "a > 0, got C %g, a %g", p0, p1);
  if (normalizer == SIGPDE_NORM_CLIP && !(p0 >= 1.0))
    return set_error(SIGPDE_EINVAL, "robust_norm_builtin: SIGPDE_NORM_CLIP needs C >= 1, got %g",
                     p0);
  if (maxit < 0) return set_error(SIGPDE_EINVAL, "robust_norm: maxit must be >= 0");
  if (int rc = check_grid(dtype, m, m, order)) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == SIGPDE_F32) {
    RobustArgs<float> a{(const float *)inc, B, inc_stride_b, (int)m, order, normalizer, maxit, tol,
                        nullptr, (float *)knorm, (float *)lambda_out, iters_out, p0, p1};
    return launch_robust(a, st);
  } else if (dtype == SIGPDE_F64) {
    RobustArgs<double> a{(const double *)inc, B, inc_stride_b, (int)m, order, normalizer, maxit,
                         tol, nullptr, (double *)knorm, (double *)lambda_out, iters_out, p0, p1};
    return launch_robust(a, st);
  }
  return set_error(SIGPDE_EINVAL, "robust_norm: unknown dtype %d", dtype);
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
    return set_error(SIGPDE_EINVAL, "robust_norm: unknown normalizer id %d (the other built-in "
                     "normalisers take parameters: sigpde_robust_norm_builtin)", normalizer);
  if (normalizer == SIGPDE_NORM_LOG)
    return sigpde_robust_norm_builtin(dtype, inc, B, m, inc_stride_b, order, normalizer, 0.0, 0.0,
                                      knorm, maxit, tol, lambda_out, iters_out, stream);
  if (!target) return set_error(SIGPDE_EINVAL, "robust_norm: SIGPDE_NORM_GIVEN needs target");
  if (maxit < 0) return set_error(SIGPDE_EINVAL, "robust_norm: maxit must be >= 0");
  if (int rc = check_grid(dtype, m, m, order)) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == SIGPDE_F32) {
    RobustArgs<float> a{(const float *)inc, B, inc_stride_b, (int)m, order, normalizer, maxit, tol,
                        (const float *)target, (float *)knorm, (float *)lambda_out, iters_out,
                        0.0, 0.0};
    return launch_robust(a, st);
  } else if (dtype == SIGPDE_F64) {
    RobustArgs<double> a{(const double *)inc, B, inc_stride_b, (int)m, order, normalizer, maxit,
                         tol, (const double *)target, (double *)knorm, (double *)lambda_out,
                         iters_out, 0.0, 0.0};
    return launch_robust(a, st);
  }
  return set_error(SIGPDE_EINVAL, "robust_norm: unknown dtype %d", dtype);
}

extern "C" int sigpde_set_spread(int groups, double wait_seconds, int test_hook) {
  clear_error();
  if (groups < 0 || groups > kXwgMaxGroups)
    return set_error(SIGPDE_EINVAL, "set_spread: groups must be 0 (automatic), 1 (never) or 2..%d",
                     kXwgMaxGroups);
  if (!(wait_seconds > 0.0 && wait_seconds <= 600.0))
    return set_error(SIGPDE_EINVAL, "set_spread: wait_seconds must be in (0, 600]");
  if (test_hook < 0 || test_hook > 1) return set_error(SIGPDE_EINVAL, "set_spread: test_hook is 0 or 1");
  std::lock_guard<std::mutex> lock(g_xwg_mutex);
  g_xwg_groups = groups;
  g_xwg_wait_s = wait_seconds;
  g_xwg_hook = test_hook;
  return SIGPDE_OK;
}
