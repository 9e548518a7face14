// Fused reference-scheme Gram for the static kernels (linear, RBF, L1-RBF): the point values and
// their second difference are formed inside the solve by PointProducer (point_producer.h), the
// wavefront is the reference's (Arith::Ref32 for float32 paths, Ref64 for float64), one wave per
// four x paths and a stream of y paths, as the fused linear Gram (fused_linear.hip). Replaces
// SigPDE._gram / _gram_sym's tile loop (sigpde/torch/kernel.py:74-171: gram_inner_product then
// the Gram solvers of sigpde/cuda_gram_kernels.py) without any increment tensor.
#pragma once

#include "point_producer.h"

namespace sigpde {

template <typename T>
struct PointGramArgs {
  PointOperands<T> op;
  const double *sx, *sy;  // optional path scales (double copies), may be null
  int64_t Bx, By;
  int Lx, Ly, order, dpad;
  int ys, tx, ty;         // y paths per wave stream; super-tile of x-quads by y-ranges
  T prm, inv4;
  int symmetric;
  T *out;
  int64_t ldo;
};

template <typename T>
struct PointArith {
  static constexpr Arith A = std::is_same<T, double>::value ? Arith::Ref64 : Arith::Ref32;
};

// LDS of one wave: the wavefront's lines and ring, the producer's line of point values (four
// pairs x Np), and for L1-RBF the band's staged x rows (four pairs x 16 rows x dpad).
template <typename T>
__host__ __device__ inline size_t point_lds_bytes(int kind, int N, int dpad) {
  const int np = Layout<16>::band_len(N);
  return r16_lds_bytes<PointArith<T>::A>(N) + (size_t)4 * np * sizeof(T) +
         (kind == kPointL1 ? (size_t)64 * dpad * sizeof(T) : 0);
}

// One wave: x paths 4 xq .. 4 xq + 3 against y paths yr ys .. (symmetric: only y <= x, each value
// also written to its mirror).
template <typename T, int KIND, int NE, int S>
__global__ __launch_bounds__(64) void k_point_gram(PointGramArgs<T> a) {
  constexpr Arith A = PointArith<T>::A;
  using Prod = PointProducer<T, KIND, NE, 1, S>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & 63;
  const int64_t nq = (a.Bx + 3) / 4;
  const int64_t nyr = (a.By + a.ys - 1) / a.ys;
  // XCD-aware order as in k_linear_gram_fused: blocks dealt round-robin to the 8 XCDs, each XCD a
  // contiguous run of units in super-tiles of tx x-quads by ty y-ranges
  const int64_t bid = blockIdx.x;
  const int64_t per_xcd = gridDim.x / 8;
  const int64_t lin = (bid % 8) * per_xcd + bid / 8;
  const int64_t tsz = (int64_t)a.tx * a.ty;
  const int64_t tile = lin / tsz, w = lin - tile * tsz;
  const int64_t ntx = (nq + a.tx - 1) / a.tx;
  const int64_t xq = (tile % ntx) * a.tx + w % a.tx;
  const int64_t yr = (tile / ntx) * a.ty + w / a.tx;
  if (xq >= nq || yr >= nyr) return;
  const int64_t y0 = yr * a.ys;
  int ys = (int)((a.By - y0) < a.ys ? (a.By - y0) : a.ys);
  if (a.symmetric) {
    const int64_t xmax = xq * 4 + 3 < a.Bx - 1 ? xq * 4 + 3 : a.Bx - 1;
    if (y0 > xmax) return;
    if (y0 + ys - 1 > xmax) ys = (int)(xmax - y0 + 1);
  }
  PairGeom<16> geo(a.Lx - 1, a.Ly - 1, a.order, ys, S);
  Prod P;
  P.op = a.op;
  P.sx = a.sx;
  P.sy = a.sy;
  P.scaled = a.sx != nullptr || a.sy != nullptr;
  P.dpad = a.dpad;
  P.Lx = a.Lx;
  P.Ly = a.Ly;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int64_t x = xq * 4 + g;
    P.xa[g] = x < a.Bx ? x : -1;
  }
  {
    const int64_t xl = xq * 4 + lane / 16;
    P.xlane = xl < a.Bx ? xl : 0;
  }
  P.ybase[0] = y0;
  P.ystep = 1;
  P.ny = a.By;
  P.prm = a.prm;
  P.inv4 = a.inv4;
  P.pline = reinterpret_cast<T *>(smem_raw + r16_lds_bytes<A>(geo.N));
  P.pls = geo.Np;
  P.xst = P.pline + 4 * geo.Np;
  const int64_t xme = xq * 4 + lane / 16;
  ring16_solve<A, S, false>(P, geo, 1.0, smem_raw, [&](int sp, double r) {
    const int64_t y = y0 + sp;
    if (xme >= a.Bx || y >= a.By) return;
    if (a.symmetric) {
      if (y > xme) return;
      a.out[xme * a.ldo + y] = (T)r;
      a.out[y * a.ldo + xme] = (T)r;
    } else {
      a.out[xme * a.ldo + y] = (T)r;
    }
  });
}

// Launch over the whole Gram (a.ys, a.tx, a.ty set here); defined per dtype and kind in
// fused_point_*.hip.
template <typename T, int KIND, int NE>
hipError_t launch_point_gram(PointGramArgs<T> a, hipStream_t st);

template <typename T, int KIND, int NE>
hipError_t launch_point_gram_impl(PointGramArgs<T> a, hipStream_t st) {
  const int64_t nq = (a.Bx + 3) / 4;
  int ys = 16;
  while (ys > 1 && nq * ((a.By + ys - 1) / ys) < 4096) ys >>= 1;
  const int S = a.order >= 3 ? 8 : 1 << a.order;
  {
    const PairGeom<16> g1(a.Lx - 1, a.Ly - 1, a.order, 1, S);
    while (ys > 1 && (int64_t)ys * g1.pair_columns() > kMaxVirtualColumns) ys >>= 1;
  }
  a.ys = ys;
  const int64_t nyr = (a.By + ys - 1) / ys;
  a.tx = (int)(nq < 8 ? nq : 8);
  a.ty = (int)(nyr < 32 ? nyr : 32);
  const int64_t nblk = (nq + a.tx - 1) / a.tx * a.tx * ((nyr + a.ty - 1) / a.ty * a.ty);
  const int64_t grid = (nblk + 7) / 8 * 8;
  const size_t lds = point_lds_bytes<T>(KIND, (a.Ly - 1) << a.order, a.dpad);
  const dim3 g3((unsigned)grid), b3(64);
  switch (S) {
    case 1: k_point_gram<T, KIND, NE, 1><<<g3, b3, lds, st>>>(a); break;
    case 2: k_point_gram<T, KIND, NE, 2><<<g3, b3, lds, st>>>(a); break;
    case 4: k_point_gram<T, KIND, NE, 4><<<g3, b3, lds, st>>>(a); break;
    default: k_point_gram<T, KIND, NE, 8><<<g3, b3, lds, st>>>(a); break;
  }
  return hipGetLastError();
}

}  // namespace sigpde
