// Increment producer of the fused reference-scheme kernels: the static kernel's point values
// k(x_i, y_j) are computed inside the solve (on MFMA for the linear and RBF kernels, on the VALU
// for L1-RBF) and turned into the 4-term second difference straight into the wavefront's LDS
// ring, so the increment tensor of sigpde/torch/kernel.py:18-21 never exists in memory.
//
// Point values, in the reference's float operations (sigpde/torch/static_kernels.py:7-54):
//   linear: k = scale * <x, y>
//   RBF:    k = exp(-((-2 <x, y>) + (|x|^2 + |y|^2)) * (1 / sigma))
//   L1-RBF: k = exp(-(sum_k |x_k - y_k|) * (1 / sigma))
// (division by the Python scalar sigma is a multiplication by its reciprocal in the tensor's
// dtype, as torch evaluates it). Increment (i, j) = (((k[i+1][j+1] + k[i][j]) - k[i+1][j]) -
// k[i][j+1]) * 4^-order, times (x_scale y_scale) when scaled (sigpde/cuda_gram_kernels.py:52,60).
// The dot products and distance sums run in another summation order than torch's GEMM and
// reductions: equal when they are exact (the quantized inputs of the parity tests), within the
// inputs' rounding otherwise, which is also how far the reference's own float32 result moves
// between summation orders.
//
// Geometry (16-lane groups, ring16.h; dyadic order o <= 3, so R = 2^o refined rows per lane and
// ring row r of band b is original increment row 16 b + r). Chunk k covers virtual columns
// v = jv0 .. jv0 + 15 of a band; column v uses increment column (v - 1) >> o, i.e. point columns
// left(v) = (v - 1) >> o and right(v) = left(v) + 1 (v = 0, the reset column, takes point
// column 0; columns past the grid take point column n).
//   * One 16 x 16 tile per pair and chunk: D[r][c] = point (16 b + r + 1, right(jv0 + c)). Lane
//     (c, h) holds rows 4h + e, e < 4: the values below its ring rows, B[e].
//   * The values on its ring rows, R[e] = point (16 b + 4h + e, right): B[e - 1] of the same
//     lane, for e = 0 B[3] of lane (c, h - 1) (ds_bpermute), for h = 0 the previous band's last
//     tile row, kept per pair in an LDS line indexed by virtual column; in band 0 it is point
//     row 0, a dot product (or distance) of x_0 with the lane's column on the VALU.
//   * Left columns: left(v) = right(v - 2^o), so the left values are the right values of lane
//     c - 2^o (DPP row_ror, the previous chunk's lanes for c < 2^o); the first chunk of a band
//     takes lane max(c - 2^o, 0) of its own tile (point column 0).
// Each point value is computed once per column repeat (2^o lanes hold the same column), always by
// the same operations, so every increment sees one value per point.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "mfma_producer.h"

namespace sigpde {

// Static kernels of the fused reference-scheme route (values of SIGPDE_STATIC_*).
constexpr int kPointLinear = 0;
constexpr int kPointRBF = 1;
constexpr int kPointL1 = 2;

// Cross-lane moves of float or double values.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  int2 a = *reinterpret_cast<int2 *>(&v);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp(0, a.x, CTRL, 0xf, 0xf, false);
  r.y = __builtin_amdgcn_update_dpp(0, a.y, CTRL, 0xf, 0xf, false);
  return *reinterpret_cast<double *>(&r);
}
// value of lane `src` (0..63)
__device__ __forceinline__ float lane_get(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ double lane_get(double v, int src) {
  int2 a = *reinterpret_cast<int2 *>(&v);
  int2 r;
  r.x = __builtin_amdgcn_ds_bpermute(src << 2, a.x);
  r.y = __builtin_amdgcn_ds_bpermute(src << 2, a.y);
  return *reinterpret_cast<double *>(&r);
}

// NE consecutive values from an address aligned to min(NE, 16 / sizeof(RT)) elements.
template <typename RT, int NE>
__device__ __forceinline__ void load_run(const RT *p, RT (&d)[NE]) {
  constexpr int W = 16 / sizeof(RT) < NE ? 16 / (int)sizeof(RT) : NE;  // elements per load
  typedef RT vec __attribute__((ext_vector_type(W)));
#pragma unroll
  for (int i = 0; i < NE / W; ++i) {
    const vec v = *reinterpret_cast<const vec *>(p + i * W);
#pragma unroll
    for (int j = 0; j < W; ++j) d[i * W + j] = v[j];
  }
}

// Operand rows of the producer: every point of every path padded to dpad elements (zero
// padding), path-major; RBF also |point|^2 per point.
template <typename RT>
struct PointOperands {
  const RT *xp, *yp;  // (path * L + i) * dpad
  const RT *xx, *yy;  // RBF: path * L + i
};

// KIND: kPointLinear / kPointRBF (MFMA tiles; NE = dpad / 4 elements of K per lane) or kPointL1
// (VALU distances; NE = dpad, the whole column per lane). NY = 1: a Gram wave, four x paths
// against one y path per stream position (ybase[0] + sp); NY = 4: four independent pairs.
// S = 2^order (the refined rows per lane and the column repeat).
template <typename RT, int KIND, int NE, int NY, int S>
struct PointProducer {
  static constexpr int G = 16;
  static constexpr int NP = 4;
  static constexpr bool kF64 = std::is_same<RT, double>::value;
  static constexpr bool kL1 = KIND == kPointL1;
  static constexpr bool kRBF = KIND == kPointRBF;
  static_assert(S == 1 || S == 2 || S == 4 || S == 8, "dyadic order <= 3");
  static_assert(NY == 1 || NY == 4, "operand sets");
  typedef typename std::conditional<kF64, double4v, float4v>::type acc_t;
  __device__ static constexpr int ys(int g) { return NY == 4 ? g : 0; }

  PointOperands<RT> op;
  const double *sx, *sy;  // optional path scales (double copies of the caller's), may be null
  int dpad, Lx, Ly;       // padded dimension, points per x / y path
  int64_t xa[NP];         // x path of each pair (-1: idle)
  int64_t xlane;          // L1: x path of the lane's own pair (lane >> 4), 0 when idle (a
                          // per-lane value: indexing xa by lane would put the producer in scratch)
  int64_t ybase[NY];      // y path of each y set at stream position 0
  int ystep;              // y path advance per stream position (Gram 1, pairwise 0)
  int64_t ny;             // valid y paths
  RT prm;                 // linear: scale; RBF, L1-RBF: 1 / sigma
  RT inv4;                // 4^-order
  RT *pline;              // LDS: per pair, the point row above the band, by virtual column
  int pls;                // its stride (elements per pair)
  RT *xst;                // L1: LDS staging of the band's x rows, [pair][16 rows][dpad]

  RT A[NP][kL1 ? 1 : NE];  // MFMA: the band's x-row slices (lane row c: point row 16b + c + 1)
  RT xn[NP][kRBF ? 4 : 1]; // RBF: |x|^2 of the lane's tile rows 16b + 4h + e + 1
  RT x0n[NP];              // RBF: |x_0|^2
  RT Bv[NY][NE];           // operands of the next chunk to multiply (y column slice / column)
  RT yn[NY];               // RBF: |y|^2 of that column
  RT nsum[NP][kRBF ? 4 : 1];  // RBF: |x|^2 + |y|^2 of the tile values of the chunk in acc (xn
                              // already holds the next band's norms when its last chunk is stored)
  RT nsum0[NP];               // RBF: |x_0|^2 + |y|^2, chunk in acc
  acc_t acc[NP];           // the chunk's tile: inner products (linear, RBF) or distances (L1)
  RT r0[NP];               // first band: point row 0's inner product / distance, chunk in acc
  RT prevB[NP][4], prevR0[NP];  // previous chunk's right values (for the left columns)
  RT scl[NP];              // x_scale y_scale of each pair (scaled only)
  bool scaled;
  bool accBand0, accStart;  // chunk in acc: first band of its pair / first chunk of its band
  int accJv;                // its first virtual column
  int64_t yrow[NY];         // first point row of the y path being fetched
  Cursor<G> cB, cC;         // next chunk to fetch operands for / to compute
  Ring16Writer<RT> W;

  __device__ __forceinline__ double lane_scale() const { return 1.0; }

  __device__ __forceinline__ int64_t ypath(int q, int sp) const {
    const int64_t yp = ybase[q] + (int64_t)sp * ystep;
    return yp < ny ? yp : ny - 1;
  }
  __device__ __forceinline__ int64_t xpath(int g) const { return xa[g] < 0 ? 0 : xa[g]; }

  // point value from an inner product (linear, RBF: ns = |x|^2 + |y|^2) or distance sum (L1-RBF)
  __device__ __forceinline__ RT point(RT v, RT ns) const {
#pragma clang fp contract(off)
    if constexpr (KIND == kPointLinear) {
      return prm * v;
    } else if constexpr (KIND == kPointRBF) {
      RT dist = (RT)-2 * v;
      dist += ns;
      return exp((-dist) * prm);
    } else {
      return exp((-v) * prm);
    }
  }

  // Rows of band `band`: MFMA x-row slices (and RBF norms), or the L1 staging in LDS.
  __device__ __forceinline__ void fetchA(int band) {
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, h = lane >> 4;
    if constexpr (kL1) {
      // lane l stages row 16 band + (l & 15) + 1 of pair l >> 4
      const int g = lane >> 4;
      int row = 16 * band + c + 1;
      row = row < Lx - 1 ? row : Lx - 1;
      RT v[NE];
      load_run<RT, NE>(op.xp + (xlane * Lx + row) * dpad, v);
      RT *dst = xst + (g * 16 + c) * dpad;
#pragma unroll
      for (int k = 0; k < NE; ++k) dst[k] = v[k];
    } else {
      // f64: D[h + 4e][c] lands in lane (c, h) element e, so lane row c carries tile row
      // 4(c & 3) + (c >> 2) (element e is then row 4h + e, as for f32)
      const int rl = kF64 ? 4 * (c & 3) + (c >> 2) : c;
      int row = 16 * band + rl + 1;
      row = row < Lx - 1 ? row : Lx - 1;
#pragma unroll
      for (int g = 0; g < NP; ++g)
        load_run<RT, NE>(op.xp + (xpath(g) * Lx + row) * dpad + h * NE, A[g]);
    }
    if constexpr (kRBF) {
#pragma unroll
      for (int g = 0; g < NP; ++g) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          int row = 16 * band + 4 * h + e + 1;
          row = row < Lx - 1 ? row : Lx - 1;
          xn[g][e] = op.xx[xpath(g) * Lx + row];
        }
        x0n[g] = op.xx[xpath(g) * Lx];
      }
    }
  }

  // operands of chunk cB, then advance cB
  __device__ __forceinline__ void fetchB(const PairGeom<G> &geo) {
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, h = lane >> 4;
    if (cB.band == 0 && cB.jv0 == 0) {
#pragma unroll
      for (int q = 0; q < NY; ++q) yrow[q] = ypath(q, cB.sp) * Ly;
    }
    int col = ((cB.jv0 + c - 1) >> geo.order) + 1;  // right(v); v = 0 gives 0
    col = col < Ly - 1 ? col : Ly - 1;
#pragma unroll
    for (int q = 0; q < NY; ++q) {
      const RT *p = op.yp + (yrow[q] + col) * dpad;
      if constexpr (kL1)
        load_run<RT, NE>(p, Bv[q]);
      else
        load_run<RT, NE>(p + h * NE, Bv[q]);
      if constexpr (kRBF) yn[q] = op.yy[yrow[q] + col];
    }
    cB.advance(geo);
  }

  // tiles of chunk cC (and, in a first band, point row 0), then advance cC; when the next chunk
  // opens a band, that band's rows are fetched now
  __device__ __forceinline__ void compute(const PairGeom<G> &geo, bool more) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int h = lane >> 4;
    accBand0 = cC.band == 0;
    accStart = cC.jv0 == 0;
    accJv = cC.jv0;
    if (scaled && accBand0 && accStart) {
#pragma unroll
      for (int g = 0; g < NP; ++g)
        scl[g] = (RT)(sx ? sx[xpath(g)] : 1.0) * (RT)(sy ? sy[ypath(ys(g), cC.sp)] : 1.0);
    }
    if constexpr (kRBF) {
#pragma unroll
      for (int g = 0; g < NP; ++g) {
#pragma unroll
        for (int e = 0; e < 4; ++e) nsum[g][e] = xn[g][e] + yn[ys(g)];
        nsum0[g] = x0n[g] + yn[ys(g)];
      }
    }
    if constexpr (kL1) {
#pragma unroll
      for (int g = 0; g < NP; ++g) {
        const RT *yb = Bv[ys(g)];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const RT *xr = xst + (g * 16 + 4 * h + e) * dpad;
          RT s = 0;
#pragma unroll
          for (int k = 0; k < NE; ++k) s += fabs(xr[k] - yb[k]);
          acc[g][e] = s;
        }
      }
      if (accBand0) {
#pragma unroll
        for (int g = 0; g < NP; ++g) {
          RT x0[NE];
          load_run<RT, NE>(op.xp + xpath(g) * Lx * dpad, x0);
          RT s = 0;
#pragma unroll
          for (int k = 0; k < NE; ++k) s += fabs(x0[k] - Bv[ys(g)][k]);
          r0[g] = s;
        }
      }
    } else {
#pragma unroll
      for (int g = 0; g < NP; ++g) {
        acc_t t = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < NE; ++j) {
          if constexpr (kF64)
            t = __builtin_amdgcn_mfma_f64_16x16x4f64(A[g][j], Bv[ys(g)][j], t, 0, 0, 0);
          else
            t = __builtin_amdgcn_mfma_f32_16x16x4f32(A[g][j], Bv[ys(g)][j], t, 0, 0, 0);
        }
        acc[g] = t;
      }
      if (accBand0) {
        // <x_0, y_col>: this lane's K slice, then the sum over the four slices (lanes c + 16h)
#pragma unroll
        for (int g = 0; g < NP; ++g) {
          RT x0[NE];
          load_run<RT, NE>(op.xp + xpath(g) * Lx * dpad + h * NE, x0);
          RT s = 0;
#pragma unroll
          for (int j = 0; j < NE; ++j) s += x0[j] * Bv[ys(g)][j];
          s = s + lane_get(s, lane ^ 16);
          s = s + lane_get(s, lane ^ 32);
          r0[g] = s;
        }
      }
    }
    cC.advance(geo);
    if (more && cC.jv0 == 0) fetchA(cC.band);
  }

  // left values of `cur` (this chunk's right values) given the previous chunk's `prev`
  __device__ __forceinline__ RT left_of(RT cur, RT prev, int lane) const {
    const int c = lane & 15;
    if (accStart) return lane_get(cur, (lane & 48) + (c >= S ? c - S : 0));
    return dpp_mov<0x120 + S>(c >= 16 - S ? prev : cur);  // row_ror:S
  }

  template <int PAR>
  __device__ __forceinline__ void store(RT *ring) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int c = lane & 15, h = lane >> 4;
    RT v[NP][4];
#pragma unroll
    for (int g = 0; g < NP; ++g) {
      RT B[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) B[e] = point(acc[g][e], kRBF ? nsum[g][e] : (RT)0);
      RT *pl = pline + g * pls + accJv + c;
      const RT fromLine = *pl;                // lanes h = 0: the previous band's last row
      const RT up = lane_get(B[3], (lane - 16) & 63);
      if (h == 3) *pl = B[3];                 // for the next band (read above first)
      RT R0;
      if (accBand0) {
        const RT p0 = point(r0[g], kRBF ? nsum0[g] : (RT)0);
        R0 = h == 0 ? p0 : up;
      } else {
        R0 = h == 0 ? fromLine : up;
      }
      RT BL[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) BL[e] = left_of(B[e], prevB[g][e], lane);
      const RT L0 = left_of(R0, prevR0[g], lane);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const RT Re = e ? B[e - 1] : R0;
        const RT Le = e ? BL[e - 1] : L0;
        RT dlt = ((B[e] + Le) - BL[e]) - Re;
        dlt = dlt * inv4;
        if (scaled) dlt = dlt * scl[g];
        v[g][e] = dlt;
        prevB[g][e] = B[e];
      }
      prevR0[g] = R0;
    }
    W.template store<PAR>(ring, v);
  }

  __device__ __forceinline__ void prologue(RT *ring, const PairGeom<G> &geo) {
    const int nchunks = geo.columns() / kChunk;
    W.init();
    cB.start(geo, 0);
    cC.start(geo, 0);
    fetchA(0);
    fetchB(geo);
    compute(geo, nchunks > 1);
    store<0>(ring);
    if (nchunks > 1) fetchB(geo);
  }
  // start of period p: compute chunk p+1, fetch chunk p+2's operands
  __device__ __forceinline__ void early(RT *, const PairGeom<G> &geo, int per) {
    const int nchunks = geo.columns() / kChunk;
    if (per + 1 < nchunks) compute(geo, per + 2 < nchunks);
    if (per + 2 < nchunks) fetchB(geo);
  }
  // between consumer groups 2 and 3 of period p: write chunk p+1 (ring parity PAR)
  template <int PAR>
  __device__ __forceinline__ void period(RT *ring, const PairGeom<G> &geo, int per) {
    const int nchunks = geo.columns() / kChunk;
    if (per + 1 < nchunks) store<PAR>(ring);
  }
};

}  // namespace sigpde
