// MFMA increment producer shared by the fused linear-kernel Gram and pairwise kernels.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "capi_util.h"
#include "ring16.h"
#include "ring16f.h"

namespace sigpde {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));
typedef double double2v __attribute__((ext_vector_type(2)));
typedef double double4v __attribute__((ext_vector_type(4)));

constexpr int kKBlock = 32;  // dimension padding unit (K of one v_mfma_f32_16x16x32_{f16,bf16})

// Operand precisions of the fused producer (values of SIGPDE_PREC_*).
constexpr int kPrecFp16x2 = 1;  // per-path-scaled fp16 hi + lo, 3 f16 MFMAs per 32-wide K block
constexpr int kPrecFp32 = 2;    // fp32 increments, v_mfma_f32_16x16x4f32
constexpr int kPrecBf16x3 = 3;  // bf16 hi + mid + lo, 6 bf16 MFMAs per 32-wide K block
constexpr int kPrecFp64 = 4;    // double increments, v_mfma_f64_16x16x4f64
constexpr int kPrecFp16x2F = 5; // the fp16x2 increments feeding a float32 wavefront (Arith::MixedF)

// ------------------------------------------------------------------ MFMA producer (G = 16)
// Lane group g (16 lanes) solves pairs (x path xa[g], y path ybase[ys(g)] + sp * ystep) of the
// stream: the wave holds four x-operand sets and NY y-operand sets (NY = 1: the Gram, four x paths
// against one y path, whose operands all four tiles share; NY = 4: four independent pairs).
// Chunk k of the stream covers 16 refined columns of one band (16 refined rows) of every group's
// pair: one 16x16 MFMA tile per group, D[row][col] with A = the band's x-increment rows and
// B = the chunk's y-increment columns (refined indices repeat original rows / columns 2^order
// times; columns outside 1..N read a clamped column, see ring_solver.h). Lane l holds
// D[4(l/16) + e][l % 16] and writes those four ring rows of column l % 16.
//   fp16x2: lane l supplies K block kb's elements 8(l/16)..+7 (one half8 per block, hi and lo);
//           the tiles go to the ring unscaled (the consumer applies the pair's scale, Mixed).
//   bf16x3: as fp16x2 with three bf16 parts per element (hi, mid, lo) and the six products
//           whose magnitude is at least 2^-16 of hi*hi, smallest first.
//   fp32:   lane l supplies elements (l/16)*dpad/4 + j, j < dpad/4, one per MFMA (any split of
//           K works as long as A and B use the same one); the tiles are scaled by
//           kscale/4^order (x_scale y_scale) in fp32 and the consumer runs Ref32.
//   fp64:   as fp32 in double on v_mfma_f64_16x16x4f64, scaled in double, consumer Ref64. That
//           instruction leaves D[h + 4e][c] in lane (c, h) element e (f32: D[4h + e][c]); lane
//           (c, h) therefore supplies A row 4(c & 3) + (c >> 2) instead of row c, which permutes
//           D's rows so that element e is again row 4h + e.
// Pipeline of period p: at its start, multiply chunk p+1 and fetch chunk p+2's B operands; between
// its consumer groups 2 and 3, write chunk p+1's tiles to the skewed ring of ring16.h (the
// accumulators never live across periods).
// After the last chunk of a band is multiplied, the next band's A operands are loaded. Operand
// rows are addressed as a wave-uniform 64-bit row base plus a 32-bit per-lane offset.
template <int KB, int NX, int NY, int PREC>
struct MfmaProducer {
  static constexpr int G = 16;
  static constexpr int NP = 4;           // pairs (lane groups) per wave
  static_assert((NX == 4 && NY == 1) || (NX == 4 && NY == 4), "operand sets");
  __device__ static constexpr int xs(int g) { return g; }
  __device__ static constexpr int ys(int g) { return NY == 4 ? g : 0; }
  static constexpr bool kF16 = PREC == kPrecFp16x2 || PREC == kPrecFp16x2F;
  static constexpr bool kB3 = PREC == kPrecBf16x3;
  static constexpr bool kF64 = PREC == kPrecFp64;
  static constexpr bool kSplit = kF16 || kB3;     // 16-bit parts, pair scale applied by Mixed
  static constexpr int kVE = kSplit ? 8 : (kF64 ? 2 : 4);  // elements per operand vector
  static constexpr int NV = KB * kKBlock / 4 / kVE;        // operand vectors per lane and set
  typedef typename std::conditional<
      kF16, half8,
      typename std::conditional<kB3, bf16x8,
                                typename std::conditional<kF64, double2v, float4v>::type>::type>::type
      vec;
  typedef typename std::conditional<
      kF16, _Float16,
      typename std::conditional<kB3, __bf16,
                                typename std::conditional<kF64, double, float>::type>::type>::type elt;
  typedef typename std::conditional<kF64, double, float>::type RT;        // ring element
  typedef typename std::conditional<kF64, double4v, float4v>::type acc_t;  // one lane's tile part
  const elt *xh, *xl, *yh, *yl;          // fp32: xh / yh only
  const elt *xm, *ym;                    // bf16x3: middle parts
  int dpad;
  int m, n;                              // original increments per x / y path
  int64_t xa[NX];                        // x path of each x set (-1: idle)
  int64_t ybase[NY];                     // y path of each y set at stream position 0
  int64_t ylane;                         // per lane: y path of the lane's pair at position 0
  int ystep;                             // y path advance per stream position
  const float *invy;
  const double *sx, *sy;                 // optional path scales (robust Gram), may be null
  int64_t ny;                            // valid y paths (ybase + sp*ystep < ny)
  double kmult;                          // fp32: kscale / 4^order
  double xsl;                            // 16-bit splits, per lane: kscale 4^-order 2^-ex (x_scale)
  double pmY;                            // 16-bit splits, per lane: 2^-ey (y_scale) of the lane's
                                         // pair multiplied last

  vec Ah[NX][NV], Al[NX][kSplit ? NV : 1], Am[NX][kB3 ? NV : 1];  // A operands of the band
  vec Bh[NY][NV], Bl[NY][kSplit ? NV : 1], Bm[NY][kB3 ? NV : 1];  // next chunk's B operands
  acc_t acc[NP];
  RT mult[NP];                           // fp32 / fp64: scale of the tiles in acc
  RT pm[NP];                             // fp32 / fp64: scale of the pair being multiplied
  int64_t yrow[NY];                      // first B row (in elements) of the pair being fetched
  Cursor<G> cB, cC;                      // next chunk to fetch B for / to multiply
  // chunk writes into the ring (the float32 wavefront's row-major ring for fp16x2f)
  typename std::conditional<PREC == kPrecFp16x2F, Ring16WriterF, Ring16Writer<RT>>::type W;

  __device__ __forceinline__ double lane_scale() const { return xsl * pmY; }

  __device__ __forceinline__ int64_t ypath(int q, int sp) const {
    const int64_t yp = ybase[q] + (int64_t)sp * ystep;
    return yp < ny ? yp : ny - 1;
  }
  // per-lane element offset of this lane's K slice in an operand row
  __device__ __forceinline__ uint32_t kslice() const {
    const int h = (threadIdx.x & 63) >> 4;
    return kSplit ? 8 * h : h * (dpad / 4);
  }
  __device__ __forceinline__ void load_row(const elt *ph, const elt *pm, const elt *pl, uint32_t lo,
                                           vec *dh, vec *dm, vec *dl) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if constexpr (kSplit) {
        dh[v] = *reinterpret_cast<const vec *>(ph + lo + v * kKBlock);
        dl[v] = *reinterpret_cast<const vec *>(pl + lo + v * kKBlock);
        if constexpr (kB3) dm[v] = *reinterpret_cast<const vec *>(pm + lo + v * kKBlock);
      } else {
        dh[v] = *reinterpret_cast<const vec *>(ph + lo + kVE * v);
      }
    }
  }

  __device__ __forceinline__ void fetchA(const PairGeom<G> &geo, int band) {
    const int lane = threadIdx.x & 63;
    const int c = lane & 15;
    const int rl = kF64 ? 4 * (c & 3) + (c >> 2) : c;  // fp64: row permutation (see above)
    int ro = (band * G * geo.R + geo.R * rl) >> geo.order;
    ro = ro > m - 1 ? m - 1 : ro;
    const uint32_t lo = (uint32_t)(ro * dpad) + kslice();
#pragma unroll
    for (int a = 0; a < NX; ++a) {
      const int64_t xrow = (xa[a] < 0 ? 0 : xa[a]) * m * (int64_t)dpad;
      load_row(xh + xrow, kB3 ? xm + xrow : nullptr, kSplit ? xl + xrow : nullptr, lo, Ah[a],
               Am[a], Al[a]);
    }
  }
  // B operands of chunk cB, then advance cB.
  __device__ __forceinline__ void fetchB(const PairGeom<G> &geo) {
    const int lane = threadIdx.x & 63;
    if (cB.band == 0 && cB.jv0 == 0) {
#pragma unroll
      for (int q = 0; q < NY; ++q) yrow[q] = ypath(q, cB.sp) * n * dpad;
    }
    int co = (cB.jv0 + (lane & 15) - 1) >> geo.order;
    co = co < 0 ? 0 : (co > n - 1 ? n - 1 : co);
    const uint32_t lo = (uint32_t)(co * dpad) + kslice();
#pragma unroll
    for (int q = 0; q < NY; ++q)
      load_row(yh + yrow[q], kB3 ? ym + yrow[q] : nullptr, kSplit ? yl + yrow[q] : nullptr, lo,
               Bh[q], Bm[q], Bl[q]);
    cB.advance(geo);
  }
  // Tiles of chunk cC, then advance cC; if the next chunk opens a band, load that band's A
  // operands now, into the registers this chunk has just finished reading (they land while the
  // wave computes).
  __device__ __forceinline__ void compute(const PairGeom<G> &geo, bool more) {
    if (cC.band == 0 && cC.jv0 == 0) {
      if constexpr (kSplit) {
        // y scale of this lane's pair (a per-lane load)
        int64_t yp = ylane + (int64_t)cC.sp * ystep;
        yp = yp < ny ? yp : ny - 1;
        pmY = sy ? (double)invy[yp] * sy[yp] : (double)invy[yp];
      } else {
#pragma unroll
        for (int g = 0; g < NP; ++g) {
          const int64_t xp = xa[xs(g)] < 0 ? 0 : xa[xs(g)];
          const int64_t yp = ypath(ys(g), cC.sp);
          double mm = kmult;
          if (sx) mm *= sx[xp];
          if (sy) mm *= sy[yp];
          pm[g] = (RT)mm;
        }
      }
    }
#pragma unroll
    for (int g = 0; g < NP; ++g) {
      const int a = xs(g), q = ys(g);
      acc_t c = {0, 0, 0, 0};
      // Products of one magnitude per accumulator, joined by float adds: adding products onto
      // an accumulator of much smaller magnitude rounds the sum downwards on this MFMA (about
      // -1 ulp on average, measured; from zero, or onto a sum of like magnitude, the rounding is
      // unbiased), and that bias, the same sign in every cell, would grow with the path length.
      if constexpr (PREC == kPrecFp16x2F) {
        // one accumulator, largest products first (hi*hi, then lo*hi + hi*lo): unbiased on this
        // MFMA (the cross terms are added onto the larger sum), and no join of accumulators
        // (whose adds would wait for the MFMA results in the middle of the period). Measured on
        // the 1024 x 1024 Gram of length 256, dim 64, with the double wavefront: 8.8e-7 against
        // 2.7e-7 with one accumulator per magnitude, below the float32 wavefront's own 2e-6.
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[a][kb], Bh[q][kb], c, 0, 0, 0);
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[a][kb], Bh[q][kb], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[a][kb], Bl[q][kb], c, 0, 0, 0);
        }
      } else if constexpr (kF16) {
        acc_t x = {0, 0, 0, 0};  // lo*hi + hi*lo (2^-11)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          x = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[a][kb], Bh[q][kb], x, 0, 0, 0);
          x = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[a][kb], Bl[q][kb], x, 0, 0, 0);
        }
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[a][kb], Bh[q][kb], c, 0, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] += x[e];
      } else if constexpr (kB3) {
        acc_t x = {0, 0, 0, 0}, y = {0, 0, 0, 0};
        // hi*lo + mid*mid + lo*hi (2^-16)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[a][kb], Bl[q][kb], x, 0, 0, 0);
          x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am[a][kb], Bm[q][kb], x, 0, 0, 0);
          x = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al[a][kb], Bh[q][kb], x, 0, 0, 0);
        }
        // hi*mid + mid*hi (2^-8)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[a][kb], Bm[q][kb], y, 0, 0, 0);
          y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am[a][kb], Bh[q][kb], y, 0, 0, 0);
        }
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[a][kb], Bh[q][kb], c, 0, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] += y[e] + x[e];
      } else if constexpr (kF64) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int e = 0; e < kVE; ++e)
            c = __builtin_amdgcn_mfma_f64_16x16x4f64(Ah[a][v][e], Bh[q][v][e], c, 0, 0, 0);
        mult[g] = pm[g];
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int e = 0; e < kVE; ++e)
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(Ah[a][v][e], Bh[q][v][e], c, 0, 0, 0);
        mult[g] = pm[g];
      }
      acc[g] = c;
    }
    cC.advance(geo);
    if (more && cC.jv0 == 0) fetchA(geo, cC.band);
  }
  template <int PAR>
  __device__ __forceinline__ void store(RT *ring) {
    RT v[NP][4];
#pragma unroll
    for (int g = 0; g < NP; ++g)
#pragma unroll
      for (int e = 0; e < 4; ++e) v[g][e] = kSplit ? acc[g][e] : acc[g][e] * mult[g];
    W.template store<PAR>(ring, v);
  }

  __device__ __forceinline__ void prologue(RT *ring, const PairGeom<G> &geo) {
    const int nchunks = geo.columns() / kChunk;
    W.init();
    cB.start(geo, 0);
    cC.start(geo, 0);
    fetchA(geo, 0);
    fetchB(geo);
    compute(geo, nchunks > 1);
    store<0>(ring);
    if (nchunks > 1) fetchB(geo);
  }
  // start of period p: multiply chunk p+1, fetch chunk p+2's B operands
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

// Arguments of the fused kernels (x / y operands from the split prepass).
template <typename TO>
struct GramFusedArgs {
  const void *xh, *xl, *yh, *yl;  // fp16x2: _Float16 hi / lo rows; fp32 / fp64: float / double
                                  // rows in xh / yh
  const void *xm, *ym;            // bf16x3: middle parts (hi / lo in xh / xl)
  const float *invx, *invy;       // fp16x2, bf16x3: per-path 2^-e
  int64_t Bx, By;
  int m, n, order, dpad, ys;
  int tx, ty;     // Gram: super-tile of x-quads by y-ranges (block order)
  double kmult;  // kscale / 4^order
  int symmetric;
  TO *out;
  int64_t ldo;
};

// Arithmetic of the wavefront for a precision mode: float32 typing for fp32, double for fp64,
// the mixed double wavefront on float increment products for the split 16-bit modes.
template <int PREC>
struct FusedArith {
  static constexpr Arith A =
      PREC == kPrecFp32 ? Arith::Ref32
                        : (PREC == kPrecFp64 ? Arith::Ref64
                                             : (PREC == kPrecFp16x2F ? Arith::MixedF : Arith::Mixed));
};

// Where a fused launch keeps its band-boundary lines: in LDS (r16_lds_bytes) unless the launch
// has more waves than the GPU holds with LDS lines and global memory lets at least 1.5 times as
// many share a CU (long rows of columns, whose lines fill LDS), then one block of r16_line_bytes
// per wave of the launch (at most kFusedLineBytes in all), with only the ring in LDS. A smaller
// gain, or a launch resident at once either way, does not pay for the global line accesses
// (measured: 12 instead of 9 waves per CU was 12% slower, 12 instead of 7 was 7% faster).
// Results are the same either way. A stream under graph capture keeps the lines in LDS (they
// always fit there below the fused column limit), so that a captured launch allocates nothing.
// k_lds / k_gl: the kernel of each placement. Returns true for lines in global memory.
constexpr size_t kFusedLineBytes = (size_t)4 << 30;
template <Arith A>
inline bool fused_lines_in_global(const void *k_lds, const void *k_gl, int N, int64_t waves,
                                  hipStream_t st) {
  const size_t lds_lines = r16_lds_bytes<A>(N);
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) != hipSuccess) {
    (void)hipGetLastError();  // a failed query only keeps the lines in LDS
    return false;
  }
  if (cap != hipStreamCaptureStatusNone) return false;
  if ((double)waves * (double)r16_line_bytes<A>(N) > (double)kFusedLineBytes) return false;
  const int occ_lds = lds_lines <= kMaxLdsBytes ? blocks_per_cu(k_lds, 64, lds_lines) : 0;
  const int occ_gl = blocks_per_cu(k_gl, 64, r16_ring_bytes<A>());
  return waves > (int64_t)cu_count() * occ_lds && occ_gl > occ_lds && 2 * occ_gl >= 3 * occ_lds;
}

// launch(lines, shmem): lines null with the lines in shmem = r16_lds_bytes bytes of LDS, or (gl)
// one block of r16_line_bytes per wave, allocated stream-ordered on st for the launch and freed
// after it on the same stream, with shmem = r16_ring_bytes. When that block cannot be allocated
// and the lines fit in LDS, the launch keeps them in LDS (same results, fewer waves per CU).
template <Arith A, class F>
inline hipError_t with_fused_lines(bool gl, int N, int64_t waves, hipStream_t st, F &&launch) {
  const size_t lds_lines = r16_lds_bytes<A>(N);
  if (gl) {
    void *lines = nullptr;
    hipError_t e = hipMallocAsync(&lines, (size_t)waves * r16_line_bytes<A>(N), st);
    if (e == hipSuccess) {
      launch(lines, r16_ring_bytes<A>());
      e = hipGetLastError();
      const hipError_t ef = hipFreeAsync(lines, st);
      return e != hipSuccess ? e : ef;
    }
    if (lds_lines > kMaxLdsBytes) return e;
    (void)hipGetLastError();  // clears the failed allocation's error before the LDS launch
  }
  launch(nullptr, lds_lines);
  return hipGetLastError();
}

// This wave's block of global boundary lines (GL launches), else null.
template <Arith A, bool GL>
__device__ __forceinline__ typename R16Bnd<A>::T *fused_wave_lines(void *gl, int N) {
  if constexpr (GL)
    return reinterpret_cast<typename R16Bnd<A>::T *>(static_cast<char *>(gl) +
                                                     (size_t)blockIdx.x * r16_line_bytes<A>(N));
  else
    return nullptr;
}

// Pairwise launcher, defined in fused_pairwise.hip (a translation unit of its own: its four
// independent B operand sets need the compiler's default accumulator placement, while the Gram
// kernel is built with MFMA results in VGPRs).
template <typename T, int KB, int PREC>
hipError_t launch_linear_pairwise_fused(const GramFusedArgs<T> &a, const double *sxd,
                                        const double *syd, hipStream_t st);

}  // namespace sigpde
