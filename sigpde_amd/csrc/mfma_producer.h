// MFMA increment producer shared by the fused linear-kernel Gram and pairwise kernels.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ring_solver.h"

namespace sigpde {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kKBlock = 32;  // K of one v_mfma_f32_16x16x32_f16

// ------------------------------------------------------------------ MFMA producer (G = 16)
// Lane group g (16 lanes) solves pairs (x path xa[g], y path ybase[g] + sp * ystep) of the stream.
// Chunk k of the stream covers 16 refined columns of one band (16 refined rows) of every group's
// pair: one 16x16 MFMA tile per group, D[row][col] with A = the band's x-increment rows and
// B = the chunk's y-increment columns (refined indices repeat original rows / columns 2^order
// times; columns outside 1..N read a clamped column, see ring_solver.h). Lane l holds
// D[4(l/16) + e][l % 16] and writes those four ring rows of column l % 16.
// Pipeline of period p: at its start, multiply chunk p+1 and fetch chunk p+2's B operands; before
// its last consumer group, write chunk p+1's tiles (the accumulators never live across periods).
// After the last chunk of a band is multiplied, the next band's A operands are loaded. Operand
// rows are addressed as a wave-uniform 64-bit row base plus a 32-bit per-lane offset.
template <int KB, bool kSharedY>
struct MfmaProducer {
  static constexpr int G = 16;
  static constexpr int NP = 4;           // pairs (lane groups) per wave
  static constexpr int NB = kSharedY ? 1 : NP;
  const _Float16 *xh, *xl, *yh, *yl;
  int dpad;
  int m, n;                              // original increments per x / y path
  int64_t xa[NP];                        // x path of each group (-1: idle)
  int64_t ybase[NP];
  int ystep;
  const float *invy;
  const double *sy;                      // optional y scales (robust Gram), may be null
  int64_t ny;                            // valid y paths (ybase + sp*ystep < ny)
  double xsl;                            // per lane: kscale 4^-order 2^-ex (x_scale) of its x path
  double ysl;                            // per lane (pairwise): 2^-ey (y_scale) of its y path
  double pmY;                            // Gram: 2^-ey (y_scale) of the pair multiplied last

  half8 Ah[NP][KB], Al[NP][KB];          // A operands of the band being multiplied
  half8 Bh[NB][KB], Bl[NB][KB];          // next chunk's B operands (fetched one period ahead)
  float4v acc[NP];
  int64_t yrow[NB];                      // first B row (in halves) of the pair being fetched
  Cursor<G> cB, cC;                      // next chunk to fetch B for / to multiply
  int slotS;                             // ring slot of the next chunk to store

  __device__ __forceinline__ double lane_scale() const { return kSharedY ? xsl * pmY : xsl * ysl; }

  __device__ __forceinline__ int64_t ypath(int g, int sp) const {
    const int64_t yp = ybase[g] + (int64_t)sp * ystep;
    return yp < ny ? yp : ny - 1;
  }

  __device__ __forceinline__ void fetchA(const PairGeom<G> &geo, int band) {
    const int lane = threadIdx.x & 63;
    int ro = (band * G + (lane & 15)) >> geo.order;
    ro = ro > m - 1 ? m - 1 : ro;
    const uint32_t lo = (uint32_t)(ro * dpad + 8 * (lane >> 4));
#pragma unroll
    for (int g = 0; g < NP; ++g) {
      const int64_t xrow = (xa[g] < 0 ? 0 : xa[g]) * m * (int64_t)dpad;
      const _Float16 *ph = xh + xrow, *pl = xl + xrow;
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        Ah[g][kb] = *reinterpret_cast<const half8 *>(ph + lo + kb * kKBlock);
        Al[g][kb] = *reinterpret_cast<const half8 *>(pl + lo + kb * kKBlock);
      }
    }
  }
  // B operands of chunk cB, then advance cB.
  __device__ __forceinline__ void fetchB(const PairGeom<G> &geo) {
    const int lane = threadIdx.x & 63;
    if (cB.band == 0 && cB.jv0 == 0) {
#pragma unroll
      for (int q = 0; q < NB; ++q) yrow[q] = ypath(q, cB.sp) * n * dpad;
    }
    int co = (cB.jv0 + (lane & 15) - 1) >> geo.order;
    co = co < 0 ? 0 : (co > n - 1 ? n - 1 : co);
    const uint32_t lo = (uint32_t)(co * dpad + 8 * (lane >> 4));
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const _Float16 *ph = yh + yrow[q], *pl = yl + yrow[q];
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        Bh[q][kb] = *reinterpret_cast<const half8 *>(ph + lo + kb * kKBlock);
        Bl[q][kb] = *reinterpret_cast<const half8 *>(pl + lo + kb * kKBlock);
      }
    }
    cB.advance(geo);
  }
  // Tiles of chunk cC, then advance cC; if the next chunk opens a band, load that band's A
  // operands now, into the registers this chunk has just finished reading (they land while the
  // wave computes).
  __device__ __forceinline__ void compute(const PairGeom<G> &geo, bool more) {
    if (kSharedY && cC.band == 0 && cC.jv0 == 0) {
      const int64_t yp = ypath(0, cC.sp);
      pmY = sy ? (double)invy[yp] * sy[yp] : (double)invy[yp];
    }
#pragma unroll
    for (int g = 0; g < NP; ++g) {
      const int q = kSharedY ? 0 : g;
      float4v a = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[g][kb], Bh[q][kb], a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[g][kb], Bl[q][kb], a, 0, 0, 0);
      }
#pragma unroll
      for (int kb = 0; kb < KB; ++kb)
        a = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[g][kb], Bh[q][kb], a, 0, 0, 0);
      acc[g] = a;
    }
    cC.advance(geo);
    if (more && cC.jv0 == 0) fetchA(geo, cC.band);
  }
  __device__ __forceinline__ void store(float *ring) {
    const int lane = threadIdx.x & 63;
    const int slot0 = slotS;
    const int c = lane & 15;
#pragma unroll
    for (int g = 0; g < NP; ++g) {
      float *dst = ring + (slot0 + c) * kRingStride + g * G + 4 * (lane >> 4);
      const float2 v0 = make_float2(acc[g][0], acc[g][1]);
      const float2 v1 = make_float2(acc[g][2], acc[g][3]);
      *reinterpret_cast<float2 *>(dst) = v0;
      *reinterpret_cast<float2 *>(dst + 2) = v1;
      if (slot0 == 0 && c < kRingMirror) {
        float *mir = ring + (Layout<G>::kRingW + c) * kRingStride + g * G + 4 * (lane >> 4);
        *reinterpret_cast<float2 *>(mir) = v0;
        *reinterpret_cast<float2 *>(mir + 2) = v1;
      }
    }
    slotS += kChunk;
    if (slotS == Layout<G>::kRingW) slotS = 0;
  }

  __device__ __forceinline__ void prologue(float *ring, const PairGeom<G> &geo) {
    const int nchunks = geo.columns() / kChunk;
    cB.start(geo, 0);
    cC.start(geo, 0);
    slotS = 0;
    fetchA(geo, 0);
    fetchB(geo);
    compute(geo, nchunks > 1);
    store(ring);
    if (nchunks > 1) fetchB(geo);
  }
  // start of period p: multiply chunk p+1, fetch chunk p+2's B operands
  __device__ __forceinline__ void early(float *, const PairGeom<G> &geo, int per) {
    const int nchunks = geo.columns() / kChunk;
    if (per + 1 < nchunks) compute(geo, per + 2 < nchunks);
    if (per + 2 < nchunks) fetchB(geo);
  }
  // before the last consumer group of period p: write chunk p+1
  __device__ __forceinline__ void period(float *ring, const PairGeom<G> &geo, int per) {
    const int nchunks = geo.columns() / kChunk;
    if (per + 1 < nchunks) store(ring);
  }
};

// Arguments of the fused kernels (x / y operands from the split prepass).
template <typename TO>
struct GramFusedArgs {
  const _Float16 *xh, *xl, *yh, *yl;
  const float *invx, *invy;
  int64_t Bx, By;
  int m, n, order, dpad, ys;
  double kmult;  // kscale / 4^order
  int symmetric;
  TO *out;
  int64_t ldo;
};

// Pairwise launcher, defined in fused_pairwise.hip (a translation unit of its own: its four
// independent B operand sets need the compiler's default accumulator placement, while the Gram
// kernel is built with MFMA results in VGPRs).
template <typename T, int KB>
hipError_t launch_linear_pairwise_mixed(const GramFusedArgs<T> &a, const double *sxd,
                                        const double *syd, size_t shmem, hipStream_t st);

}  // namespace sigpde
