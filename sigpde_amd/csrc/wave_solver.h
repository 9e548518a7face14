// Device-side Goursat PDE solver for one path pair on one gfx950 wavefront (64 lanes).
//
// Scheme (stojl/SigPDE sigpde/cuda_device_functions.py:3-5, sigpde/cuda_pairwise_kernels.py:8-39):
//   k[i][0] = k[0][j] = 1, and for 1 <= i <= M, 1 <= j <= N
//   k[i][j] = (k[i-1][j] + k[i][j-1]) * (1 + c/2 + c^2/12) - k[i-1][j-1] * (1 - c^2/12)
//   with c = Delta[(i-1) >> order][(j-1) >> order] (times a scale); the result is k[M][N].
//
// Mapping (lane-systolic; no barriers, no global scratch):
//   * Rows are cut into bands of 64*H rows. Inside a band, lane r owns the H consecutive rows
//     band*64H + r*H + 1 .. + H and keeps their previous-column values in registers.
//   * Lane r works on column j at step t = j - 1 + r (+ band offset): the value above its first
//     row comes from lane r-1's last row one step earlier through a DPP wave_shr:1 move, so a
//     wave sweeps anti-diagonals without any synchronisation.
//   * Bands run back to back: lane 0 of band b+1 starts as soon as it has finished band b. The
//     last row of band b is parked in an LDS buffer of N values by lane 63 and read back by lane 0
//     64 or more steps later (each band is padded to Np = max(N, 64) steps to keep that distance).
//   * Increments are read from global memory (L1/L2 resident per pair) when a lane enters a new
//     original column; rows sharing an original row reuse the loaded value.
// Every cell reads only its three neighbours, so any dependency-respecting order reproduces the
// reference's anti-diagonal sweep value for value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sigpde {

constexpr int kWave = 64;

// DPP wave_shr:1 -- lane l receives lane l-1's value; lane 0 receives `fill`.
__device__ __forceinline__ int dpp_shr1_i32(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ float wave_shr1(float v, float fill) {
  return __int_as_float(dpp_shr1_i32(__float_as_int(v), __float_as_int(fill)));
}
__device__ __forceinline__ double wave_shr1(double v, double fill) {
  int2 a = *reinterpret_cast<int2 *>(&v);
  int2 f = *reinterpret_cast<int2 *>(&fill);
  int2 r;
  r.x = dpp_shr1_i32(a.x, f.x);
  r.y = dpp_shr1_i32(a.y, f.y);
  return *reinterpret_cast<double *>(&r);
}

__device__ __forceinline__ float readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane(double v, int l) {
  int2 a = *reinterpret_cast<int2 *>(&v);
  int2 r;
  r.x = __builtin_amdgcn_readlane(a.x, l);
  r.y = __builtin_amdgcn_readlane(a.y, l);
  return *reinterpret_cast<double *>(&r);
}

// Per-increment coefficients A = 1 + c/2 + c^2/12 and B = 1 - c^2/12, in double.
//
// Storage float ("FP32 reference semantics"): c is a float (increment times a float scale) and
// c*c is computed in float, everything else in double (Numba's typing of pde_step for float32
// arrays); the stored cell is rounded to float.
// Storage double: all double.
// kDoubleC: c is formed in double as double(inc) * scale -- the root-find's inc * xt**2
// (sigpde/cuda_pairwise_kernels.py:114), whatever the storage type.
template <typename T, bool kDoubleC>
struct Coef;

template <>
struct Coef<float, false> {
  using Scale = float;
  __device__ __forceinline__ static void get(float inc, float s, double &A, double &B) {
    float c = inc * s;
    float c2 = c * c;
    double cd = (double)c, c2d = (double)c2;
    A = (1.0 + 0.5 * cd) + (1.0 / 12) * c2d;
    B = 1.0 - (1.0 / 12) * c2d;
  }
};
template <bool D>
struct CoefD {
  using Scale = double;
  template <typename I>
  __device__ __forceinline__ static void get(I inc, double s, double &A, double &B) {
    double c = (double)inc * s;
    double c2 = c * c;
    A = (1.0 + 0.5 * c) + (1.0 / 12) * c2;
    B = 1.0 - (1.0 / 12) * c2;
  }
};
template <>
struct Coef<float, true> : CoefD<true> {};
template <>
struct Coef<double, false> : CoefD<false> {};
template <>
struct Coef<double, true> : CoefD<true> {};

template <typename T>
__device__ __forceinline__ T store_round(double v) { return (T)v; }

// Solves one pair on the calling wave and returns k[M][N] in every lane.
//   inc: m x n increments with row stride ld (elements); bnd: LDS scratch of >= N elements of T
//   (only touched when the grid needs more than one band).
template <typename T, int H, bool kDoubleC>
__device__ T wave_solve(const T *__restrict__ inc, int64_t ld, int m, int n, int order,
                        typename Coef<T, kDoubleC>::Scale scale, T *bnd) {
  const int lane = threadIdx.x & (kWave - 1);
  const int R = 1 << order;
  const int M = m << order, N = n << order;
  const int rowsPerBand = kWave * H;
  const int nb = (M + rowsPerBand - 1) / rowsPerBand;
  const int Np = N > kWave ? N : kWave;
  const int nsteps = nb * Np + (kWave - 1);
  // Rows of one lane share original rows when H >= R (load every R-th row) or all sit in one
  // original row when H < R (H divides R): this flag is uniform.
  const bool hGeR = H >= R;

  T kprev[H];
  double A[H], B[H];
#pragma unroll
  for (int e = 0; e < H; ++e) {
    kprev[e] = (T)1;
    A[e] = 1.0;
    B[e] = 1.0;
  }
  T upPrev = (T)1;  // k[first row - 1][j - 1]
  T bottom = (T)1;  // k[last row][j], read by lane + 1 on the next step
  T result = (T)0;

  // Position of this lane in the (band, column) stream; lane r starts at step r.
  int band = 0, col = 0;  // col is 1-based once started
  const T *rowBase = inc;
  // Result cell (M, N): band, lane, row-in-lane.
  const int bandM = (M - 1) / rowsPerBand;
  const int laneM = ((M - 1) % rowsPerBand) / H;
  const int eM = (M - 1) % H;

  for (int t = 0; t < nsteps; ++t) {
    T up = wave_shr1(bottom, (T)1);
    if (t >= lane) {
      // advance to the next virtual column
      ++col;
      if (col > Np) {
        col = 1;
        ++band;
      }
      if (band < nb && col <= N) {
        if (col == 1) {
#pragma unroll
          for (int e = 0; e < H; ++e) kprev[e] = (T)1;
          upPrev = (T)1;
          // first refined row of this lane in this band (0-based), clamped original row
          int r0 = band * rowsPerBand + lane * H;
          int o0 = r0 >> order;
          rowBase = inc + (int64_t)(o0 < m ? o0 : m - 1) * ld;
        }
        if (lane == 0 && band > 0) up = bnd[col - 1];
        if (((col - 1) & (R - 1)) == 0) {
          const int jo = (col - 1) >> order;
          const int r0 = band * rowsPerBand + lane * H;
          const int o0 = r0 >> order;
#pragma unroll
          for (int e = 0; e < H; ++e) {
            if (e == 0 || (hGeR && (e & (R - 1)) == 0)) {
              int oe = o0 + (hGeR ? (e >> order) : 0);
              const T *p = inc + (int64_t)(oe < m ? oe : m - 1) * ld + jo;
              Coef<T, kDoubleC>::get(*p, scale, A[e], B[e]);
            } else {
              A[e] = A[e - 1];
              B[e] = B[e - 1];
            }
          }
        }
        // one column of H cells
        T k00 = upPrev, k01 = up;
#pragma unroll
        for (int e = 0; e < H; ++e) {
          double r = ((double)k01 + (double)kprev[e]) * A[e] - (double)k00 * B[e];
          T kn = store_round<T>(r);
          k00 = kprev[e];
          k01 = kn;
          kprev[e] = kn;
        }
        upPrev = up;
        bottom = kprev[H - 1];
        if (lane == kWave - 1 && band + 1 < nb) bnd[col - 1] = bottom;
        if (band == bandM && col == N && lane == laneM) {
#pragma unroll
          for (int e = 0; e < H; ++e)
            if (e == eM) result = kprev[e];
        }
      }
    }
  }
  (void)rowBase;
  return readlane(result, laneM);
}

}  // namespace sigpde
