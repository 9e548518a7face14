// Issue-rate probe of the current GPU: cycles per wave64 instruction of FP32 FMA, FP64 FMA and
// v_mfma_f32_16x16x32_f16 on one SIMD, and whether MFMA and VALU work overlap on it. The
// roofline of the fused Gram kernels (MFMA increments feeding a VALU wavefront on the same
// SIMDs) is priced with these figures beside the published peaks.
//
// Every kernel runs two waves per SIMD on every CU (blocks of one wave, 8 x CU count blocks);
// each wave times a loop of 6 x (NM independent MFMA accumulations + NV x 8 independent FMA
// chains) with clock64 and writes one value per lane (so nothing is optimised away).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "capi_util.h"

namespace sigpde {

typedef _Float16 rate_half8 __attribute__((ext_vector_type(8)));
typedef float rate_float4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float rate_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double rate_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <int NM, int NV, typename V>
__global__ __launch_bounds__(64) void k_issue_rate(const rate_half8 *in, float *out,
                                                   long long *ticks, int iters) {
  const int l = threadIdx.x;
  const rate_half8 a = in[l], b = in[64 + l];
  rate_float4 c[NM > 0 ? NM : 1];
#pragma unroll
  for (int j = 0; j < (NM > 0 ? NM : 1); ++j) c[j] = rate_float4{0, 0, 0, 0};
  V v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (V)(l * 1e-3f + j);
  const V m = (V)0.999f, s = (V)1e-4f;
  const long long t0 = clock64();
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 6; ++r) {
#pragma unroll
      for (int j = 0; j < NM; ++j) c[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c[j], 0, 0, 0);
#pragma unroll
      for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = rate_fma(v[j], m, s);
    }
  }
  const long long t1 = clock64();
  float t = 0.0f;
#pragma unroll
  for (int j = 0; j < (NM > 0 ? NM : 1); ++j) t += c[j][0] + c[j][1] + c[j][2] + c[j][3];
#pragma unroll
  for (int j = 0; j < 8; ++j) t += (float)v[j];
  out[blockIdx.x * 64 + l] = t;
  if (l == 0) ticks[blockIdx.x] = t1 - t0;
}

// Median per-wave tick count per loop iteration of one probe launch over `blocks` one-wave
// blocks, after one untimed launch of the same probe (clocks and caches warm).
template <int NM, int NV, typename V>
static hipError_t probe(const rate_half8 *in, float *out, long long *ticks, long long *host,
                        int blocks, int iters, double *median) {
  hipError_t e = hipSuccess;
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
    k_issue_rate<NM, NV, V><<<dim3(blocks), dim3(64)>>>(in, out, ticks, iters);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  if (e == hipSuccess) e = hipMemcpy(host, ticks, sizeof(long long) * blocks, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  std::sort(host, host + blocks);
  *median = (double)host[blocks / 2] / iters;
  return hipSuccess;
}

}  // namespace sigpde

using namespace sigpde;

extern "C" int sigpde_issue_rates(double *out, int n) {
  clear_error();
  if (!out || n < 7) return set_error(SIGPDE_EINVAL, "issue_rates: need room for 7 values");
  const int cus = cu_count();
  if (cus <= 0) return set_error(SIGPDE_EDEVICE, "issue_rates: no device");
  const int blocks = 8 * cus;  // two one-wave blocks per SIMD
  const int iters = 2000;
  rate_half8 *in = nullptr;
  float *buf = nullptr;
  long long *ticks = nullptr;
  long long *host = new long long[blocks];
  hipError_t e = hipMalloc((void **)&in, 128 * sizeof(rate_half8));
  if (e == hipSuccess) e = hipMemset(in, 0, 128 * sizeof(rate_half8));
  if (e == hipSuccess) e = hipMalloc((void **)&buf, (size_t)blocks * 64 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&ticks, (size_t)blocks * sizeof(long long));
  double t_mfma = 0, t_f32 = 0, t_f64 = 0, t_mix = 0;
  if (e == hipSuccess) e = probe<4, 0, float>(in, buf, ticks, host, blocks, iters, &t_mfma);
  if (e == hipSuccess) e = probe<0, 4, float>(in, buf, ticks, host, blocks, iters, &t_f32);
  if (e == hipSuccess) e = probe<0, 4, double>(in, buf, ticks, host, blocks, iters, &t_f64);
  if (e == hipSuccess) e = probe<4, 4, float>(in, buf, ticks, host, blocks, iters, &t_mix);
  int clock_khz = 0;
  if (e == hipSuccess) {
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, dev);
  }
  if (in) (void)hipFree(in);
  if (buf) (void)hipFree(buf);
  if (ticks) (void)hipFree(ticks);
  delete[] host;
  if (e != hipSuccess) return hip_status(e, "issue_rates");
  // per loop iteration a wave issues 24 MFMA and / or 192 FMA; two waves share each SIMD
  out[0] = t_f32 / (2 * 192.0);   // ticks per wave64 FP32 FMA, per SIMD
  out[1] = t_f64 / (2 * 192.0);   // ticks per wave64 FP64 FMA, per SIMD
  out[2] = t_mfma / (2 * 24.0);   // ticks per v_mfma_f32_16x16x32_f16, per SIMD
  out[3] = t_mix / (t_mfma + t_f32);  // MFMA + FP32 FMA in the same waves, against their sum
  out[4] = (double)clock_khz * 1e3;   // the device's reported clock (ticks of clock64 per second)
  out[5] = (double)cus;
  out[6] = (double)blocks;
  return SIGPDE_OK;
}
