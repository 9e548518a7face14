// Issue-rate probe of the current GPU: cycles per wave64 instruction of FP32 FMA, FP64 FMA and
// v_mfma_f32_16x16x32_f16 on one SIMD, whether MFMA and VALU work overlap on it (issued by the
// same waves, and by different waves of one SIMD), the rate of an FP32 add with a DPP operand
// (the wavefront's cross-lane shift) and the dependent-issue latency of FP32 FMA and DPP add.
// The roofline of the fused Gram kernels (MFMA increments feeding a VALU wavefront on the same
// SIMDs) is priced with these figures beside the published peaks.
//
// The rate kernels run two waves per SIMD on every CU (blocks of one wave, 8 x CU count blocks);
// each wave times a loop of 6 x (NM independent MFMA accumulations + NV x 8 independent FMA
// chains) with clock64 and writes one value per lane (so nothing is optimised away). The
// split kernels run blocks of four waves whose waves take one role each (MFMA only or FMA only);
// a CU's blocks place their waves on its four SIMDs in turn, so with role = (wave + block) & 1
// every SIMD holds as many MFMA waves as FMA waves.
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

// One role per wave: MODE 0 every wave MFMA only (6 x 4 independent accumulations per
// iteration), MODE 1 every wave FP32 FMA only (12 x 8 independent chains), MODE 2 alternate
// roles, role = (wave + block) & 1. ticks[w] = 2 x (clock64 ticks of the loop) + role.
template <int MODE>
__global__ __launch_bounds__(256) void k_issue_split(const rate_half8 *in, float *out,
                                                     long long *ticks, int iters) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int role = MODE == 2 ? ((w + (int)blockIdx.x) & 1) : MODE;
  const rate_half8 a = in[l], b = in[64 + l];
  rate_float4 c[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) c[j] = rate_float4{0, 0, 0, 0};
  float v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (float)a[0] + l * 1e-3f + j;
  const float m = 0.999f, s = 1e-4f;
  const long long t0 = clock64();
  if (role == 0) {
    for (int it = 0; it < iters; ++it) {
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int j = 0; j < 4; ++j) c[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c[j], 0, 0, 0);
    }
  } else {
    for (int it = 0; it < iters; ++it) {
#pragma unroll
      for (int r = 0; r < 12; ++r)
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = rate_fma(v[j], m, s);
    }
  }
  const long long t1 = clock64();
  float t = 0.0f;
#pragma unroll
  for (int j = 0; j < 4; ++j) t += c[j][0] + c[j][1] + c[j][2] + c[j][3];
#pragma unroll
  for (int j = 0; j < 8; ++j) t += v[j];
  out[((int)blockIdx.x * 4 + w) * 64 + l] = t;
  if (l == 0) ticks[blockIdx.x * 4 + w] = (t1 - t0) * 2 + role;
}

// Median ticks per loop iteration of the MFMA-role and FMA-role waves of one split launch of
// `blocks` four-wave blocks (after an untimed launch). host holds 4 x blocks values.
template <int MODE>
static hipError_t probe_split(const rate_half8 *in, float *out, long long *ticks, long long *host,
                              int blocks, int iters, double *t_mfma, double *t_fma) {
  hipError_t e = hipSuccess;
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
    k_issue_split<MODE><<<dim3(blocks), dim3(256)>>>(in, out, ticks, iters);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  const int nw = 4 * blocks;
  if (e == hipSuccess) e = hipMemcpy(host, ticks, sizeof(long long) * nw, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  std::sort(host, host + nw);  // role bit is below the tick count: sorts by ticks
  long long *r0 = host + nw, *r1 = host + 2 * nw;  // scratch halves of host
  int n0 = 0, n1 = 0;
  for (int i = 0; i < nw; ++i) {
    if (host[i] & 1) r1[n1++] = host[i] >> 1;
    else r0[n0++] = host[i] >> 1;
  }
  *t_mfma = n0 ? (double)r0[n0 / 2] / iters : 0.0;
  *t_fma = n1 ? (double)r1[n1 / 2] / iters : 0.0;
  return hipSuccess;
}

// FP32 add with a DPP row_shr:1 operand (the wavefront's shift), 8 independent chains (rate) or
// one chain on one wave per SIMD (dependent-issue latency); KIND 1: FP32 FMA, one chain.
template <int KIND, int NCH>
__global__ __launch_bounds__(64) void k_issue_dep(const rate_half8 *in, float *out,
                                                  long long *ticks, int iters) {
  const int l = threadIdx.x;
  float v[8], z[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    v[j] = (float)in[l][0] + l * 1e-3f + j;
    z[j] = (float)in[64 + l][0] + 0.5f * j;
  }
  const float m = 0.999f;
  const long long t0 = clock64();
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int j = 0; j < NCH; ++j) {
        if constexpr (KIND == 0)
          v[j] = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v[j]), 0x111, 0xf,
                                                            0xf, true)) + z[j];
        else
          v[j] = __builtin_fmaf(v[j], m, z[j]);
      }
  }
  const long long t1 = clock64();
  float t = 0.0f;
#pragma unroll
  for (int j = 0; j < 8; ++j) t += v[j];
  out[blockIdx.x * 64 + l] = t;
  if (l == 0) ticks[blockIdx.x] = t1 - t0;
}

template <int KIND, int NCH>
static hipError_t probe_dep(const rate_half8 *in, float *out, long long *ticks, long long *host,
                            int blocks, int iters, double *median) {
  hipError_t e = hipSuccess;
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
    k_issue_dep<KIND, NCH><<<dim3(blocks), dim3(64)>>>(in, out, ticks, iters);
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
  const int sblocks = 4 * cus; // four-wave blocks: two waves per SIMD (one of each role)
  const int iters = 2000;
  rate_half8 *in = nullptr;
  float *buf = nullptr;
  long long *ticks = nullptr;
  // host: the rate probes read `blocks` values, the split probes 4 x sblocks plus two halves
  const int nhost = 3 * 4 * sblocks > blocks ? 3 * 4 * sblocks : blocks;
  long long *host = new long long[nhost];
  const size_t nout = (size_t)(4 * sblocks > blocks ? 4 * sblocks : blocks) * 64;
  const size_t nticks = (size_t)(4 * sblocks > blocks ? 4 * sblocks : blocks);
  hipError_t e = hipMalloc((void **)&in, 128 * sizeof(rate_half8));
  if (e == hipSuccess) e = hipMemset(in, 0, 128 * sizeof(rate_half8));
  if (e == hipSuccess) e = hipMalloc((void **)&buf, nout * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&ticks, nticks * sizeof(long long));
  double t_mfma = 0, t_f32 = 0, t_f64 = 0, t_mix = 0;
  if (e == hipSuccess) e = probe<4, 0, float>(in, buf, ticks, host, blocks, iters, &t_mfma);
  if (e == hipSuccess) e = probe<0, 4, float>(in, buf, ticks, host, blocks, iters, &t_f32);
  if (e == hipSuccess) e = probe<0, 4, double>(in, buf, ticks, host, blocks, iters, &t_f64);
  if (e == hipSuccess) e = probe<4, 4, float>(in, buf, ticks, host, blocks, iters, &t_mix);
  // split roles: one MFMA wave beside one FMA wave per SIMD, and two of each
  double sm1 = 0, sf1 = 0, sx1m = 0, sx1f = 0, sm2 = 0, sf2 = 0, sx2m = 0, sx2f = 0, dummy = 0;
  const bool more = n >= 12;
  if (more) {
    if (e == hipSuccess) e = probe_split<0>(in, buf, ticks, host, cus, iters, &sm1, &dummy);
    if (e == hipSuccess) e = probe_split<1>(in, buf, ticks, host, cus, iters, &dummy, &sf1);
    if (e == hipSuccess) e = probe_split<2>(in, buf, ticks, host, 2 * cus, iters, &sx1m, &sx1f);
    if (e == hipSuccess) e = probe_split<0>(in, buf, ticks, host, 2 * cus, iters, &sm2, &dummy);
    if (e == hipSuccess) e = probe_split<1>(in, buf, ticks, host, 2 * cus, iters, &dummy, &sf2);
    if (e == hipSuccess) e = probe_split<2>(in, buf, ticks, host, sblocks, iters, &sx2m, &sx2f);
  }
  double t_dpp = 0, l_fma = 0, l_dpp = 0;
  if (more) {
    if (e == hipSuccess) e = probe_dep<0, 8>(in, buf, ticks, host, blocks, iters, &t_dpp);
    if (e == hipSuccess) e = probe_dep<1, 1>(in, buf, ticks, host, 4 * cus, iters, &l_fma);
    if (e == hipSuccess) e = probe_dep<0, 1>(in, buf, ticks, host, 4 * cus, iters, &l_dpp);
  }
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
  if (more) {
    // MFMA waves beside FMA waves of one SIMD: the slower role's time over the sum of the two
    // roles' times alone with the same waves per SIMD (1: no overlap; max/sum: full overlap)
    const double sum1 = sm1 + sf1, sum2 = sm2 + sf2;
    out[7] = sum1 > 0 ? (sx1m > sx1f ? sx1m : sx1f) / sum1 : 0.0;   // one wave of each role
    out[8] = sum1 > 0 ? (sm1 > sf1 ? sm1 : sf1) / sum1 : 0.0;       // the full-overlap ratio
    out[9] = sum2 > 0 ? (sx2m > sx2f ? sx2m : sx2f) / sum2 : 0.0;   // two waves of each role
    out[10] = t_dpp / (2 * 64.0);      // ticks per FP32 add with a DPP operand, per SIMD
    out[11] = l_fma / 8.0;             // dependent FP32 FMA: ticks from issue to dependent issue
    if (n >= 13) out[12] = l_dpp / 8.0;  // the same for the DPP add
  }
  return SIGPDE_OK;
}
