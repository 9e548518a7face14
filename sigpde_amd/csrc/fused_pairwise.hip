// Fused linear-kernel pairwise solve (precisions "fp16x2" and "fp32"): four pairs (x_b, y_b) per
// wave, the increment Gram tiles produced on MFMA into the wavefront's LDS ring. Same arithmetic
// as the Gram kernel in fused_linear.hip (reference: sigpde/torch/kernel.py:13-16 with
// sigpde/torch/static_kernels.py:7-8, solved as sigpde/cuda_pairwise_kernels.py:8-85).
#include "mfma_producer.h"

namespace sigpde {

// Pairwise: a wave solves four pairs (x_b, y_b). GL: boundary lines in global memory at gl (one
// block per wave of the launch), else in LDS.
template <typename TO, int KB, int PREC, int R, bool GL>
__global__ __launch_bounds__(64) void k_linear_pairwise_fused(GramFusedArgs<TO> a,
                                                              const double *sxd,
                                                              const double *syd, void *gl) {
  using Prod = MfmaProducer<KB, 4, 4, PREC>;
  using E = typename Prod::elt;
  constexpr Arith A = FusedArith<PREC>::A;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int G = 16;
  const int lane = threadIdx.x & 63;
  PairGeom<G> geo(a.m, a.n, a.order, 1, R);
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
  P.ystep = 0;
  P.invy = a.invy;
  P.sx = sxd;
  P.sy = syd;
  P.kmult = a.kmult;
  P.ny = a.By;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int64_t b = (int64_t)blockIdx.x * 4 + g;
    P.xa[g] = b < a.Bx ? b : -1;
    P.ybase[g] = b < a.By ? b : a.By - 1;
  }
  const int64_t bme = (int64_t)blockIdx.x * 4 + lane / G;
  {
    const int64_t xc = bme < a.Bx ? bme : a.Bx - 1;
    P.ylane = bme < a.By ? bme : a.By - 1;
    P.xsl = Prod::kSplit ? a.kmult * (double)a.invx[xc] * (sxd ? sxd[xc] : 1.0) : 1.0;
    P.pmY = 1.0;  // the y scale is read when the pair's first chunk is multiplied
  }
  auto write = [&](int, double r) {
    if (bme < a.Bx) a.out[bme] = (TO)r;
  };
  if constexpr (A == Arith::MixedF)
    ring16f_solve<R, GL>(P, geo, smem_raw, write, fused_wave_lines<A, GL>(gl, geo.N));
  else
    ring16_solve<A, R, GL>(P, geo, 1.0, smem_raw, write, fused_wave_lines<A, GL>(gl, geo.N));
}

template <typename T, int KB, int PREC>
hipError_t launch_linear_pairwise_fused(const GramFusedArgs<T> &a, const double *sxd,
                                        const double *syd, hipStream_t st) {
  constexpr Arith A = FusedArith<PREC>::A;
  const int64_t waves = (a.Bx + 3) / 4;
  const dim3 grid((unsigned)waves), block(64);
  const int N = a.n << a.order;
  auto go = [&](auto rc) -> hipError_t {
    constexpr int R = decltype(rc)::value;
    const bool gl = fused_lines_in_global<A>(
        reinterpret_cast<const void *>(&k_linear_pairwise_fused<T, KB, PREC, R, false>),
        reinterpret_cast<const void *>(&k_linear_pairwise_fused<T, KB, PREC, R, true>), N, waves,
        st);
    return with_fused_lines<A>(gl, N, waves, st, [&](void *lines, size_t shmem) {
      if (lines)
        k_linear_pairwise_fused<T, KB, PREC, R, true><<<grid, block, shmem, st>>>(a, sxd, syd,
                                                                                  lines);
      else
        k_linear_pairwise_fused<T, KB, PREC, R, false><<<grid, block, shmem, st>>>(a, sxd, syd,
                                                                                   nullptr);
    });
  };
  switch (r16_rows(a.order)) {
    case 1: return go(std::integral_constant<int, 1>());
    case 2: return go(std::integral_constant<int, 2>());
    case 4: return go(std::integral_constant<int, 4>());
    default: return go(std::integral_constant<int, 8>());
  }
}

#define SIGPDE_PAIRWISE_INST(T, KB)                                                               \
  template hipError_t launch_linear_pairwise_fused<T, KB, kPrecFp16x2F>(                         \
      const GramFusedArgs<T> &, const double *, const double *, hipStream_t);                      \
  template hipError_t launch_linear_pairwise_fused<T, KB, kPrecFp16x2>(                          \
      const GramFusedArgs<T> &, const double *, const double *, hipStream_t);          \
  template hipError_t launch_linear_pairwise_fused<T, KB, kPrecFp32>(                            \
      const GramFusedArgs<T> &, const double *, const double *, hipStream_t);         \
  template hipError_t launch_linear_pairwise_fused<T, KB, kPrecBf16x3>(                          \
      const GramFusedArgs<T> &, const double *, const double *, hipStream_t);
SIGPDE_PAIRWISE_INST(float, 1)
SIGPDE_PAIRWISE_INST(float, 2)
SIGPDE_PAIRWISE_INST(float, 3)
SIGPDE_PAIRWISE_INST(float, 4)
SIGPDE_PAIRWISE_INST(double, 1)
SIGPDE_PAIRWISE_INST(double, 2)
SIGPDE_PAIRWISE_INST(double, 3)
SIGPDE_PAIRWISE_INST(double, 4)
#undef SIGPDE_PAIRWISE_INST
template hipError_t launch_linear_pairwise_fused<float, 1, kPrecFp64>(
    const GramFusedArgs<float> &, const double *, const double *, hipStream_t);
template hipError_t launch_linear_pairwise_fused<float, 2, kPrecFp64>(
    const GramFusedArgs<float> &, const double *, const double *, hipStream_t);
template hipError_t launch_linear_pairwise_fused<double, 1, kPrecFp64>(
    const GramFusedArgs<double> &, const double *, const double *, hipStream_t);
template hipError_t launch_linear_pairwise_fused<double, 2, kPrecFp64>(
    const GramFusedArgs<double> &, const double *, const double *, hipStream_t);

}  // namespace sigpde
