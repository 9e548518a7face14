// Fused reference-scheme Gram, float64 paths: linear and RBF point values on
// v_mfma_f64_16x16x4f64 (fused_point.h).
#include "fused_point.h"

namespace sigpde {

template <typename T, int KIND, int NE>
hipError_t launch_point_gram(PointGramArgs<T> a, hipStream_t st) {
  return launch_point_gram_impl<T, KIND, NE>(a, st);
}

#define SIGPDE_POINT_INST(KIND, NE) \
  template hipError_t launch_point_gram<double, KIND, NE>(PointGramArgs<double>, hipStream_t);
SIGPDE_POINT_INST(kPointLinear, 2)
SIGPDE_POINT_INST(kPointLinear, 4)
SIGPDE_POINT_INST(kPointLinear, 8)
SIGPDE_POINT_INST(kPointRBF, 2)
SIGPDE_POINT_INST(kPointRBF, 4)
SIGPDE_POINT_INST(kPointRBF, 8)
#undef SIGPDE_POINT_INST

}  // namespace sigpde
