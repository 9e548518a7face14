// Fused reference-scheme Gram, L1-RBF: distance sums on the VALU from x rows staged in LDS
// (fused_point.h).
#include "fused_point.h"

namespace sigpde {

template <typename T, int KIND, int NE>
hipError_t launch_point_gram(PointGramArgs<T> a, hipStream_t st) {
  return launch_point_gram_impl<T, KIND, NE>(a, st);
}

template hipError_t launch_point_gram<float, kPointL1, 8>(PointGramArgs<float>, hipStream_t);
template hipError_t launch_point_gram<float, kPointL1, 16>(PointGramArgs<float>, hipStream_t);
template hipError_t launch_point_gram<double, kPointL1, 8>(PointGramArgs<double>, hipStream_t);

}  // namespace sigpde
