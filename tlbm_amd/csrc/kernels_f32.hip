// kernels_f32.hip -- FP32 storage and FP16 (scaled f - w) storage, both
// with FP32 arithmetic.
#include "kernels_impl.h"

namespace tlbm {
const LaunchOps &ops_f32() { return Launch<ST_F32>::table(); }
const LaunchOps &ops_f16() { return Launch<ST_F16>::table(); }
}  // namespace tlbm
