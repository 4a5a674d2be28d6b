// kernels_f32x.hip -- FP32 storage (f - w) with FP64 arithmetic.
#include "kernels_impl.h"

namespace tlbm {
const LaunchOps &ops_f32x() { return Launch<ST_F32X>::table(); }
}  // namespace tlbm
