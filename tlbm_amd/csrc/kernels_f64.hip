// kernels_f64.hip -- FP64 storage and arithmetic.  Built with
// -ffp-contract=off: every multiply and add rounds on its own, as in the
// reference built with real=double on x86-64 (no FMA), so the FP64 path
// reproduces the reference's results.
#include "kernels_impl.h"

namespace tlbm {
const LaunchOps &ops_f64() { return Launch<ST_F64>::table(); }
}  // namespace tlbm
