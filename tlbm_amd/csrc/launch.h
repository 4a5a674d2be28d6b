// launch.h -- host-side launchers of the kernels, one table per storage.
// The tables are defined in kernels_f64.hip (FP64, compiled without
// floating-point contraction so it keeps the reference's rounding) and
// kernels_f32.hip (FP32 and FP16 storage) and kernels_f32x.hip (FP32
// storage with FP64 arithmetic).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_kernels.h"

namespace tlbm {

struct LaunchOps {
    hipError_t (*lbm)(int lattice, int mode, const void *fin, void *fout, const uint8_t *nd,
                      const uint8_t *chunk_type, void *macro, const void *uy_head,
                      const StepArgs &a, hipStream_t s);
    hipError_t (*init_rest)(int lattice, void *fout, const StepArgs &a, hipStream_t s);
    hipError_t (*init_macros)(int lattice, void *fout, const double *mac, const StepArgs &a,
                              hipStream_t s);
    hipError_t (*get_populations)(int lattice, const void *fin, double *f_aos, const StepArgs &a,
                                  hipStream_t s);
    hipError_t (*set_populations)(int lattice, void *fout, const double *f_aos,
                                  const StepArgs &a, hipStream_t s);
    hipError_t (*forces)(int lattice, const void *fin, const uint8_t *nd, double *F,
                         const StepArgs &a, hipStream_t s);
    hipError_t (*reduce)(int lattice, const void *cur, const void *prev, double *partial,
                         int nparts, double *out, const StepArgs &a, hipStream_t s);
    hipError_t (*box_filter)(const void *macro, double *out, const StepArgs &a, hipStream_t s);
    int macro_bytes;  // size of one macro value (the compute type)
};

const LaunchOps &ops_f64();
const LaunchOps &ops_f32();
const LaunchOps &ops_f16();
const LaunchOps &ops_f32x();

}  // namespace tlbm
