/*
 * ref_fp64_defs.h -- selects the reference's compile-time precision switch
 * (include/TLBM_definitions.h:11,14: `using real = float;` and
 * `#define MPI_DTYPE MPI_FLOAT`) as double for the FP64 reference build.
 * Force-included ahead of the reference sources, it defines that header's
 * include guard so the float definitions are skipped; no reference file is
 * edited.  TEST INFRASTRUCTURE ONLY.
 */
#define INCLUDE_TLBM_DEFINTIONS_H_
using real = double;
#define MPI_DTYPE MPI_DOUBLE
