/*
 * lbm_oracle.c -- CPU oracle for the tLBM partition time step.
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against,
 * never part of the product.  See lbm_oracle.h for the contract and
 * lbm_oracle_impl.h for the restated algorithm with its reference citations.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"

/* ORACLE_PLANT selects one deliberately wrong variant of the restatement
 * (tests/test_oracle.py builds them to show the golden fixtures catch each
 * mistake); 0, the default, is the faithful restatement. */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
#include "oracle_lattice.h"

void mlp_infer_bf16(const float *params, const float *x, int n, double *z_out);

#define REAL double
#define FN(name) name##_f64
#include "lbm_oracle_impl.h"
#undef REAL
#undef FN

#define REAL float
#define FN(name) name##_f32
#include "lbm_oracle_impl.h"
#undef REAL
#undef FN

void oracle_get_weights_f64(const oracle_f64 *o, double *w)
{
    for (int s = 0; s < o->Q; ++s) w[s] = o->w[s];
}
