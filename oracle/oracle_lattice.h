/*
 * oracle_lattice.h -- velocity sets, weights and boundary index lists of
 * the three lattices, as the reference defines them.  TEST INFRASTRUCTURE.
 *
 * The orderings, the float-literal weights of D3Q15/D3Q19 and the
 * double-literal weights of D3Q27 must be the reference's own, so these
 * tables follow:
 *   D3Q15: include/D3Q15LatticeStructure.hpp:39-45 (tables),
 *          :77-96 / :99-110 (inlet / outlet unknown-direction pairs),
 *          :113-134 / :137-153 (inlet / outlet macroscopic sums)
 *   D3Q19: include/D3Q19LatticeStructure.hpp:40-47, 80-104, 108-156
 *   D3Q27: include/D3Q27LatticeStructure.hpp:40-52, 84-109, 112-168
 * Note that D3Q15's non-equilibrium partners are the z-mirror images
 * (7<->11, 8<->12, ...), not the opposite directions; D3Q19/D3Q27 use
 * opposites.  The inlet "known" sum of D3Q27 is not in ascending order and
 * multiplies by the int 2; both are kept because they change float rounding.
 */
#ifndef TLBM_ORACLE_LATTICE_H
#define TLBM_ORACLE_LATTICE_H

typedef struct oracle_lattice {
    int Q;
    int ex[27], ey[27], ez[27], bb[27];
    float wf[27];      /* float-literal weights (D3Q15/D3Q19) */
    double wd[27];     /* double-literal weights (D3Q27) or exact fractions */
    int w_is_float;    /* 1: the reference writes the weights as float literals */
    int n_unk;         /* number of unknown directions at the inlet/outlet */
    int in_unk[9], in_par[9];    /* inlet: f[in_unk] = feq + f[in_par] - feq[in_par] */
    int out_unk[9], out_par[9];  /* outlet */
    int n_zero;
    int zero_list[9];            /* e_z = 0 directions, summed in this order */
    int in_list[9];              /* e_z = -1 directions as the inlet sums them */
    int out_list[9];             /* e_z = +1 directions as the outlet sums them */
    int in_int_two;              /* inlet doubles the sum with the int 2 */
} oracle_lattice;

static const oracle_lattice ORACLE_D3Q15 = {
    15,
    {0,1,-1,0,0,0,0,1,-1,1,-1,1,-1,1,-1},
    {0,0,0,1,-1,0,0,1,1,-1,-1,1,1,-1,-1},
    {0,0,0,0,0,1,-1,1,1,1,1,-1,-1,-1,-1},
    {0,2,1,4,3,6,5,14,13,12,11,10,9,8,7},
    {2.f/9.f, 1.f/9.f, 1.f/9.f, 1.f/9.f, 1.f/9.f, 1.f/9.f, 1.f/9.f,
     1.f/72.f, 1.f/72.f, 1.f/72.f, 1.f/72.f, 1.f/72.f, 1.f/72.f, 1.f/72.f, 1.f/72.f},
    {2./9., 1./9., 1./9., 1./9., 1./9., 1./9., 1./9.,
     1./72., 1./72., 1./72., 1./72., 1./72., 1./72., 1./72., 1./72.},
    1,
    5,
    {5,7,8,9,10}, {6,11,12,13,14},
    {6,11,12,13,14}, {5,7,8,9,10},
    5,
    {0,1,2,3,4},
    {6,11,12,13,14},
    {5,7,8,9,10},
    0
};

static const oracle_lattice ORACLE_D3Q19 = {
    19,
    {0,1,-1,0,0,0,0,1,-1,1,-1,1,-1,1,-1,0,0,0,0},
    {0,0,0,1,-1,0,0,1,1,-1,-1,0,0,0,0,1,-1,1,-1},
    {0,0,0,0,0,1,-1,0,0,0,0,1,1,-1,-1,1,1,-1,-1},
    {0,2,1,4,3,6,5,10,9,8,7,14,13,12,11,18,17,16,15},
    {3.f/9.f, 1.f/18.f, 1.f/18.f, 1.f/18.f, 1.f/18.f, 1.f/18.f, 1.f/18.f,
     1.f/36.f, 1.f/36.f, 1.f/36.f, 1.f/36.f, 1.f/36.f, 1.f/36.f,
     1.f/36.f, 1.f/36.f, 1.f/36.f, 1.f/36.f, 1.f/36.f, 1.f/36.f},
    {3./9., 1./18., 1./18., 1./18., 1./18., 1./18., 1./18.,
     1./36., 1./36., 1./36., 1./36., 1./36., 1./36.,
     1./36., 1./36., 1./36., 1./36., 1./36., 1./36.},
    1,
    5,
    {5,11,12,15,16}, {6,14,13,18,17},
    {6,14,13,18,17}, {5,11,12,15,16},
    9,
    {0,1,2,3,4,7,8,9,10},
    {6,13,14,17,18},
    {5,11,12,15,16},
    0
};

static const oracle_lattice ORACLE_D3Q27 = {
    27,
    {0,1,-1,0,0,0,0,1,1,-1,-1,1,1,-1,-1,0,0,0,0,1,1,1,1,-1,-1,-1,-1},
    {0,0,0,1,-1,0,0,1,-1,1,-1,0,0,0,0,1,1,-1,-1,1,1,-1,-1,1,1,-1,-1},
    {0,0,0,0,0,1,-1,0,0,0,0,1,-1,1,-1,1,-1,1,-1,1,-1,1,-1,1,-1,1,-1},
    {0,2,1,4,3,6,5,10,9,8,7,14,13,12,11,18,17,16,15,26,25,24,23,22,21,20,19},
    {0},
    {8./27., 2./27., 2./27., 2./27., 2./27., 2./27., 2./27.,
     1./54., 1./54., 1./54., 1./54., 1./54., 1./54.,
     1./54., 1./54., 1./54., 1./54., 1./54., 1./54.,
     1./216., 1./216., 1./216., 1./216., 1./216., 1./216., 1./216., 1./216.},
    0,
    9,
    {5,11,13,15,17,19,21,23,25}, {6,14,12,18,16,26,24,22,20},
    {6,14,12,18,16,26,24,22,20}, {5,11,13,15,17,19,21,23,25},
    9,
    {0,1,2,3,4,7,8,9,10},
    {6,14,12,18,16,26,24,22,20},
    {5,11,13,15,17,19,21,23,25},
    1
};

static const oracle_lattice *oracle_lattice_for(int q)
{
    switch (q) {
    case 15: return &ORACLE_D3Q15;
    case 19: return &ORACLE_D3Q19;
    case 27: return &ORACLE_D3Q27;
    default: return 0;
    }
}

#endif
