/*
 * lbm_oracle.h -- CPU oracle for the tLBM partition time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * It restates, on one CPU thread and in the reference's own node order,
 * what src/TLBM_Partition.cpp:738-837 and 910-947 do with the per-node
 * math of include/LatticeStructure.hpp:69-177 and
 * include/D3Q{15,19,27}LatticeStructure.hpp for a single MPI rank
 * (the whole lattice is one partition, every node is "interior",
 * nodes are processed in ascending global index g = x + nx*y + nx*ny*z).
 *
 * Two precisions are compiled from one source (lbm_oracle_impl.h):
 *   *_f64 : every quantity is double (the reference built with real=double)
 *   *_f32 : real=float with the reference's double-literal promotions kept
 *           expression by expression (the reference as shipped,
 *           include/TLBM_definitions.h:11).
 */
#ifndef TLBM_ORACLE_H
#define TLBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_params {
    int    lattice;       /* 15, 19 or 27 */
    int    nx, ny, nz;
    int    dynamics;      /* 1 = BGK relax; any other value: no relaxation */
    double cs;            /* Smagorinsky constant, active when > 0 */
    double rho_lbm;       /* outlet density and initial density */
    double u_lbm;         /* inlet / prescribed u_z */
    double omega;         /* BGK relaxation rate 1/tau */
    int    exact_weights; /* 0: the reference's float-literal weights */
} oracle_params;

typedef struct oracle_f64 oracle_f64;
typedef struct oracle_f32 oracle_f32;

/* ndType holds nx*ny*nz ints in global order.  Returns NULL on a bad
 * lattice, bad size or a node type outside {0,1,2,3,5}. */
oracle_f64 *oracle_create_f64(const oracle_params *p, const int *ndType);
oracle_f32 *oracle_create_f32(const oracle_params *p, const int *ndType);
void oracle_destroy_f64(oracle_f64 *o);
void oracle_destroy_f32(oracle_f32 *o);

/* Run nsteps time steps (TLBM_Partition::take_LBM_time_step). */
void oracle_step_f64(oracle_f64 *o, int nsteps);
void oracle_step_f32(oracle_f32 *o, int nsteps);

/* Moments of the last step, as write_data() dumps them: ux, uy, uz, rho. */
void oracle_get_macros_f64(const oracle_f64 *o, double *ux, double *uy, double *uz, double *rho);
void oracle_get_macros_f32(const oracle_f32 *o, float *ux, float *uy, float *uz, float *rho);

/* Momentum-exchange force of the last step (TLBM_Partition.cpp:143-208). */
void oracle_get_forces_f64(const oracle_f64 *o, double *fx, double *fy, double *fz);
void oracle_get_forces_f32(const oracle_f32 *o, float *fx, float *fy, float *fz);

/* Populations the next step will read (reference fIn after the swap),
 * node-major: f[g*Q + s]. */
void oracle_get_populations_f64(const oracle_f64 *o, double *f);
void oracle_get_populations_f32(const oracle_f32 *o, float *f);

/* Set the populations the next step reads (node-major f[g*Q + s]) in both
 * buffers: set_data_member (TLBM_Partition.cpp:866-870) for every member. */
void oracle_set_populations_f64(oracle_f64 *o, const double *f);
void oracle_set_populations_f32(oracle_f32 *o, const float *f);

/* Learned closure in place of Smagorinsky: the network (1345 floats, the
 * layout of tlbm_gpu.h) on the stress features times cl_in, inferred as the
 * collide kernel does (bf16), tau = 1/omega + cl_out*softplus(z).
 * params = NULL switches it off. */
void oracle_set_closure_f64(oracle_f64 *o, const float *params, double cl_in, double cl_out);
void oracle_set_closure_f32(oracle_f32 *o, const float *params, double cl_in, double cl_out);

/* Weight table the oracle uses (Q entries), for table checks. */
void oracle_get_weights_f64(const oracle_f64 *o, double *w);

#ifdef __cplusplus
}
#endif
#endif
