/*
 * tlbm_gpu.h -- C-ABI of the MI355X lattice-Boltzmann time step.
 *
 * One handle holds one whole lattice on one GPU (the reference's
 * TLBM_Partition for a single partition).  Plain pointers and sizes only;
 * host arrays are in the reference's global node order
 * g = x + nx*y + nx*ny*z (TLBM_Partition.cpp:730-737).
 *
 * Every function returns TLBM_OK (0) or a negative TLBM_E* code and leaves
 * a message for tlbm_gpu_last_error() on the calling thread.  Steps are
 * queued on the handle's HIP stream and return before the GPU finishes;
 * every function that hands results to the host synchronises that stream.
 * A handle is used by one host thread at a time.
 */
#ifndef TLBM_GPU_H
#define TLBM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TLBM_OK 0
#define TLBM_EINVAL (-1)   /* bad argument: lattice, size, node type, mode */
#define TLBM_EHIP (-2)     /* a HIP runtime call failed */
#define TLBM_ENOMEM (-3)   /* device or host allocation failed */
#define TLBM_ESTATE (-4)   /* call not valid in the handle's current state */
#define TLBM_ENONFINITE (-5) /* the state holds NaN or infinite populations (tlbm_gpu_reduce) */

/* Population storage in HBM.  F64 stores f_s and computes in FP64 with the
 * reference's operation order.  F32 stores d_s = f_s - w_s and F16 the
 * scaled deviation (f_s - w_s)/w_s (each step's store rounded
 * stochastically); both compute in FP32 on d_s.  F32 relaxes fluid nodes
 * keeping the exact-arithmetic mass and momentum of f - feq to rounding;
 * F16 corrects its float-weight equilibrium to the balance of the weights
 * where the two differ (D3Q27, exact weights).  F32X stores d_s as F32 does and computes in FP64 with
 * the FP64 weights (no reference counterpart). */
#define TLBM_STORAGE_F64 0
#define TLBM_STORAGE_F32 1
#define TLBM_STORAGE_F16 2
#define TLBM_STORAGE_F32X 3

typedef struct tlbm_gpu tlbm_gpu;

/* The fields of Problem (include/Problem.h:10-45) the time step reads,
 * plus the storage choice. */
typedef struct tlbm_params {
    int lattice;       /* 15, 19 or 27: Problem::latticeType "D3Q15"/"D3Q19"/"D3Q27" */
    int dynamics;      /* Problem::dynamics; 1 = BGK, any other value: no relaxation */
    int nx, ny, nz;    /* Problem::nx, ny, nz */
    int storage;       /* TLBM_STORAGE_* */
    int exact_weights; /* 0: the reference's float-literal D3Q15/D3Q19 weights */
    int reserved;      /* must be 0 */
    double cs;         /* Problem::cs, Smagorinsky when > 0 */
    double rho_lbm;    /* Problem::rhoLBM: initial and outlet density */
    double u_lbm;      /* Problem::uLBM: inlet / prescribed u_z */
    double omega;      /* Problem::omega */
} tlbm_params;

/* Allocate the lattice on the current device and set f = w*rho_lbm.
 * nd_type: nx*ny*nz node types in {0,1,2,3,5}.
 * Replaces TLBM_Partition::TLBM_Partition / tlbm_initialize
 * (src/TLBM_Partition.cpp:10-14, 77-118) with load_ndType (120-141),
 * create_adj_matrix (254-271), allocate_arrays (421-446) and
 * initialize_data_arrays (448-479) on one partition.  Node types outside
 * {0,1,2,3,5} are rejected (the reference relaxes them against a stale fEq),
 * and a type-5 node with index below Q-1 is rejected (its uy[spd] term
 * would read values the reference has not computed yet). */
int tlbm_gpu_create(const tlbm_params *p, const uint8_t *nd_type, tlbm_gpu **out);

/* Free every device buffer.  ~TLBM_Partition (src/TLBM_Partition.cpp:16-45). */
int tlbm_gpu_destroy(tlbm_gpu *h);

/* Queue nsteps time steps.  TLBM_Partition::take_LBM_time_step
 * (src/TLBM_Partition.cpp:910-947) called nsteps times. */
int tlbm_gpu_step(tlbm_gpu *h, int nsteps);

/* Run nsteps steps between two HIP events on the handle's stream and
 * return their elapsed time in milliseconds (synchronises). */
int tlbm_gpu_step_timed(tlbm_gpu *h, int nsteps, double *ms);

/* Wait for queued work. */
int tlbm_gpu_sync(tlbm_gpu *h);

/* ux, uy, uz, rho of the last step, as TLBM_Partition::write_data dumps
 * them (src/TLBM_Partition.cpp:608-650): moments after the boundary rule
 * (solid u = 0, inlet u = (0,0,u_lbm), outlet rho = rho_lbm).  Each array
 * holds nx*ny*nz doubles.  Needs at least one step since create/set. */
int tlbm_gpu_get_macros(tlbm_gpu *h, double *ux, double *uy, double *uz, double *rho);

/* Momentum-exchange force on solid nodes next to fluid nodes, after the
 * last step.  TLBM_Partition::make_force_calc_map + calc_force
 * (src/TLBM_Partition.cpp:143-208) on one partition. */
int tlbm_gpu_get_forces(tlbm_gpu *h, double *fx, double *fy, double *fz);

/* Populations the next step will read, node-major f[g*Q + s]: the
 * reference's fOut after the last step, read one member at a time by
 * TLBM_Partition::get_data_member (src/TLBM_Partition.cpp:860-864). */
int tlbm_gpu_get_populations(tlbm_gpu *h, double *f);

/* Set the populations the next step will read (node-major f[g*Q + s]).
 * TLBM_Partition::set_data_member (src/TLBM_Partition.cpp:866-870) for
 * every member at once. */
int tlbm_gpu_set_populations(tlbm_gpu *h, const double *f);

/* f = f_eq(rho, u) per node: TLBM_Partition::load_restart_data
 * (src/TLBM_Partition.cpp:481-540) after the HDF5 read. */
int tlbm_gpu_init_from_macros(tlbm_gpu *h, const double *ux, const double *uy, const double *uz,
                              const double *rho);

/* FP64 reductions over the populations the next step will read:
 * out[0] = mass sum f, out[1..3] = momentum sum f*e, out[4] = relative
 * change of the state over the last step, ||B_n - B_{n-1}|| / ||B_n||
 * (0 before the first step).  The same pass counts populations that are
 * NaN or infinite: when there are any, out[] is still filled and the call
 * returns TLBM_ENONFINITE with the count in the message (a diverged run).
 * No reference counterpart: the reference has no conservation, residual
 * or divergence checks (its loop, src/tLBM.cpp:45-62, steps on). */
int tlbm_gpu_reduce(tlbm_gpu *h, double out[5]);

/* The sums behind tlbm_gpu_reduce, written to device memory and queued on
 * the handle's stream: out_dev (device pointer, TLBM_REDUCE_SUMS doubles)
 * gets sum f, sum f*e (x, y, z), sum (B_n - B_{n-1})^2 (0 before the first
 * step), sum B_n^2 and the number of NaN or infinite populations.  It
 * allocates nothing and does not synchronise (read out_dev after
 * tlbm_gpu_sync): its partial sums live in a fixed device array of
 * 1024 x 7 doubles (56 KiB per GPU, held until the process ends), and a
 * lattice whose reduction needs more partials is rejected (TLBM_EINVAL;
 * none does today: the partial count is capped at 1024).  Every handle on
 * a GPU shares that GPU's array, so calls whose reductions may run at the
 * same time on one GPU (other threads, other handles' streams) race.
 * No counterpart in the reference: for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops). */
#define TLBM_REDUCE_SUMS 7
int tlbm_gpu_reduce_device(tlbm_gpu *h, double *out_dev);

/* enable = 1: zero the time-average sums and from then on add ux, uy, uz
 * and rho of every step to them on the device; enable = 0: stop adding.
 * TLBM_Partition::update_time_avg (src/TLBM_Partition.cpp:815-824) driven
 * by Problem::timeAvgFlag (the reference reads an uninitialised flag).
 * The sums are FP64. */
int tlbm_gpu_time_average(tlbm_gpu *h, int enable);

/* The time-average sums (not divided by the step count, as
 * TLBM_Partition::write_time_avg_data writes them, src/TLBM_Partition.cpp:
 * 564-606) and the number of steps summed.  Arrays of nx*ny*nz doubles. */
int tlbm_gpu_get_time_average(tlbm_gpu *h, double *u, double *v, double *w, double *rho,
                              int64_t *nsteps);

/* Inlet velocity per inlet column: type-2 nodes at (x, y) take
 * u_z = uz[x + nx*y] (nx*ny doubles, stored as float) instead of u_lbm
 * (uz = NULL: back to u_lbm).  No reference counterpart: the reference's
 * inlet is uniform (D3Q19LatticeStructure.hpp:108-132). */
int tlbm_gpu_set_inlet_profile(tlbm_gpu *h, const double *uz);

/* ---- learned closure (no reference counterpart) --------------------------
 * A network of TLBM_MLP_NPARAM floats -- W1[32][7], b1[32], W2[32][32], b2[32],
 * w3[32], b3 -- maps a node's non-equilibrium stress (xx, yy, zz, xy, xz, yz
 * and its norm, the inputs of the Smagorinsky rule of
 * include/LatticeStructure.hpp:140-153), each times feature_scale, through
 * relu(32) -> relu(32) -> z to the relaxation time
 * tau = 1/omega + output_scale * softplus(z).  It runs inside the collide
 * kernel on bf16 MFMA (FP32 accumulation). */
#define TLBM_MLP_NPARAM 1345

/* Use the network in place of Smagorinsky for the following steps
 * (params = NULL: back to the Smagorinsky / BGK rule).  FP32/FP16 storage. */
int tlbm_gpu_set_closure(tlbm_gpu *h, const float *params, double feature_scale,
                         double output_scale);

/* 2x2x2 box filter of the last step's moments onto the (nx/2, ny/2, nz/2)
 * lattice (all even): TLBM_FILTER_PLANES planes of doubles -- rho, ux, uy,
 * uz, uxux, uyuy, uzuz, uxuy, uxuz, uyuz, then the viscous stress
 * sigma = -(1 - omega/2) Pi_neq (xx, yy, zz, xy, xz, yz; omega of the node's
 * BGK or Smagorinsky relaxation, zero on nodes that do not relax) -- means
 * of the 8 fine nodes, coarse global order.  Not available while a closure
 * network is set (TLBM_ESTATE).  No reference counterpart: it makes the
 * closure's training data. */
#define TLBM_FILTER_PLANES 16
int tlbm_gpu_box_filter(tlbm_gpu *h, double *out);

typedef struct tlbm_mlp_trainer tlbm_mlp_trainer;

/* A training set of n samples (x: n*7 floats, t: n targets of
 * softplus(z)) on the device, from initial parameters; batches of at most
 * max_batch samples. */
int tlbm_mlp_trainer_create(const float *x, const float *t, int64_t n, const float *init,
                            uint64_t seed, int max_batch, tlbm_mlp_trainer **out);
/* steps Adam steps (beta 0.9/0.999, eps 1e-8) on the mean of
 * (softplus(z) - t)^2 / 2 over batches of `batch` hashed sample indices;
 * last_loss (may be NULL) gets the last batch's loss. */
int tlbm_mlp_trainer_step(tlbm_mlp_trainer *tr, int steps, int batch, float lr,
                          double *last_loss);
/* Loss and gradient of samples [first, first+count) at the current
 * parameters (no update). */
int tlbm_mlp_trainer_loss_grad(tlbm_mlp_trainer *tr, int64_t first, int64_t count, double *loss,
                               float *grad);
int tlbm_mlp_trainer_params(tlbm_mlp_trainer *tr, float *params);
int tlbm_mlp_trainer_destroy(tlbm_mlp_trainer *tr);

/* z of n feature vectors through the collide kernel's wavefront inference
 * (the same device code), for checking it. */
int tlbm_mlp_infer(const float *params, const float *x, int64_t n, float *z);

/* Bytes of device memory the handle holds. */
int tlbm_gpu_memory_bytes(const tlbm_gpu *h, uint64_t *bytes);

/* Message of the last failed call on this thread ("" if none). */
const char *tlbm_gpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
