// TLBM_Partition.h -- the reference's partition object, backed by one GPU.
//
// Drop-in for the class of include/TLBM_Partition.h:33-133 as src/tLBM.cpp
// drives it: construct (reads params.lbm, ndType.lbm, parts.lbm and, with
// restartFlag, restart.h5 from the working directory), take_LBM_time_step
// once per step, write_data on dump steps, write_time_avg_data at the end,
// and the same getters.  The whole lattice is one partition resident on the
// current GPU through the C-ABI of include/tlbm_gpu.h, so there is no MPI
// communicator and no halo.
//
// Storage of the populations (the reference's compile-time `real`) is chosen
// at run time by the environment variable TLBM_STORAGE = f64 | f32 | f16
// (default f32, the reference's real = float); .b_dat dumps are written as
// double for f64 and as float otherwise.  TLBM_EXACT_WEIGHTS=1 selects the
// exact D3Q15/D3Q19 weights instead of the reference's float literals.
// TLBM_HDF5=1 also writes out<k>.h5 / data<k>.xmf on every dump and
// timeAvg.h5 / timeAvg.xmf with the time average (the files
// python/processTLBM.py makes from the dumps), when built with HDF5.
// TLBM_CHECK_FINITE=0 turns off the driver's divergence guard.
#pragma once

#include <cstdint>
#include <set>
#include <string>
#include <vector>

#include "Problem.h"
#include "tlbm_gpu.h"

struct LatticeIndex {
    // a 3-tuple of lattice indices, x fastest (TLBM_Partition.h:24-31)
    int X;
    int Y;
    int Z;
};

class TLBM_Partition {
public:
    Problem thisProblem;

    // rank/size replace (rank, size, MPI_Comm) of TLBM_Partition.cpp:10-14;
    // one GPU holds the whole lattice, so size must be 1.
    explicit TLBM_Partition(int rank = 0, int size = 1);
    ~TLBM_Partition();
    TLBM_Partition(const TLBM_Partition &) = delete;
    TLBM_Partition &operator=(const TLBM_Partition &) = delete;

    LatticeIndex get_xyz_index(int gInd);
    int get_tgt_index(int gInd, int ex, int ey, int ez);
    int get_gInd(int x, int y, int z);
    int get_gInd(LatticeIndex myXYZ);
    int get_cut_size();  // populations crossing partitions per step: 0
    int get_num_ts();
    int get_ts_rep_freq();
    int get_plot_freq();
    int get_warmupTs();
    int get_num_global_nodes();
    int is_restart();
    int get_time_avg_flag();
    // The reference's accessors of a node-major population array
    // (include/TLBM_Partition.h:51-52, src/TLBM_Partition.cpp:860-870):
    // f[getIDx(Q, nd, spd)].  Its own arrays are private and the state here
    // lives in HBM, so a caller's array is filled by copy_populations and
    // given back by load_populations.  Real is double or float (the
    // reference's compile-time `real`).
    template <class Real>
    Real get_data_member(const Real *f, const int nd, const int spd) {
        return f[getIDx(Q_, nd, spd)];
    }
    template <class Real>
    void set_data_member(Real *f, const Real val, const int nd, const int spd) {
        f[getIDx(Q_, nd, spd)] = val;
    }
    // include/TLBM_Partition.h:57-60: node-major index of a population.
    static inline unsigned getIDx(int nSpd, int nIdx, int spd) {
        return unsigned(nIdx) * unsigned(nSpd) + unsigned(spd);
    }
    // The populations the next step reads, node-major (N * Q values), and
    // their replacement (the reference's fOut buffer).
    void copy_populations(double *f);
    void load_populations(const double *f);
    // Population spd of node nd that the next step reads, one at a time
    // (a convenience over the two calls above).
    double get_data_member(int nd, int spd);
    void set_data_member(double val, int nd, int spd);
    // include/TLBM_Partition.h:56: the reference updates the nodes of a
    // list (boundary nodes, then interior nodes) from fIn into fOut.  One
    // kernel updates every node of the lattice per step here, and the
    // buffers are in HBM, so a partial update has no counterpart: it throws
    // std::logic_error.  take_LBM_time_step is the whole step.
    void process_node_list(double *fOut, double *fIn, const std::set<int> &nodeList);
    void process_node_list(float *fOut, float *fIn, const std::set<int> &nodeList);
    // One time step (TLBM_Partition.cpp:910-947).  isEven only names which
    // of the two buffers the reference reads; the GPU handle keeps its own
    // ping-pong.  Queued on the GPU: returns before the step completes.
    void take_LBM_time_step(bool isEven);
    // ux, uy, uz, density, Fx, Fy, Fz of the last step into <field><k>.b_dat
    // (TLBM_Partition.cpp:608-683).
    void write_data();
    // uAvg, vAvg, wAvg, rhoAvg sums into *.b_dat (TLBM_Partition.cpp:564-606).
    void write_time_avg_data();
    void print_adjacency(const int nIdx);

    int num_speeds() const { return Q_; }
    int storage() const { return storage_; }
    void synchronize();
    // Divergence guard (no reference counterpart: its loop, src/tLBM.cpp:
    // 45-62, steps on through NaN).  Throws std::runtime_error naming
    // time step `ts` when any population of the state is NaN or infinite.
    void check_finite(int ts);

private:
    void load_ndType(std::vector<uint8_t> &nd);
    void check_parts();
    void write_node_ordering();
    void load_restart_data();
    void sync_populations_to_device();
    template <class Real>
    void write_field(const std::string &name, const std::vector<double> &v);

    int rank_ = 0;
    int Q_ = 0;
    int storage_ = TLBM_STORAGE_F32;
    int dataWriteNum = 0;
    tlbm_gpu *gpu_ = nullptr;
    std::vector<int> ex_, ey_, ez_;
    std::vector<double> pop_cache_;  // host copy for get/set_data_member
    bool pop_cached_ = false, pop_dirty_ = false;
};
