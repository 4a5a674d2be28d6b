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
// TLBM_HDF5=1 also writes out<k>.h5 / data<k>.xmf on every dump (the files
// python/processTLBM.py makes from the dumps), when built with HDF5.
#pragma once

#include <cstdint>
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
    // Population spd of node nd that the next step reads (the reference's
    // get_data_member / set_data_member on its fOut buffer, which its
    // callers reach through private pointers).
    double get_data_member(int nd, int spd);
    void set_data_member(double val, int nd, int spd);
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
