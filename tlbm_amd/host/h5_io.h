// h5_io.h -- HDF5 files of a run: the per-dump out<k>.h5 / data<k>.xmf that
// python/processTLBM.py + python/hdf5Helper.py make from the .b_dat dumps,
// and the restart.h5 that src/TLBM_Partition.cpp:481-540 reads.  Built with
// the HDF5 C++ library when TLBM_WITH_HDF5 is defined; without it both
// functions throw std::runtime_error.
#pragma once

#include <string>
#include <vector>

#include "Problem.h"

namespace tlbm_host {

// Datasets velo_group/{x_velo,y_velo,z_velo,velmag}, force_group/{x,y,z}_force
// and pres_group/presmag in physical units (hdf5Helper.py:4-31 with the
// conversions of processTLBM.py:70-113), float32, plus the XDMF description
// (hdf5Helper.py:58-139).
void write_dump_h5(const std::string &h5name, const std::string &xmfname, const Problem &p,
                   const std::vector<double> &ux, const std::vector<double> &uy,
                   const std::vector<double> &uz, const std::vector<double> &rho,
                   const std::vector<double> &fx, const std::vector<double> &fy,
                   const std::vector<double> &fz);

// timeAvg.h5 / timeAvg.xmf from the time-average sums: each sum divided by
// Num_ts, then velo_group/{x_velo,y_velo,z_velo,velmag} and
// pres_group/presmag in physical units as for a dump, no forces
// (processTLBM.py:134-186, hdf5Helper.py:33-56, 140-200).
void write_time_avg_h5(const std::string &h5name, const std::string &xmfname, const Problem &p,
                       const std::vector<double> &u_sum, const std::vector<double> &v_sum,
                       const std::vector<double> &w_sum, const std::vector<double> &rho_sum);

// /velocity/{x,y,z} and /density/rho, float32, global node order
// (python/tLBM_create_restart.py:55-66), read as float32 and widened.
void read_restart(const std::string &path, size_t n, std::vector<double> &ux,
                  std::vector<double> &uy, std::vector<double> &uz, std::vector<double> &rho);

bool hdf5_available();

}  // namespace tlbm_host
