// lattice.h -- velocity sets, weights and boundary index lists of the
// D3Q15, D3Q19 and D3Q27 lattices, usable on host and device.
//
// The orderings and boundary lists have to be the reference's because the
// population index s is part of the drop-in contract (the same s appears in
// get_data_member/set_data_member and in any restart), so they follow:
//   D3Q15: include/D3Q15LatticeStructure.hpp:39-45, 77-110, 113-153
//   D3Q19: include/D3Q19LatticeStructure.hpp:40-47, 80-104, 108-156
//   D3Q27: include/D3Q27LatticeStructure.hpp:40-52, 84-109, 112-168
// The reference writes the D3Q15/D3Q19 weights as float literals (1.f/18.f),
// so in double precision they are the float-rounded values; `w_ref` keeps
// those, `w_exact` holds the exact fractions (opt-in, changes mass balance).
// D3Q15 pairs each unknown inlet/outlet direction with its z-mirror image
// (7<->11, 8<->12, 9<->13, 10<->14), D3Q19/D3Q27 with the opposite one; the
// D3Q27 inlet sums its e_z=-1 populations in the order 6,14,12,18,16,...
// which changes double rounding, so the lists keep that order.
#pragma once

namespace tlbm {

struct D3Q15 {
    static constexpr int Q = 15;
    static constexpr int ex[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
    static constexpr int ey[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int ez[Q] = {0, 0, 0, 0, 0, 1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
    static constexpr int bb[Q] = {0, 2, 1, 4, 3, 6, 5, 14, 13, 12, 11, 10, 9, 8, 7};
    static constexpr double w_ref[Q] = {
        double(2.f / 9.f),  double(1.f / 9.f),  double(1.f / 9.f),  double(1.f / 9.f),
        double(1.f / 9.f),  double(1.f / 9.f),  double(1.f / 9.f),  double(1.f / 72.f),
        double(1.f / 72.f), double(1.f / 72.f), double(1.f / 72.f), double(1.f / 72.f),
        double(1.f / 72.f), double(1.f / 72.f), double(1.f / 72.f)};
    static constexpr double w_exact[Q] = {2. / 9.,  1. / 9.,  1. / 9.,  1. / 9.,  1. / 9.,
                                          1. / 9.,  1. / 9.,  1. / 72., 1. / 72., 1. / 72.,
                                          1. / 72., 1. / 72., 1. / 72., 1. / 72., 1. / 72.};
    static constexpr int n_unk = 5;
    static constexpr int in_unk[n_unk] = {5, 7, 8, 9, 10};
    static constexpr int in_par[n_unk] = {6, 11, 12, 13, 14};
    static constexpr int out_unk[n_unk] = {6, 11, 12, 13, 14};
    static constexpr int out_par[n_unk] = {5, 7, 8, 9, 10};
    static constexpr int in_list[n_unk] = {6, 11, 12, 13, 14};
    static constexpr int out_list[n_unk] = {5, 7, 8, 9, 10};
    static constexpr int n_zero = 5;
    static constexpr int zero_list[n_zero] = {0, 1, 2, 3, 4};
};

struct D3Q19 {
    static constexpr int Q = 19;
    static constexpr int ex[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int ey[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int ez[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int bb[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
    static constexpr double w_ref[Q] = {
        double(3.f / 9.f),  double(1.f / 18.f), double(1.f / 18.f), double(1.f / 18.f),
        double(1.f / 18.f), double(1.f / 18.f), double(1.f / 18.f), double(1.f / 36.f),
        double(1.f / 36.f), double(1.f / 36.f), double(1.f / 36.f), double(1.f / 36.f),
        double(1.f / 36.f), double(1.f / 36.f), double(1.f / 36.f), double(1.f / 36.f),
        double(1.f / 36.f), double(1.f / 36.f), double(1.f / 36.f)};
    static constexpr double w_exact[Q] = {
        3. / 9.,  1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
        1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
        1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
    static constexpr int n_unk = 5;
    static constexpr int in_unk[n_unk] = {5, 11, 12, 15, 16};
    static constexpr int in_par[n_unk] = {6, 14, 13, 18, 17};
    static constexpr int out_unk[n_unk] = {6, 14, 13, 18, 17};
    static constexpr int out_par[n_unk] = {5, 11, 12, 15, 16};
    static constexpr int in_list[n_unk] = {6, 13, 14, 17, 18};
    static constexpr int out_list[n_unk] = {5, 11, 12, 15, 16};
    static constexpr int n_zero = 9;
    static constexpr int zero_list[n_zero] = {0, 1, 2, 3, 4, 7, 8, 9, 10};
};

struct D3Q27 {
    static constexpr int Q = 27;
    static constexpr int ex[Q] = {0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1,
                                  -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1};
    static constexpr int ey[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0,
                                  0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int ez[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1,
                                  -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1};
    static constexpr int bb[Q] = {0,  2,  1,  4,  3,  6,  5,  10, 9,  8,  7,  14, 13, 12,
                                  11, 18, 17, 16, 15, 26, 25, 24, 23, 22, 21, 20, 19};
    static constexpr double w_ref[Q] = {
        8. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,
        1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,
        1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 216., 1. / 216.,
        1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216.};
    static constexpr double w_exact[Q] = {
        8. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,
        1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,
        1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 216., 1. / 216.,
        1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216.};
    static constexpr int n_unk = 9;
    static constexpr int in_unk[n_unk] = {5, 11, 13, 15, 17, 19, 21, 23, 25};
    static constexpr int in_par[n_unk] = {6, 14, 12, 18, 16, 26, 24, 22, 20};
    static constexpr int out_unk[n_unk] = {6, 14, 12, 18, 16, 26, 24, 22, 20};
    static constexpr int out_par[n_unk] = {5, 11, 13, 15, 17, 19, 21, 23, 25};
    static constexpr int in_list[n_unk] = {6, 14, 12, 18, 16, 26, 24, 22, 20};
    static constexpr int out_list[n_unk] = {5, 11, 13, 15, 17, 19, 21, 23, 25};
    static constexpr int n_zero = 9;
    static constexpr int zero_list[n_zero] = {0, 1, 2, 3, 4, 7, 8, 9, 10};
};

}  // namespace tlbm
