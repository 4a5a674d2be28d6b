// lbm_aux.h -- kernels around the time step: initial and restart state,
// population import/export, the momentum-exchange force and the FP64
// conservation / residual reductions.  All of them address the pull state
// B (post-collision, not yet streamed; see lbm_kernels.h): the populations
// the next step reads are A[g,s] = B[s][g - e_s].
#pragma once

#include "lbm_kernels.h"

namespace tlbm {

// Periodic coordinate shift by d (any sign, |d| <= 2n).
__device__ inline uint32_t wrap(int v, int n) {
    v %= n;
    return uint32_t(v < 0 ? v + n : v);
}

struct Coord {
    int x, y, z;
    __device__ inline Coord(uint32_t g, const StepArgs &a) {
        z = int(g / a.nxny);
        const uint32_t r = g - uint32_t(z) * a.nxny;
        y = int(r / a.nx);
        x = int(r - uint32_t(y) * a.nx);
    }
    __device__ inline uint32_t shifted(int dx, int dy, int dz, const StepArgs &a) const {
        return wrap(x + dx, int(a.nx)) + wrap(y + dy, int(a.ny)) * a.nx +
               wrap(z + dz, int(a.nz)) * a.nxny;
    }
};

// f_eq of one direction, LatticeStructure.hpp:92-103 in the reference's
// operation order.
template <class L, class T>
__device__ inline T feq_one(int s, T w, T rho, T ux, T uy, T uz) {
    const T usq = ux * ux + uy * uy + uz * uz;
    const T cu = (s == 0) ? T(0) : T(3) * edot<T>(L::ex[s], L::ey[s], L::ez[s], ux, uy, uz);
    return w * rho * (T(1) + cu + T(0.5) * (cu * cu) - T(1.5) * usq);
}

// B[s][h] = f_eq_s(rho, u at node h + e_s): the pull state whose next step
// reads f_eq(rho, u) at every node (restart, TLBM_Partition.cpp:527-532).
// mac holds ux, uy, uz, rho planes of N doubles.  FP64 storage evaluates
// f_eq as the reference does; FP32/FP16 store f_eq - w, evaluated in double
// as w (delta + rho (cu + cu^2/2 - 3u^2/2)).
template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_init_macros(typename StorageT<ST>::st *__restrict__ fout, const double *__restrict__ mac,
              StepArgs a) {
    using S = StorageT<ST>;
    using T = typename S::T;
    const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
    if (h >= a.N) return;
    const Coord c(h, a);
#pragma unroll
    for (int s = 0; s < L::Q; ++s) {
        const uint32_t src = c.shifted(L::ex[s], L::ey[s], L::ez[s], a);
        const double ux = mac[src], uy = mac[size_t(a.N) + src], uz = mac[2 * size_t(a.N) + src],
                     rho = mac[3 * size_t(a.N) + src];
        T v;
        if constexpr (S::kShift) {
            const double w = S::shift(s, a);
            const double usq = ux * ux + uy * uy + uz * uz;
            const double cu = (s == 0) ? 0.0 : 3.0 * edot<double>(L::ex[s], L::ey[s], L::ez[s], ux, uy, uz);
            v = T(w * ((rho - 1.0) + rho * (cu + 0.5 * (cu * cu) - 1.5 * usq)));
        } else {
            v = feq_one<L, T>(s, S::weight(s, a), T(rho), T(ux), T(uy), T(uz));
        }
        S::store(fout + fidx<L::Q>(s, h, a), s, v, a);
    }
}

// Uniform rest state f = w * rho (initialize_data_arrays,
// TLBM_Partition.cpp:458-465); it is its own stream, so B = A.
template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_init_rest(typename StorageT<ST>::st *__restrict__ fout, StepArgs a) {
    using S = StorageT<ST>;
    using T = typename S::T;
    const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
    if (h >= a.N) return;
#pragma unroll
    for (int s = 0; s < L::Q; ++s) {
        T v;
        if constexpr (S::kShift)
            v = T(S::shift(s, a) * a.rho_bc - S::shift(s, a));
        else
            v = S::weight(s, a) * T(a.rho_bc);
        S::store(fout + fidx<L::Q>(s, h, a), s, v, a);
    }
}

// f_aos[g*Q + s] = A[g,s] = B[s][g - e_s]
template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_get_populations(const typename StorageT<ST>::st *__restrict__ fin, double *__restrict__ f_aos,
                  StepArgs a) {
    using S = StorageT<ST>;
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= a.N) return;
    const Coord c(g, a);
#pragma unroll
    for (int s = 0; s < L::Q; ++s) {
        const uint32_t src = c.shifted(-L::ex[s], -L::ey[s], -L::ez[s], a);
        f_aos[size_t(g) * L::Q + s] = S::absolute(S::load(fin + fidx<L::Q>(s, src, a), s, a), s, a);
    }
}

// B[s][h] = f_aos[(h + e_s)*Q + s]
template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_set_populations(typename StorageT<ST>::st *__restrict__ fout, const double *__restrict__ f_aos,
                  StepArgs a) {
    using S = StorageT<ST>;
    const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
    if (h >= a.N) return;
    const Coord c(h, a);
#pragma unroll
    for (int s = 0; s < L::Q; ++s) {
        const uint32_t src = c.shifted(L::ex[s], L::ey[s], L::ez[s], a);
        S::store(fout + fidx<L::Q>(s, h, a), s, S::from_absolute(f_aos[size_t(src) * L::Q + s], s, a), a);
    }
}

// calc_force (TLBM_Partition.cpp:174-208) over the map of
// make_force_calc_map (143-172): for a solid node g and each direction spd
// (ascending) whose neighbour g+e_spd is a fluid (type 0) node,
//   f1 = A[g, spd], f2 = A[g+e_spd, bb(spd)],  F += e_bb(spd) * (f1 + f2).
// In the pull state f1 = B[spd][g - e_spd] and f2 = B[bb][g + 2 e_spd].
// F holds Fx, Fy, Fz planes of N doubles.
template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_forces(const typename StorageT<ST>::st *__restrict__ fin, const uint8_t *__restrict__ nd_type,
         double *__restrict__ F, StepArgs a) {
    using S = StorageT<ST>;
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= a.N) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (nd_type[g] == 1) {
        const Coord c(g, a);
#pragma unroll
        for (int s = 0; s < L::Q; ++s) {
            const uint32_t ngb = c.shifted(L::ex[s], L::ey[s], L::ez[s], a);
            if (nd_type[ngb] != 0) continue;
            const int b = L::bb[s];
            const double f1 = S::absolute(
                S::load(fin + fidx<L::Q>(s, c.shifted(-L::ex[s], -L::ey[s], -L::ez[s], a), a), s, a), s, a);
            const double f2 = S::absolute(
                S::load(fin + fidx<L::Q>(b, c.shifted(2 * L::ex[s], 2 * L::ey[s], 2 * L::ez[s], a), a), b, a), b, a);
            const double sum = f1 + f2;
            if (L::ex[b] > 0) fx += sum; else if (L::ex[b] < 0) fx -= sum;
            if (L::ey[b] > 0) fy += sum; else if (L::ey[b] < 0) fy -= sum;
            if (L::ez[b] > 0) fz += sum; else if (L::ez[b] < 0) fz -= sum;
        }
    }
    F[g] = fx;
    F[size_t(a.N) + g] = fy;
    F[2 * size_t(a.N) + g] = fz;
}

// 2x2x2 box filter of one step's moments and stresses onto the
// (nx/2, ny/2, nz/2) lattice.  macro: the 10 planes MODE_STRESS writes
// (ux, uy, uz, rho, sigma xx, yy, zz, xy, xz, yz).  out: kFilterPlanes
// planes of means, rho, ux, uy, uz, uxux, uyuy, uzuz, uxuy, uxuz, uyuz,
// sxx, syy, szz, sxy, sxz, syz.  Sums in double, in a fixed order.

template <class T>
__global__ void __launch_bounds__(kBlock)
k_box_filter(const T *__restrict__ macro, double *__restrict__ out, StepArgs a) {
    const uint32_t cx_n = a.nx / 2, cy_n = a.ny / 2, cz_n = a.nz / 2;
    const uint32_t Nc = cx_n * cy_n * cz_n;
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= Nc) return;
    const uint32_t cz = c / (cx_n * cy_n), cr = c - cz * cx_n * cy_n, cy = cr / cx_n, cx = cr - cy * cx_n;
    double acc[kFilterPlanes];
#pragma unroll
    for (int k = 0; k < kFilterPlanes; ++k) acc[k] = 0.0;
    for (uint32_t dz = 0; dz < 2; ++dz)
        for (uint32_t dy = 0; dy < 2; ++dy)
            for (uint32_t dx = 0; dx < 2; ++dx) {
                const uint32_t g = (2 * cx + dx) + a.nx * (2 * cy + dy) + a.nxny * (2 * cz + dz);
                const double ux = double(macro[g]), uy = double(macro[size_t(a.N) + g]),
                             uz = double(macro[2 * size_t(a.N) + g]),
                             rho = double(macro[3 * size_t(a.N) + g]);
                acc[0] += rho;
                acc[1] += ux;
                acc[2] += uy;
                acc[3] += uz;
                acc[4] += ux * ux;
                acc[5] += uy * uy;
                acc[6] += uz * uz;
                acc[7] += ux * uy;
                acc[8] += ux * uz;
                acc[9] += uy * uz;
#pragma unroll
                for (int k = 0; k < 6; ++k) acc[10 + k] += double(macro[size_t(4 + k) * a.N + g]);
            }
#pragma unroll
    for (int k = 0; k < kFilterPlanes; ++k) out[size_t(k) * Nc + c] = acc[k] * 0.125;
}

// Sum over the 64 lanes of a wavefront: a fixed-order butterfly of lane
// exchanges (no LDS); lane 0 holds the result.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The kRed sums of a BS-thread workgroup: wavefront sums, then the
// BS/64 wave totals folded in wave order by threads 0..kRed-1, which
// return them (other threads return zeros).  Deterministic.
template <int BS>
__device__ inline void block_sums(double (&acc)[kRed]) {
    constexpr int W = BS / 64;
    __shared__ double sh[kRed][W];
    const int lane = int(threadIdx.x & 63u), wave = int(threadIdx.x >> 6);
#pragma unroll
    for (int k = 0; k < kRed; ++k) {
        const double v = wave_sum(acc[k]);
        if (lane == 0) sh[k][wave] = v;
    }
    __syncthreads();
    const int k = int(threadIdx.x);
    double t = 0.0;
    if (k < kRed) {
#pragma unroll
        for (int w = 0; w < W; ++w) t += sh[k][w];
    }
    acc[0] = t;
}

// Per-workgroup FP64 partial sums over the state: sum f, sum f e_x,
// sum f e_y, sum f e_z, sum (B_n - B_{n-1})^2, sum B_n^2 and the number of
// populations that are NaN or infinite (kRed values).  Workgroups of
// kRedBlock threads: each thread streams 2 Q values per node through a few
// FP64 chains, and the loads of one node only overlap with other waves',
// so the partials run as many waves as their registers allow (D3Q19: 8 per
// SIMD; 1024x512x512 FP16: 18.5 ms with 256-thread workgroups, 10.4 ms
// with 1024).
constexpr int kRedBlock = 1024;
template <class L, int ST>
__global__ void __launch_bounds__(kRedBlock)
k_reduce_partial(const typename StorageT<ST>::st *__restrict__ cur,
                 const typename StorageT<ST>::st *__restrict__ prev, double *__restrict__ partial,
                 StepArgs a) {
    using S = StorageT<ST>;
    double acc[kRed] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t g = blockIdx.x * kRedBlock + threadIdx.x; g < a.N; g += gridDim.x * kRedBlock) {
#pragma unroll
        for (int s = 0; s < L::Q; ++s) {
            const double v = S::absolute(S::load(cur + fidx<L::Q>(s, g, a), s, a), s, a);
            if (!__builtin_isfinite(v)) acc[6] += 1.0;
            acc[0] += v;
            acc[1] += L::ex[s] * v;
            acc[2] += L::ey[s] * v;
            acc[3] += L::ez[s] * v;
            acc[5] += v * v;
            if (prev) {
                const double d = v - S::absolute(S::load(prev + fidx<L::Q>(s, g, a), s, a), s, a);
                acc[4] += d * d;
            }
        }
    }
    block_sums<kRedBlock>(acc);
    if (threadIdx.x < kRed) partial[size_t(blockIdx.x) * kRed + threadIdx.x] = acc[0];
}

// One workgroup folds the partials in a fixed order (deterministic).
// Templated on the storage only so each kernel file owns its copy.
template <int ST>
__global__ void __launch_bounds__(kBlock)
k_reduce_final(const double *__restrict__ partial, int nparts, double *__restrict__ out) {
    double acc[kRed] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < nparts; i += kBlock) {
#pragma unroll
        for (int k = 0; k < kRed; ++k) acc[k] += partial[size_t(i) * kRed + k];
    }
    block_sums<kBlock>(acc);
    if (threadIdx.x < kRed) out[threadIdx.x] = acc[0];
}

}  // namespace tlbm
