// kernels_impl.h -- launcher bodies, instantiated per storage by
// kernels_f64.hip and kernels_f32.hip.
#pragma once

#include "launch.h"
#include "lbm_aux.h"

namespace tlbm {

inline unsigned blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

template <int ST>
struct Launch {
    using st = typename StorageT<ST>::st;
    using T = typename StorageT<ST>::T;

    template <class L>
    static hipError_t lbm_l(int mode, const void *fin, void *fout, const uint8_t *nd,
                            const uint8_t *chunk_type, void *macro, const void *uy_head,
                            const StepArgs &a, hipStream_t s) {
        const st *in = static_cast<const st *>(fin);
        st *out = static_cast<st *>(fout);
        void *mac = macro;
        const T *uyh = static_cast<const T *>(uy_head);
        using TU = Tune<ST, L::Q>;
        switch (mode) {
        case MODE_STEP:
            if (a.closure) {
                if constexpr (ST == ST_F64 || ST == ST_F32X) {
                    return hipErrorInvalidValue;
                } else {
                    constexpr int CBS = kClosureBlock;
                    hipLaunchKernelGGL((k_lbm_closure<L, ST, CBS, kClosureNTLoad<ST>>),
                                       dim3((a.N + CBS - 1) / CBS), dim3(CBS), 0, s, in,
                                       out, nd, chunk_type, uyh, a);
                }
            } else {
                hipLaunchKernelGGL((k_lbm<L, ST, MODE_STEP, false, TU::kNTLoad, true, TU::kBS>),
                                   dim3((a.N + TU::kBS - 1) / TU::kBS), dim3(TU::kBS), 0, s, in, out,
                                   nd, chunk_type, mac, uyh, a);
            }
            break;
        case MODE_MACROS:
            hipLaunchKernelGGL((k_lbm<L, ST, MODE_MACROS>), dim3(blocks_for(a.N)), dim3(kBlock), 0,
                               s, in, out, nd, nullptr, mac, uyh, a);
            break;
        case MODE_AVG:
            hipLaunchKernelGGL((k_lbm<L, ST, MODE_AVG>), dim3(blocks_for(a.N)), dim3(kBlock), 0,
                               s, in, out, nd, nullptr, mac, uyh, a);
            break;
        case MODE_STRESS:
            hipLaunchKernelGGL((k_lbm<L, ST, MODE_STRESS>), dim3(blocks_for(a.N)), dim3(kBlock), 0,
                               s, in, out, nd, nullptr, mac, uyh, a);
            break;
        case MODE_HEAD:
            hipLaunchKernelGGL((k_lbm<L, ST, MODE_HEAD>), dim3(1), dim3(64), 0, s, in, out, nd,
                               nullptr, mac, uyh, a);
            break;
        default:
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }

#define TLBM_BY_LATTICE(fn, ...)                              \
    switch (lattice) {                                        \
    case 15: return fn<D3Q15>(__VA_ARGS__);                   \
    case 19: return fn<D3Q19>(__VA_ARGS__);                   \
    case 27: return fn<D3Q27>(__VA_ARGS__);                   \
    default: return hipErrorInvalidValue;                     \
    }

    static hipError_t lbm(int lattice, int mode, const void *fin, void *fout, const uint8_t *nd,
                          const uint8_t *chunk_type, void *macro, const void *uy_head,
                          const StepArgs &a, hipStream_t s) {
        TLBM_BY_LATTICE(lbm_l, mode, fin, fout, nd, chunk_type, macro, uy_head, a, s)
    }

    template <class L>
    static hipError_t init_rest_l(void *fout, const StepArgs &a, hipStream_t s) {
        hipLaunchKernelGGL((k_init_rest<L, ST>), dim3(blocks_for(a.N)), dim3(kBlock), 0, s,
                           static_cast<st *>(fout), a);
        return hipGetLastError();
    }
    static hipError_t init_rest(int lattice, void *fout, const StepArgs &a, hipStream_t s) {
        TLBM_BY_LATTICE(init_rest_l, fout, a, s)
    }

    template <class L>
    static hipError_t init_macros_l(void *fout, const double *mac, const StepArgs &a, hipStream_t s) {
        hipLaunchKernelGGL((k_init_macros<L, ST>), dim3(blocks_for(a.N)), dim3(kBlock), 0, s,
                           static_cast<st *>(fout), mac, a);
        return hipGetLastError();
    }
    static hipError_t init_macros(int lattice, void *fout, const double *mac, const StepArgs &a,
                                  hipStream_t s) {
        TLBM_BY_LATTICE(init_macros_l, fout, mac, a, s)
    }

    template <class L>
    static hipError_t get_pop_l(const void *fin, double *f, const StepArgs &a, hipStream_t s) {
        hipLaunchKernelGGL((k_get_populations<L, ST>), dim3(blocks_for(a.N)), dim3(kBlock), 0, s,
                           static_cast<const st *>(fin), f, a);
        return hipGetLastError();
    }
    static hipError_t get_populations(int lattice, const void *fin, double *f, const StepArgs &a,
                                      hipStream_t s) {
        TLBM_BY_LATTICE(get_pop_l, fin, f, a, s)
    }

    template <class L>
    static hipError_t set_pop_l(void *fout, const double *f, const StepArgs &a, hipStream_t s) {
        hipLaunchKernelGGL((k_set_populations<L, ST>), dim3(blocks_for(a.N)), dim3(kBlock), 0, s,
                           static_cast<st *>(fout), f, a);
        return hipGetLastError();
    }
    static hipError_t set_populations(int lattice, void *fout, const double *f, const StepArgs &a,
                                      hipStream_t s) {
        TLBM_BY_LATTICE(set_pop_l, fout, f, a, s)
    }

    template <class L>
    static hipError_t forces_l(const void *fin, const uint8_t *nd, double *F, const StepArgs &a,
                               hipStream_t s) {
        hipLaunchKernelGGL((k_forces<L, ST>), dim3(blocks_for(a.N)), dim3(kBlock), 0, s,
                           static_cast<const st *>(fin), nd, F, a);
        return hipGetLastError();
    }
    static hipError_t forces(int lattice, const void *fin, const uint8_t *nd, double *F,
                             const StepArgs &a, hipStream_t s) {
        TLBM_BY_LATTICE(forces_l, fin, nd, F, a, s)
    }

    template <class L>
    static hipError_t reduce_l(const void *cur, const void *prev, double *partial, int nparts,
                               double *out, const StepArgs &a, hipStream_t s) {
        hipLaunchKernelGGL((k_reduce_partial<L, ST>), dim3(nparts), dim3(kRedBlock), 0, s,
                           static_cast<const st *>(cur), static_cast<const st *>(prev), partial, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_reduce_final<ST>), dim3(1), dim3(kBlock), 0, s, partial, nparts, out);
        return hipGetLastError();
    }
    static hipError_t reduce(int lattice, const void *cur, const void *prev, double *partial,
                             int nparts, double *out, const StepArgs &a, hipStream_t s) {
        TLBM_BY_LATTICE(reduce_l, cur, prev, partial, nparts, out, a, s)
    }
#undef TLBM_BY_LATTICE

    static hipError_t box_filter(const void *macro, double *out, const StepArgs &a, hipStream_t s) {
        const uint32_t nc = (a.nx / 2) * (a.ny / 2) * (a.nz / 2);
        hipLaunchKernelGGL((k_box_filter<T>), dim3(blocks_for(nc)), dim3(kBlock), 0, s,
                           static_cast<const T *>(macro), out, a);
        return hipGetLastError();
    }

    static const LaunchOps &table() {
        static const LaunchOps t = {&lbm, &init_rest, &init_macros, &get_populations,
                                    &set_populations, &forces, &reduce, &box_filter, int(sizeof(T))};
        return t;
    }
};

}  // namespace tlbm
