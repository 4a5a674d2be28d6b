// tlbm_gpu.cpp -- the C-ABI of include/tlbm_gpu.h over the HIP kernels.
#include "tlbm_gpu.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "launch.h"

namespace {
thread_local std::string g_err;

// Partial sums of tlbm_gpu_reduce_device: a fixed device array, so that
// entry point allocates nothing.  Each GPU holds its own copy until the
// process ends; every handle on that GPU shares it.  In host code the
// array's name is not a device address: its address on each device comes
// from hipGetSymbolAddress (g_scratch_ptr, resolved once per device).
constexpr int kScratchParts = 1024;
static_assert(TLBM_REDUCE_SUMS == tlbm::kRed, "sums of the FP64 reduction");
__device__ __align__(256) double scratch[kScratchParts * tlbm::kRed];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

namespace tlbm {
// Message for tlbm_gpu_last_error(), shared with the network entry points.
void set_last_error(const std::string &msg) { g_err = msg; }
}  // namespace tlbm

namespace {

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define TLBM_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(TLBM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

template <class L>
void fill_weights(tlbm::StepArgs &a, bool exact, int storage) {
    for (int s = 0; s < L::Q; ++s) {
        a.w[s] = exact ? L::w_exact[s] : L::w_ref[s];
        a.wf[s] = float(a.w[s]);
        // FP16 storage keeps (f - w)/w: a relative deviation, O(0.1) at Ma ~ 0.1
        a.h_scale[s] = float(1.0 / double(a.wf[s]));
        a.h_unscale[s] = a.wf[s];
    }
    // constants of the deviation form, from the float weights the FP32
    // arithmetic uses, evaluated in double
    double sum = 0, s_in = 0, s_out = 0, s_zero = 0;
    for (int s = 0; s < L::Q; ++s) sum += double(a.wf[s]);
    for (int i = 0; i < L::n_unk; ++i) {
        s_in += double(a.wf[L::in_list[i]]);
        s_out += double(a.wf[L::out_list[i]]);
    }
    for (int i = 0; i < L::n_zero; ++i) s_zero += double(a.wf[L::zero_list[i]]);
    a.dw = float(sum - 1.0);
    // moments of f - feq in exact arithmetic with the weights w themselves
    // (D3Q15/D3Q19 reference weights: the float values; D3Q27 and exact
    // weights: the double values, which sum to 1)
    double sum_w = 0, w2 = 0;
    for (int s = 0; s < L::Q; ++s) {
        sum_w += a.w[s];
        w2 += a.w[s] * double(L::ex[s] * L::ex[s]);
    }
    a.k0 = float(sum_w - 1.0);
    a.k1 = float(1.0 + (sum_w - 1.0) - 3.0 * w2);
    a.k2 = float(1.0 - 3.0 * w2);
    // FP16: equilibrium corrections from the float weights' balance to that
    // of w (zero where the two coincide: the reference's D3Q15/D3Q19 weights)
    a.dt3 = 0.0f;
    a.dk0 = a.dk1 = 0.0f;
    if (storage == TLBM_STORAGE_F16) {
        double w2f = 0;
        for (int s = 0; s < L::Q; ++s) w2f += double(a.wf[s]) * double(L::ex[s] * L::ex[s]);
        const double k1f = 1.0 + (sum - 1.0) - 3.0 * w2f;
        // the deviation of the scale 3 W2 / W2_f from 3 (it fits a float;
        // the scale itself rounds to 3.0f)
        a.dt3 = float(3.0 * ((w2 - w2f) / w2f));
        a.dk0 = float((sum - 1.0) - (sum_w - 1.0));
        a.dk1 = float(k1f - (1.0 + (sum_w - 1.0) - 3.0 * w2));
    }
    a.c_in = float(2.0 * s_in + s_zero - (1.0 - a.u_bc));
    {
        double sum_d = 0, in_d = 0, out_d = 0, zero_d = 0;
        for (int s = 0; s < L::Q; ++s) sum_d += a.w[s];
        for (int i = 0; i < L::n_unk; ++i) {
            in_d += a.w[L::in_list[i]];
            out_d += a.w[L::out_list[i]];
        }
        for (int i = 0; i < L::n_zero; ++i) zero_d += a.w[L::zero_list[i]];
        a.dw_d = sum_d - 1.0;
        a.c_in_d = 2.0 * in_d + zero_d - (1.0 - a.u_bc);
        a.c_out_d = 2.0 * out_d + zero_d - a.rho_bc;
    }
    a.c_out = float(2.0 * s_out + s_zero - a.rho_bc);
    a.drho_bc = float(a.rho_bc - 1.0);
}

// Chunk of the layout: 32 KiB of one direction plane (8192 FP32 values),
// the size that ran the D3Q19 gather/scatter fastest on MI355X.
uint32_t chunk_bits(size_t value_bytes) {
    uint32_t cb = 0;
    while ((size_t(1) << (cb + 1)) * value_bytes <= 32768) ++cb;
    return cb;
}

int lattice_q(int lattice) {
    return (lattice == 15 || lattice == 19 || lattice == 27) ? lattice : 0;
}

const tlbm::LaunchOps &ops_for(int storage) {
    switch (storage) {
    case TLBM_STORAGE_F64: return tlbm::ops_f64();
    case TLBM_STORAGE_F32: return tlbm::ops_f32();
    case TLBM_STORAGE_F32X: return tlbm::ops_f32x();
    default: return tlbm::ops_f16();
    }
}

size_t storage_bytes(int storage) {
    return storage == TLBM_STORAGE_F64 ? 8 : storage == TLBM_STORAGE_F16 ? 2 : 4;
}

}  // namespace

struct tlbm_gpu {
    tlbm_params p{};
    int Q = 0;
    uint32_t N = 0;
    tlbm::StepArgs args{};
    const tlbm::LaunchOps *ops = nullptr;
    void *f[2] = {nullptr, nullptr};
    int cur = 0;              // f[cur] holds the state the next step reads from
    uint8_t *nd = nullptr;
    uint8_t *chunk_type = nullptr;  // per chunk: common node type or 0xFF
    void *macro = nullptr;    // 4*N compute-type values (dumps)
    void *filt_in = nullptr;  // 10*N compute-type values (box filter input)
    void *uy_head = nullptr;  // Q compute-type values (type-5 rule)
    double *staging = nullptr;  // staging for forces / populations / macros (lazy)
    size_t staging_bytes = 0;
    double *partial = nullptr;  // reduction partials
    int nparts = 0;
    bool has_lid = false;
    long long steps_since_set = 0;
    uint64_t step_count = 0;  // steps since creation (FP16 rounding stream)
    hipStream_t stream = nullptr;
    double *fixed_partial = nullptr;  // this device's copy of `scratch`
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t bytes = 0;
    std::vector<uint8_t> chunk_types;  // host copy, kept until the upload completes
    tlbm::MlpPacked *mlp = nullptr;    // closure network, device copy
    float *inlet = nullptr;            // nx*ny inlet u_z profile
    double *avg = nullptr;             // 4*N FP64 time-average sums
    bool avg_on = false;
    int64_t avg_steps = 0;
};

namespace {

int ensure_staging(tlbm_gpu *h, size_t bytes) {
    if (h->staging_bytes >= bytes) return TLBM_OK;
    if (h->staging) {
        TLBM_HIP(hipFree(h->staging));
        h->bytes -= h->staging_bytes;
        h->staging = nullptr;
        h->staging_bytes = 0;
    }
    TLBM_HIP(hipMalloc(&h->staging, bytes));
    h->staging_bytes = bytes;
    h->bytes += bytes;
    return TLBM_OK;
}

int check(const tlbm_gpu *h) {
    if (!h) return fail(TLBM_EINVAL, "null handle");
    return TLBM_OK;
}

// The current device's address of `scratch`, resolved with
// hipGetSymbolAddress on the first handle created on that device.
int fixed_partials(double **out) {
    int dev = 0;
    TLBM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TLBM_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        TLBM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(TLBM_EHIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<double *>(g_scratch_ptr[dev]);
    return TLBM_OK;
}

}  // namespace

extern "C" {

const char *tlbm_gpu_last_error(void) { return g_err.c_str(); }

int tlbm_gpu_create(const tlbm_params *p, const uint8_t *nd_type, tlbm_gpu **out) {
    g_err.clear();
    if (!p || !nd_type || !out) return fail(TLBM_EINVAL, "null argument");
    *out = nullptr;
    const int Q = lattice_q(p->lattice);
    if (!Q) return fail(TLBM_EINVAL, "lattice must be 15, 19 or 27 (D3Q15/D3Q19/D3Q27)");
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0)
        return fail(TLBM_EINVAL, "nx, ny, nz must be positive");
    const uint64_t N64 = uint64_t(p->nx) * uint64_t(p->ny) * uint64_t(p->nz);
    if (N64 >= (uint64_t(1) << 31) - 4096)
        return fail(TLBM_EINVAL, "lattice has 2^31 - 4096 nodes or more");
    if (p->storage < TLBM_STORAGE_F64 || p->storage > TLBM_STORAGE_F32X)
        return fail(TLBM_EINVAL, "unknown storage");
    if (p->reserved != 0) return fail(TLBM_EINVAL, "reserved field must be 0");
    if (!(p->omega > 0.0) || !std::isfinite(p->omega)) return fail(TLBM_EINVAL, "omega must be > 0");
    if (!(1.0 - p->u_lbm != 0.0)) return fail(TLBM_EINVAL, "u_lbm must differ from 1");
    if (!(p->rho_lbm != 0.0)) return fail(TLBM_EINVAL, "rho_lbm must be nonzero");
    const uint32_t N = uint32_t(N64);
    bool has_lid = false;
    for (uint32_t g = 0; g < N; ++g) {
        const uint8_t t = nd_type[g];
        if (!(t == 0 || t == 1 || t == 2 || t == 3 || t == 5)) {
            char buf[128];
            std::snprintf(buf, sizeof buf, "node %u has type %u; valid types are 0,1,2,3,5", g,
                          unsigned(t));
            return fail(TLBM_EINVAL, buf);
        }
        if (t == 5) {
            has_lid = true;
            if (g + 1 < uint32_t(Q)) {
                char buf[160];
                std::snprintf(buf, sizeof buf,
                              "type-5 node %u precedes node %d, whose uy its rule reads", g, Q - 1);
                return fail(TLBM_EINVAL, buf);
            }
        }
    }
    tlbm_gpu *h = new (std::nothrow) tlbm_gpu;
    if (!h) return fail(TLBM_ENOMEM, "host allocation failed");
    h->p = *p;
    h->Q = Q;
    h->N = N;
    h->has_lid = has_lid;
    h->ops = &ops_for(p->storage);
    tlbm::StepArgs &a = h->args;
    a.nx = uint32_t(p->nx);
    a.ny = uint32_t(p->ny);
    a.nz = uint32_t(p->nz);
    a.nxny = a.nx * a.ny;
    a.N = N;
    a.dynamics = p->dynamics;
    a.smag = p->cs > 0.0;
    a.cs = p->cs;
    a.omega = p->omega;
    a.nu = ((1. / p->omega) - 0.5) / 3.0;  // LatticeStructure.hpp:145
    a.nu2 = a.nu * a.nu;
    a.u_bc = p->u_lbm;
    a.rho_bc = p->rho_lbm;
    a.inv_1mu = 1. / (1. - p->u_lbm);
    a.inv_rho_bc = 1. / p->rho_lbm;
    a.cb = chunk_bits(storage_bytes(p->storage));
    a.div_nx = tlbm::udiv_magic(a.nx);
    a.div_nxny = tlbm::udiv_magic(a.nxny);
    {
        const uint64_t CH = uint64_t(1) << a.cb;
        // a pull source lies less than 2 nx ny nodes before its node (the
        // y wrap of the plane behind); then r < 2 kb CH
        a.kb = uint32_t((2 * uint64_t(a.nxny) + CH - 1) / CH + 1);
        a.rel = (2 * uint64_t(a.kb) + 2) * CH * uint64_t(Q) * storage_bytes(p->storage) <
                (uint64_t(1) << 32);
        // TLBM_GENERIC_ADDRESSING=1: the 64-bit gather everywhere, the path
        // of lattices too large for 32-bit relative offsets (for tests)
        const char *gen = std::getenv("TLBM_GENERIC_ADDRESSING");
        if (gen && gen[0] == '1') a.rel = 0;
        // rows never straddle a chunk: the source row's chunk is the source's
        // (TLBM_ROW_ADDRESSING=0: the per-population relative gather, for tests)
        const char *rowa = std::getenv("TLBM_ROW_ADDRESSING");
        a.row_chunks = a.rel && (CH % a.nx == 0) && !(rowa && rowa[0] == '0');
    }
    switch (Q) {
    case 15: fill_weights<tlbm::D3Q15>(a, p->exact_weights != 0, p->storage); break;
    case 19: fill_weights<tlbm::D3Q19>(a, p->exact_weights != 0, p->storage); break;
    default: fill_weights<tlbm::D3Q27>(a, p->exact_weights != 0, p->storage); break;
    }
    auto cleanup_fail = [&](int code) {
        tlbm_gpu_destroy(h);
        return code;
    };
    if (int rc = fixed_partials(&h->fixed_partial)) return cleanup_fail(rc);
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
        return cleanup_fail(fail(TLBM_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
    if ((e = hipEventCreate(&h->ev0)) != hipSuccess || (e = hipEventCreate(&h->ev1)) != hipSuccess)
        return cleanup_fail(fail(TLBM_EHIP, std::string("hipEventCreate: ") + hipGetErrorString(e)));
    const size_t nchunks = (size_t(N) + (size_t(1) << a.cb) - 1) >> a.cb;
    const size_t fbytes = (nchunks * Q << a.cb) * storage_bytes(p->storage);
    a.fbytes = fbytes;
    for (int i = 0; i < 2; ++i) {
        if ((e = hipMalloc(&h->f[i], fbytes)) != hipSuccess)
            return cleanup_fail(fail(TLBM_ENOMEM, std::string("hipMalloc populations: ") +
                                                      hipGetErrorString(e)));
        h->bytes += fbytes;
    }
    if ((e = hipMalloc(&h->nd, N)) != hipSuccess)
        return cleanup_fail(fail(TLBM_ENOMEM, std::string("hipMalloc ndType: ") + hipGetErrorString(e)));
    h->bytes += N;
    if ((e = hipMalloc(&h->uy_head, 32 * sizeof(double))) != hipSuccess)
        return cleanup_fail(fail(TLBM_ENOMEM, "hipMalloc uy_head"));
    h->nparts = int(std::min<uint32_t>(1024u, (N + tlbm::kBlock - 1) / tlbm::kBlock));
    if ((e = hipMalloc(&h->partial, (size_t(h->nparts) * tlbm::kRed + 8) * sizeof(double))) != hipSuccess)
        return cleanup_fail(fail(TLBM_ENOMEM, "hipMalloc reduction partials"));
    if ((e = hipMemcpyAsync(h->nd, nd_type, N, hipMemcpyHostToDevice, h->stream)) != hipSuccess)
        return cleanup_fail(fail(TLBM_EHIP, std::string("upload ndType: ") + hipGetErrorString(e)));
    {
        // per-chunk common node type (0xFF: mixed), read once per workgroup
        const uint32_t CH = 1u << a.cb;
        h->chunk_types.assign(nchunks, 0xFF);
        for (size_t c = 0; c < nchunks; ++c) {
            const uint32_t g0 = uint32_t(c * CH), g1 = std::min<uint32_t>(N, g0 + CH);
            uint8_t t0 = nd_type[g0];
            for (uint32_t g = g0 + 1; g < g1 && t0 != 0xFF; ++g)
                if (nd_type[g] != t0) t0 = 0xFF;
            h->chunk_types[c] = t0;
        }
        if ((e = hipMalloc(&h->chunk_type, nchunks)) != hipSuccess)
            return cleanup_fail(fail(TLBM_ENOMEM, "hipMalloc chunk types"));
        h->bytes += nchunks;
        if ((e = hipMemcpyAsync(h->chunk_type, h->chunk_types.data(), nchunks, hipMemcpyHostToDevice,
                                h->stream)) != hipSuccess)
            return cleanup_fail(fail(TLBM_EHIP, std::string("upload chunk types: ") + hipGetErrorString(e)));
    }
    for (int i = 0; i < 2; ++i)
        if ((e = h->ops->init_rest(Q, h->f[i], a, h->stream)) != hipSuccess)
            return cleanup_fail(fail(TLBM_EHIP, std::string("init: ") + hipGetErrorString(e)));
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess)
        return cleanup_fail(fail(TLBM_EHIP, std::string("init sync: ") + hipGetErrorString(e)));
    h->cur = 0;
    h->steps_since_set = 0;
    *out = h;
    return TLBM_OK;
}

int tlbm_gpu_destroy(tlbm_gpu *h) {
    if (!h) return TLBM_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (int i = 0; i < 2; ++i)
        if (h->f[i]) (void)hipFree(h->f[i]);
    if (h->nd) (void)hipFree(h->nd);
    if (h->chunk_type) (void)hipFree(h->chunk_type);
    if (h->avg) (void)hipFree(h->avg);
    if (h->mlp) (void)hipFree(h->mlp);
    if (h->inlet) (void)hipFree(h->inlet);
    if (h->macro) (void)hipFree(h->macro);
    if (h->filt_in) (void)hipFree(h->filt_in);
    if (h->uy_head) (void)hipFree(h->uy_head);
    if (h->staging) (void)hipFree(h->staging);
    if (h->partial) (void)hipFree(h->partial);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return TLBM_OK;
}

static int queue_steps(tlbm_gpu *h, int nsteps) {
    for (int k = 0; k < nsteps; ++k) {
        const void *fin = h->f[h->cur];
        void *fout = h->f[1 - h->cur];
        if (h->has_lid) {
            hipError_t e = h->ops->lbm(h->Q, tlbm::MODE_HEAD, fin, nullptr, h->nd, nullptr,
                                       h->uy_head, nullptr, h->args, h->stream);
            if (e != hipSuccess) return fail(TLBM_EHIP, std::string("type-5 helper: ") + hipGetErrorString(e));
        }
        h->args.step = uint32_t(h->step_count++);
        hipError_t e = h->ops->lbm(h->Q, tlbm::MODE_STEP, fin, fout, h->nd, h->chunk_type, nullptr,
                                   h->uy_head, h->args, h->stream);
        if (e != hipSuccess) return fail(TLBM_EHIP, std::string("stream-collide: ") + hipGetErrorString(e));
        h->cur = 1 - h->cur;
        ++h->steps_since_set;
        if (h->avg_on) {
            // the step just read f[1-cur]; its moments go into the sums
            e = h->ops->lbm(h->Q, tlbm::MODE_AVG, h->f[1 - h->cur], nullptr, h->nd, nullptr, h->avg,
                            nullptr, h->args, h->stream);
            if (e != hipSuccess) return fail(TLBM_EHIP, std::string("time average: ") + hipGetErrorString(e));
            ++h->avg_steps;
        }
    }
    return TLBM_OK;
}

int tlbm_gpu_step(tlbm_gpu *h, int nsteps) {
    if (int rc = check(h)) return rc;
    if (nsteps < 0) return fail(TLBM_EINVAL, "nsteps < 0");
    return queue_steps(h, nsteps);
}

int tlbm_gpu_step_timed(tlbm_gpu *h, int nsteps, double *ms) {
    if (int rc = check(h)) return rc;
    if (nsteps < 0 || !ms) return fail(TLBM_EINVAL, "nsteps < 0 or null ms");
    TLBM_HIP(hipEventRecord(h->ev0, h->stream));
    if (int rc = queue_steps(h, nsteps)) return rc;
    TLBM_HIP(hipEventRecord(h->ev1, h->stream));
    TLBM_HIP(hipEventSynchronize(h->ev1));
    float t = 0.f;
    TLBM_HIP(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = double(t);
    return TLBM_OK;
}

int tlbm_gpu_sync(tlbm_gpu *h) {
    if (int rc = check(h)) return rc;
    TLBM_HIP(hipStreamSynchronize(h->stream));
    return TLBM_OK;
}

// Copy `count` compute-type values from device `src` into host doubles.
static int download_as_double(tlbm_gpu *h, const void *src, size_t count, double *const *dst,
                              int nplanes) {
    const size_t esz = size_t(h->ops->macro_bytes);
    std::vector<unsigned char> tmp(esz * count * nplanes);
    TLBM_HIP(hipMemcpyAsync(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost, h->stream));
    TLBM_HIP(hipStreamSynchronize(h->stream));
    for (int k = 0; k < nplanes; ++k) {
        if (!dst[k]) continue;
        if (esz == 8) {
            std::memcpy(dst[k], tmp.data() + size_t(k) * count * 8, count * 8);
        } else {
            const float *fp = reinterpret_cast<const float *>(tmp.data()) + size_t(k) * count;
            for (size_t i = 0; i < count; ++i) dst[k][i] = double(fp[i]);
        }
    }
    return TLBM_OK;
}

int tlbm_gpu_get_macros(tlbm_gpu *h, double *ux, double *uy, double *uz, double *rho) {
    if (int rc = check(h)) return rc;
    if (h->steps_since_set < 1)
        return fail(TLBM_ESTATE, "moments exist only after a step (the reference computes them inside the step)");
    if (!h->macro) {
        const size_t b = size_t(4) * h->N * size_t(h->ops->macro_bytes);
        TLBM_HIP(hipMalloc(&h->macro, b));
        h->bytes += b;
    }
    // the last step read f[1-cur]; it is still intact (pull writes the other buffer)
    hipError_t e = h->ops->lbm(h->Q, tlbm::MODE_MACROS, h->f[1 - h->cur], nullptr, h->nd, nullptr,
                               h->macro, nullptr, h->args, h->stream);
    if (e != hipSuccess) return fail(TLBM_EHIP, std::string("moments: ") + hipGetErrorString(e));
    double *dst[4] = {ux, uy, uz, rho};
    return download_as_double(h, h->macro, h->N, dst, 4);
}

int tlbm_gpu_get_forces(tlbm_gpu *h, double *fx, double *fy, double *fz) {
    if (int rc = check(h)) return rc;
    if (!fx || !fy || !fz) return fail(TLBM_EINVAL, "null output");
    if (int rc = ensure_staging(h, size_t(3) * h->N * sizeof(double))) return rc;
    hipError_t e = h->ops->forces(h->Q, h->f[h->cur], h->nd, h->staging, h->args, h->stream);
    if (e != hipSuccess) return fail(TLBM_EHIP, std::string("forces: ") + hipGetErrorString(e));
    double *dst[3] = {fx, fy, fz};
    for (int k = 0; k < 3; ++k)
        TLBM_HIP(hipMemcpyAsync(dst[k], h->staging + size_t(k) * h->N, size_t(h->N) * 8,
                                hipMemcpyDeviceToHost, h->stream));
    TLBM_HIP(hipStreamSynchronize(h->stream));
    return TLBM_OK;
}

int tlbm_gpu_get_populations(tlbm_gpu *h, double *f) {
    if (int rc = check(h)) return rc;
    if (!f) return fail(TLBM_EINVAL, "null output");
    const size_t b = size_t(h->N) * h->Q * sizeof(double);
    if (int rc = ensure_staging(h, b)) return rc;
    hipError_t e = h->ops->get_populations(h->Q, h->f[h->cur], h->staging, h->args, h->stream);
    if (e != hipSuccess) return fail(TLBM_EHIP, std::string("get populations: ") + hipGetErrorString(e));
    TLBM_HIP(hipMemcpyAsync(f, h->staging, b, hipMemcpyDeviceToHost, h->stream));
    TLBM_HIP(hipStreamSynchronize(h->stream));
    return TLBM_OK;
}

int tlbm_gpu_set_populations(tlbm_gpu *h, const double *f) {
    if (int rc = check(h)) return rc;
    if (!f) return fail(TLBM_EINVAL, "null input");
    const size_t b = size_t(h->N) * h->Q * sizeof(double);
    if (int rc = ensure_staging(h, b)) return rc;
    TLBM_HIP(hipMemcpyAsync(h->staging, f, b, hipMemcpyHostToDevice, h->stream));
    for (int i = 0; i < 2; ++i) {
        hipError_t e = h->ops->set_populations(h->Q, h->f[i], h->staging, h->args, h->stream);
        if (e != hipSuccess) return fail(TLBM_EHIP, std::string("set populations: ") + hipGetErrorString(e));
    }
    TLBM_HIP(hipStreamSynchronize(h->stream));
    h->steps_since_set = 0;
    return TLBM_OK;
}

int tlbm_gpu_init_from_macros(tlbm_gpu *h, const double *ux, const double *uy, const double *uz,
                              const double *rho) {
    if (int rc = check(h)) return rc;
    if (!ux || !uy || !uz || !rho) return fail(TLBM_EINVAL, "null input");
    const size_t plane = size_t(h->N) * sizeof(double);
    if (int rc = ensure_staging(h, 4 * plane)) return rc;
    const double *src[4] = {ux, uy, uz, rho};
    for (int k = 0; k < 4; ++k)
        TLBM_HIP(hipMemcpyAsync(reinterpret_cast<char *>(h->staging) + k * plane, src[k], plane,
                                hipMemcpyHostToDevice, h->stream));
    for (int i = 0; i < 2; ++i) {
        hipError_t e = h->ops->init_macros(h->Q, h->f[i], h->staging, h->args, h->stream);
        if (e != hipSuccess) return fail(TLBM_EHIP, std::string("init from macros: ") + hipGetErrorString(e));
    }
    TLBM_HIP(hipStreamSynchronize(h->stream));
    h->steps_since_set = 0;
    return TLBM_OK;
}

int tlbm_gpu_reduce(tlbm_gpu *h, double out[5]) {
    if (int rc = check(h)) return rc;
    if (!out) return fail(TLBM_EINVAL, "null output");
    const void *prev = h->steps_since_set >= 1 ? h->f[1 - h->cur] : nullptr;
    double *res = h->partial + size_t(h->nparts) * tlbm::kRed;
    hipError_t e = h->ops->reduce(h->Q, h->f[h->cur], prev, h->partial, h->nparts, res, h->args,
                                  h->stream);
    if (e != hipSuccess) return fail(TLBM_EHIP, std::string("reduce: ") + hipGetErrorString(e));
    double r[tlbm::kRed];
    TLBM_HIP(hipMemcpyAsync(r, res, sizeof r, hipMemcpyDeviceToHost, h->stream));
    TLBM_HIP(hipStreamSynchronize(h->stream));
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
    out[3] = r[3];
    out[4] = (prev && r[5] > 0.0) ? std::sqrt(r[4] / r[5]) : 0.0;
    if (r[6] > 0.0) {
        char buf[160];
        std::snprintf(buf, sizeof buf,
                      "%.0f populations are NaN or infinite after %lld steps: the run diverged",
                      r[6], (long long)h->step_count);
        return fail(TLBM_ENONFINITE, buf);
    }
    return TLBM_OK;
}

int tlbm_gpu_reduce_device(tlbm_gpu *h, double *out_dev) {
    if (int rc = check(h)) return rc;
    if (!out_dev) return fail(TLBM_EINVAL, "null output");
    if (h->nparts > kScratchParts)
        return fail(TLBM_EINVAL, "the reduction needs " + std::to_string(h->nparts) +
                                     " partial sums; the fixed array holds " +
                                     std::to_string(kScratchParts));
    // the partial kernel writes every one of the nparts partials before the
    // final kernel reads them, so nothing needs clearing between calls
    const void *prev = h->steps_since_set >= 1 ? h->f[1 - h->cur] : nullptr;
    hipError_t e = h->ops->reduce(h->Q, h->f[h->cur], prev, h->fixed_partial, h->nparts, out_dev,
                                  h->args, h->stream);
    if (e != hipSuccess) return fail(TLBM_EHIP, std::string("reduce: ") + hipGetErrorString(e));
    return TLBM_OK;
}

int tlbm_gpu_time_average(tlbm_gpu *h, int enable) {
    if (int rc = check(h)) return rc;
    if (enable) {
        const size_t b = size_t(4) * h->N * sizeof(double);
        if (!h->avg) {
            TLBM_HIP(hipMalloc(&h->avg, b));
            h->bytes += b;
        }
        TLBM_HIP(hipMemsetAsync(h->avg, 0, b, h->stream));
        h->avg_steps = 0;
    }
    h->avg_on = enable != 0;
    return TLBM_OK;
}

int tlbm_gpu_get_time_average(tlbm_gpu *h, double *u, double *v, double *w, double *rho,
                              int64_t *nsteps) {
    if (int rc = check(h)) return rc;
    if (!h->avg) return fail(TLBM_ESTATE, "time averaging was never enabled");
    double *dst[4] = {u, v, w, rho};
    for (int k = 0; k < 4; ++k)
        if (dst[k])
            TLBM_HIP(hipMemcpyAsync(dst[k], h->avg + size_t(k) * h->N, size_t(h->N) * sizeof(double),
                                    hipMemcpyDeviceToHost, h->stream));
    TLBM_HIP(hipStreamSynchronize(h->stream));
    if (nsteps) *nsteps = h->avg_steps;
    return TLBM_OK;
}

int tlbm_gpu_set_closure(tlbm_gpu *h, const float *params, double feature_scale,
                         double output_scale) {
    if (int rc = check(h)) return rc;
    if (!params) {
        h->args.closure = 0;
        return TLBM_OK;
    }
    if (h->p.storage == TLBM_STORAGE_F64 || h->p.storage == TLBM_STORAGE_F32X)
        return fail(TLBM_EINVAL, "the learned closure runs with FP32 or FP16 storage");
    if (!(feature_scale > 0.0) || !(output_scale > 0.0) || !std::isfinite(feature_scale) ||
        !std::isfinite(output_scale))
        return fail(TLBM_EINVAL, "closure scales must be positive and finite");
    for (int k = 0; k < tlbm::kMlpParams; ++k)
        if (!std::isfinite(params[k])) return fail(TLBM_EINVAL, "closure parameters must be finite");
    tlbm::MlpPacked packed;
    tlbm::mlp_pack(params, &packed);
    if (!h->mlp) {
        TLBM_HIP(hipMalloc(&h->mlp, sizeof packed));
        h->bytes += sizeof packed;
    }
    TLBM_HIP(hipMemcpyAsync(h->mlp, &packed, sizeof packed, hipMemcpyHostToDevice, h->stream));
    TLBM_HIP(hipStreamSynchronize(h->stream));
    h->args.closure = 1;
    h->args.cl_in = float(feature_scale);
    h->args.cl_out = float(output_scale);
    h->args.tau0 = float(1.0 / h->p.omega);
    h->args.mlp = h->mlp;
    return TLBM_OK;
}

int tlbm_gpu_set_inlet_profile(tlbm_gpu *h, const double *uz) {
    if (int rc = check(h)) return rc;
    if (!uz) {
        h->args.inlet_uz = nullptr;
        return TLBM_OK;
    }
    const size_t n = size_t(h->args.nxny);
    std::vector<float> v(n);
    for (size_t i = 0; i < n; ++i) {
        if (!(uz[i] < 1.0) || !std::isfinite(uz[i])) return fail(TLBM_EINVAL, "inlet u_z must be finite and < 1");
        v[i] = float(uz[i]);
    }
    if (!h->inlet) {
        TLBM_HIP(hipMalloc(&h->inlet, n * sizeof(float)));
        h->bytes += n * sizeof(float);
    }
    TLBM_HIP(hipMemcpyAsync(h->inlet, v.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    TLBM_HIP(hipStreamSynchronize(h->stream));
    h->args.inlet_uz = h->inlet;
    return TLBM_OK;
}

int tlbm_gpu_box_filter(tlbm_gpu *h, double *out) {
    if (int rc = check(h)) return rc;
    if (!out) return fail(TLBM_EINVAL, "null output");
    if ((h->args.nx | h->args.ny | h->args.nz) & 1u)
        return fail(TLBM_EINVAL, "the box filter needs even nx, ny, nz");
    if (h->steps_since_set < 1) return fail(TLBM_ESTATE, "the box filter reads the last step's moments");
    if (h->args.closure)
        return fail(TLBM_ESTATE, "the box filter's stress planes need the BGK or Smagorinsky relaxation");
    if (!h->filt_in) {
        const size_t b = size_t(10) * h->N * size_t(h->ops->macro_bytes);
        TLBM_HIP(hipMalloc(&h->filt_in, b));
        h->bytes += b;
    }
    // the state the last step read is still intact; uy_head holds that
    // step's type-5 values
    hipError_t e = h->ops->lbm(h->Q, tlbm::MODE_STRESS, h->f[1 - h->cur], nullptr, h->nd, nullptr,
                               h->filt_in, h->has_lid ? h->uy_head : nullptr, h->args, h->stream);
    if (e != hipSuccess) return fail(TLBM_EHIP, std::string("stresses: ") + hipGetErrorString(e));
    const size_t nc = size_t(h->args.nx / 2) * (h->args.ny / 2) * (h->args.nz / 2);
    static_assert(tlbm::kFilterPlanes == TLBM_FILTER_PLANES, "box filter planes");
    const size_t ob = size_t(tlbm::kFilterPlanes) * nc * sizeof(double);
    if (int rc = ensure_staging(h, ob)) return rc;
    e = h->ops->box_filter(h->filt_in, h->staging, h->args, h->stream);
    if (e != hipSuccess) return fail(TLBM_EHIP, std::string("box filter: ") + hipGetErrorString(e));
    TLBM_HIP(hipMemcpyAsync(out, h->staging, ob, hipMemcpyDeviceToHost, h->stream));
    TLBM_HIP(hipStreamSynchronize(h->stream));
    return TLBM_OK;
}

int tlbm_gpu_memory_bytes(const tlbm_gpu *h, uint64_t *bytes) {
    if (!h || !bytes) return fail(TLBM_EINVAL, "null argument");
    *bytes = h->bytes;
    return TLBM_OK;
}

}  // extern "C"
