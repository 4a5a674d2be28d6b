// cabi_stub.cpp -- a host-only test double of the C-ABI (include/tlbm_gpu.h)
// for driving the drop-in driver's control flow on a machine without a GPU.
// It computes nothing: steps are counted, moments and forces are zeros, and
// tlbm_gpu_reduce reports a non-finite state from the step named by the
// environment variable TLBM_STUB_NONFINITE_FROM on (never when unset), the
// way the real library reports a diverged run.  Only the tests link it.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tlbm_gpu.h"

struct tlbm_gpu {
    size_t N = 0;
    int Q = 0;
    long long steps = 0;
    std::vector<double> f;
};

namespace {
std::string g_err;
int fail(int code, const char *msg) {
    g_err = msg;
    return code;
}
}  // namespace

extern "C" {

const char *tlbm_gpu_last_error(void) { return g_err.c_str(); }

int tlbm_gpu_create(const tlbm_params *p, const uint8_t *, tlbm_gpu **out) {
    if (!p || !out) return fail(TLBM_EINVAL, "null argument");
    tlbm_gpu *h = new tlbm_gpu;
    h->N = size_t(p->nx) * p->ny * p->nz;
    h->Q = p->lattice;
    h->f.assign(h->N * h->Q, 0.0);
    *out = h;
    return TLBM_OK;
}

int tlbm_gpu_destroy(tlbm_gpu *h) {
    delete h;
    return TLBM_OK;
}

int tlbm_gpu_step(tlbm_gpu *h, int nsteps) {
    h->steps += nsteps;
    return TLBM_OK;
}

int tlbm_gpu_sync(tlbm_gpu *) { return TLBM_OK; }

int tlbm_gpu_get_macros(tlbm_gpu *h, double *ux, double *uy, double *uz, double *rho) {
    for (size_t i = 0; i < h->N; ++i) {
        ux[i] = uy[i] = uz[i] = 0.0;
        rho[i] = 1.0;
    }
    return TLBM_OK;
}

int tlbm_gpu_get_forces(tlbm_gpu *h, double *fx, double *fy, double *fz) {
    for (size_t i = 0; i < h->N; ++i) fx[i] = fy[i] = fz[i] = 0.0;
    return TLBM_OK;
}

int tlbm_gpu_get_populations(tlbm_gpu *h, double *f) {
    std::memcpy(f, h->f.data(), h->f.size() * sizeof(double));
    return TLBM_OK;
}

int tlbm_gpu_set_populations(tlbm_gpu *h, const double *f) {
    std::memcpy(h->f.data(), f, h->f.size() * sizeof(double));
    return TLBM_OK;
}

int tlbm_gpu_init_from_macros(tlbm_gpu *, const double *, const double *, const double *,
                              const double *) {
    return TLBM_OK;
}

int tlbm_gpu_reduce(tlbm_gpu *h, double out[5]) {
    for (int k = 0; k < 5; ++k) out[k] = 0.0;
    out[0] = double(h->N);
    const char *from = std::getenv("TLBM_STUB_NONFINITE_FROM");
    if (from && h->steps >= std::atoll(from)) {
        out[0] = NAN;
        char buf[128];
        std::snprintf(buf, sizeof buf, "%zu populations are NaN or infinite after %lld steps",
                      h->N, h->steps);
        return fail(TLBM_ENONFINITE, buf);
    }
    return TLBM_OK;
}

int tlbm_gpu_time_average(tlbm_gpu *, int) { return TLBM_OK; }

int tlbm_gpu_get_time_average(tlbm_gpu *h, double *u, double *v, double *w, double *rho,
                              int64_t *nsteps) {
    for (size_t i = 0; i < h->N; ++i) u[i] = v[i] = w[i] = rho[i] = 0.0;
    if (nsteps) *nsteps = 0;
    return TLBM_OK;
}

}  // extern "C"
