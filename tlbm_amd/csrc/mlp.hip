// mlp.hip -- training and stand-alone inference of the closure network
// (mlp.h), and their C-ABI (include/tlbm_gpu.h, tlbm_mlp_*).
//
// Training: each step draws `batch` samples (a counter-based hash of seed,
// step and position picks each index), one thread per sample computes the
// forward and backward pass in FP32 with the parameters in LDS, and the
// workgroup turns its 128 samples' activations and deltas into the weight
// gradient as small LDS matrix products.  Per-workgroup partial gradients
// are folded in a fixed order by a second kernel, which applies Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mlp.h"
#include "tlbm_gpu.h"

namespace tlbm {

constexpr int kTB = 128;            // samples per training workgroup
constexpr int kGradLen = kMlpParams + 1;  // + loss

struct ParamView {
    const float *W1, *b1, *W2, *b2, *w3, *b3;
    __device__ ParamView(const float *p)
        : W1(p), b1(p + kMlpH * kMlpIn), W2(b1 + kMlpH), b2(W2 + kMlpH * kMlpH), w3(b2 + kMlpH),
          b3(w3 + kMlpH) {}
};

__device__ inline uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// mode 0: sample i of the batch is index first + i; mode 1: a hashed index.
__global__ void __launch_bounds__(kTB)
k_mlp_grad(const float *__restrict__ X, const float *__restrict__ Tg, int64_t M, int64_t first,
           int mode, uint64_t seed, int64_t step, int64_t batch, const float *__restrict__ params,
           float *__restrict__ partial) {
    __shared__ float sp[kMlpParams];
    __shared__ float sa[kTB][kMlpH + 1];
    __shared__ float sb[kTB][kMlpH + 1];
    __shared__ float sred[kTB];
    for (int k = threadIdx.x; k < kMlpParams; k += kTB) sp[k] = params[k];
    __syncthreads();
    const ParamView P(sp);
    const int tid = threadIdx.x;
    const int64_t pos = int64_t(blockIdx.x) * kTB + tid;
    const bool live = pos < batch;
    int64_t idx = 0;
    if (live) idx = mode == 0 ? first + pos : int64_t(mix64(seed ^ mix64(uint64_t(step) * 0x100000001ull + uint64_t(pos))) % uint64_t(M));
    float x[kMlpIn], h1[kMlpH], h2[kMlpH];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) x[i] = live ? X[idx * kMlpIn + i] : 0.f;
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) {
        float a = P.b1[j];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) a += P.W1[j * kMlpIn + i] * x[i];
        h1[j] = fmaxf(a, 0.f);
    }
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) {
        float a = P.b2[j];
#pragma unroll
        for (int i = 0; i < kMlpH; ++i) a += P.W2[j * kMlpH + i] * h1[i];
        h2[j] = fmaxf(a, 0.f);
    }
    float z = P.b3[0];
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) z += P.w3[j] * h2[j];
    const float y = z > 20.f ? z : log1pf(expf(z));
    const float r = live ? y - Tg[idx] : 0.f;
    const float loss = 0.5f * r * r;
    const float dz = r / (1.f + expf(-z)) / float(batch);
    float d2[kMlpH];
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) d2[j] = h2[j] > 0.f ? dz * P.w3[j] : 0.f;
    // phase A: dW2 = sum_s d2 h1^T, db2
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) {
        sa[tid][j] = h1[j];
        sb[tid][j] = d2[j];
    }
    __syncthreads();
    float *out = partial + size_t(blockIdx.x) * kGradLen;
    float *gW1 = out, *gb1 = gW1 + kMlpH * kMlpIn, *gW2 = gb1 + kMlpH, *gb2 = gW2 + kMlpH * kMlpH,
          *gw3 = gb2 + kMlpH, *gb3 = gw3 + kMlpH;
    for (int e = tid; e < kMlpH * kMlpH; e += kTB) {
        const int j = e / kMlpH, i = e % kMlpH;
        float acc = 0.f;
        for (int s = 0; s < kTB; ++s) acc += sb[s][j] * sa[s][i];
        gW2[e] = acc;
    }
    if (tid < kMlpH) {
        float acc = 0.f;
        for (int s = 0; s < kTB; ++s) acc += sb[s][tid];
        gb2[tid] = acc;
    }
    // d1 = (W2^T d2) masked by h1 > 0
    float d1[kMlpH];
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) {
        float a = 0.f;
#pragma unroll
        for (int j = 0; j < kMlpH; ++j) a += P.W2[j * kMlpH + i] * d2[j];
        d1[i] = h1[i] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // phase B: dW1 = sum_s d1 x^T, db1
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) sa[tid][j] = d1[j];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) sb[tid][i] = x[i];
    __syncthreads();
    for (int e = tid; e < kMlpH * kMlpIn; e += kTB) {
        const int j = e / kMlpIn, i = e % kMlpIn;
        float acc = 0.f;
        for (int s = 0; s < kTB; ++s) acc += sa[s][j] * sb[s][i];
        gW1[e] = acc;
    }
    if (tid < kMlpH) {
        float acc = 0.f;
        for (int s = 0; s < kTB; ++s) acc += sa[s][tid];
        gb1[tid] = acc;
    }
    __syncthreads();
    // phase C: dw3 = sum_s dz h2, db3, loss
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) sa[tid][j] = h2[j];
    sb[tid][0] = dz;
    sred[tid] = loss;
    __syncthreads();
    if (tid < kMlpH) {
        float acc = 0.f;
        for (int s = 0; s < kTB; ++s) acc += sb[s][0] * sa[s][tid];
        gw3[tid] = acc;
    }
    if (tid == kMlpH) {
        float acc = 0.f;
        for (int s = 0; s < kTB; ++s) acc += sb[s][0];
        gb3[0] = acc;
    }
    if (tid == kMlpH + 1) {
        float acc = 0.f;
        for (int s = 0; s < kTB; ++s) acc += sred[s];
        out[kMlpParams] = acc;
    }
}

// Fold the partial gradients in block order; with adam != 0 apply Adam.
// grad_out (may be null) receives the folded gradient, loss_out the mean loss.
__global__ void __launch_bounds__(256)
k_mlp_fold_adam(const float *__restrict__ partial, int nblocks, int64_t batch, int adam,
                float *__restrict__ params, float *__restrict__ mom, float *__restrict__ vel,
                float lr, float c1, float c2, float *__restrict__ grad_out,
                double *__restrict__ loss_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == kMlpParams) {
        double acc = 0.0;
        for (int b = 0; b < nblocks; ++b) acc += double(partial[size_t(b) * kGradLen + kMlpParams]);
        *loss_out = acc / double(batch);
        return;
    }
    if (k > kMlpParams) return;
    float g = 0.f;
    for (int b = 0; b < nblocks; ++b) g += partial[size_t(b) * kGradLen + k];
    if (grad_out) grad_out[k] = g;
    if (adam) {
        const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
        const float m = beta1 * mom[k] + (1.f - beta1) * g;
        const float v = beta2 * vel[k] + (1.f - beta2) * g * g;
        mom[k] = m;
        vel[k] = v;
        params[k] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    }
}

__global__ void __launch_bounds__(256)
k_mlp_infer(const float *__restrict__ X, int64_t n, const MlpPacked *__restrict__ m,
            float *__restrict__ z) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int64_t src = i < n ? i : n - 1;  // whole wavefronts take part
    float x[kMlpIn];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) x[k] = X[src * kMlpIn + k];
    const float r = mlp_wave(x, m);
    if (i < n) z[i] = r;
}

}  // namespace tlbm

namespace tlbm {
void set_last_error(const std::string &msg);  // tlbm_gpu.cpp
}

namespace {

int mfail(int code, const std::string &msg) {
    tlbm::set_last_error(msg);
    return code;
}

#define MLP_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return mfail(TLBM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

}  // namespace

struct tlbm_mlp_trainer {
    int64_t n = 0;
    float *X = nullptr, *T = nullptr, *params = nullptr, *mom = nullptr, *vel = nullptr;
    float *partial = nullptr, *grad = nullptr;
    double *loss = nullptr;
    int max_blocks = 0;
    int64_t steps = 0;
    uint64_t seed = 0;
    hipStream_t stream = nullptr;
};

extern "C" {

int tlbm_mlp_trainer_create(const float *x, const float *t, int64_t n, const float *init,
                            uint64_t seed, int max_batch, tlbm_mlp_trainer **out) {
    if (!x || !t || !init || !out || n <= 0 || max_batch <= 0)
        return mfail(TLBM_EINVAL, "bad trainer arguments");
    *out = nullptr;
    tlbm_mlp_trainer *tr = new (std::nothrow) tlbm_mlp_trainer;
    if (!tr) return mfail(TLBM_ENOMEM, "host allocation failed");
    tr->n = n;
    tr->seed = seed;
    tr->max_blocks = (max_batch + tlbm::kTB - 1) / tlbm::kTB;
    auto bail = [&](int rc) {
        tlbm_mlp_trainer_destroy(tr);
        return rc;
    };
    const size_t P = tlbm::kMlpParams;
    if (hipStreamCreateWithFlags(&tr->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&tr->X, sizeof(float) * size_t(n) * tlbm::kMlpIn) != hipSuccess ||
        hipMalloc(&tr->T, sizeof(float) * size_t(n)) != hipSuccess ||
        hipMalloc(&tr->params, sizeof(float) * P) != hipSuccess ||
        hipMalloc(&tr->mom, sizeof(float) * P) != hipSuccess ||
        hipMalloc(&tr->vel, sizeof(float) * P) != hipSuccess ||
        hipMalloc(&tr->grad, sizeof(float) * P) != hipSuccess ||
        hipMalloc(&tr->loss, sizeof(double)) != hipSuccess ||
        hipMalloc(&tr->partial, sizeof(float) * size_t(tr->max_blocks) * tlbm::kGradLen) != hipSuccess)
        return bail(mfail(TLBM_ENOMEM, "device allocation failed"));
    if (hipMemcpy(tr->X, x, sizeof(float) * size_t(n) * tlbm::kMlpIn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(tr->T, t, sizeof(float) * size_t(n), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(tr->params, init, sizeof(float) * P, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(tr->mom, 0, sizeof(float) * P) != hipSuccess ||
        hipMemset(tr->vel, 0, sizeof(float) * P) != hipSuccess)
        return bail(mfail(TLBM_EHIP, "upload failed"));
    *out = tr;
    return TLBM_OK;
}

int tlbm_mlp_trainer_destroy(tlbm_mlp_trainer *tr) {
    if (!tr) return TLBM_OK;
    if (tr->stream) (void)hipStreamSynchronize(tr->stream);
    for (void *p : {(void *)tr->X, (void *)tr->T, (void *)tr->params, (void *)tr->mom,
                    (void *)tr->vel, (void *)tr->partial, (void *)tr->grad, (void *)tr->loss})
        if (p) (void)hipFree(p);
    if (tr->stream) (void)hipStreamDestroy(tr->stream);
    delete tr;
    return TLBM_OK;
}

int tlbm_mlp_trainer_step(tlbm_mlp_trainer *tr, int steps, int batch, float lr, double *last_loss) {
    if (!tr || steps < 0 || batch <= 0) return mfail(TLBM_EINVAL, "bad step arguments");
    const int nb = (batch + tlbm::kTB - 1) / tlbm::kTB;
    if (nb > tr->max_blocks) return mfail(TLBM_EINVAL, "batch larger than the trainer's max_batch");
    for (int k = 0; k < steps; ++k) {
        ++tr->steps;
        hipLaunchKernelGGL(tlbm::k_mlp_grad, dim3(nb), dim3(tlbm::kTB), 0, tr->stream, tr->X, tr->T,
                           tr->n, int64_t(0), 1, tr->seed, tr->steps, int64_t(batch), tr->params,
                           tr->partial);
        MLP_HIP(hipGetLastError());
        const float c1 = 1.f - std::pow(0.9f, float(tr->steps));
        const float c2 = 1.f - std::pow(0.999f, float(tr->steps));
        hipLaunchKernelGGL(tlbm::k_mlp_fold_adam, dim3((tlbm::kMlpParams + 1 + 255) / 256), dim3(256),
                           0, tr->stream, tr->partial, nb, int64_t(batch), 1, tr->params, tr->mom,
                           tr->vel, lr, c1, c2, (float *)nullptr, tr->loss);
        MLP_HIP(hipGetLastError());
    }
    if (last_loss) {
        MLP_HIP(hipMemcpyAsync(last_loss, tr->loss, sizeof(double), hipMemcpyDeviceToHost, tr->stream));
    }
    MLP_HIP(hipStreamSynchronize(tr->stream));
    return TLBM_OK;
}

int tlbm_mlp_trainer_loss_grad(tlbm_mlp_trainer *tr, int64_t first, int64_t count, double *loss,
                               float *grad) {
    if (!tr || first < 0 || count <= 0 || first + count > tr->n || !loss || !grad)
        return mfail(TLBM_EINVAL, "bad loss_grad arguments");
    const int nb = int((count + tlbm::kTB - 1) / tlbm::kTB);
    if (nb > tr->max_blocks) return mfail(TLBM_EINVAL, "count larger than the trainer's max_batch");
    hipLaunchKernelGGL(tlbm::k_mlp_grad, dim3(nb), dim3(tlbm::kTB), 0, tr->stream, tr->X, tr->T, tr->n,
                       first, 0, tr->seed, int64_t(0), count, tr->params, tr->partial);
    MLP_HIP(hipGetLastError());
    hipLaunchKernelGGL(tlbm::k_mlp_fold_adam, dim3((tlbm::kMlpParams + 1 + 255) / 256), dim3(256), 0,
                       tr->stream, tr->partial, nb, count, 0, tr->params, tr->mom, tr->vel, 0.f, 1.f,
                       1.f, tr->grad, tr->loss);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(grad, tr->grad, sizeof(float) * tlbm::kMlpParams, hipMemcpyDeviceToHost, tr->stream));
    MLP_HIP(hipMemcpyAsync(loss, tr->loss, sizeof(double), hipMemcpyDeviceToHost, tr->stream));
    MLP_HIP(hipStreamSynchronize(tr->stream));
    return TLBM_OK;
}

int tlbm_mlp_trainer_params(tlbm_mlp_trainer *tr, float *params) {
    if (!tr || !params) return mfail(TLBM_EINVAL, "null argument");
    MLP_HIP(hipMemcpyAsync(params, tr->params, sizeof(float) * tlbm::kMlpParams, hipMemcpyDeviceToHost, tr->stream));
    MLP_HIP(hipStreamSynchronize(tr->stream));
    return TLBM_OK;
}

int tlbm_mlp_infer(const float *params, const float *x, int64_t n, float *z) {
    if (!params || !x || !z || n <= 0) return mfail(TLBM_EINVAL, "bad inference arguments");
    tlbm::MlpPacked host;
    tlbm::mlp_pack(params, &host);
    tlbm::MlpPacked *dm = nullptr;
    float *dx = nullptr, *dz = nullptr;
    int rc = TLBM_OK;
    if (hipMalloc(&dm, sizeof host) != hipSuccess || hipMalloc(&dx, sizeof(float) * size_t(n) * tlbm::kMlpIn) != hipSuccess ||
        hipMalloc(&dz, sizeof(float) * size_t(n)) != hipSuccess) {
        rc = mfail(TLBM_ENOMEM, "device allocation failed");
    } else if (hipMemcpy(dm, &host, sizeof host, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dx, x, sizeof(float) * size_t(n) * tlbm::kMlpIn, hipMemcpyHostToDevice) != hipSuccess) {
        rc = mfail(TLBM_EHIP, "upload failed");
    } else {
        hipLaunchKernelGGL(tlbm::k_mlp_infer, dim3(unsigned((n + 255) / 256)), dim3(256), 0, 0, dx, n, dm, dz);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(z, dz, sizeof(float) * size_t(n), hipMemcpyDeviceToHost) != hipSuccess)
            rc = mfail(TLBM_EHIP, "inference failed");
    }
    if (dm) (void)hipFree(dm);
    if (dx) (void)hipFree(dx);
    if (dz) (void)hipFree(dz);
    return rc;
}

}  // extern "C"
