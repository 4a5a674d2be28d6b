// mlp.h -- the learned closure network and its wavefront-level inference.
//
// Network: x (7 features) -> 32 relu -> 32 relu -> z -> softplus.
// Features of a node: the non-equilibrium stress Pi = sum_s e e (f - feq)
// (xx, yy, zz, xy, xz, yz) and its norm
// Qn = sqrt(Pi_xx^2 + Pi_yy^2 + Pi_zz^2 + 2 (Pi_xy^2 + Pi_xz^2 + Pi_yz^2)),
// the quantities the Smagorinsky closure reads
// (include/LatticeStructure.hpp:105-122, 140-153), each times a feature scale.
// The relaxation time becomes tau0 + output_scale * softplus(z).
//
// Parameters, float, flat: W1 [32][7], b1 [32], W2 [32][32], b2 [32],
// w3 [32], b3 -- kMlpParams values.
//
// Inference runs on a whole wavefront: its 64 nodes form two blocks of 32
// (lanes 0-31 and 32-63) and both hidden layers are 32x32x8 bf16 MFMA
// (v_mfma_f32_32x32x8_bf16, FP32 accumulation) with the weights as the A
// operand and 32 nodes as the columns:
//   layer 1: H1^T = W1 (32x8, input 7 padded) x X^T (8 x 32 nodes), 1 MFMA,
//   layer 2: H2^T = W2 (32x32) x H1^T, 4 MFMAs (K = 8 each).
// In the MFMA layout lane l holds column l%32 and the rows
// 8q + 4(l/32) + r (q, r = 0..3) of the output, which are exactly the K
// rows layer 2 needs from lane l, so layer 1's output feeds layer 2 in
// registers.  The only cross-lane moves are the 7 features (lane l <->
// l^32) and the final sum of the two lane halves.  Weights, features and
// h1 are rounded to bf16 (round to nearest even); biases, h2 and the last
// layer are FP32.  The whole wavefront must execute mlp_wave together.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tlbm {

constexpr int kMlpIn = 7;
constexpr int kMlpH = 32;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;

// Weights laid out per lane for the MFMA operands (built by mlp_pack).
struct MlpPacked {
    uint32_t a1[64][2];     // W1[l%32][4(l/32)+t], t = 0..3, bf16 pairs
    uint32_t a2[4][64][2];  // W2[l%32][8kb+4(l/32)+t]
    float b1[2][16];        // b1 of unit 8q+4h+r at [h][4q+r]
    float b2[2][16];
    float w3[2][16];        // w3 rounded to bf16
    float b3;
};

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0x007fffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    const uint32_t u = uint32_t(h) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// Host-side packing of the flat float parameters.
inline void mlp_pack(const float *p, MlpPacked *m) {
    const float *W1 = p, *b1 = W1 + kMlpH * kMlpIn, *W2 = b1 + kMlpH, *b2 = W2 + kMlpH * kMlpH,
                *w3 = b2 + kMlpH, *b3 = w3 + kMlpH;
    auto pair = [](float a, float b) { return uint32_t(f32_to_bf16(a)) | (uint32_t(f32_to_bf16(b)) << 16); };
    for (int l = 0; l < 64; ++l) {
        const int i = l % 32, h = l / 32;
        float v[4];
        for (int t = 0; t < 4; ++t) {
            const int k = 4 * h + t;
            v[t] = k < kMlpIn ? W1[i * kMlpIn + k] : 0.f;
        }
        m->a1[l][0] = pair(v[0], v[1]);
        m->a1[l][1] = pair(v[2], v[3]);
        for (int kb = 0; kb < 4; ++kb) {
            for (int t = 0; t < 4; ++t) v[t] = W2[i * kMlpH + 8 * kb + 4 * h + t];
            m->a2[kb][l][0] = pair(v[0], v[1]);
            m->a2[kb][l][1] = pair(v[2], v[3]);
        }
    }
    for (int h = 0; h < 2; ++h)
        for (int q = 0; q < 4; ++q)
            for (int r = 0; r < 4; ++r) {
                const int u = 8 * q + 4 * h + r;
                m->b1[h][4 * q + r] = b1[u];
                m->b2[h][4 * q + r] = b2[u];
                m->w3[h][4 * q + r] = bf16_to_f32(f32_to_bf16(w3[u]));
            }
    m->b3 = b3[0];
}

__device__ inline bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = short(f32_to_bf16(a));
    r[1] = short(f32_to_bf16(b));
    r[2] = short(f32_to_bf16(c));
    r[3] = short(f32_to_bf16(d));
    return r;
}

__device__ inline bf16x4 load_pair(const uint32_t (&w)[2]) {
    bf16x4 r;
    r[0] = short(w[0] & 0xffffu);
    r[1] = short(w[0] >> 16);
    r[2] = short(w[1] & 0xffffu);
    r[3] = short(w[1] >> 16);
    return r;
}

// z of this lane's node.  All 64 lanes must call it (idle lanes pass any
// finite features and ignore the result).
__device__ inline float mlp_wave(const float (&x)[kMlpIn], const MlpPacked *__restrict__ m) {
    const int l = int(threadIdx.x & 63u);
    const int h = l >> 5;
    float xo[kMlpIn];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) xo[k] = __shfl_xor(x[k], 32);
    // layer-1 B operand: block A = nodes of lanes 0-31, block B = lanes 32-63;
    // lane l supplies inputs 4h..4h+3 of node l%32 of each block
    const bf16x4 xa = h ? pack4(xo[4], xo[5], xo[6], 0.f) : pack4(x[0], x[1], x[2], x[3]);
    const bf16x4 xb = h ? pack4(x[4], x[5], x[6], 0.f) : pack4(xo[0], xo[1], xo[2], xo[3]);
    const bf16x4 a1 = load_pair(m->a1[l]);
    const f32x16 zero = {0.f};
    const f32x16 ca = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, xa, zero, 0, 0, 0);
    const f32x16 cb = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, xb, zero, 0, 0, 0);
    f32x16 da = zero, db = zero;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        // register 4kb+t holds unit 8kb+4h+t: the K rows of block kb
        float ha[4], hb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float bias = m->b1[h][4 * kb + t];
            ha[t] = fmaxf(ca[4 * kb + t] + bias, 0.f);
            hb[t] = fmaxf(cb[4 * kb + t] + bias, 0.f);
        }
        const bf16x4 a2 = load_pair(m->a2[kb][l]);
        da = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2, pack4(ha[0], ha[1], ha[2], ha[3]), da, 0, 0, 0);
        db = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2, pack4(hb[0], hb[1], hb[2], hb[3]), db, 0, 0, 0);
    }
    float pa = 0.f, pb = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const float w = m->w3[h][r];
        const float bias = m->b2[h][r];
        pa += w * fmaxf(da[r] + bias, 0.f);
        pb += w * fmaxf(db[r] + bias, 0.f);
    }
    pa += __shfl_xor(pa, 32);
    pb += __shfl_xor(pb, 32);
    return (h ? pb : pa) + m->b3;
}

__device__ inline float softplus_f(float z) { return z > 20.f ? z : log1pf(__expf(z)); }

}  // namespace tlbm
