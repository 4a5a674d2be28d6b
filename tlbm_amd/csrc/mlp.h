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
// (lanes 0-31 and 32-63) and both hidden layers are bf16 MFMA with FP32
// accumulation, the weights as the A operand and 32 nodes as the columns:
//   layer 1: H1^T = W1 (32x8, input 7 padded) x X^T (8 x 32 nodes), one
//            v_mfma_f32_32x32x8_bf16,
//   layer 2: H2^T = W2 (32x32) x H1^T, two v_mfma_f32_32x32x16_bf16 (K = 16
//            each; the full-rate bf16 instruction of gfx950).
// In the 32x32 output layout lane l holds column l%32 and, in register
// 4q + r, row 8q + 4(l/32) + r (q, r = 0..3).  A K = 16 MFMA multiplies
// item i of lane group g (= l/32) of A with item i of lane group g of B,
// so any assignment of the 16 K indices to (g, i) gives the same sum as
// long as A and B agree on it: K block kb takes, as its 8 items in lane l,
// layer 1's registers 8kb .. 8kb+7 (units 16kb + 4g + r and
// 16kb + 8 + 4g + r, r = 0..3), and W2 is packed in that order.  So
// layer 1's output feeds layer 2 in registers.  The only cross-lane moves
// are four v_permlane32_swap of the features (they build both blocks' B
// operands at once) and one for the final sum of the two lane halves; no
// LDS round trip.  Weights, features and h1 are rounded to bf16 (round to nearest
// even); biases, h2 and the last layer are FP32.  The whole wavefront must
// execute mlp_wave together.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tlbm {

constexpr int kMlpIn = 7;
constexpr int kMlpH = 32;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;

// Weights laid out per lane for the MFMA operands (built by mlp_pack).  The
// values a lane reads from LDS sit at small constant offsets from one base
// per lane half (hb[h]) or per lane (a1[l], a2[kb][l]), so the reads take
// their offsets as immediates instead of an address computation each.
struct alignas(16) MlpHalf {
    float b1[16];  // b1 of unit 8q+4h+r at [4q+r]
    float b2[16];
    float w3[16];  // w3 rounded to bf16
};
struct alignas(16) MlpPacked {
    MlpHalf hb[2];
    float b3, pad[3];
    uint32_t a1[64][2];     // W1[l%32][4(l/32)+t], t = 0..3, bf16 pairs
    uint32_t a2[2][64][4];  // W2[l%32][u], u = 16kb + 8(t/4) + 4(l/32) + t%4, t = 0..7
};

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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
        for (int kb = 0; kb < 2; ++kb) {
            float u8[8];
            for (int t = 0; t < 8; ++t) u8[t] = W2[i * kMlpH + 16 * kb + 8 * (t / 4) + 4 * h + t % 4];
            for (int p2 = 0; p2 < 4; ++p2) m->a2[kb][l][p2] = pair(u8[2 * p2], u8[2 * p2 + 1]);
        }
    }
    for (int h = 0; h < 2; ++h)
        for (int q = 0; q < 4; ++q)
            for (int r = 0; r < 4; ++r) {
                const int u = 8 * q + 4 * h + r;
                m->hb[h].b1[4 * q + r] = b1[u];
                m->hb[h].b2[4 * q + r] = b2[u];
                m->hb[h].w3[4 * q + r] = bf16_to_f32(f32_to_bf16(w3[u]));
            }
    m->b3 = b3[0];
    m->pad[0] = m->pad[1] = m->pad[2] = 0.f;
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Two floats to a bf16 pair, round to nearest even (v_cvt_pk_bf16_f32).
__device__ inline uint32_t cvt_pk_bf16(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

__device__ inline bf16x4 pack4(float a, float b, float c, float d) {
    const u32x2 w = {cvt_pk_bf16(a, b), cvt_pk_bf16(c, d)};
    return __builtin_bit_cast(bf16x4, w);
}

__device__ inline bf16x4 load_pair(const uint32_t (&w)[2]) {
    const u32x2 v = {w[0], w[1]};
    return __builtin_bit_cast(bf16x4, v);
}

// max(x, 0) as a signed-integer max of the bits: negative floats (sign bit
// set) are negative integers; no NaN canonicalisation is inserted.
__device__ inline float relu(float x) {
    return __builtin_bit_cast(float, max(__builtin_bit_cast(int, x), 0));
}

// relu of both halves of a bf16 pair (v_pk_max_i16 with 0); rounding and
// relu commute, so relu(round(x)) = round(relu(x)).
typedef short i16x2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t relu_bf16x2(uint32_t v) {
    const i16x2 z = {0, 0};
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2, v), z));
}

__device__ inline bf16x8 pack8_relu(float a, float b, float c, float d, float e, float f,
                                    float g, float h) {
    const uint32_t w[4] = {relu_bf16x2(cvt_pk_bf16(a, b)), relu_bf16x2(cvt_pk_bf16(c, d)),
                           relu_bf16x2(cvt_pk_bf16(e, f)), relu_bf16x2(cvt_pk_bf16(g, h))};
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = {w[0], w[1], w[2], w[3]};
    return __builtin_bit_cast(bf16x8, v);
}

__device__ inline bf16x8 load_quad(const uint32_t (&w)[4]) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = {w[0], w[1], w[2], w[3]};
    return __builtin_bit_cast(bf16x8, v);
}

// 0, as a value the compiler cannot see through (an empty asm statement
// owns it), so that loads through it are not merged with the same loads
// elsewhere.
__device__ inline uint32_t opaque_zero() {
    uint32_t z = 0;
    __asm__ volatile("" : "+s"(z));
    return z;
}

// v_permlane32_swap_b32: lanes 32-63 of a trade places with lanes 0-31 of b.
// Afterwards lane l < 32 of b holds a's lane l + 32 and lane l >= 32 of a
// holds b's lane l - 32; the other halves stay.
__device__ inline void swap_halves(float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, a),
                                                    __builtin_bit_cast(uint32_t, b), false, false);
    // the elements go through named scalars: this clang reads
    // __builtin_bit_cast(float, r[1]) from r's first element
    const uint32_t r0 = r[0], r1 = r[1];
    a = __builtin_bit_cast(float, r0);
    b = __builtin_bit_cast(float, r1);
}

__device__ inline f32x16 load16(const float (&v)[16]) {
    f32x16 r;
#pragma unroll
    for (int k = 0; k < 16; ++k) r[k] = v[k];
    return r;
}

// z of this lane's node.  All 64 lanes must call it (idle lanes pass any
// finite features and ignore the result).  The two 32-node blocks run one
// after the other so that one set of accumulators is live at a time; the
// biases are the accumulators' initial values (bias first, as the sums of
// the network's definition run).  The output layer is FP32 on the vector
// ALU: each lane sums w3 relu(h2) over its 16 units, the two lane halves
// add up.
__device__ inline float mlp_wave(const float (&x)[kMlpIn], const MlpPacked *__restrict__ m) {
    const int l = int(threadIdx.x & 63u);
    const int h = l >> 5;
    // layer-1 B operand: block 0 = nodes of lanes 0-31, block 1 = lanes 32-63;
    // lane l supplies inputs 4h..4h+3 of node l%32 of each block.  Swapping
    // the upper half of (x0..x3) with the lower half of (x4..x6, 0) leaves
    // block 0's operand in the first and block 1's in the second.
    float lo[4] = {x[0], x[1], x[2], x[3]}, hi[4] = {x[4], x[5], x[6], 0.f};
#pragma unroll
    for (int k = 0; k < 4; ++k) swap_halves(lo[k], hi[k]);
    const bf16x4 xb[2] = {pack4(lo[0], lo[1], lo[2], lo[3]), pack4(hi[0], hi[1], hi[2], hi[3])};
    const bf16x4 a1 = load_pair(m->a1[l]);
    float part[2];
#pragma unroll
    for (int blk = 0; blk < 2; ++blk) {
        // each block reads its biases afresh (opaque_zero): they go straight
        // into the accumulators that the MFMAs then overwrite, instead of
        // staying in registers for the other block
        const MlpHalf &H = m->hb[h + int(opaque_zero())];
        const f32x16 c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, xb[blk], load16(H.b1), 0, 0, 0);
        f32x16 d = load16(H.b2);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            // registers 8kb .. 8kb+7: the 8 items of K block kb in this lane
            const bf16x8 hk = pack8_relu(c[8 * kb], c[8 * kb + 1], c[8 * kb + 2], c[8 * kb + 3],
                                         c[8 * kb + 4], c[8 * kb + 5], c[8 * kb + 6], c[8 * kb + 7]);
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(load_quad(m->a2[kb][l]), hk, d, 0, 0, 0);
        }
        float p = 0.f;
#pragma unroll
        for (int r = 0; r < 16; ++r) p += H.w3[r] * relu(d[r]);
        part[blk] = p;
    }
    // block 0's halves meet in the lower lanes, block 1's in the upper ones
    swap_halves(part[0], part[1]);
    return (part[0] + part[1]) + m->b3;
}

// softplus with the hardware exp2/log2 (v_exp_f32, v_log_f32, 1 ulp each):
// the argument of the logarithm is at least 1, so neither needs range
// handling; a value below float resolution of 1 + e^z is dropped (it moves
// the relaxation time by less than 1e-7).
__device__ inline float softplus_f(float z) {
    return z > 20.f ? z
                    : __builtin_amdgcn_logf(1.f + __builtin_amdgcn_exp2f(z * 1.44269504f)) * 0.693147181f;
}

}  // namespace tlbm
