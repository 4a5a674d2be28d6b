// lbm_kernels.h -- fused pull stream-collide kernel for gfx950 and the
// per-node code it shares with the moment (dump) and type-5 helper passes.
//
// Layout in HBM: chunked structure of arrays.  Nodes g = x + nx*y + nx*ny*z
// (TLBM_Partition.cpp:730-737) are grouped in chunks of CH = 2^cb
// consecutive nodes; a chunk stores its Q direction planes back to back:
//   f(s, g) at ((g >> cb) * Q + s) * CH + (g & (CH - 1)).
// Within a chunk the access is that of a plain structure of arrays (each
// direction a contiguous, coalesced run); chunking keeps the 2*Q streams of
// a step inside a few nearby DRAM regions, which the HBM serves closer to
// copy speed than Q planes N values apart (measured on MI355X).
// Two state buffers ping-pong.  A state buffer holds post-collision,
// not-yet-streamed populations B_n; step n+1 gathers
// A[g,s] = B_n[g - e_s (periodic), s], which is exactly what the reference's
// push stream (TLBM_Partition.cpp:826-837) would have delivered to g,
// collides at g and writes B_{n+1}[s][g] aligned.  The periodic wrap in x, y
// and z (TLBM_Partition.cpp:691-712) is index arithmetic in the gather, so
// no halo or adjacency table exists.
//
// Per-node order and arithmetic follow TLBM_Partition.cpp:742-809 and
// include/LatticeStructure.hpp:69-177.  In double precision every operation
// is the reference's, in the same order (the FP64 kernels are compiled
// without floating-point contraction), so the FP64 path reproduces the
// reference's rho and u.  Terms that multiply a population or velocity by a
// zero velocity component are left out: they add +-0, which leaves every
// nonzero sum unchanged.
//
// FP32 and FP16 storage keep d_s = f_s - w_s (FP16 scaled by 1/w_s, each
// step's store rounded stochastically) and the FP32 arithmetic works on d
// and on delta = rho - 1 throughout: the same formulas rewritten exactly
// (w is symmetric under bounce-back, sum w e = 0), without the cancellation
// of O(1) values that costs FP32 its accuracy.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.h"
#include "mlp.h"
#include "udiv.h"

namespace tlbm {

enum Storage : int { ST_F64 = 0, ST_F32 = 1, ST_F16 = 2, ST_F32X = 3 };
enum Mode : int { MODE_STEP = 0, MODE_MACROS = 1, MODE_HEAD = 2, MODE_AVG = 3, MODE_STRESS = 4 };

constexpr int kBlock = 256;  // 4 wave64 per workgroup
constexpr int kRed = 7;      // values per FP64 reduction partial
constexpr int kFilterPlanes = 16;  // box filter output planes
// workgroup of the closure collide.  With the weights staged by direct
// global-to-LDS loads, 256 threads beat 512 (128x128x512, us per step, same
// box: FP16 D3Q19 137.9 -> 132.6, D3Q15 121.6 -> 118.5, D3Q27 184.3 ->
// 177.5; FP32 D3Q19 205.1 -> 201.9, D3Q15 162.3 -> 160.1, D3Q27 291.6 ->
// 289.7): a smaller workgroup waits at the weights' barrier for fewer waves.
constexpr int kClosureBlock = 256;
// non-temporal population loads in the closure collide: FP16 only (FP32:
// 221.8 against 217.6 us per step with them)
template <int ST> constexpr bool kClosureNTLoad = (ST == ST_F16);
// waves per SIMD the closure collide is register-allocated for.  The
// compiler alone stops at 4 (3 for D3Q27); more waves hide the latency of
// the network's MFMA chain.  Measured on 128x128x512 (us per step, 4 waves ->
// this): FP16 D3Q19 156.5 -> 149.1 (6), D3Q15 141.9 -> 130.8 (6), D3Q27
// 316.8 -> 197.6 (5); FP32 D3Q19 216.5 -> 214.5 (5), D3Q15 181.9 -> 167.8
// (6).  FP32 D3Q27 stays at 4: at 5 it spills (325 against 299 us).
template <int ST, int Q>
constexpr int kClosureWaves = Q == 27 ? (ST == ST_F16 ? 5 : 4) : (ST == ST_F32 && Q == 19 ? 5 : 6);
// Steps with kTypeAhead, and the closure collide in every storage, load the
// chunk's common node type (or the node's own) ahead of their gathers, so
// that it lands with them; loaded after them it costs each wavefront one
// more memory round trip before the collision.  That pays where the step is
// not HBM-bound: FP16 (128x128x512, us per step, same box: D3Q19 108.7 ->
// 104.9, D3Q15 90.9 -> 82.9, D3Q27 155.5 -> 152.0) and the closure collide
// (FP16 D3Q19 145.5 -> 141.5, D3Q15 127.1 -> 124.9, D3Q27 188.8 -> 189.7;
// FP32 D3Q19 206.9 -> 205.4).  The FP32 and FP64 steps hide that round trip
// already and ran 0.1-0.9 % slower with the load ahead.
template <int ST> constexpr bool kTypeAhead = (ST == ST_F16);
// The step raises its wavefront's priority (s_setprio 3) while it issues
// its gathers and drops it once they are issued, so that wavefronts still
// issuing loads win the arbiter over those computing.  Where it pays is a
// matter of measurement (128x128x512, us per step, same box, two runs each):
// D3Q27 FP32 292.8 -> 289.4, FP16 151.4 -> 146.8, FP64 594.2 -> 587.3; D3Q19
// FP16 104.9 -> 103.0.  It cost D3Q19 FP32 (201.8 -> 205.1), D3Q19 FP64
// (401.0 -> 418.2) and D3Q15 FP16 (81.9 -> 84.2), and left D3Q15 FP32/FP64
// unchanged, so those keep the default priority.  A second box confirmed
// it, FP32X D3Q27 included (291.9 -> 289.4).
template <int ST, int Q> constexpr bool kGatherPrio = (Q == 27) || (ST == ST_F16 && Q == 19);

// The first of the node-type loads: the chunk's common type (0xFF where its
// nodes differ) when the step has the chunk table, else the node's type.
template <int MODE>
__device__ inline uint32_t type_first(const uint8_t *nd_type, const uint8_t *chunk_type, uint32_t chunk,
                                      uint32_t g) {
    return (MODE == MODE_STEP && chunk_type) ? uint32_t(chunk_type[chunk]) : uint32_t(nd_type[g]);
}

// The node's type from the first load: the node's own where its chunk is mixed.
template <int MODE>
__device__ inline int type_final(uint32_t t0, const uint8_t *nd_type, const uint8_t *chunk_type, uint32_t g) {
    return (MODE == MODE_STEP && chunk_type && t0 == 0xFFu) ? int(nd_type[g]) : int(t0);
}

// Launch arguments, passed by value (kernarg segment, scalar registers).
struct StepArgs {
    uint32_t nx, ny, nz, nxny, N;
    uint32_t cb;       // log2 of the chunk size CH (nodes)
    UDiv div_nx, div_nxny;  // g / nxny and r / nx without an integer divide
    uint32_t kb;       // chunks a pull source lies at most before its node's chunk
    int rel;           // 1: PullRel offsets fit 32 bits on this lattice
    int row_chunks;    // 1: rel, and every row lies in one chunk (CH a multiple of nx): PullRow
    uint32_t step;     // steps taken since creation: the stream of FP16 stochastic rounding
    uint64_t fbytes;   // bytes of one state buffer
    int dynamics;      // 1: BGK relaxation, other values: none (TLBM_Partition.cpp:795-807)
    int smag;          // cs > 0
    double cs, omega, nu, nu2, u_bc, rho_bc, inv_1mu, inv_rho_bc;
    double w[27];      // weights (the reference's float-literal values unless exact)
    float wf[27];      // the same weights as float (FP32/FP16 compute)
    float h_scale[27];     // FP16 storage: h = d * h_scale (= d / w)
    float h_unscale[27];   // FP16 storage: d = h * h_unscale (= h * w)
    // deviation form (FP32/FP16), each computed in double on the host
    float dw;          // sum_s w_s - 1
    float c_in;        // 2*sum_{e_z=-1} w + sum_{e_z=0} w - (1 - u_bc)
    float c_out;       // 2*sum_{e_z=+1} w + sum_{e_z=0} w - rho_bc
    float drho_bc;     // rho_bc - 1
    // the same deviation constants from the weights w in double (FP32X,
    // whose arithmetic and shift are FP64)
    double dw_d, c_in_d, c_out_d;
    // FP16: equilibrium corrections that give the float weights the
    // exact-arithmetic mass and momentum balance of the weights w (D3Q27 and
    // exact weights): 3 e.u is scaled to (3 + dt3) e.u (dt3 = 3 (W2 / W2_f - 1),
    // about -2.2e-8: 3 + dt3 itself rounds to 3 in float, so the kernel
    // adds dt3 rho u to the unrounded product 3 rho u in one FMA) and the
    // rest population gains rho (1.5 dk1 u^2 - dk0); dt3 = dk0 = dk1 = 0
    // where the two coincide and in the other storages
    float dt3, dk0, dk1;
    // exact-arithmetic moments of f - feq with the weights w (the conserving
    // relax; its float weights may sum differently): sum_s (f - feq) =
    // rho (1.5 k1 u^2 - k0), sum_s e (f - feq) = k2 rho u; k0 = sum_s w - 1,
    // k1 = 1 + k0 - 3 W2, k2 = 1 - 3 W2, W2 = sum_s w e_x^2
    float k0, k1, k2;
    // learned closure (FP32/FP16 storage): tau = tau0 + cl_out * softplus(z),
    // z = network(features * cl_in)
    int closure;
    float cl_in, cl_out, tau0;
    const MlpPacked *mlp;
    // per-node inlet velocity u_z(x, y) (nx*ny values), or null for u_bc
    const float *inlet_uz;
};

// Offset of population s of node g in a state buffer (chunked layout).
template <int Q>
__device__ inline size_t fidx(int s, uint32_t g, const StepArgs &a) {
    return ((size_t(g >> a.cb) * Q + size_t(s)) << a.cb) + (g & ((1u << a.cb) - 1u));
}

// Raw value <-> compute value.  kShift: the compute value is d = f - w.
template <int ST> struct StorageT;
template <> struct StorageT<ST_F64> {
    using st = double;
    static constexpr int kCB = 12;  // 32 KiB chunks
    using T = double;
    static constexpr bool kShift = false;
    static constexpr bool kScaled = false;
    __device__ static inline T in(st r, int, const StepArgs &) { return r; }
    __device__ static inline st out(T v, int, const StepArgs &) { return v; }
    __device__ static inline T load(const st *p, int, const StepArgs &) { return *p; }
    __device__ static inline void store(st *p, int, T v, const StepArgs &) {
        __builtin_nontemporal_store(v, p);
    }
    __device__ static inline st step_out(T v, int, const StepArgs &, uint32_t) { return v; }
    __device__ static inline T weight(int s, const StepArgs &a) { return a.w[s]; }
    __device__ static inline double absolute(T v, int, const StepArgs &) { return v; }
    __device__ static inline T from_absolute(double f, int, const StepArgs &) { return f; }
};
template <> struct StorageT<ST_F32> {
    using st = float;
    static constexpr int kCB = 13;  // 32 KiB chunks
    using T = float;
    static constexpr bool kShift = true;
    static constexpr bool kScaled = false;
    // the weight the stored deviation is taken from
    __device__ static inline double shift(int s, const StepArgs &a) { return double(a.wf[s]); }
    __device__ static inline T in(st r, int, const StepArgs &) { return r; }
    __device__ static inline st out(T v, int, const StepArgs &) { return v; }
    __device__ static inline T load(const st *p, int, const StepArgs &) { return *p; }
    __device__ static inline void store(st *p, int, T v, const StepArgs &) {
        __builtin_nontemporal_store(v, p);
    }
    __device__ static inline st step_out(T v, int, const StepArgs &, uint32_t) { return v; }
    __device__ static inline T weight(int s, const StepArgs &a) { return a.wf[s]; }
    __device__ static inline double absolute(T v, int s, const StepArgs &a) {
        return double(a.wf[s]) + double(v);
    }
    __device__ static inline T from_absolute(double f, int s, const StepArgs &a) {
        return T(f - double(a.wf[s]));
    }
};
// FP32 storage of d = f - w with FP64 arithmetic and the FP64 weights w
// (for D3Q27 their float roundings sum to 1 + 7.45e-9, a mass source): the
// loads widen, the stores round once.
template <> struct StorageT<ST_F32X> {
    using st = float;
    static constexpr int kCB = 13;  // 32 KiB chunks
    using T = double;
    static constexpr bool kShift = true;
    static constexpr bool kScaled = false;
    __device__ static inline double shift(int s, const StepArgs &a) { return a.w[s]; }
    __device__ static inline T in(st r, int, const StepArgs &) { return T(r); }
    __device__ static inline st out(T v, int, const StepArgs &) { return st(v); }
    __device__ static inline T load(const st *p, int, const StepArgs &) { return T(*p); }
    __device__ static inline void store(st *p, int, T v, const StepArgs &) {
        __builtin_nontemporal_store(st(v), p);
    }
    __device__ static inline st step_out(T v, int, const StepArgs &, uint32_t) { return st(v); }
    __device__ static inline T weight(int s, const StepArgs &a) { return a.w[s]; }
    __device__ static inline double absolute(T v, int s, const StepArgs &a) {
        return a.w[s] + double(v);
    }
    __device__ static inline T from_absolute(double f, int s, const StepArgs &a) {
        return T(st(f - a.w[s]));
    }
};
template <> struct StorageT<ST_F16> {
    using st = _Float16;
    static constexpr int kCB = 14;  // 32 KiB chunks
    using T = float;
    static constexpr bool kShift = true;
    // The time step computes in the stored scale itself, on h = d / w (see
    // node_prepare): the gathered value is widened as it is, and the step's
    // store rounds h without a rescale.
    static constexpr bool kScaled = true;
    __device__ static inline double shift(int s, const StepArgs &a) { return double(a.wf[s]); }
    __device__ static inline T in(st r, int, const StepArgs &) { return float(r); }
    __device__ static inline st out(T v, int s, const StepArgs &a) { return _Float16(v * a.h_scale[s]); }
    __device__ static inline T load(const st *p, int s, const StepArgs &a) {
        return float(*p) * a.h_unscale[s];
    }
    __device__ static inline void store(st *p, int s, T v, const StepArgs &a) {
        __builtin_nontemporal_store(_Float16(v * a.h_scale[s]), p);
    }
    // The time step's store rounds stochastically (v_cvt_sr_f16_f32, random
    // bits r): round to nearest repeats the same error every step wherever
    // the flow is steady, and that bias accumulates (D3Q19 cylinder array:
    // velocity L2 against FP64 1.0e-2 after 20000 steps, against 8.0e-4
    // rounded stochastically).
    __device__ static inline st step_out(T v, int, const StepArgs &, uint32_t r) {
        typedef __fp16 h2 __attribute__((ext_vector_type(2)));
        const h2 z = {0, 0};
        const h2 x = __builtin_amdgcn_cvt_sr_f16_f32(z, v, r, false);
        return __builtin_bit_cast(_Float16, x[0]);
    }
    __device__ static inline T weight(int s, const StepArgs &a) { return a.wf[s]; }
    __device__ static inline double absolute(T v, int s, const StepArgs &a) {
        return double(a.wf[s]) + double(v);
    }
    __device__ static inline T from_absolute(double f, int s, const StepArgs &a) {
        return T(f - double(a.wf[s]));
    }
};

// Random bits of node g's stores at step `step`: an integer hash of the
// pair (multiply-xorshift rounds), per direction s offset by s times the
// golden-ratio constant.  Reproducible: the same run rounds the same way.
__device__ inline uint32_t round_stream(uint32_t g, uint32_t step) {
    uint32_t h = g * 0x9E3779B1u + step * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h;
}
constexpr uint32_t kStreamStep = 0x9E3779B9u;

// Launch choices per storage and lattice, measured on MI355X: one node per
// lane everywhere.
// * FP64 D3Q19 (128x128x512): 1024-thread workgroups with non-temporal
//   loads, ahead of 512 threads and of a two-nodes-per-lane 16-byte
//   variant (20.9 against 19.8 GLUPS, and 13.5 against 12.0 for D3Q27).
// * FP64 D3Q27 (porous 384^3, 133 VGPRs, 3 waves per SIMD): 512-thread
//   workgroups with plain loads, 4.12 ms per step against 4.38 for 1024
//   threads with non-temporal loads (4.16 for 512 with them, 4.53 for 256).
//   Forced to 4 waves (128 VGPRs and 28 bytes of spills) it took 4.46 ms;
//   FP32 D3Q27 forced from 4 to 5 waves, 2.057 against 2.043 ms.
// * FP32: 256-thread workgroups (D3Q19: 203 against 205 us with 512; D3Q27
//   porous: 2.05 ms, the same as 512 and 1024).
// * FP16: 256-thread workgroups, since its node-type load went ahead of the
//   gathers (D3Q19 101.6 against 103.1 us with 512, D3Q27 145.0 against
//   147.0; D3Q15 82.3 against 82.2).
// FP32 and FP16 use plain loads: FP16 101.6 against 105.4 us per step with
// non-temporal loads (D3Q27 145.0 against 153.8); FP32 D3Q19 202.9 against 205.0, D3Q27 2.05 against
// 2.17 ms.  Every kernel stores
// non-temporally.
template <int ST, int Q> struct Tune;
template <int Q> struct Tune<ST_F64, Q> {
    static constexpr int kBS = Q == 27 ? 512 : 1024;
    static constexpr bool kNTLoad = Q != 27;
};
template <int Q> struct Tune<ST_F32, Q> {
    static constexpr int kBS = 256;
    static constexpr bool kNTLoad = false;
};
template <int Q> struct Tune<ST_F32X, Q> {
    static constexpr int kBS = 256;
    static constexpr bool kNTLoad = false;
};
template <int Q> struct Tune<ST_F16, Q> {
    static constexpr int kBS = 256;
    static constexpr bool kNTLoad = false;
};

// Workgroup -> node block.  The dispatcher deals workgroups round-robin to
// the 8 XCDs; this hands XCD k one contiguous range of node blocks so the
// cache lines two neighbouring blocks share stay in one L2.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t per = nb >> 3, rem = nb & 7u;
    const uint32_t k = b & 7u, slot = b >> 3;
    return k * per + (k < rem ? k : rem) + slot;
}

// Pull source offsets of node (x,y,z): index e+1 holds the coordinate of
// the node at -e (periodic), scaled to a global-index offset.
struct Gather {
    uint32_t xs[3], ys[3], zs[3];
    __device__ inline void init(uint32_t x, uint32_t y, uint32_t z, const StepArgs &a) {
        xs[0] = (x + 1 == a.nx) ? 0 : x + 1;  // e_x = -1 comes from x+1
        xs[1] = x;
        xs[2] = (x == 0) ? a.nx - 1 : x - 1;  // e_x = +1 comes from x-1
        ys[0] = ((y + 1 == a.ny) ? 0 : y + 1) * a.nx;
        ys[1] = y * a.nx;
        ys[2] = ((y == 0) ? a.ny - 1 : y - 1) * a.nx;
        zs[0] = ((z + 1 == a.nz) ? 0 : z + 1) * a.nxny;
        zs[1] = z * a.nxny;
        zs[2] = ((z == 0) ? a.nz - 1 : z - 1) * a.nxny;
    }
    __device__ inline uint32_t src(int ex, int ey, int ez) const {
        return xs[ex + 1] + ys[ey + 1] + zs[ez + 1];
    }
};

// Buffer-resource access (raw buffer loads and stores: a scalar 128-bit
// descriptor, a scalar offset per direction and one 32-bit lane offset).
// kRsrcWord3 is the gfx9-family descriptor word (raw, untyped access);
// kAuxNT the non-temporal cache bit of gfx942/gfx950 buffer instructions.
constexpr int kRsrcWord3 = 0x00020000;
constexpr int kAuxNT = 2;

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(uintptr_t base, uint64_t bytes) {
    const int num = bytes >= 0xFFFFFFFFull ? -1 : int(uint32_t(bytes));
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(base), 0, num, kRsrcWord3);
}

template <class stt> struct RawBuf;
template <> struct RawBuf<float> {
    template <int AUX> __device__ static float load(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(v), int(s), AUX));
    }
    template <int AUX> __device__ static void store(float x, __amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), r, int(v), int(s), AUX);
    }
};
template <> struct RawBuf<_Float16> {
    template <int AUX> __device__ static _Float16 load(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, int(v), int(s), AUX));
    }
    template <int AUX> __device__ static void store(_Float16 x, __amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, x), r, int(v), int(s), AUX);
    }
};
template <> struct RawBuf<double> {
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    template <int AUX> __device__ static double load(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, int(v), int(s), AUX));
    }
    template <int AUX> __device__ static void store(double x, __amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x), r, int(v), int(s), AUX);
    }
};

// Pull addressing of a workgroup whose BS nodes (first node g0, all in one
// chunk c0, BS dividing CH) lie in planes 1 <= z <= nz - 2, so that no
// source wraps in z.  A source node g' lies less than 2 nx ny nodes before
// its node (the periodic y wrap of the plane behind), so with
// kb = ceil(2 nx ny / CH) + 1 it lies after chunk c0 - kb, and its position relative to
// chunk c0 - kb, r = g' - (c0 - kb) CH >= 0, gives the population's byte
// offset from that chunk's start,
//   (r + (r >> CB) (Q - 1) CH) * sizeof(st) + s CH sizeof(st),
// a 32-bit lane offset (one add, one shift and one 24-bit multiply-add per
// population) plus a scalar offset per direction on a buffer descriptor
// based at chunk c0 - kb, where the general chunked index costs 64-bit
// arithmetic per population.
template <class L, int ST>
struct PullRel {
    using st = typename StorageT<ST>::st;
    static constexpr uint32_t kSize = sizeof(st);
    static constexpr int kLog = kSize == 8 ? 3 : (kSize == 4 ? 2 : 1);
    static constexpr int CB = StorageT<ST>::kCB;
    static constexpr uint32_t CH = 1u << CB;
    static_assert((uint64_t(L::Q - 1) * CH * kSize) < (1u << 24), "24-bit multiplier");
    __amdgpu_buffer_rsrc_t rsrc;    // based at chunk c0 - kb
    uint32_t xs[3], ys[3], zs[3];   // byte offsets, summing to r * sizeof(st)
    __device__ inline void init(const st *fin, uint32_t g0, uint32_t x, uint32_t y,
                                uint32_t zplane, const StepArgs &a) {
        const uint32_t c0 = g0 >> CB;
        const intptr_t back = (intptr_t(c0) - intptr_t(a.kb)) * intptr_t(L::Q * CH * kSize);
        rsrc = make_rsrc(uintptr_t(fin) + back, uint64_t(intptr_t(a.fbytes) - back));
        xs[0] = ((x + 1 == a.nx) ? 0 : x + 1) << kLog;
        xs[1] = x << kLog;
        xs[2] = ((x == 0) ? a.nx - 1 : x - 1) << kLog;
        ys[0] = (((y + 1 == a.ny) ? 0 : y + 1) * a.nx) << kLog;
        ys[1] = (y * a.nx) << kLog;
        ys[2] = (((y == 0) ? a.ny - 1 : y - 1) * a.nx) << kLog;
        const uint32_t zr = zplane - (c0 - a.kb) * CH;  // modulo 2^32; the sums are exact
        zs[0] = (zr + a.nxny) << kLog;
        zs[1] = zr << kLog;
        zs[2] = (zr - a.nxny) << kLog;
    }
    template <bool NT>
    __device__ inline st load(int s, int ex, int ey, int ez) const {
        const uint32_t rb = xs[ex + 1] + ys[ey + 1] + zs[ez + 1];
        const uint32_t off = rb + __umul24(rb >> (CB + kLog), (L::Q - 1) * CH * kSize);
        return RawBuf<st>::template load<NT ? kAuxNT : 0>(rsrc, off, uint32_t(s) * CH * kSize);
    }
};

// PullRel on a lattice whose rows never straddle a chunk (CH a multiple of
// nx): the chunk of a source node is the chunk of its row's first node, so
// the chunk-crossing term of the offset, (r >> CB) (Q - 1) CH, is formed
// once per source row -- the 9 combinations of e_y and e_z -- and each
// population adds only its source's x offset to its row's offset: one
// VALU add per population instead of an add, a shift and a multiply-add.
template <class L, int ST>
struct PullRow {
    using st = typename StorageT<ST>::st;
    static constexpr uint32_t kSize = sizeof(st);
    static constexpr int kLog = kSize == 8 ? 3 : (kSize == 4 ? 2 : 1);
    static constexpr int CB = StorageT<ST>::kCB;
    static constexpr uint32_t CH = 1u << CB;
    static_assert((uint64_t(L::Q - 1) * CH * kSize) < (1u << 24), "24-bit multiplier");
    __amdgpu_buffer_rsrc_t rsrc;  // based at chunk c0 - kb, as PullRel's
    uint32_t xs[3];               // byte offset of the source's x (e_x = -1, 0, +1)
    uint32_t row[3][3];           // [e_y + 1][e_z + 1]: byte offset of the source row
    __device__ inline void init(const st *fin, uint32_t g0, uint32_t x, uint32_t y,
                                uint32_t zplane, const StepArgs &a) {
        const uint32_t c0 = g0 >> CB;
        const intptr_t back = (intptr_t(c0) - intptr_t(a.kb)) * intptr_t(L::Q * CH * kSize);
        rsrc = make_rsrc(uintptr_t(fin) + back, uint64_t(intptr_t(a.fbytes) - back));
        xs[0] = ((x + 1 == a.nx) ? 0 : x + 1) << kLog;
        xs[1] = x << kLog;
        xs[2] = ((x == 0) ? a.nx - 1 : x - 1) << kLog;
        const uint32_t ys[3] = {(((y + 1 == a.ny) ? 0 : y + 1) * a.nx) << kLog, (y * a.nx) << kLog,
                                (((y == 0) ? a.ny - 1 : y - 1) * a.nx) << kLog};
        const uint32_t zr = (zplane - (c0 - a.kb) * CH) << kLog;  // modulo 2^32; the sums are exact
        const uint32_t zs[3] = {zr + (a.nxny << kLog), zr, zr - (a.nxny << kLog)};
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const uint32_t rb = ys[j] + zs[k];
                row[j][k] = rb + __umul24(rb >> (CB + kLog), (L::Q - 1) * CH * kSize);
            }
    }
    template <bool NT>
    __device__ inline st load(int s, int ex, int ey, int ez) const {
        return RawBuf<st>::template load<NT ? kAuxNT : 0>(rsrc, row[ey + 1][ez + 1] + xs[ex + 1],
                                                          uint32_t(s) * CH * kSize);
    }
};

// Whether the workgroup of nodes [g0, g0 + BS) may use PullRel.
template <int BS>
__device__ inline bool pull_rel_ok(uint32_t g0, const StepArgs &a) {
    return a.rel && g0 >= a.nxny && g0 + BS + a.nxny <= a.N;
}

// Aligned output of a workgroup inside chunk c0: a descriptor at the
// chunk's start, a scalar offset per direction, one lane offset.
template <class L, int ST>
struct PushOut {
    using st = typename StorageT<ST>::st;
    static constexpr int CB = StorageT<ST>::kCB;
    static constexpr uint32_t CH = 1u << CB;
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t off;
    __device__ inline PushOut(st *fout, uint32_t c0, uint32_t g) {
        rsrc = make_rsrc(uintptr_t(fout) + uintptr_t(c0) * L::Q * CH * sizeof(st),
                         uint64_t(L::Q) * CH * sizeof(st));
        off = (g & (CH - 1u)) * uint32_t(sizeof(st));
    }
    template <bool NT>
    __device__ inline void store(int s, st v) const {
        RawBuf<st>::template store<NT ? kAuxNT : 0>(v, rsrc, off, uint32_t(s) * CH * uint32_t(sizeof(st)));
    }
};

// a * b + c rounded once.
__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// sum_k e_k u_k over the nonzero components, in x, y, z order.
template <class T>
__device__ inline T edot(int ex, int ey, int ez, T ux, T uy, T uz) {
    T d = T(-0.0);
    if (ex > 0) d += ux; else if (ex < 0) d -= ux;
    if (ey > 0) d += uy; else if (ey < 0) d -= uy;
    if (ez > 0) d += uz; else if (ez < 0) d -= uz;
    return d;
}

// The reference's node update on registers, in phases.  f[] holds the
// gathered pre-collision populations on entry (deviations d = f - w when
// SHIFT).  node_prepare computes the moments and applies the node type's
// rule; rho/ux/uy/uz end as the values write_data() would dump for this
// step.  With COLLIDE it also applies the type-5 rule and the inlet/outlet
// reconstruction, fills feq and returns true for nodes that relax; a solid
// node is bounced back in place (feq = f) and returns false.
// SC (scaled, with SHIFT): f[] and feq[] hold h = d / w instead of d.  Every
// rule is the same rewritten in that scale: sums of populations weigh each
// h by its w (one FMA per term), the equilibrium loses its factor w, the
// type-5 increment its w, and the bounce-back and inlet/outlet
// reconstructions keep their form because a direction and its partner have
// the same weight.
template <class L, class T, bool SHIFT, bool COLLIDE, bool SC = false>
__device__ inline bool node_prepare(T (&f)[L::Q], T (&feq)[L::Q], int t, T &rho, T &ux, T &uy,
                                    T &uz, const StepArgs &a, const T *__restrict__ uy_head,
                                    const T (&w)[L::Q], uint32_t g) {
    constexpr int Q = L::Q;
    static_assert(!SC || SHIFT, "the scaled form is a deviation form");
    T dlt = T(0);  // rho - 1 (SHIFT only)
    // compute_macroscopic_data, LatticeStructure.hpp:69-90
    {
        T r = T(-0.0), mx = T(-0.0), my = T(-0.0), mz = T(-0.0);
#pragma unroll
        for (int s = 0; s < Q; ++s) {
            if constexpr (SC) {
                r = fma_t(w[s], f[s], r);
                if (L::ex[s] > 0) mx = fma_t(w[s], f[s], mx); else if (L::ex[s] < 0) mx = fma_t(-w[s], f[s], mx);
                if (L::ey[s] > 0) my = fma_t(w[s], f[s], my); else if (L::ey[s] < 0) my = fma_t(-w[s], f[s], my);
                if (L::ez[s] > 0) mz = fma_t(w[s], f[s], mz); else if (L::ez[s] < 0) mz = fma_t(-w[s], f[s], mz);
            } else {
                r += f[s];
                if (L::ex[s] > 0) mx += f[s]; else if (L::ex[s] < 0) mx -= f[s];
                if (L::ey[s] > 0) my += f[s]; else if (L::ey[s] < 0) my -= f[s];
                if (L::ez[s] > 0) mz += f[s]; else if (L::ez[s] < 0) mz -= f[s];
            }
        }
        if constexpr (SHIFT) {
            dlt = r + (sizeof(T) == 8 ? T(a.dw_d) : T(a.dw));
            rho = T(1) + dlt;
            // a correctly rounded 1/rho: a 1-ulp reciprocal biases u and
            // doubles the FP32 velocity error against FP64 over 1000 steps
            const T inv = T(1) / rho;
            ux = mx * inv;
            uy = my * inv;
            uz = mz * inv;
        } else {
            rho = r;
            ux = mx / r;
            uy = my / r;
            uz = mz / r;
        }
    }
    if (t == 1) {
        // solid node: u = 0, whole-vector bounce-back, no relaxation
        // (TLBM_Partition.cpp:750-756, D3Q19LatticeStructure.hpp:65-76)
        ux = T(0); uy = T(0); uz = T(0);
        if (COLLIDE) {
            T g[Q];
#pragma unroll
            for (int s = 0; s < Q; ++s) g[s] = f[L::bb[s]];
#pragma unroll
            for (int s = 0; s < Q; ++s) {
                f[s] = g[s];
                feq[s] = g[s];
            }
        }
        return false;
    }
    if (t == 2) {
        // set_inletW_bc_macro (D3Q19LatticeStructure.hpp:108-132)
        T s_in = T(-0.0), sz = T(-0.0);
#pragma unroll
        for (int i = 0; i < L::n_unk; ++i)
            s_in = SC ? fma_t(w[L::in_list[i]], f[L::in_list[i]], s_in) : s_in + f[L::in_list[i]];
#pragma unroll
        for (int i = 0; i < L::n_zero; ++i)
            sz = SC ? fma_t(w[L::zero_list[i]], f[L::zero_list[i]], sz) : sz + f[L::zero_list[i]];
        ux = T(0); uy = T(0);
        const T c_in0 = sizeof(T) == 8 ? T(a.c_in_d) : T(a.c_in);
        T inv_1mu = T(a.inv_1mu), c_in = c_in0;
        if (a.inlet_uz) {
            // a prescribed inlet profile: 1/(1-u) and the deviation constant per node
            const T u = T(a.inlet_uz[g % a.nxny]);
            uz = u;
            inv_1mu = T(1) / (T(1) - u);
            c_in = c_in0 + (u - T(a.u_bc));
        } else {
            uz = T(a.u_bc);
        }
        if constexpr (SHIFT) {
            dlt = inv_1mu * (c_in + (T(2) * s_in + sz));
            rho = T(1) + dlt;
        } else {
            rho = inv_1mu * (T(2) * s_in + sz);
        }
    } else if (t == 3) {
        // set_outletE_bc_macro (D3Q19LatticeStructure.hpp:135-156); ux, uy kept
        T s_out = T(-0.0), sz = T(-0.0);
#pragma unroll
        for (int i = 0; i < L::n_unk; ++i)
            s_out = SC ? fma_t(w[L::out_list[i]], f[L::out_list[i]], s_out) : s_out + f[L::out_list[i]];
#pragma unroll
        for (int i = 0; i < L::n_zero; ++i)
            sz = SC ? fma_t(w[L::zero_list[i]], f[L::zero_list[i]], sz) : sz + f[L::zero_list[i]];
        rho = T(a.rho_bc);
        if constexpr (SHIFT) {
            dlt = sizeof(T) == 8 ? T(a.rho_bc - 1.0) : T(a.drho_bc);
            uz = T(a.inv_rho_bc) * ((sizeof(T) == 8 ? T(a.c_out_d) : T(a.c_out)) + (T(2) * s_out + sz));
        } else {
            uz = T(-1) + T(a.inv_rho_bc) * (T(2) * s_out + sz);
        }
    }
    if (!COLLIDE) return true;
    if (t == 5) {
        // set_uz_bc (LatticeStructure.hpp:166-177); the y term reads uy of
        // node s of this step, as the reference's uy[spd] does
        const T ubc = a.inlet_uz ? T(a.inlet_uz[g % a.nxny]) : T(a.u_bc);
#pragma unroll
        for (int s = 1; s < Q; ++s) {
            T br = T(-0.0);
            if (L::ez[s] > 0) br += (ubc - uz); else if (L::ez[s] < 0) br -= (ubc - uz);
            if (L::ex[s] > 0) br -= ux; else if (L::ex[s] < 0) br += ux;
            if (L::ey[s] > 0) br -= uy_head[s]; else if (L::ey[s] < 0) br += uy_head[s];
            f[s] = SC ? f[s] + T(3) * rho * br : f[s] + T(3) * rho * w[s] * br;
        }
    }
    // compute_equilibrium, LatticeStructure.hpp:92-103
    {
        const T usq = ux * ux + uy * uy + uz * uz;
        if constexpr (SHIFT) {
            // f_eq - w = w (delta + rho (cu + cu^2/2 - 3u^2/2)), cu = 3 e.u,
            // rewritten with t = e.(3 rho u) = rho cu:  w (A + t + t^2 / (2 rho)),
            // A = delta - 3/2 rho u^2 (the same polynomial, fewer operations)
            const T A = dlt - T(1.5) * rho * usq;
            T h;
            if constexpr (sizeof(T) == 8)
                h = T(0.5) / rho;  // FP32X: every operation in FP64
            else
                h = T(0.5) * __builtin_amdgcn_rcpf(rho);
            // FP16 (SC): U = (3 + dt3) rho u, as fma(3 rho, u, dt3 rho u),
            // which rounds the exact product plus the correction once, so
            // the correction survives rounding on average (3 + dt3 itself
            // rounds to 3).  The other storages have dt3 = 0 and form
            // 3 rho u as a plain product, which leaves the compiler free to
            // contract it into the sums of t below.
            T Ux, Uy, Uz;
            if constexpr (SC) {
                const T r3 = T(3) * rho, rd = T(a.dt3) * rho;
                Ux = fma_t(r3, ux, rd * ux);
                Uy = fma_t(r3, uy, rd * uy);
                Uz = fma_t(r3, uz, rd * uz);
            } else {
                const T t3 = T(3);
                Ux = t3 * rho * ux;
                Uy = t3 * rho * uy;
                Uz = t3 * rho * uz;
            }
#pragma unroll
            for (int s = 0; s < Q; ++s) {
                const T t = (s == 0) ? T(0) : edot<T>(L::ex[s], L::ey[s], L::ez[s], Ux, Uy, Uz);
                feq[s] = SC ? A + t * (T(1) + t * h) : w[s] * (A + t * (T(1) + t * h));
            }
            const T corr = rho * (T(1.5) * T(a.dk1) * usq - T(a.dk0));
            feq[0] += SC ? corr * T(a.h_scale[0]) : corr;
        } else {
#pragma unroll
            for (int s = 0; s < Q; ++s) {
                const T cu = (s == 0) ? T(0) : T(3) * edot<T>(L::ex[s], L::ey[s], L::ez[s], ux, uy, uz);
                feq[s] = w[s] * rho * (T(1) + cu + T(0.5) * (cu * cu) - T(1.5) * usq);
            }
        }
    }
    if (t == 2) {
#pragma unroll
        for (int i = 0; i < L::n_unk; ++i)
            f[L::in_unk[i]] = feq[L::in_unk[i]] + f[L::in_par[i]] - feq[L::in_par[i]];
    } else if (t == 3) {
#pragma unroll
        for (int i = 0; i < L::n_unk; ++i)
            f[L::out_unk[i]] = feq[L::out_unk[i]] + f[L::out_par[i]] - feq[L::out_par[i]];
    }
    return true;
}

// Non-equilibrium stress Pi = sum_s e e (f - feq) (compute_strain_tensor,
// LatticeStructure.hpp:105-122): xx, yy, zz, xy, xz, yz.  SC: f and feq
// hold h = d / w, and each difference is weighed by its w; a direction and
// its partner (bounce-back table) have the same e e and the same w, so
// their differences are added first and each pair enters a sum once (about
// a third fewer operations than one term per direction).
template <class L, class T, bool SC = false>
__device__ inline void node_stress(const T (&f)[L::Q], const T (&feq)[L::Q], T (&pi)[6],
                                   const T (&w)[L::Q]) {
    T sxx = T(-0.0), syy = T(-0.0), szz = T(-0.0), sxy = T(-0.0), sxz = T(-0.0), syz = T(-0.0);
#pragma unroll
    for (int s = 0; s < L::Q; ++s) {
        // the rest direction adds to no sum; the others come with their partner
        if (SC && (s == 0 || L::bb[s] < s)) continue;
        const T fn = SC ? ((f[s] - feq[s]) + (f[L::bb[s]] - feq[L::bb[s]])) * w[s] : f[s] - feq[s];
        if (L::ex[s] != 0) sxx += fn;
        if (L::ey[s] != 0) syy += fn;
        if (L::ez[s] != 0) szz += fn;
        if (L::ex[s] * L::ey[s] > 0) sxy += fn; else if (L::ex[s] * L::ey[s] < 0) sxy -= fn;
        if (L::ex[s] * L::ez[s] > 0) sxz += fn; else if (L::ex[s] * L::ez[s] < 0) sxz -= fn;
        if (L::ey[s] * L::ez[s] > 0) syz += fn; else if (L::ey[s] * L::ez[s] < 0) syz -= fn;
    }
    pi[0] = sxx; pi[1] = syy; pi[2] = szz; pi[3] = sxy; pi[4] = sxz; pi[5] = syz;
}

// |Pi| as apply_turbulence_model forms it (LatticeStructure.hpp:147).
template <class T>
__device__ inline T stress_norm(const T (&pi)[6]) {
    return sqrt(pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                T(2) * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5]));
}

// apply_turbulence_model, LatticeStructure.hpp:140-153.  FAST (FP16
// storage, whose own rounding is ~1e-3 of a population): the hardware square
// roots and reciprocal, 1 ulp each, in place of the correctly rounded ones.
template <class T, bool FAST = false>
__device__ inline T smagorinsky_omega(const T (&pi)[6], const StepArgs &a) {
    if constexpr (FAST) {
        static_assert(sizeof(T) == 4, "FP32 arithmetic");
        const float n2 = pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                         2.f * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5]);
        const float P = __builtin_amdgcn_sqrtf(__builtin_amdgcn_sqrtf(n2) * float(a.cs) + float(a.nu2)) -
                        float(a.nu);
        return __builtin_amdgcn_rcpf(3.f * (float(a.nu) + P * (1.f / 6.f)) + 0.5f);
    }
    T P = stress_norm<T>(pi);
    P *= T(a.cs);
    P = sqrt(P + T(a.nu2)) - T(a.nu);
    const T nu_e = P / T(6);
    return T(1) / (T(3) * (T(a.nu) + nu_e) + T(0.5));
}

// relax, LatticeStructure.hpp:155-163.
template <class L, class T>
__device__ inline void node_relax(T (&f)[L::Q], const T (&feq)[L::Q], T om) {
#pragma unroll
    for (int s = 0; s < L::Q; ++s) f[s] = f[s] - om * (f[s] - feq[s]);
}

// relax of a fluid node in FP32 storage, conserving to the rounding of
// f - feq.  In exact arithmetic the mass and momentum of f - feq are small
// values set by the weights (the reference's D3Q15/D3Q19 float weights do
// not sum to 1); rounded FP32 equilibria miss them by a few ulp of feq, the
// same amount at every node of a steady flow, and that residue accumulates
// step after step (D3Q19 cylinder array: velocity L2 against FP64 1.6e-5
// after 20000 steps).  The values are those of the weights w, not of their
// float roundings: the D3Q27 weights sum to 1 in double but to 1 + 7.5e-9
// in float, a mass source FP64 does not have.  The residue of mass m0 and
// momentum m is taken out of f - feq as w_s (m0 + 3 e_s.m), which carries
// that mass and momentum.
template <class L>
__device__ inline void node_relax_conserving(float (&f)[L::Q], const float (&feq)[L::Q],
                                             const float (&w)[L::Q], float om, float rho,
                                             float ux, float uy, float uz, const StepArgs &a) {
    constexpr int Q = L::Q;
    float nq[Q], m0 = -0.0f, mx = -0.0f, my = -0.0f, mz = -0.0f;
#pragma unroll
    for (int s = 0; s < Q; ++s) {
        nq[s] = f[s] - feq[s];
        m0 += nq[s];
        if (L::ex[s] > 0) mx += nq[s]; else if (L::ex[s] < 0) mx -= nq[s];
        if (L::ey[s] > 0) my += nq[s]; else if (L::ey[s] < 0) my -= nq[s];
        if (L::ez[s] > 0) mz += nq[s]; else if (L::ez[s] < 0) mz -= nq[s];
    }
    m0 -= rho * (1.5f * a.k1 * (ux * ux + uy * uy + uz * uz) - a.k0);
    const float rk = rho * a.k2;
    mx = 3.0f * (mx - rk * ux);
    my = 3.0f * (my - rk * uy);
    mz = 3.0f * (mz - rk * uz);
#pragma unroll
    for (int s = 0; s < Q; ++s) {
        const float c = (s == 0) ? m0 : m0 + edot<float>(L::ex[s], L::ey[s], L::ez[s], mx, my, mz);
        f[s] = f[s] - om * (nq[s] - w[s] * c);
    }
}

// Whether a step kernel relaxes fluid nodes with node_relax_conserving:
// FP32 storage.  FP16's own rounding is stochastic, no bias; its float
// weights get the mass and momentum balance of w from the equilibrium
// corrections dt3, dk0, dk1 instead, at a fraction of the cost (the
// conserving relax took 13 % of the FP16 D3Q27 step).
template <int ST, int Q>
constexpr bool kConserve = ST == ST_F32;

// The whole reference update of one node (TLBM_Partition.cpp:742-807).
// CONSERVE: fluid nodes relax with node_relax_conserving.  SC: the scaled
// form of node_prepare (the relaxation is the same in either scale).
template <class L, class T, bool SHIFT, bool COLLIDE, bool CONSERVE = false, bool SC = false>
__device__ inline void node_update(T (&f)[L::Q], int t, T &rho, T &ux, T &uy, T &uz,
                                   const StepArgs &a, const T *__restrict__ uy_head,
                                   const T (&w)[L::Q], uint32_t g) {
    static_assert(!(SC && CONSERVE), "the conserving relax works on d");
    T feq[L::Q];
    if (!node_prepare<L, T, SHIFT, COLLIDE, SC>(f, feq, t, rho, ux, uy, uz, a, uy_head, w, g)) return;
    if (!COLLIDE) return;
    T om = T(a.omega);
    if (a.smag) {
        T pi[6];
        node_stress<L, T, SC>(f, feq, pi, w);
        om = smagorinsky_omega<T, SC>(pi, a);
    }
    if (a.dynamics == 1) {
        if constexpr (CONSERVE) {
            if (t == 0) {
                node_relax_conserving<L>(f, feq, w, om, rho, ux, uy, uz, a);
                return;
            }
        }
        node_relax<L, T>(f, feq, om);
    }
}

// Has the step kernels' prologue load, in one batch of scalar loads, the
// arguments that the node's coordinates and the gather addresses depend on.
// Left to itself the compiler loads them in five dependent rounds (each a
// scalar-cache round trip) split by the early-exit and division branches,
// before the first gather can go out.  The empty asm only consumes the values.
// Same box, µs per step (D3Q19 128x128x512): FP32 203.0-203.1 -> 202.7-202.8,
// FP16 110.0 -> 109.4-109.6; closure FP16 D3Q27 192.6 -> 190.3-190.9, FP32
// D3Q19 209.1 -> 207.3-207.8.
__device__ inline void load_args_early(const StepArgs &a, const void *fin) {
    __asm__ volatile("" ::"s"(a.nx), "s"(a.ny), "s"(a.nxny), "s"(a.N), "s"(a.cb), "s"(a.div_nx.m),
                     "s"(a.div_nx.s), "s"(a.div_nxny.m), "s"(a.div_nxny.s), "s"(a.kb), "s"(a.rel),
                     "s"(a.row_chunks), "s"(a.fbytes), "s"(fin));
}

// One lattice update per thread.  MODE_STEP streams and collides
// fin -> fout; MODE_MACROS recomputes the moments of the step that read fin
// (ux, uy, uz, rho planes of N compute-type values in `out`); MODE_AVG adds
// those moments to four planes of N doubles in `out` (time average);
// MODE_HEAD writes uy of nodes 0..Q-1 (the values the type-5 rule reads) to
// `out`; MODE_STRESS writes ux, uy, uz, rho and the viscous stress
// sigma = -(1 - omega/2) Pi (xx, yy, zz, xy, xz, yz; omega of the BGK or
// Smagorinsky relaxation, zero where the node does not relax) as 10 planes
// of N compute-type values.
// chunk_type (MODE_STEP, may be null): per chunk, the node type shared by
// all its nodes, or 0xFF when they differ; a workgroup inside a uniform
// chunk skips the per-node type load.
template <class L, int ST, int MODE, bool REMAP = true, bool NTLOAD = false, bool NTSTORE = true,
          int BS = kBlock>
__global__ void __launch_bounds__(BS)
k_lbm(const typename StorageT<ST>::st *__restrict__ fin, typename StorageT<ST>::st *__restrict__ fout,
      const uint8_t *__restrict__ nd_type, const uint8_t *__restrict__ chunk_type,
      void *__restrict__ out, const typename StorageT<ST>::T *__restrict__ uy_head, StepArgs a) {
    using S = StorageT<ST>;
    using T = typename S::T;
    using st = typename S::st;
    constexpr int Q = L::Q;
    if (MODE == MODE_STEP) load_args_early(a, fin);
    if constexpr (MODE == MODE_STEP && kGatherPrio<ST, Q>) __builtin_amdgcn_s_setprio(3);
    uint32_t g, blk = 0;
    if (MODE == MODE_HEAD) {
        g = threadIdx.x;
        if (g >= (uint32_t)Q || g >= a.N) return;
    } else {
        blk = REMAP ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
        g = blk * BS + threadIdx.x;
        if (g >= a.N) return;
    }
    const uint32_t z = udiv(g, a.div_nxny);
    const uint32_t r = g - z * a.nxny;
    const uint32_t y = udiv(r, a.div_nx);
    const uint32_t x = r - y * a.nx;
    const uint32_t chunk = (blk * BS) >> a.cb;  // BS divides the chunk
    uint32_t t0 = 0;
    if constexpr (kTypeAhead<ST>) t0 = type_first<MODE>(nd_type, chunk_type, chunk, g);
    T f[Q];
    if (MODE == MODE_STEP && a.row_chunks && pull_rel_ok<BS>(blk * BS, a)) {
        PullRow<L, ST> pr;
        pr.init(fin, blk * BS, x, y, g - r, a);
#pragma unroll
        for (int s = 0; s < Q; ++s) f[s] = S::in(pr.template load<NTLOAD>(s, L::ex[s], L::ey[s], L::ez[s]), s, a);
    } else if (MODE == MODE_STEP && pull_rel_ok<BS>(blk * BS, a)) {
        PullRel<L, ST> pr;
        pr.init(fin, blk * BS, x, y, g - r, a);
#pragma unroll
        for (int s = 0; s < Q; ++s) f[s] = S::in(pr.template load<NTLOAD>(s, L::ex[s], L::ey[s], L::ez[s]), s, a);
    } else {
        Gather gat;
        gat.init(x, y, z, a);
#pragma unroll
        for (int s = 0; s < Q; ++s) {
            const st *p = fin + fidx<Q>(s, gat.src(L::ex[s], L::ey[s], L::ez[s]), a);
            f[s] = S::in(NTLOAD ? __builtin_nontemporal_load(p) : *p, s, a);
        }
    }
    if constexpr (MODE == MODE_STEP && kGatherPrio<ST, Q>) __builtin_amdgcn_s_setprio(0);
    T w[Q];
#pragma unroll
    for (int s = 0; s < Q; ++s) w[s] = S::weight(s, a);
    if constexpr (!kTypeAhead<ST>) t0 = type_first<MODE>(nd_type, chunk_type, chunk, g);
    const int t = type_final<MODE>(t0, nd_type, chunk_type, g);
    T rho, ux, uy, uz;
    if (MODE == MODE_STEP) {
        node_update<L, T, S::kShift, true, kConserve<ST, Q>, S::kScaled>(f, t, rho, ux, uy, uz, a,
                                                                         uy_head, w, g);
        const PushOut<L, ST> po(fout, (blk * BS) >> StorageT<ST>::kCB, g);
        const uint32_t rs = ST == ST_F16 ? round_stream(g, a.step) : 0u;
#pragma unroll
        for (int s = 0; s < Q; ++s)
            po.template store<NTSTORE>(s, S::step_out(f[s], s, a, rs + uint32_t(s) * kStreamStep));
    } else if (MODE == MODE_STRESS) {
        T feq[Q], pi[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        T c = T(0);
        if (node_prepare<L, T, S::kShift, true, S::kScaled>(f, feq, t, rho, ux, uy, uz, a, uy_head,
                                                            w, g) &&
            a.dynamics == 1) {
            node_stress<L, T, S::kScaled>(f, feq, pi, w);
            const T om = a.smag ? smagorinsky_omega<T>(pi, a) : T(a.omega);
            c = T(0.5) * om - T(1);
        }
        T *macro = static_cast<T *>(out);
        macro[g] = ux;
        macro[size_t(a.N) + g] = uy;
        macro[2 * size_t(a.N) + g] = uz;
        macro[3 * size_t(a.N) + g] = rho;
#pragma unroll
        for (int k = 0; k < 6; ++k) macro[size_t(4 + k) * a.N + g] = c * pi[k];
    } else {
        node_update<L, T, S::kShift, false, false, S::kScaled>(f, t, rho, ux, uy, uz, a, uy_head, w, g);
        if (MODE == MODE_MACROS) {
            T *macro = static_cast<T *>(out);
            macro[g] = ux;
            macro[size_t(a.N) + g] = uy;
            macro[2 * size_t(a.N) + g] = uz;
            macro[3 * size_t(a.N) + g] = rho;
        } else if (MODE == MODE_AVG) {
            double *sum = static_cast<double *>(out);
            sum[g] += double(ux);
            sum[size_t(a.N) + g] += double(uy);
            sum[2 * size_t(a.N) + g] += double(uz);
            sum[3 * size_t(a.N) + g] += double(rho);
        } else {
            static_cast<T *>(out)[g] = uy;
        }
    }
}

// Step with the learned closure in place of Smagorinsky (FP32/FP16 storage).
// One node per lane as k_lbm; every lane of a wavefront runs the network
// together (lanes past the last node work on the last node and store
// nothing).  A solid node's stress is zero and its result unused.
template <class L, int ST, int BS, bool NTLOAD>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(kClosureWaves<ST, L::Q>)))
k_lbm_closure(const typename StorageT<ST>::st *__restrict__ fin, typename StorageT<ST>::st *__restrict__ fout,
              const uint8_t *__restrict__ nd_type, const uint8_t *__restrict__ chunk_type,
              const typename StorageT<ST>::T *__restrict__ uy_head, StepArgs a) {
    using S = StorageT<ST>;
    using T = typename S::T;
    using st = typename S::st;
    constexpr int Q = L::Q;
    // The network's packed weights, staged once per workgroup by direct
    // global-to-LDS loads (global_load_lds_dwordx4: lane i of a wavefront
    // writes 16 bytes at the wavefront's LDS base + 16 i).  They go out ahead
    // of the gathers and hold no registers; a load into registers followed by
    // an LDS write made every wavefront wait one L2 round trip before issuing
    // its gathers (FP16 D3Q19 128x128x512: 147 -> 143 us per step; FP32
    // 217 -> 207).  A copy per wavefront, which needs no workgroup barrier,
    // was slower (148 us).
    load_args_early(a, fin);
    __asm__ volatile("" ::"s"(a.mlp));  // in the same batch
    __shared__ MlpPacked smlp;
    static_assert(sizeof(MlpPacked) % 16 == 0 && sizeof(MlpPacked) <= 16u * BS, "one 16-byte load per lane");
    if (threadIdx.x < sizeof(MlpPacked) / 16) {
        using gptr = const __attribute__((address_space(1))) void *;
        using lptr = __attribute__((address_space(3))) void *;
        const uint32_t wave_base = (threadIdx.x & ~63u) * 16;  // uniform per wavefront
        __builtin_amdgcn_global_load_lds((gptr)(reinterpret_cast<const char *>(a.mlp) + threadIdx.x * 16),
                                         (lptr)(reinterpret_cast<char *>(&smlp) + wave_base), 16, 0, 0);
    }
    const uint32_t g0 = blockIdx.x * BS + threadIdx.x;
    const bool live = g0 < a.N;
    const uint32_t g = live ? g0 : a.N - 1;
    const uint32_t z = udiv(g, a.div_nxny);
    const uint32_t r = g - z * a.nxny;
    const uint32_t y = udiv(r, a.div_nx);
    const uint32_t x = r - y * a.nx;
    // the node type is loaded ahead of the gathers, in every storage
    const uint32_t t0 = type_first<MODE_STEP>(nd_type, chunk_type, (blockIdx.x * BS) >> a.cb, g);
    T f[Q];
    if (a.row_chunks && pull_rel_ok<BS>(blockIdx.x * BS, a)) {
        PullRow<L, ST> pr;
        pr.init(fin, blockIdx.x * BS, x, y, g - r, a);
#pragma unroll
        for (int s = 0; s < Q; ++s) f[s] = S::in(pr.template load<NTLOAD>(s, L::ex[s], L::ey[s], L::ez[s]), s, a);
    } else if (pull_rel_ok<BS>(blockIdx.x * BS, a)) {
        PullRel<L, ST> pr;
        pr.init(fin, blockIdx.x * BS, x, y, g - r, a);
#pragma unroll
        for (int s = 0; s < Q; ++s) f[s] = S::in(pr.template load<NTLOAD>(s, L::ex[s], L::ey[s], L::ez[s]), s, a);
    } else {
        Gather gat;
        gat.init(x, y, z, a);
#pragma unroll
        for (int s = 0; s < Q; ++s) {
            const st *p = fin + fidx<Q>(s, gat.src(L::ex[s], L::ey[s], L::ez[s]), a);
            f[s] = S::in(NTLOAD ? __builtin_nontemporal_load(p) : *p, s, a);
        }
    }
    T w[Q];
#pragma unroll
    for (int s = 0; s < Q; ++s) w[s] = S::weight(s, a);
    const int t = type_final<MODE_STEP>(t0, nd_type, chunk_type, g);
    T feq[Q], rho, ux, uy, uz;
    const bool relaxes = node_prepare<L, T, S::kShift, true, S::kScaled>(f, feq, t, rho, ux, uy, uz,
                                                                         a, uy_head, w, g);
    T pi[6];
    node_stress<L, T, S::kScaled>(f, feq, pi, w);
    float feat[kMlpIn];
#pragma unroll
    for (int k = 0; k < 6; ++k) feat[k] = float(pi[k]) * a.cl_in;
    // the norm feature is rounded to bf16: the hardware square root (1 ulp) suffices
    feat[6] = __builtin_amdgcn_sqrtf(float(pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                                           T(2) * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5]))) *
              a.cl_in;

    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wavefront's load into smlp has landed
    __syncthreads();                     // smlp complete (every lane of the workgroup gets here)
    const float zz = mlp_wave(feat, &smlp);
    const float om = __builtin_amdgcn_rcpf(a.tau0 + a.cl_out * softplus_f(zz));  // 1 ulp
    if (a.dynamics == 1) {
        if (kConserve<ST, Q> && t == 0)
            node_relax_conserving<L>(f, feq, w, om, rho, ux, uy, uz, a);
        else
            // the one node type that does not relax, the solid node, has
            // feq = f, so the rate 0 leaves it exactly as it is; no branch
            node_relax<L, T>(f, feq, relaxes ? T(om) : T(0));
    }
    if (!live) return;
    const PushOut<L, ST> po(fout, (blockIdx.x * BS) >> StorageT<ST>::kCB, g);
    const uint32_t rs = ST == ST_F16 ? round_stream(g, a.step) : 0u;
#pragma unroll
    for (int s = 0; s < Q; ++s)
        po.template store<true>(s, S::step_out(f[s], s, a, rs + uint32_t(s) * kStreamStep));
}

}  // namespace tlbm
