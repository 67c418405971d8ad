// Far-field MLP: 17 -> 64 -> 64 -> 64 -> 64 -> 5, GELU (tanh form) on the
// hidden layers, BF16 weights and activations, FP32 accumulation on the
// gfx950 matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Inputs (per cone, after k marched steps): the last sample position (2x-1,
// texture space), the cone direction, the sun direction, and (density, light)
// sampled at LODs 3..6 at the distances along the cone where its radius
// matches those LODs.  Outputs: the far field's accumulated (r, g, b, a) in
// units of the remaining transmittance, and an estimate of its own error.
//
// A wave evaluates 64 rows (one per lane) as Y^T = W X^T: weights are the A
// operand (rows = output units), activations the B operand (columns = rows of
// the batch).  With K permuted inside each 16-wide step as
//     slot q of lane-half h  <->  feature 16 s + 8 (q / 4) + 4 h + (q % 4)
// the accumulator registers 8 (s % 2) .. 8 (s % 2) + 7 of M-tile s / 2 are
// exactly the B operand of step s of the next layer, so layers chain in
// registers.  Weight blocks are pre-packed in that order (one 16-byte load per
// lane per MFMA).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cv {

constexpr int kMlpIn = 17;
constexpr int kMlpHid = 64;
constexpr int kMlpOut = 5;
constexpr int kMlpLayers = 5;
constexpr int kMlpInPad = 32;   // K of layer 0
constexpr int kMlpOutPad = 32;  // M of layer 4

// flat FP32 parameter vector: W0 (64x17), b0, W1 (64x64), b1, W2, b2, W3, b3, W4 (5x64), b4
constexpr int kMlpW0 = 0;
constexpr int kMlpB0 = kMlpW0 + kMlpHid * kMlpIn;
constexpr int kMlpW1 = kMlpB0 + kMlpHid;
constexpr int kMlpB1 = kMlpW1 + kMlpHid * kMlpHid;
constexpr int kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;
constexpr int kMlpW3 = kMlpB2 + kMlpHid;
constexpr int kMlpB3 = kMlpW3 + kMlpHid * kMlpHid;
constexpr int kMlpW4 = kMlpB3 + kMlpHid;
constexpr int kMlpB4 = kMlpW4 + kMlpOut * kMlpHid;
constexpr int kMlpParams = kMlpB4 + kMlpOut;  // 13957
constexpr int kMlpParamsPad = 13960;

__host__ __device__ constexpr int mlp_in(int l) { return l == 0 ? kMlpIn : kMlpHid; }
__host__ __device__ constexpr int mlp_out(int l) { return l == 4 ? kMlpOut : kMlpHid; }
__host__ __device__ constexpr int mlp_kpad(int l) { return l == 0 ? kMlpInPad : kMlpHid; }
__host__ __device__ constexpr int mlp_mpad(int l) { return l == 4 ? kMlpOutPad : kMlpHid; }
__host__ __device__ constexpr int mlp_woff(int l) {
    return l == 0 ? kMlpW0 : l == 1 ? kMlpW1 : l == 2 ? kMlpW2 : l == 3 ? kMlpW3 : kMlpW4;
}
__host__ __device__ constexpr int mlp_boff(int l) {
    return l == 0 ? kMlpB0 : l == 1 ? kMlpB1 : l == 2 ? kMlpB2 : l == 3 ? kMlpB3 : kMlpB4;
}
// forward A blocks (W, M = outputs, K = inputs), 512 bf16 each: layer l, m-tile mt, k-step s
__host__ __device__ constexpr int mlp_fblk0(int l) { return l == 0 ? 0 : l == 1 ? 4 : l == 2 ? 12 : l == 3 ? 20 : 28; }
constexpr int kMlpFwdBlocks = 32;
// backward A blocks (W^T, M = inputs, K = outputs) for layers 1..4
__host__ __device__ constexpr int mlp_bblk0(int l) { return l == 1 ? 0 : l == 2 ? 8 : l == 3 ? 16 : 24; }
constexpr int kMlpBwdBlocks = 28;
// padded biases: 64, 64, 64, 64, 32
__host__ __device__ constexpr int mlp_bpad0(int l) { return l * 64; }
constexpr int kMlpBiasPad = 288;

// K permutation inside a 16-wide step: (lane half h, slot q) -> feature offset
__host__ __device__ constexpr int mlp_kperm(int h, int q) { return 8 * (q >> 2) + 4 * h + (q & 3); }
// accumulator register j of lane half h -> row offset inside a 32-row tile
__host__ __device__ constexpr int mlp_crow(int h, int j) { return 8 * (j >> 2) + 4 * h + (j & 3); }

constexpr float kGeluC1 = 1.5957691216057308f;                  // 2 sqrt(2/pi)
constexpr float kGeluC2 = 1.5957691216057308f * 0.044715f;      // 2 sqrt(2/pi) 0.044715

struct MlpDev {
    const uint16_t* fwd;   // kMlpFwdBlocks * 512 bf16
    const uint16_t* bwd;   // kMlpBwdBlocks * 512 bf16
    const float* bias;     // kMlpBiasPad
};

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
typedef __bf16 cv_bf16x8 __attribute__((ext_vector_type(8)));
typedef float cv_f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int cv_u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t f2bf_bits(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);  // inf / nan
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf2f(uint32_t b) { return __uint_as_float(b << 16); }
__device__ __forceinline__ uint32_t pack_bf2(float lo, float hi) { return f2bf_bits(lo) | (f2bf_bits(hi) << 16); }

__device__ __forceinline__ cv_bf16x8 as_bf16x8(cv_u32x4 v) { return __builtin_bit_cast(cv_bf16x8, v); }

__device__ __forceinline__ cv_f32x16 mfma_bf16(cv_u32x4 a, cv_u32x4 b, cv_f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a), as_bf16x8(b), c, 0, 0, 0);
}

// GELU, tanh form: x * sigmoid(2 sqrt(2/pi) (x + 0.044715 x^3))
__device__ __forceinline__ float gelu_f(float x) {
    float u = x * fmaf(kGeluC2, x * x, kGeluC1);
    float e = __builtin_amdgcn_exp2f(-u * 1.4426950408889634f);
    return x * __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu_grad(float x) {
    float x2 = x * x;
    float u = x * fmaf(kGeluC2, x2, kGeluC1);
    float e = __builtin_amdgcn_exp2f(-u * 1.4426950408889634f);
    float sg = __builtin_amdgcn_rcpf(1.0f + e);
    return fmaf(x * sg * (1.0f - sg), fmaf(3.0f * kGeluC2, x2, kGeluC1), sg);
}

// accumulator registers 8 (s % 2) .. + 7 as the bf16 B operand of step s
__device__ __forceinline__ cv_u32x4 chain_b(const cv_f32x16& acc, int half) {
    cv_u32x4 r;
    const int o = 8 * half;
    r.x = pack_bf2(acc[o + 0], acc[o + 1]);
    r.y = pack_bf2(acc[o + 2], acc[o + 3]);
    r.z = pack_bf2(acc[o + 4], acc[o + 5]);
    r.w = pack_bf2(acc[o + 6], acc[o + 7]);
    return r;
}

__device__ __forceinline__ cv_u32x4 load_blk(const uint16_t* __restrict__ p, int blk, int lane) {
    return *(const cv_u32x4*)(p + (size_t)blk * 512 + lane * 8);
}

__device__ __forceinline__ cv_f32x16 bias_init(const float* __restrict__ b, int row0, int h) {
    cv_f32x16 c;
#pragma unroll
    for (int j = 0; j < 16; ++j) c[j] = b[row0 + mlp_crow(h, j)];
    return c;
}

// Layer-0 B operand of row n (k-step s) from a row-major bf16 feature tile
// xs[64][kMlpInPad] in LDS.
__device__ __forceinline__ cv_u32x4 x_operand(const uint16_t* xs, int n, int s, int h) {
    const uint16_t* r = xs + n * kMlpInPad + 16 * s + 4 * h;
    uint2 lo = *(const uint2*)r;
    uint2 hi = *(const uint2*)(r + 8);
    cv_u32x4 v;
    v.x = lo.x; v.y = lo.y; v.z = hi.x; v.w = hi.y;
    return v;
}

// One hidden layer l (1..3) or the first layer (from B operands), 64 outputs.
// in[s][nt]: B operands, KS steps; out: 2 x 2 accumulators (pre-activation).
template <int KS>
__device__ __forceinline__ void mlp_layer64(const MlpDev& w, int l, const cv_u32x4 (&in)[4][2], cv_f32x16 (&acc)[2][2],
                                            int lane) {
    const int h = lane >> 5;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        cv_f32x16 c0 = bias_init(w.bias, mlp_bpad0(l) + 32 * mt, h);
        cv_f32x16 c1 = c0;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            cv_u32x4 a = load_blk(w.fwd, mlp_fblk0(l) + mt * KS + s, lane);
            c0 = mfma_bf16(a, in[s][0], c0);
            c1 = mfma_bf16(a, in[s][1], c1);
        }
        acc[mt][0] = c0;
        acc[mt][1] = c1;
    }
}

// Forward pass of one 32-row half (N-tile nt) of the wave's batch: layers
// chained in registers.  Returns the accumulator of the output layer (rows
// 0..4 = outputs, column = row 32 nt + lane % 32).
__device__ __forceinline__ cv_f32x16 mlp_forward_tile(const MlpDev& w, const uint16_t* xs, int nt, int lane) {
    const int h = lane >> 5, c = lane & 31;
    cv_u32x4 B[4];
    B[0] = x_operand(xs, 32 * nt + c, 0, h);
    B[1] = x_operand(xs, 32 * nt + c, 1, h);
    cv_f32x16 acc[2];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int KS = l == 0 ? 2 : 4;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            cv_f32x16 a = bias_init(w.bias, mlp_bpad0(l) + 32 * mt, h);
#pragma unroll
            for (int s = 0; s < KS; ++s) a = mfma_bf16(load_blk(w.fwd, mlp_fblk0(l) + mt * KS + s, lane), B[s], a);
            acc[mt] = a;
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[mt][j] = gelu_f(acc[mt][j]);
#pragma unroll
        for (int s = 0; s < 4; ++s) B[s] = chain_b(acc[s >> 1], s & 1);
        // keep each layer's weight loads inside the layer (bounds live registers)
        __builtin_amdgcn_sched_barrier(0);
    }
    cv_f32x16 y = bias_init(w.bias, mlp_bpad0(4), h);
#pragma unroll
    for (int s = 0; s < 4; ++s) y = mfma_bf16(load_blk(w.fwd, mlp_fblk0(4) + s, lane), B[s], y);
    return y;
}

// Forward pass for the 64 rows of a wave.  xs: LDS tile [64][kMlpInPad] bf16
// (row n = lane n's features, zero padded).  Every lane must be active.
// Returns the lane's own 5 outputs in out[]: row r of column n sits in
// register r (r < 4) of lane n % 32, and in register 0 of lane n % 32 + 32
// for r = 4, in the accumulator of N-tile n / 32.
__device__ __forceinline__ void mlp_forward_wave(const MlpDev& w, const uint16_t* xs, int lane, float out[kMlpOut]) {
    const int src = lane & 31;
    const bool hi = lane >= 32;
#pragma unroll 1
    for (int nt = 0; nt < 2; ++nt) {
        cv_f32x16 y = mlp_forward_tile(w, xs, nt, lane);
        float v0 = __shfl(y[0], src, 64), v1 = __shfl(y[1], src, 64), v2 = __shfl(y[2], src, 64),
              v3 = __shfl(y[3], src, 64), v4 = __shfl(y[0], src + 32, 64);
        if (hi == (nt == 1)) {
            out[0] = v0; out[1] = v1; out[2] = v2; out[3] = v3; out[4] = v4;
        }
    }
}
#endif

}  // namespace cv
