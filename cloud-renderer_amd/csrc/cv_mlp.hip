// Far-field MLP kernels: batched inference, training step (forward with saved
// activations, backward, per-workgroup gradient partials), deterministic
// reduction + Adam, and weight packing for the matrix cores.
//
// Training loss per row (weight w = remaining transmittance 1 - a_k):
//     w^2 * sum_c (y_c - t_c)^2   +   (y_4 - e)^2,
//     e = max_c |clamp(y_c, 0, 1) - t_c|   (treated as a constant),
// averaged over the batch.  The first term is the squared error the far
// field adds to the fragment colour; the second trains the error estimate
// that decides the fallback.
#include "cv_mlp.h"
#include "cv_mlp_launch.h"

namespace cv {

// ---------------------------------------------------------------------------
// batched inference: rows X[n][17] -> Y[n][5], one wave per 64 rows
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_mlp_forward(MlpDev w, const float* __restrict__ X, int n,
                                                    float* __restrict__ Y) {
    __shared__ __align__(16) uint16_t xs[64 * kMlpInPad];
    const int lane = threadIdx.x;
    const int row = blockIdx.x * 64 + lane;
    const bool valid = row < n;
#pragma unroll
    for (int k = 0; k < kMlpInPad; ++k) {
        float v = (valid && k < kMlpIn) ? X[(size_t)row * kMlpIn + k] : 0.0f;
        xs[lane * kMlpInPad + k] = (uint16_t)f2bf_bits(v);
    }
    __syncthreads();
    float out[kMlpOut];
    mlp_forward_wave(w, xs, lane, out);
    if (valid) {
#pragma unroll
        for (int q = 0; q < kMlpOut; ++q) Y[(size_t)row * kMlpOut + q] = out[q];
    }
}

// ---------------------------------------------------------------------------
// training: one wave per 64 rows; writes every parameter's partial gradient
// of this workgroup (no atomics) and its partial loss
// ---------------------------------------------------------------------------
struct TrainLds {
    uint16_t xs[64 * kMlpInPad];        // layer-0 input, row-major
    uint16_t hs[5][64 * 64];            // H_l as [unit][row] (l = 0 uses 32 units)
    uint16_t gs[64 * 64];               // dL/dZ_l as [unit][row]
    float zs[4][4][16][64];             // Z_l (l = 0..3) in accumulator order [mt*2+nt][j][lane]
};

__device__ __forceinline__ cv_u32x4 lds_row8(const uint16_t* base) { return *(const cv_u32x4*)base; }

// sum over the 32 lanes of each half (rows of the batch) of each register
__device__ __forceinline__ float half_sum(float v) {
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// G (accumulators of layer l's outputs, MT m-tiles) -> bias partials, LDS copy
template <int MT>
__device__ __forceinline__ void store_grad(const cv_f32x16 (&g)[2][2], int l, TrainLds& S, float* __restrict__ part,
                                           int lane) {
    const int h = lane >> 5, c = lane & 31;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int unit = 32 * mt + mlp_crow(h, j);
            S.gs[unit * 64 + c] = (uint16_t)f2bf_bits(g[mt][0][j]);
            S.gs[unit * 64 + 32 + c] = (uint16_t)f2bf_bits(g[mt][1][j]);
            float sm = half_sum(g[mt][0][j] + g[mt][1][j]);
            if (c == 0 && unit < mlp_out(l)) part[mlp_boff(l) + unit] = sm;
        }
    }
}

// dW_l = G H_l^T over the 64 rows (K = rows), written to the partials
template <int MT, int NT>
__device__ __forceinline__ void weight_grad(int l, const TrainLds& S, float* __restrict__ part, int lane) {
    const int h = lane >> 5, c = lane & 31;
    const int nin = mlp_in(l), nout = mlp_out(l);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            cv_f32x16 acc;
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                cv_u32x4 a = lds_row8(S.gs + (32 * mt + c) * 64 + 16 * s + 8 * h);
                cv_u32x4 b = lds_row8(S.hs[l] + (32 * nt + c) * 64 + 16 * s + 8 * h);
                acc = mfma_bf16(a, b, acc);
            }
            const int in = 32 * nt + c;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int out = 32 * mt + mlp_crow(h, j);
                if (out < nout && in < nin) part[mlp_woff(l) + out * nin + in] = acc[j];
            }
        }
    }
}

// dH_l = W_l^T G_l (layers 1..4), then G_{l-1} = dH_l * GELU'(Z_{l-1})
template <int KS>
__device__ __forceinline__ void back_layer(const MlpDev& w, int l, cv_f32x16 (&g)[2][2], const TrainLds& S,
                                           int lane) {
    cv_u32x4 B[4][2];
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) B[s][nt] = chain_b(g[s >> 1][nt], s & 1);
    cv_f32x16 d[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        cv_f32x16 c0, c1;
#pragma unroll
        for (int j = 0; j < 16; ++j) { c0[j] = 0.0f; c1[j] = 0.0f; }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            cv_u32x4 a = load_blk(w.bwd, mlp_bblk0(l) + mt * KS + s, lane);
            c0 = mfma_bf16(a, B[s][0], c0);
            c1 = mfma_bf16(a, B[s][1], c1);
        }
        d[mt][0] = c0;
        d[mt][1] = c1;
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int j = 0; j < 16; ++j) g[mt][nt][j] = d[mt][nt][j] * gelu_grad(S.zs[l - 1][mt * 2 + nt][j][lane]);
}

__global__ void __launch_bounds__(64) k_mlp_train(MlpDev w, const float* __restrict__ X, const float* __restrict__ T,
                                                  int n, float inv_n, float* __restrict__ partial,
                                                  float* __restrict__ loss_part) {
    __shared__ __align__(16) TrainLds S;
    const int lane = threadIdx.x;
    const int h = lane >> 5, c = lane & 31;
    const int row0 = blockIdx.x * 64;
    float* part = partial + (size_t)blockIdx.x * kMlpParamsPad;
    {
        const int row = row0 + lane;
        const bool valid = row < n;
#pragma unroll
        for (int k = 0; k < kMlpInPad; ++k) {
            float v = (valid && k < kMlpIn) ? X[(size_t)row * kMlpIn + k] : 0.0f;
            uint16_t b = (uint16_t)f2bf_bits(v);
            S.xs[lane * kMlpInPad + k] = b;
            S.hs[0][k * 64 + lane] = b;
        }
    }
    __syncthreads();

    // ---- forward, saving Z_l and H_{l+1}
    cv_u32x4 B[4][2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int s = 0; s < 2; ++s) B[s][nt] = x_operand(S.xs, 32 * nt + c, s, h);
    cv_f32x16 acc[2][2];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        if (l == 0) mlp_layer64<2>(w, 0, B, acc, lane);
        else mlp_layer64<4>(w, l, B, acc, lane);
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    float z = acc[mt][nt][j];
                    S.zs[l][mt * 2 + nt][j][lane] = z;
                    float a = gelu_f(z);
                    acc[mt][nt][j] = a;
                    S.hs[l + 1][(32 * mt + mlp_crow(h, j)) * 64 + 32 * nt + c] = (uint16_t)f2bf_bits(a);
                }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) B[s][nt] = chain_b(acc[s >> 1][nt], s & 1);
    }
    cv_f32x16 y[2];
    {
        cv_f32x16 y0 = bias_init(w.bias, mlp_bpad0(4), h);
        cv_f32x16 y1 = y0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            cv_u32x4 a = load_blk(w.fwd, mlp_fblk0(4) + s, lane);
            y0 = mfma_bf16(a, B[s][0], y0);
            y1 = mfma_bf16(a, B[s][1], y1);
        }
        y[0] = y0;
        y[1] = y1;
    }

    // ---- loss gradient in accumulator layout: rows 0..3 in registers 0..3
    // of half 0, row 4 in register 0 of half 1
    cv_f32x16 g[2][2];
    float lsum = 0.0f;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int row = row0 + 32 * nt + c;
        const bool valid = row < n;
        float e = 0.0f, wt = 0.0f;
        float d4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (h == 0 && valid) {
            const float* t = T + (size_t)row * kMlpOut;
            wt = t[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                d4[q] = y[nt][q] - t[q];
                e = fmaxf(e, fabsf(fminf(fmaxf(y[nt][q], 0.0f), 1.0f) - t[q]));
            }
        }
        float e_hi = __shfl(e, c, 64);
#pragma unroll
        for (int j = 0; j < 16; ++j) g[0][nt][j] = 0.0f;
        if (valid) {
            if (h == 0) {
                const float w2 = wt * wt;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    g[0][nt][q] = 2.0f * w2 * d4[q] * inv_n;
                    lsum = fmaf(w2 * d4[q], d4[q], lsum);
                }
            } else {
                const float de = y[nt][0] - e_hi;
                g[0][nt][0] = 2.0f * de * inv_n;
                lsum = fmaf(de, de, lsum);
            }
        }
    }
    {
        float tot = lsum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) tot += __shfl_xor(tot, off, 64);
        if (lane == 0) loss_part[blockIdx.x] = tot * inv_n;
    }

    // ---- backward
    store_grad<1>(g, 4, S, part, lane);
    __syncthreads();
    weight_grad<1, 2>(4, S, part, lane);
    back_layer<2>(w, 4, g, S, lane);
    __syncthreads();
#pragma unroll
    for (int l = 3; l >= 1; --l) {
        store_grad<2>(g, l, S, part, lane);
        __syncthreads();
        weight_grad<2, 2>(l, S, part, lane);
        back_layer<4>(w, l, g, S, lane);
        __syncthreads();
    }
    store_grad<2>(g, 0, S, part, lane);
    __syncthreads();
    weight_grad<2, 1>(0, S, part, lane);
}

// ---------------------------------------------------------------------------
// gradient reduction (fixed order over workgroups) + Adam
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mlp_adam(const float* __restrict__ partial, int nwg,
                                                  const float* __restrict__ loss_part, float* __restrict__ master,
                                                  float* __restrict__ m1, float* __restrict__ m2,
                                                  float* __restrict__ grad_out, float* __restrict__ loss_out,
                                                  AdamStep st) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss_out) {
        float s = 0.0f;
        for (int i = 0; i < nwg; ++i) s += loss_part[i];
        *loss_out = s;
    }
    if (p >= kMlpParams) return;
    float gsum = 0.0f;
    for (int i = 0; i < nwg; ++i) gsum += partial[(size_t)i * kMlpParamsPad + p];
    if (grad_out) grad_out[p] = gsum;
    if (st.apply) {
        float a = fmaf(st.beta1, m1[p], (1.0f - st.beta1) * gsum);
        float b = fmaf(st.beta2, m2[p], (1.0f - st.beta2) * gsum * gsum);
        m1[p] = a;
        m2[p] = b;
        float mh = a / st.bc1, vh = b / st.bc2;
        master[p] = master[p] - st.lr * mh / (sqrtf(vh) + st.eps);
    }
}

// ---------------------------------------------------------------------------
// packing: FP32 master -> bf16 A blocks (forward W, backward W^T), padded biases
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mlp_pack(const float* __restrict__ master, uint16_t* __restrict__ fwd,
                                                  uint16_t* __restrict__ bwd, float* __restrict__ bias) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = kMlpFwdBlocks * 512, nb = kMlpBwdBlocks * 512;
    if (t < nf) {
        const int blk = t >> 9, lane = (t >> 3) & 63, q = t & 7;
        int l = 0;
        while (l < 4 && blk >= mlp_fblk0(l + 1)) ++l;
        const int KS = mlp_kpad(l) / 16;
        const int local = blk - mlp_fblk0(l), mt = local / KS, s = local % KS;
        const int m = 32 * mt + (lane & 31);
        const int k = 16 * s + mlp_kperm(lane >> 5, q);
        float v = (m < mlp_out(l) && k < mlp_in(l)) ? master[mlp_woff(l) + m * mlp_in(l) + k] : 0.0f;
        fwd[t] = (uint16_t)f2bf_bits(v);
    } else if (t < nf + nb) {
        const int u = t - nf;
        const int blk = u >> 9, lane = (u >> 3) & 63, q = u & 7;
        int l = 1;
        while (l < 4 && blk >= mlp_bblk0(l + 1)) ++l;
        const int KS = mlp_mpad(l) / 16;
        const int local = blk - mlp_bblk0(l), mt = local / KS, s = local % KS;
        const int m = 32 * mt + (lane & 31);           // input unit of layer l
        const int k = 16 * s + mlp_kperm(lane >> 5, q);  // output unit of layer l
        float v = (k < mlp_out(l) && m < mlp_in(l)) ? master[mlp_woff(l) + k * mlp_in(l) + m] : 0.0f;
        bwd[u] = (uint16_t)f2bf_bits(v);
    } else if (t < nf + nb + kMlpBiasPad) {
        const int u = t - nf - nb;
        const int l = u / 64, r = u % 64;
        bias[u] = (l < 5 && r < mlp_out(l)) ? master[mlp_boff(l) + r] : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
void launch_mlp_pack(const float* master, uint16_t* fwd, uint16_t* bwd, float* bias, hipStream_t s) {
    const int total = kMlpFwdBlocks * 512 + kMlpBwdBlocks * 512 + kMlpBiasPad;
    hipLaunchKernelGGL(k_mlp_pack, dim3((total + 255) / 256), dim3(256), 0, s, master, fwd, bwd, bias);
}

void launch_mlp_forward(const MlpDev& w, const float* X, int n, float* Y, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + 63) / 64), dim3(64), 0, s, w, X, n, Y);
}

int mlp_train_groups(int n) { return (n + 63) / 64; }

void launch_mlp_train(const MlpDev& w, const float* X, const float* T, int n, float* partial, float* loss_part,
                      hipStream_t s) {
    const int nwg = mlp_train_groups(n);
    hipLaunchKernelGGL(k_mlp_train, dim3(nwg), dim3(64), 0, s, w, X, T, n, 1.0f / (float)n, partial,
                       loss_part);
}

void launch_mlp_adam(const float* partial, int nwg, const float* loss_part, float* master, float* m1, float* m2,
                     float* grad_out, float* loss_out, const AdamStep& st, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, partial, nwg, loss_part, master,
                       m1, m2, grad_out, loss_out, st);
}

}  // namespace cv
