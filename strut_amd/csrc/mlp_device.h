// Device side of the learned Newton predictor: a 4 x 256 SiLU MLP with BF16
// weights and activations on the gfx950 matrix cores (FP32 accumulate).
//
// Layer l (1..5) maps K_l inputs to N_l outputs: K_1 = k0 (the padded
// feature count), N_1..N_4 = K_2..K_5 = 256, N_5 = n5 (the padded output
// count).  Hidden layers apply SiLU; the output layer is linear.  Padding
// rows and columns of every weight matrix are zero and stay zero (their
// gradients are zero).
//
// Matrix-core tile (v_mfma_f32_16x16x16_bf16): for a 16 x 16 x 16 product
// D = A B, lane l holds A[l % 16][4 (l / 16) + e] and B[4 (l / 16) + e][l % 16]
// (e = 0..3) and receives D[4 (l / 16) + e][l % 16].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgpu {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kMlpHidden = 256;
constexpr int kMlpLayers = 5;
constexpr int kMlpPad = 4;  // LDS row skew of activation tiles, in bf16 elements

// Round-to-nearest-even FP32 -> BF16 (finite inputs), and back.
__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

__device__ __forceinline__ float silu_f(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu_grad_f(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// Device view of one network.
struct MlpDev {
    int32_t n_in, n_out;  // valid features / outputs
    int32_t k0, n5;       // padded to multiples of 16
    const uint16_t *w[kMlpLayers];   // bf16 [N_l][K_l]
    const uint16_t *wt[kMlpLayers];  // bf16 [K_l][N_l]
    const float *b[kMlpLayers];      // [N_l]
    const float *x_mean, *x_scale;   // feature f: (x - x_mean) * x_scale
    const float *y_mean, *y_scale;   // output j: y_n * y_scale + y_mean
};

__host__ __device__ __forceinline__ int mlp_k(const MlpDev &m, int l) { return l == 0 ? m.k0 : kMlpHidden; }
__host__ __device__ __forceinline__ int mlp_n(const MlpDev &m, int l) { return l == kMlpLayers - 1 ? m.n5 : kMlpHidden; }

// C[16][N] = A[16][K] W^T for one 16-row tile: A is bf16 in LDS (row stride
// lda elements; rows >= `rows` are taken as zero and never read), W is bf16
// [N][K] in global memory, K and N multiples of 16.  Each wavefront of the
// block takes 16-column tiles in turn; epi(i, j, c) receives every element.
template <class Epi>
__device__ __forceinline__ void mlp_tile_gemm(const uint16_t *A, int lda, int rows, const uint16_t *W, int N, int K,
                                              Epi epi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const bool arow = r < rows;
    const uint16_t *ap = A + r * lda + 4 * g;
    for (int jt = wave; jt < N / 16; jt += nw) {
        const uint16_t *wp = W + (size_t)(jt * 16 + r) * K + 4 * g;
        f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        int kb = 0;
        for (; kb + 64 <= K; kb += 64) {  // four independent weight loads in flight
            bf16x4_t b0 = *reinterpret_cast<const bf16x4_t *>(wp + kb);
            bf16x4_t b1 = *reinterpret_cast<const bf16x4_t *>(wp + kb + 16);
            bf16x4_t b2 = *reinterpret_cast<const bf16x4_t *>(wp + kb + 32);
            bf16x4_t b3 = *reinterpret_cast<const bf16x4_t *>(wp + kb + 48);
            bf16x4_t a0 = {0, 0, 0, 0}, a1 = a0, a2 = a0, a3 = a0;
            if (arow) {
                a0 = *reinterpret_cast<const bf16x4_t *>(ap + kb);
                a1 = *reinterpret_cast<const bf16x4_t *>(ap + kb + 16);
                a2 = *reinterpret_cast<const bf16x4_t *>(ap + kb + 32);
                a3 = *reinterpret_cast<const bf16x4_t *>(ap + kb + 48);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b2, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, b3, acc, 0, 0, 0);
        }
        for (; kb < K; kb += 16) {
            bf16x4_t b0 = *reinterpret_cast<const bf16x4_t *>(wp + kb);
            bf16x4_t a0 = {0, 0, 0, 0};
            if (arow) a0 = *reinterpret_cast<const bf16x4_t *>(ap + kb);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc, 0, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) epi(4 * g + e, jt * 16 + r, acc[e]);
    }
}

// One-row product for in-step inference (C[0][N] = A[0][K] W^T), the
// latency-bound case: a wavefront issues all weight loads of a 256-wide K
// chunk of its 16-column tile before the matrix operations, so a layer costs
// a few memory latencies.  Lanes of row 0 read A from LDS (the other rows
// are zero).
template <class Epi>
__device__ __forceinline__ void mlp_row_gemm(const uint16_t *A, const uint16_t *W, int N, int K, Epi epi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    for (int jt = wave; jt < N / 16; jt += nw) {
        const uint16_t *wp = W + (size_t)(jt * 16 + r) * K + 4 * g;
        f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int kc = 0; kc < K; kc += 256) {
            const int nk = (K - kc) / 16 < 16 ? (K - kc) / 16 : 16;
            bf16x4_t b[16];
#pragma unroll
            for (int t = 0; t < 16; ++t)
                if (t < nk) b[t] = *reinterpret_cast<const bf16x4_t *>(wp + kc + 16 * t);
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                if (t < nk) {
                    bf16x4_t a = {0, 0, 0, 0};
                    if (r == 0) a = *reinterpret_cast<const bf16x4_t *>(A + 4 * g + kc + 16 * t);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[t], acc, 0, 0, 0);
                }
            }
        }
        if (g == 0) epi(jt * 16 + r, acc[0]);  // row 0 of D: lanes 0..15, element 0
    }
}

// Inference of one sample by the whole block.  feat (LDS, k0 + kMlpPad bf16)
// holds the normalised, padded features; h0/h1 (LDS, kMlpHidden + kMlpPad
// bf16 each) are scratch; out (LDS, n5 floats) receives the de-normalised
// outputs.  Called by every thread; ends with a barrier.
__device__ __forceinline__ void mlp_infer_one(const MlpDev &m, const uint16_t *feat, uint16_t *h0, uint16_t *h1,
                                              float *out) {
    const uint16_t *a = feat;
    uint16_t *hb[2] = {h0, h1};
    for (int l = 0; l < kMlpLayers - 1; ++l) {
        uint16_t *h = hb[l & 1];
        const float *bias = m.b[l];
        mlp_row_gemm(a, m.w[l], kMlpHidden, mlp_k(m, l),
                     [&](int j, float c) { h[j] = f32_to_bf16(silu_f(c + bias[j])); });
        __syncthreads();
        a = h;
    }
    const float *bias = m.b[kMlpLayers - 1];
    mlp_row_gemm(a, m.w[kMlpLayers - 1], m.n5, kMlpHidden, [&](int j, float c) {
        if (j < m.n_out) out[j] = (c + bias[j]) * m.y_scale[j] + m.y_mean[j];
    });
    __syncthreads();
}

}  // namespace sgpu
