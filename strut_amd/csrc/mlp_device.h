// Device side of the learned Newton predictor: a 4 x 256 SiLU MLP with BF16
// weights and activations on the gfx950 matrix cores (FP32 accumulate).
//
// Layer l (1..5) maps K_l inputs to N_l outputs: K_1 = k0 (the padded
// feature count), N_1..N_4 = K_2..K_5 = 256, N_5 = n5 (the padded output
// count).  Hidden layers apply SiLU; the output layer is linear.  Padding
// rows and columns of every weight matrix are zero and stay zero (their
// gradients are zero).
//
// Matrix-core tiles.  v_mfma_f32_16x16x32_bf16 (gfx950) takes 32 of the
// inner dimension per instruction: for D = A B (16 x 16 x 32), lane l holds
// A[l % 16][8 (l / 16) + e] and B[8 (l / 16) + e][l % 16] (e = 0..7); an
// inner dimension that ends in 16 takes one v_mfma_f32_16x16x16_bf16 step,
// where lane l holds A[l % 16][4 (l / 16) + e] and B[4 (l / 16) + e][l % 16]
// (e = 0..3).  Both leave D[4 (l / 16) + e][l % 16] (e = 0..3) in lane l.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgpu {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef short bf16x8s_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// One 16 x 16 x 32 step from two 4-element halves of each operand (8-byte
// loads: activation rows in LDS are only 8-byte aligned).
__device__ __forceinline__ f32x4_t mfma_16x16x32(bf16x4_t a_lo, bf16x4_t a_hi, bf16x4_t b_lo, bf16x4_t b_hi,
                                                 f32x4_t acc) {
    const bf16x8s_t a = __builtin_shufflevector(a_lo, a_hi, 0, 1, 2, 3, 4, 5, 6, 7);
    const bf16x8s_t b = __builtin_shufflevector(b_lo, b_hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                  acc, 0, 0, 0);
}

__device__ __forceinline__ bf16x4_t load4(const uint16_t *p) { return *reinterpret_cast<const bf16x4_t *>(p); }

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
    const uint16_t *ap = A + r * lda + 8 * g;  // 16 x 16 x 32 steps
    const uint16_t *ap16 = A + r * lda + 4 * g;  // the 16 x 16 x 16 tail
    for (int jt = wave; jt < N / 16; jt += nw) {
        const uint16_t *wp = W + (size_t)(jt * 16 + r) * K + 8 * g;
        const uint16_t *wp16 = W + (size_t)(jt * 16 + r) * K + 4 * g;
        f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        int kb = 0;
        for (; kb + 64 <= K; kb += 64) {  // four independent weight loads in flight
            const bf16x4_t b0 = load4(wp + kb), b1 = load4(wp + kb + 4);
            const bf16x4_t b2 = load4(wp + kb + 32), b3 = load4(wp + kb + 36);
            bf16x4_t a0 = {0, 0, 0, 0}, a1 = a0, a2 = a0, a3 = a0;
            if (arow) {
                a0 = load4(ap + kb);
                a1 = load4(ap + kb + 4);
                a2 = load4(ap + kb + 32);
                a3 = load4(ap + kb + 36);
            }
            acc = mfma_16x16x32(a0, a1, b0, b1, acc);
            acc = mfma_16x16x32(a2, a3, b2, b3, acc);
        }
        for (; kb + 32 <= K; kb += 32) {
            const bf16x4_t b0 = load4(wp + kb), b1 = load4(wp + kb + 4);
            bf16x4_t a0 = {0, 0, 0, 0}, a1 = a0;
            if (arow) {
                a0 = load4(ap + kb);
                a1 = load4(ap + kb + 4);
            }
            acc = mfma_16x16x32(a0, a1, b0, b1, acc);
        }
        if (kb < K) {  // K = 16 mod 32
            const bf16x4_t b0 = load4(wp16 + kb);
            bf16x4_t a0 = {0, 0, 0, 0};
            if (arow) a0 = load4(ap16 + kb);
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
        const uint16_t *wp = W + (size_t)(jt * 16 + r) * K + 8 * g;
        const uint16_t *wp16 = W + (size_t)(jt * 16 + r) * K + 4 * g;
        f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int kc = 0; kc < K; kc += 256) {
            const int nk = (K - kc) / 32 < 8 ? (K - kc) / 32 : 8;  // 16 x 16 x 32 steps of this chunk
            bf16x4_t bl[8], bh[8];
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (t < nk) {
                    bl[t] = load4(wp + kc + 32 * t);
                    bh[t] = load4(wp + kc + 32 * t + 4);
                }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                if (t < nk) {
                    bf16x4_t al = {0, 0, 0, 0}, ah = al;
                    if (r == 0) {
                        al = load4(A + 8 * g + kc + 32 * t);
                        ah = load4(A + 8 * g + kc + 32 * t + 4);
                    }
                    acc = mfma_16x16x32(al, ah, bl[t], bh[t], acc);
                }
            }
            const int kt = kc + 32 * nk;
            if (nk < 8 && kt < K) {  // K = 16 mod 32: one 16 x 16 x 16 step
                const bf16x4_t b0 = load4(wp16 + kt);
                bf16x4_t a0 = {0, 0, 0, 0};
                if (r == 0) a0 = load4(A + 4 * g + kt);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc, 0, 0, 0);
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
