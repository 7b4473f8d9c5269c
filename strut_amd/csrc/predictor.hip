// Training and batched inference kernels of the learned Newton predictor
// (the network is described in mlp_device.h).  One minibatch step is:
//   mlp_prep_kernel     gather + normalise the minibatch (bf16 X, X^T; fp32 targets)
//   mlp_fwdbwd_kernel   per 16-sample tile: forward through the five layers
//                       with activations in LDS, squared-error loss, and the
//                       backward pass down to layer 1's pre-activation gradient
//   mlp_wgrad_kernel    per 16 x 16 weight tile: dW = dZ^T H over the minibatch
//                       (and the bias gradients)
//   mlp_adam_kernel     Adam on the FP32 master weights, refreshing the bf16
//                       copies W and W^T used by the matrix cores
#include <hip/hip_runtime.h>

#include "mlp_device.h"
#include "mlp_types.h"

namespace sgpu {

__global__ void mlp_prep_kernel(MlpPrepArgs P) {
    const MlpDev &m = P.m;
    const int width = m.k0 > m.n5 ? m.k0 : m.n5;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)P.batch * width) return;
    const int b = (int)(idx / width), f = (int)(idx % width);
    const bool valid = b < P.rows;
    const int64_t s = valid ? (P.perm ? (int64_t)P.perm[P.first + b] : P.first + b) : 0;
    if (f < m.k0) {
        float xn = 0.0f;
        if (valid && f < m.n_in) xn = (P.x[s * m.n_in + f] - m.x_mean[f]) * m.x_scale[f];
        const uint16_t h = f32_to_bf16(xn);
        P.xb[(int64_t)b * m.k0 + f] = h;
        P.xt[(int64_t)f * P.batch + b] = h;
    }
    if (f < m.n5 && P.t) {
        float tn = 0.0f;
        if (valid && P.y && f < m.n_out) tn = (P.y[s * m.n_out + f] - m.y_mean[f]) / m.y_scale[f];
        P.t[(int64_t)b * m.n5 + f] = tn;
    }
}

// Dynamic LDS of mlp_fwdbwd_kernel.
size_t mlp_fwdbwd_lds(int k0, int n5) {
    const int nz = n5 > kMlpHidden ? n5 : kMlpHidden;
    return 2 * (size_t)16 * (k0 + kMlpPad) + 2 * 2 * (size_t)16 * (kMlpHidden + kMlpPad) +
           4 * (size_t)4 * 16 * kMlpHidden + 2 * 2 * (size_t)16 * (nz + kMlpPad);
}

__global__ __launch_bounds__(256) void mlp_fwdbwd_kernel(MlpTrainArgs T) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ float red[4];
    const MlpDev &m = T.m;
    const int m0 = blockIdx.x * 16;
    const int lda0 = m.k0 + kMlpPad, ldh = kMlpHidden + kMlpPad;
    const int nz = m.n5 > kMlpHidden ? m.n5 : kMlpHidden, ldz = nz + kMlpPad;
    uint16_t *xa = reinterpret_cast<uint16_t *>(smem);
    uint16_t *ha0 = xa + 16 * lda0;
    uint16_t *ha1 = ha0 + 16 * ldh;
    float *z = reinterpret_cast<float *>(ha1 + 16 * ldh);
    uint16_t *dz0 = reinterpret_cast<uint16_t *>(z + 4 * 16 * kMlpHidden);
    uint16_t *dz1 = dz0 + 16 * ldz;
    for (int idx = threadIdx.x; idx < 16 * m.k0; idx += blockDim.x) {
        const int i = idx / m.k0, k = idx % m.k0;
        xa[i * lda0 + k] = T.xb[(int64_t)(m0 + i) * m.k0 + k];
    }
    __syncthreads();
    const uint16_t *a = xa;
    int lda = lda0;
    uint16_t *hb[2] = {ha0, ha1};
    for (int l = 0; l < kMlpLayers - 1; ++l) {
        uint16_t *h = hb[l & 1];
        float *zl = z + l * 16 * kMlpHidden;
        const float *bias = m.b[l];
        uint16_t *ht = T.ht[l];
        const bool keep = T.backward != 0;
        mlp_tile_gemm(a, lda, 16, m.w[l], kMlpHidden, mlp_k(m, l), [&](int i, int j, float c) {
            const float zz = c + bias[j];
            zl[i * kMlpHidden + j] = zz;
            const uint16_t hv = f32_to_bf16(silu_f(zz));
            h[i * ldh + j] = hv;
            if (keep) ht[(int64_t)j * T.batch + m0 + i] = hv;
        });
        __syncthreads();
        a = h;
        lda = ldh;
    }
    float lsum = 0.0f;
    {
        const float *bias = m.b[kMlpLayers - 1];
        mlp_tile_gemm(a, lda, 16, m.w[kMlpLayers - 1], m.n5, kMlpHidden, [&](int i, int j, float c) {
            const float yv = c + bias[j];
            const int row = m0 + i;
            if (T.y_raw && row < T.rows && j < m.n_out)
                T.y_raw[(int64_t)row * m.n_out + j] = yv * m.y_scale[j] + m.y_mean[j];
            if (T.backward) {
                float d = 0.0f;
                if (j < m.n_out && row < T.rows) {
                    const float e = yv - T.t[(int64_t)row * m.n5 + j];
                    lsum += e * e;
                    d = T.grad_scale * e;
                }
                const uint16_t db = f32_to_bf16(d);
                dz0[i * ldz + j] = db;
                T.dzt[kMlpLayers - 1][(int64_t)j * T.batch + row] = db;
            }
        });
    }
    if (!T.backward) return;
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
        T.loss_acc[blockIdx.x] += s;
    }
    uint16_t *dzb[2] = {dz0, dz1};
    int cur = 0;
    for (int l = kMlpLayers - 1; l >= 1; --l) {
        // dH_{l-1} = dZ_l W_l, then dZ_{l-1} = dH_{l-1} * SiLU'(Z_{l-1})
        const uint16_t *A = dzb[cur];
        uint16_t *D = dzb[cur ^ 1];
        const float *zp = z + (l - 1) * 16 * kMlpHidden;
        uint16_t *dzt = T.dzt[l - 1];
        mlp_tile_gemm(A, ldz, 16, m.wt[l], kMlpHidden, mlp_n(m, l), [&](int i, int k, float c) {
            const uint16_t db = f32_to_bf16(c * silu_grad_f(zp[i * kMlpHidden + k]));
            D[i * ldz + k] = db;
            dzt[(int64_t)k * T.batch + m0 + i] = db;
        });
        __syncthreads();
        cur ^= 1;
    }
}

// dW_l[n][k] = sum_m dZ_l[m][n] H_{l-1}[m][k] (H_{-1} = X) and
// db_l[n] = sum_m dZ_l[m][n]; one wavefront per 16 x 16 tile of dW_l.
__global__ __launch_bounds__(256) void mlp_wgrad_kernel(MlpTrainArgs T, MlpTiles tiles) {
    const MlpDev &m = T.m;
    const int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (t >= tiles.base[kMlpLayers]) return;
    int l = 0;
    while (t >= tiles.base[l + 1]) ++l;
    const int K = mlp_k(m, l), N = mlp_n(m, l);
    const int tt = t - tiles.base[l];
    const int nt = tt / (K / 16), kt = tt % (K / 16);
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int B = T.batch;
    const uint16_t *dz_row = T.dzt[l] + (int64_t)(nt * 16 + r) * B;
    const uint16_t *hh_row = (l == 0 ? T.xt : T.ht[l - 1]) + (int64_t)(kt * 16 + r) * B;
    f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
    float bs = 0.0f;
    int mb = 0;
    for (; mb + 32 <= B; mb += 32) {  // the batch is the inner dimension: 16 x 16 x 32 steps
        const bf16x4_t a0 = load4(dz_row + mb + 8 * g), a1 = load4(dz_row + mb + 8 * g + 4);
        const bf16x4_t b0 = load4(hh_row + mb + 8 * g), b1 = load4(hh_row + mb + 8 * g + 4);
        acc = mfma_16x16x32(a0, a1, b0, b1, acc);
        if (kt == 0) {
#pragma unroll
            for (int e = 0; e < 4; ++e) bs += bf16_to_f32((uint16_t)a0[e]) + bf16_to_f32((uint16_t)a1[e]);
        }
    }
    if (mb < B) {  // B = 16 mod 32
        const bf16x4_t a = load4(dz_row + mb + 4 * g);
        const bf16x4_t b = load4(hh_row + mb + 4 * g);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        if (kt == 0) {
#pragma unroll
            for (int e = 0; e < 4; ++e) bs += bf16_to_f32((uint16_t)a[e]);
        }
    }
    float *gw = T.g + tiles.off_w[l];
#pragma unroll
    for (int e = 0; e < 4; ++e) gw[(int64_t)(nt * 16 + 4 * g + e) * K + kt * 16 + r] = acc[e];
    if (kt == 0) {
        bs += __shfl_xor(bs, 16);
        bs += __shfl_xor(bs, 32);
        if (g == 0) T.g[tiles.off_b[l] + nt * 16 + r] = bs;
    }
    (void)N;
}

__global__ void mlp_adam_kernel(MlpAdamArgs P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const float g = P.g[i];
    const float mo = P.b1 * P.mom[i] + (1.0f - P.b1) * g;
    const float ve = P.b2 * P.vel[i] + (1.0f - P.b2) * (g * g);
    P.mom[i] = mo;
    P.vel[i] = ve;
    const float p = P.p[i] - P.lr_t * mo / (sqrtf(ve) + P.eps_t);
    P.p[i] = p;
    for (int l = 0; l < kMlpLayers; ++l) {
        const int64_t rel = i - P.off_w[l];
        if (rel >= 0 && rel < (int64_t)P.N[l] * P.K[l]) {
            const int n = (int)(rel / P.K[l]), k = (int)(rel % P.K[l]);
            const uint16_t h = f32_to_bf16(p);
            P.w[P.off_bf[l] + rel] = h;
            P.wt[P.off_bf[l] + (int64_t)k * P.N[l] + n] = h;
        }
    }
}

// Column sums and sums of squares of a row-major [rows][cols] FP32 matrix,
// in FP64, per chunk of rows (partial[chunk][col][2]).
__global__ void mlp_colstats_kernel(const float *x, int64_t rows, int cols, int64_t rows_per_chunk,
                                    double *partial) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= cols) return;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
    const int64_t r1 = r0 + rows_per_chunk < rows ? r0 + rows_per_chunk : rows;
    double s = 0.0, q = 0.0;
    for (int64_t r = r0; r < r1; ++r) {
        const double v = x[r * cols + col];
        s += v;
        q += v * v;
    }
    partial[((int64_t)blockIdx.y * cols + col) * 2] = s;
    partial[((int64_t)blockIdx.y * cols + col) * 2 + 1] = q;
}

hipError_t launch_mlp_prep(hipStream_t s, const MlpPrepArgs &P) {
    const int width = P.m.k0 > P.m.n5 ? P.m.k0 : P.m.n5;
    const int64_t n = (int64_t)P.batch * width;
    hipLaunchKernelGGL(mlp_prep_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P);
    return hipGetLastError();
}

hipError_t launch_mlp_fwdbwd(hipStream_t s, const MlpTrainArgs &T) {
    const size_t lds = mlp_fwdbwd_lds(T.m.k0, T.m.n5);
    hipLaunchKernelGGL(mlp_fwdbwd_kernel, dim3(T.batch / 16), dim3(256), lds, s, T);
    return hipGetLastError();
}

hipError_t launch_mlp_wgrad(hipStream_t s, const MlpTrainArgs &T, const MlpTiles &tiles) {
    const int n = tiles.base[kMlpLayers];
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((n + 3) / 4), dim3(256), 0, s, T, tiles);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(hipStream_t s, const MlpAdamArgs &P) {
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, s, P);
    return hipGetLastError();
}

hipError_t launch_mlp_colstats(hipStream_t s, const float *x, int64_t rows, int cols, int chunks, double *partial) {
    const int64_t per = (rows + chunks - 1) / chunks;
    hipLaunchKernelGGL(mlp_colstats_kernel, dim3((cols + 255) / 256, chunks), dim3(256), 0, s, x, rows, cols, per,
                       partial);
    return hipGetLastError();
}

hipError_t mlp_set_lds(int k0, int n5) {
    return hipFuncSetAttribute((const void *)mlp_fwdbwd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)mlp_fwdbwd_lds(k0, n5));
}

}  // namespace sgpu
