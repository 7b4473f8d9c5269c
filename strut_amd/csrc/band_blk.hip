// Blocked BandGeneral factorization and solve with the reference's row
// swaps, for bands too wide for one workgroup's registers, on gfx950.
//
// The reference's factorization (linear_solver_backend.mojo:631-679, with
// the dense partial-pivot LU of :175-214 where a pivot search picks another
// row) does, per step k: pick the first row of largest magnitude in column
// k, swap it with row k, form each multiplier l_ik = a_ik / a_kk, and
// subtract l_ik a_kj from every a_ij of the window, one rounding per
// product and per difference, steps in ascending k.  Here the steps are
// taken sixteen columns at a time (a panel) and every entry still receives
// the same operations in the same order:
//   * panel: one workgroup per system, one wavefront per panel column, the
//     panel's rows (k0 .. k0 + 15 + hbw) in registers; per column, the
//     owning wavefront searches, swaps, divides and publishes the
//     multipliers in LDS (one barrier), every other wavefront swaps the two
//     rows in its column and the later ones update theirs.  The multipliers
//     are stored where they were formed (the reference's band LU keeps them
//     so, and its solves swap the right-hand side step by step); the
//     trailing work takes them as the rows stand after the panel's last
//     swap (rows moved by a later swap carry their multipliers along), which
//     gives each row the same products in the same order;
//   * the panel's swaps applied in order to the trailing columns (k0 + 16 ..
//     k0 + 15 + 2 hbw) and the right-hand side, then the panel's pivot rows
//     finished there (u_kj -= l_kk' u_k'j, k' ascending), in the same
//     launch, one thread per column;
//   * the trailing update a_ij -= l_ik u_kj over the panel's sixteen k in
//     ascending order, tiles of 64 x 64 entries over as many workgroups as
//     the window holds (the right-hand side is one more column, so the
//     forward substitution rides along);
//   * the back substitution by blocks of 64 rows from the bottom: the
//     block's sums over the solved unknowns in parallel, then the block's
//     triangle in one wavefront (its sums run in another order than the
//     reference's, as the other band kernels' back substitutions do).
// Storage: `work` holds each row at columns i - hbw .. i + 2 hbw (n x
// (3 hbw + 1)), followed by the n pivot rows (int32) and the panel's
// multipliers as the rows stand (16 x (hbw + 16)); x holds the forward
// values and then the solution.  Singular (a column's largest magnitude at
// or below 1e-14): status 2 and no solution, as the other kernels.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/strut_gpu.h"
#include "ctx.h"

namespace sgpu {

namespace {

constexpr int kBkNb = 16;                 // panel columns, one wavefront each
constexpr int kBkThreads = 64 * kBkNb;    // 1024
constexpr int kBkMaxM = 20;               // panel rows per lane: hbw + 16 <= 1280
constexpr int kBkTile = 64;               // trailing update tile (rows and columns)
constexpr int kBkBack = 64;               // back substitution block rows
constexpr double kBkPivotTol = 1.0e-14;

struct BkSys {
    double *work, *x, *status;
    int32_t *piv;
    double *lt;  // the panel's multipliers with its later swaps applied: 16 columns x (hbw + 16) rows
};

__device__ __forceinline__ BkSys bk_sys(const int32_t *recs, int n, int hbw, double *work, double *x,
                                        double *status) {
    const int r = recs ? recs[blockIdx.y] : 0;
    const size_t V = 3 * (size_t)hbw + 1;
    const size_t per = (size_t)n * V + ((size_t)n + 1) / 2 + 1 + (size_t)kBkNb * (hbw + kBkNb);  // per system
    BkSys s;
    s.work = work + (size_t)r * per;
    s.piv = (int32_t *)(s.work + (size_t)n * V);
    s.lt = s.work + (size_t)n * V + ((size_t)n + 1) / 2 + 1;
    s.x = x ? x + (size_t)r * n : nullptr;  // null: factor only
    s.status = status + 4 * (size_t)r;
    return s;
}

__device__ __forceinline__ double lane_read(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// value of panel row i (relative index q = i - k0) held in a[q / 64] of lane q % 64
__device__ __forceinline__ double bk_get(const double (&a)[kBkMaxM], int q) {
    const int m = q >> 6, l = q & 63;
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < kBkMaxM; ++t)
        if (t == m) v = lane_read(a[t], l);
    return v;
}

__device__ __forceinline__ void bk_set(double (&a)[kBkMaxM], int q, double v, int lane) {
    const int m = q >> 6, l = q & 63;
#pragma unroll
    for (int t = 0; t < kBkMaxM; ++t)
        if (t == m && lane == l) a[t] = v;
}

}  // namespace

// rows copied into the wider rows of `work`, the right-hand side into x,
// status 0
__global__ void bk_init_kernel(const int32_t *recs, int n, int hbw, const double *band, const double *rhs,
                               double *work, double *x, double *status) {
    const BkSys s = bk_sys(recs, n, hbw, work, x, status);
    const int r = recs ? recs[blockIdx.y] : 0;
    const int W = 2 * hbw + 1, V = 3 * hbw + 1;
    band += (size_t)r * n * W;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < (int64_t)n * V) {
        const int i = (int)(q / V), o = (int)(q % V);
        s.work[q] = o < W ? band[(size_t)i * W + o] : 0.0;
    }
    if (rhs && q < n) s.x[q] = rhs[(size_t)r * n + q];
    if (q == 0) s.status[0] = 0.0;
}

// Panel k0 .. k0 + kb - 1 of every listed system, then its swaps and pivot
// rows over the trailing columns and the right-hand side.
__global__ __launch_bounds__(kBkThreads) void bk_panel_kernel(const int32_t *recs, int n, int hbw, int k0,
                                                              double *work, double *x, double *status) {
    const BkSys s = bk_sys(recs, n, hbw, work, x, status);
    if (s.status[0] == 2.0) return;  // singular earlier: uniform exit
    const int V = 3 * hbw + 1;
    auto at = [&](int i, int j) -> double & { return s.work[(size_t)i * V + (j - i + hbw)]; };
    __shared__ double s_l[2][64 * kBkMaxM];
    __shared__ double s_l11[kBkNb][kBkNb];
    __shared__ int s_p[kBkNb];
    __shared__ int s_swapped;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int kb = min(kBkNb, n - k0);
    const int r_end = min(k0 + kb - 1 + hbw, n - 1);
    const int col = k0 + w;
    if (tid == 0) s_swapped = 0;
    double a[kBkMaxM];
#pragma unroll
    for (int m = 0; m < kBkMaxM; ++m) {
        const int i = k0 + lane + 64 * m;
        a[m] = (w < kb && i <= r_end && i - col <= hbw && col - i <= 2 * hbw) ? at(i, col) : 0.0;
    }
    __syncthreads();
    for (int jj = 0; jj < kb; ++jj) {
        const int k = k0 + jj, kmax = min(k + hbw, n - 1), buf = jj & 1;
        if (w == jj) {
            double best = -1.0;
            int bi = 0x7fffffff;
#pragma unroll
            for (int m = 0; m < kBkMaxM; ++m) {
                const int i = k0 + lane + 64 * m;
                if (i >= k && i <= kmax) {
                    const double v = fabs(a[m]);
                    if (v > best) {  // rows ascend within a lane: the first of equals stays
                        best = v;
                        bi = i;
                    }
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ov > best || (ov == best && oi < bi)) {
                    best = ov;
                    bi = oi;
                }
            }
            if (!(best > kBkPivotTol)) {
                if (lane == 0) s_p[jj] = -1;
            } else {
                const int p = bi;
                const double vk = bk_get(a, k - k0), vp = bk_get(a, p - k0);
                if (p != k) {
                    bk_set(a, k - k0, vp, lane);
                    bk_set(a, p - k0, vk, lane);
                }
                const double diag = vp;
#pragma unroll
                for (int m = 0; m < kBkMaxM; ++m) {
                    const int i = k0 + lane + 64 * m;
                    if (i > k && i <= kmax) {
                        a[m] = a[m] / diag;
                        s_l[buf][i - k0] = a[m];
                        at(i, k) = a[m];  // kept where it was formed (the interleaved-swap convention)
                    }
                }
                if (lane == 0) {
                    s_p[jj] = p;
                    if (p != k) s_swapped = 1;
                }
            }
        }
        __syncthreads();
        const int p = s_p[jj];
        if (p < 0) {
            if (tid == 0) s.status[0] = 2.0;
            return;  // uniform: every thread read the same s_p
        }
        if (w != jj && w < kb && p != k) {  // earlier columns too: their multipliers follow the rows
            const double vk = bk_get(a, k - k0), vp = bk_get(a, p - k0);
            bk_set(a, k - k0, vp, lane);
            bk_set(a, p - k0, vk, lane);
        }
        if (w > jj && w < kb) {
            const double u = bk_get(a, k - k0);
#pragma unroll
            for (int m = 0; m < kBkMaxM; ++m) {
                const int i = k0 + lane + 64 * m;
                if (i > k && i <= kmax) a[m] = a[m] - s_l[buf][i - k0] * u;
            }
        }
    }
    // U rows of the panel back to memory (the later swaps never reach them);
    // the multipliers as the rows finally stand, for the trailing work
#pragma unroll
    for (int m = 0; m < kBkMaxM; ++m) {
        const int i = k0 + lane + 64 * m;
        if (w < kb && i <= col && col - i <= 2 * hbw) at(i, col) = a[m];
        if (w < kb && i > col && i <= r_end) s.lt[(size_t)w * (hbw + kBkNb) + (i - k0)] = a[m];
    }
    if (w < kb && lane < kBkNb) s_l11[lane][w] = (lane < kb && lane > w) ? a[0] : 0.0;
    if (tid < kb) s.piv[k0 + tid] = s_p[tid];
    __syncthreads();
    if (tid == 0 && s_swapped && s.status[0] == 0.0) s.status[0] = 1.0;
    // trailing columns k0 + kb .. jmax and the right-hand side (c == nc)
    const int j0 = k0 + kb, jmax = min(k0 + kb - 1 + 2 * hbw, n - 1);
    const int nc = jmax >= j0 ? jmax - j0 + 1 : 0;
    const int ncr = nc + (s.x ? 1 : 0);  // the right-hand side is column nc
    for (int c = tid; c < ncr; c += kBkThreads) {
        const bool rhs = c == nc;
        const int j = j0 + c;
        for (int q = 0; q < kb; ++q) {  // the swaps in order (row k holds columns up to k + 2 hbw)
            const int k = k0 + q, p = s_p[q];
            if (p == k) continue;
            if (rhs) {
                const double t = s.x[k];
                s.x[k] = s.x[p];
                s.x[p] = t;
            } else if (j <= k + 2 * hbw) {
                const double t = at(k, j);
                at(k, j) = at(p, j);
                at(p, j) = t;
            }
        }
        double u[kBkNb];
#pragma unroll
        for (int q = 0; q < kBkNb; ++q) {
            const int k = k0 + q;
            u[q] = 0.0;
            if (q < kb && (rhs || j <= k + 2 * hbw)) {
                double t = rhs ? s.x[k] : at(k, j);
#pragma unroll
                for (int q2 = 0; q2 < kBkNb; ++q2)
                    if (q2 < q && (rhs || j <= k0 + q2 + 2 * hbw)) t = t - s_l11[q][q2] * u[q2];
                u[q] = t;
                if (rhs)
                    s.x[k] = t;
                else
                    at(k, j) = t;
            }
        }
    }
}

// a_ij -= sum over the panel's k (ascending) of l_ik u_kj, rows k0 + kb ..
// r_end, columns k0 + kb .. jmax and the right-hand side; 64 x 64 tiles
__global__ __launch_bounds__(256) void bk_update_kernel(const int32_t *recs, int n, int hbw, int k0, double *work,
                                                        double *x, double *status) {
    const BkSys s = bk_sys(recs, n, hbw, work, x, status);
    if (s.status[0] == 2.0) return;
    const int V = 3 * hbw + 1;
    auto at = [&](int i, int j) -> double & { return s.work[(size_t)i * V + (j - i + hbw)]; };
    const int kb = min(kBkNb, n - k0);
    const int i0 = k0 + kb, r_end = min(k0 + kb - 1 + hbw, n - 1);
    const int j0 = k0 + kb, jmax = min(k0 + kb - 1 + 2 * hbw, n - 1);
    const bool has_rhs = s.x != nullptr;
    const int nr = r_end - i0 + 1, nc = (jmax >= j0 ? jmax - j0 + 1 : 0) + (has_rhs ? 1 : 0);
    if (nr <= 0 || nc <= 0) return;
    const int tiles_c = (nc + kBkTile - 1) / kBkTile;
    const int ti = blockIdx.x / tiles_c, tj = blockIdx.x % tiles_c;
    __shared__ double s_lt[kBkTile][kBkNb + 1], s_ut[kBkNb][kBkTile];
    const int tid = threadIdx.x;
    for (int e = tid; e < kBkTile * kBkNb; e += 256) {
        const int r = e / kBkNb, q = e % kBkNb;
        const int i = i0 + ti * kBkTile + r;
        s_lt[r][q] = (q < kb && i <= r_end) ? s.lt[(size_t)q * (hbw + kBkNb) + (i - k0)] : 0.0;
    }
    for (int e = tid; e < kBkTile * kBkNb; e += 256) {
        const int q = e / kBkTile, c = e % kBkTile;
        const int cc = tj * kBkTile + c, j = j0 + cc, k = k0 + q;
        double v = 0.0;
        if (q < kb && cc < nc) {
            if (has_rhs && cc == nc - 1)
                v = s.x[k];
            else if (j <= k + 2 * hbw)
                v = at(k, j);
        }
        s_ut[q][c] = v;
    }
    __syncthreads();
    const int tr = (tid / 16) * 4, tc = (tid % 16) * 4;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int cc = tj * kBkTile + tc + b;
        if (cc >= nc) continue;
        const bool rhs = has_rhs && cc == nc - 1;
        const int j = j0 + cc;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int i = i0 + ti * kBkTile + tr + a;
            if (i > r_end) continue;
            double t = rhs ? s.x[i] : at(i, j);
#pragma unroll
            for (int q = 0; q < kBkNb; ++q)
                if (q < kb) t = t - s_lt[tr + a][q] * s_ut[q][tc + b];
            if (rhs)
                s.x[i] = t;
            else
                at(i, j) = t;
        }
    }
}

// x_i = (y_i - sum_{j > i} u_ij x_j) / u_ii by blocks of 64 rows from the
// bottom; U row i at work columns i .. i + 2 hbw
__global__ __launch_bounds__(kBkThreads) void bk_back_kernel(const int32_t *recs, int n, int hbw, double *work,
                                                             double *x, double *status) {
    const BkSys s = bk_sys(recs, n, hbw, work, x, status);
    if (s.status[0] == 2.0) return;
    const int V = 3 * hbw + 1;
    auto at = [&](int i, int j) -> double { return s.work[(size_t)i * V + (j - i + hbw)]; };
    __shared__ double s_s[kBkBack];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int r1 = n; r1 > 0; r1 -= kBkBack) {
        const int r0 = max(r1 - kBkBack, 0);
        // each wavefront: rows r0 + w, r0 + w + 16, ...; lanes over the solved unknowns j >= r1
        for (int i = r0 + w; i < r1; i += kBkNb) {
            const int jend = min(i + 2 * hbw, n - 1);
            double part = 0.0;
            for (int j = r1 + lane; j <= jend; j += 64) part += at(i, j) * s.x[j];
            for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
            if (lane == 0) s_s[i - r0] = part;
        }
        __syncthreads();
        if (w == 0) {
            const int q = r0 + lane;
            double t = q < r1 ? s.x[q] - s_s[lane] : 0.0;
            for (int i = r1 - 1; i >= r0; --i) {
                double xi = 0.0;
                if (lane == i - r0) xi = t / at(i, i);
                xi = lane_read(xi, i - r0);
                if (lane == i - r0) s.x[i] = xi;
                if (q < i && i <= q + 2 * hbw) t = t - at(q, i) * xi;
            }
        }
        __syncthreads();
    }
}

// The factors in the other band kernels' formats: U row k at band slots
// k W + (j mod W) for j = k .. k + 2 hbw (0 past n); the multiplier of step
// k for row i at Lmul[k hbw + i - k - 1]; the row swapped with k at ipiv[k].
__global__ void bk_store_kernel(const int32_t *recs, int n, int hbw, double *work, double *status, double *band,
                                double *Lmul, int32_t *ipiv) {
    const BkSys s = bk_sys(recs, n, hbw, work, nullptr, status);
    if (s.status[0] == 2.0) return;
    const int r = recs ? recs[blockIdx.y] : 0;
    const int W = 2 * hbw + 1, V = 3 * hbw + 1;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)n * W) return;
    const int k = (int)(q / W), o = (int)(q % W);
    const int j = k + o;  // column of slot o of U row k
    band[(size_t)r * n * W + (size_t)k * W + j % W] = j < n ? s.work[(size_t)k * V + (j - k + hbw)] : 0.0;
    if (Lmul && o < hbw) {
        const int i = k + 1 + o;
        Lmul[(size_t)r * n * (hbw > 0 ? hbw : 1) + (size_t)k * hbw + o] = i < n ? s.work[(size_t)i * V + (k - i + hbw)] : 0.0;
    }
    if (ipiv && o == 0) ipiv[(size_t)r * n + k] = s.piv[k];
}

bool band_blk_fits(int n, int hbw) { return n > 0 && hbw >= 0 && hbw + kBkNb <= 64 * kBkMaxM; }

size_t band_blk_work_doubles(int n, int hbw) {
    return (size_t)n * (3 * (size_t)hbw + 1) + ((size_t)n + 1) / 2 + 1 + (size_t)kBkNb * (hbw + kBkNb);
}

// Factor and solve `count` systems (recs: their record indices; nullptr for
// one system at the given pointers) on the context's stream.
// d_rhs == nullptr: factor only; d_L: keep the factors (U in the band, L,
// ipiv) for later solves.
int launch_band_blk(sgpu_ctx *ctx, int n, int hbw, double *d_band, const double *d_rhs, double *d_x,
                    double *d_status, double *d_L, int32_t *d_ipiv, double *d_work, const int32_t *recs, int count) {
    if (!band_blk_fits(n, hbw))
        return fail(ctx, SGPU_ERR_ARG, "half bandwidth %d exceeds the blocked band solve (%d)", hbw,
                    64 * kBkMaxM - kBkNb);
    if (!d_work || (d_rhs && !d_x)) return fail(ctx, SGPU_ERR_ARG, "the blocked band solve needs work (and x)");
    if (!d_rhs) d_x = nullptr;
    hipStream_t st = ctx->stream;
    const int64_t cells = (int64_t)n * (3 * hbw + 1);
    hipLaunchKernelGGL(bk_init_kernel, dim3((unsigned)((cells + 255) / 256), count), dim3(256), 0, st, recs, n, hbw,
                       d_band, d_rhs, d_work, d_x, d_status);
    for (int k0 = 0; k0 < n; k0 += kBkNb) {
        hipLaunchKernelGGL(bk_panel_kernel, dim3(1, count), dim3(kBkThreads), 0, st, recs, n, hbw, k0, d_work, d_x,
                           d_status);
        const int kb = std::min(kBkNb, n - k0);
        const int nr = std::min(k0 + kb - 1 + hbw, n - 1) - (k0 + kb) + 1;
        if (nr > 0) {
            const int jmax = std::min(k0 + kb - 1 + 2 * hbw, n - 1);
            const int nc = (jmax >= k0 + kb ? jmax - (k0 + kb) + 1 : 0) + 1;
            const int tiles = ((nr + kBkTile - 1) / kBkTile) * ((nc + kBkTile - 1) / kBkTile);
            hipLaunchKernelGGL(bk_update_kernel, dim3(tiles, count), dim3(256), 0, st, recs, n, hbw, k0, d_work, d_x,
                               d_status);
        }
    }
    if (d_x)
        hipLaunchKernelGGL(bk_back_kernel, dim3(1, count), dim3(kBkThreads), 0, st, recs, n, hbw, d_work, d_x,
                           d_status);
    if (d_L || !d_x) {
        const int64_t slots = (int64_t)n * (2 * hbw + 1);
        hipLaunchKernelGGL(bk_store_kernel, dim3((unsigned)((slots + 255) / 256), count), dim3(256), 0, st, recs, n,
                           hbw, d_work, d_status, d_band, d_L, d_ipiv);
    }
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

}  // namespace sgpu
