// Modal analysis of a 3D frame on gfx950: run_modal_eigen
// (solver/run_case/analysis/modal_eigen.mojo:143-362) with
// jacobi_symmetric_eigen (solver/eigen.mojo:15-79), on the frame a context
// holds (sgpu_frame3d_create).
//
// The reference forms dense K, M, L (M = L L^T), L^-1, A = L^-1 K L^-T and
// V, then runs classical Jacobi: every pass scans the whole upper triangle
// for its largest entry and rotates one pair.  Here:
//   * K_ff is gathered on the device, one workgroup per free row, from the
//     beam-columns' collapsed initial tangents and the shells' collapsed
//     stiffness (the free-DOF gather lists of the transient); rows are
//     filled entry by entry, so the result does not depend on timing;
//   * the reduced mass couples at most the few retained DOFs of one
//     diaphragm, so M_ff is block diagonal with tiny blocks: its Cholesky
//     factor and inverse are those blocks (the reference's dense products
//     only add zeros outside them), formed on the host with the reference's
//     loops and uploaded as sparse rows;
//   * A = L^-1 K L^-T, one thread per entry, sums over the blocks in the
//     reference's order (its first product skips the zero factors, its
//     second adds zero products, which changes no sum);
//   * each Jacobi pass is three launches and no host round trip: a
//     grid-wide search of the upper triangle for the largest magnitude (the
//     first in row-major order among equals, as the reference's strict
//     comparison keeps), a one-workgroup final reduction that also latches
//     convergence (largest magnitude at or below tol), and the rotation of
//     rows and columns p, q of A and columns p, q of V;
//   * the eigenvalues (A's diagonal) come back to the host, whose
//     selection sort (the reference's, run only for the positions asked
//     for) picks the modes; their V columns come back for the block
//     back-substitution and the expansion to every DOF.
// Dense A and V need 16 n^2 bytes of device memory (33,480 free DOFs: 18 GB).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/strut_gpu.h"
#include "ctx.h"
#include "frame3d_host.h"

namespace sgpu {

namespace {

constexpr int kSearchBlocks = 1024;
constexpr int kSearchThreads = 256;

// K_ff row r: every gather entry of r (beam-column (element, position)
// entries, then shell entries) adds its collapsed row into the dense row
__global__ __launch_bounds__(64) void modal_gather_k_kernel(int n, int n_el, const int32_t *dof_off,
                                                            const int32_t *dof_src, const int32_t *ret,
                                                            const int32_t *ret_n, const double *kinit_hat,
                                                            const int32_t *sh_ret, const int32_t *sh_ret_n,
                                                            const double *sh_khat, double *K) {
    const int r = blockIdx.x, lane = threadIdx.x;
    double *row = K + (size_t)r * n;
    for (int t = dof_off[r]; t < dof_off[r + 1]; ++t) {
        const int s = dof_src[t];
        if (s >= 0) {
            const int e = s / 12, a = s % 12;
            if (lane < ret_n[e]) row[ret[e * 12 + lane]] += kinit_hat[(size_t)e * 144 + a * 12 + lane];
        } else {
            const int q = -s - 1, e = q / 24, a = q % 24;
            if (lane < sh_ret_n[e]) row[sh_ret[e * 24 + lane]] += sh_khat[(size_t)e * 576 + a * 24 + lane];
        }
        __syncthreads();  // the next entry may add to the same columns
    }
}

__global__ void modal_diag_kernel(int n, const double *A, double *d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = A[(size_t)i * n + i];
}

// A_ij = sum_{k in row j of L^-1} (sum_{a in row i of L^-1} Li_ia K_ak) Li_jk
__global__ void modal_form_a_kernel(int n, const int32_t *li_off, const int32_t *li_col, const double *li_val,
                                    const double *K, double *A) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    double value = 0.0;
    for (int tk = li_off[j]; tk < li_off[j + 1]; ++tk) {
        const int k = li_col[tk];
        double p = 0.0;
        for (int ta = li_off[i]; ta < li_off[i + 1]; ++ta) p += li_val[ta] * K[(size_t)li_col[ta] * n + k];
        value += p * li_val[tk];
    }
    A[idx] = value;
}

__global__ void modal_identity_kernel(int n, double *V) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n) return;
    V[idx] = (idx / n == idx % n) ? 1.0 : 0.0;
}

// larger magnitude, then the smaller row-major key
__device__ __forceinline__ void better(double &v, int64_t &k, double ov, int64_t ok) {
    if (ov > v || (ov == v && ok < k)) {
        v = ov;
        k = ok;
    }
}

// state[0] largest magnitude, [1] p, [2] q, [3] app, [4] aqq, [5] apq,
// [6] done (1 once the largest magnitude is at or below tol), [7] passes made
__global__ __launch_bounds__(kSearchThreads) void modal_search_kernel(int n, const double *A, const double *state,
                                                                      double *part_v, int64_t *part_k) {
    __shared__ double sv[kSearchThreads];
    __shared__ int64_t sk[kSearchThreads];
    const int tid = threadIdx.x;
    double v = -1.0;
    int64_t k = INT64_MAX;
    if (state[6] == 0.0)
        for (int i = blockIdx.x; i < n - 1; i += gridDim.x) {
            const double *row = A + (size_t)i * n;
            for (int j = i + 1 + tid; j < n; j += kSearchThreads) better(v, k, fabs(row[j]), (int64_t)i * n + j);
        }
    sv[tid] = v;
    sk[tid] = k;
    __syncthreads();
    for (int w = kSearchThreads / 2; w > 0; w >>= 1) {
        if (tid < w) {
            double a = sv[tid];
            int64_t b = sk[tid];
            better(a, b, sv[tid + w], sk[tid + w]);
            sv[tid] = a;
            sk[tid] = b;
        }
        __syncthreads();
    }
    if (tid == 0) {
        part_v[blockIdx.x] = sv[0];
        part_k[blockIdx.x] = sk[0];
    }
}

__global__ __launch_bounds__(1024) void modal_pick_kernel(int n, int parts, const double *A, const double *part_v,
                                                          const int64_t *part_k, double tol, double *state) {
    __shared__ double sv[1024];
    __shared__ int64_t sk[1024];
    const int tid = threadIdx.x;
    if (state[6] != 0.0) return;  // converged earlier: uniform exit
    double v = -1.0;
    int64_t k = INT64_MAX;
    for (int b = tid; b < parts; b += 1024) better(v, k, part_v[b], part_k[b]);
    sv[tid] = v;
    sk[tid] = k;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (tid < w) {
            double a = sv[tid];
            int64_t b = sk[tid];
            better(a, b, sv[tid + w], sk[tid + w]);
            sv[tid] = a;
            sk[tid] = b;
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int p = (int)(sk[0] / n), q = (int)(sk[0] % n);
        state[0] = sv[0];
        state[1] = p;
        state[2] = q;
        state[3] = A[(size_t)p * n + p];
        state[4] = A[(size_t)q * n + q];
        state[5] = A[(size_t)p * n + q];
        if (sv[0] <= tol) state[6] = 1.0;  // eigen.mojo:45-46
    }
}

// eigen.mojo:48-74: one rotation of rows / columns p, q
__global__ void modal_rotate_kernel(int n, double *A, double *V, double *state) {
    if (state[6] != 0.0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = (int)state[1], q = (int)state[2];
    const double app = state[3], aqq = state[4], apq = state[5];
    const double phi = 0.5 * atan2(2.0 * apq, aqq - app);
    const double cs = cos(phi), sn = sin(phi);
    if (i >= n) return;
    if (i != p && i != q) {
        const double aip = A[(size_t)i * n + p], aiq = A[(size_t)i * n + q];
        const double np_ = cs * aip - sn * aiq, nq = sn * aip + cs * aiq;
        A[(size_t)i * n + p] = np_;
        A[(size_t)p * n + i] = np_;
        A[(size_t)i * n + q] = nq;
        A[(size_t)q * n + i] = nq;
    } else if (i == p) {
        A[(size_t)p * n + p] = cs * cs * app - 2.0 * sn * cs * apq + sn * sn * aqq;
        A[(size_t)q * n + q] = sn * sn * app + 2.0 * sn * cs * apq + cs * cs * aqq;
        A[(size_t)p * n + q] = 0.0;
        A[(size_t)q * n + p] = 0.0;
        state[7] += 1.0;
    }
    const double vip = V[(size_t)i * n + p], viq = V[(size_t)i * n + q];
    V[(size_t)i * n + p] = cs * vip - sn * viq;
    V[(size_t)i * n + q] = sn * vip + cs * viq;
}

// columns cols[m] of V into out[m n ..]
__global__ void modal_columns_kernel(int n, int m, const int32_t *cols, const double *V, double *out) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * m) return;
    const int k = (int)(idx / n), i = (int)(idx % n);
    out[idx] = V[(size_t)i * n + cols[k]];
}

inline dim3 blocks_for(int64_t count, int b = 256) { return dim3((unsigned)((count + b - 1) / b)); }

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace
}  // namespace sgpu

using namespace sgpu;

int sgpu_frame3d_modal(sgpu_ctx *ctx, int num_modes, int max_rotations, double tol, double *eigenvalues,
                       double *modes, int32_t *rotations) {
    if (!ctx || num_modes < 0 || !eigenvalues || !modes) return SGPU_ERR_ARG;
    Frame3dHost *hp = ctx->frame3d;
    if (!hp) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame3d_create has not run");
    Frame3dHost &h = *hp;
    if (max_rotations <= 0) max_rotations = 100;  // jacobi_symmetric_eigen's max_sweeps
    const int n = h.n, total = h.total;
    if (num_modes > n) return fail(ctx, SGPU_ERR_ARG, "modal_eigen num_modes exceeds free dof count");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    int rc;
    if ((rc = frame3d_initialize(ctx, h))) return rc;
    hipStream_t st = ctx->stream;
    const size_t nn = (size_t)n * n;
    DevBuf bK, bA, bV, bPart, bState, bDiag, bLi, bCols, bOut;
    if (hipMalloc(&bK.p, sizeof(double) * nn) != hipSuccess || hipMalloc(&bA.p, sizeof(double) * nn) != hipSuccess ||
        hipMalloc(&bV.p, sizeof(double) * nn) != hipSuccess ||
        hipMalloc(&bPart.p, (sizeof(double) + sizeof(int64_t)) * kSearchBlocks) != hipSuccess ||
        hipMalloc(&bState.p, sizeof(double) * 8) != hipSuccess || hipMalloc(&bDiag.p, sizeof(double) * n) != hipSuccess)
        return fail(ctx, SGPU_ERR_HIP, "modal analysis of %d free DOFs: %zu bytes of device memory", n,
                    3 * nn * sizeof(double));
    double *K = (double *)bK.p, *A = (double *)bA.p, *V = (double *)bV.p, *dg = (double *)bDiag.p;
    double *part_v = (double *)bPart.p, *state = (double *)bState.p;
    int64_t *part_k = (int64_t *)(part_v + kSearchBlocks);
    // K_ff at rest (record 0's initial tangents, the shells' stiffness)
    HIP_OK(ctx, hipMemsetAsync(K, 0, sizeof(double) * nn, st));
    hipLaunchKernelGGL(modal_gather_k_kernel, dim3(n), dim3(64), 0, st, n, h.n_el, h.d_dof_off, h.d_dof_src, h.d_ret,
                       h.d_ret_n, h.d_kinit_hat, h.d_sh_ret, h.d_sh_ret_n, h.d_sh_khat, K);
    hipLaunchKernelGGL(modal_diag_kernel, blocks_for(n), dim3(256), 0, st, n, K, dg);
    HIP_OK(ctx, hipGetLastError());
    std::vector<double> kdiag(n);
    HIP_OK(ctx, hipMemcpyAsync(kdiag.data(), dg, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIP_OK(ctx, hipStreamSynchronize(st));
    // modal_eigen.mojo:239-262: reduced mass with its floors
    std::vector<std::vector<std::pair<int, double>>> mrow(n);
    double max_mass = 0.0;
    bool has_mass = false;
    for (int r = 0; r < n; ++r)
        for (int t = h.h_moff[r]; t < h.h_moff[r + 1]; ++t) {
            mrow[r].push_back({h.h_mcol[t], h.h_mval[t]});
            if (h.h_mval[t] != 0.0) has_mass = true;
            if (h.h_mcol[t] == r && h.h_mval[t] > max_mass) max_mass = h.h_mval[t];
        }
    if (!has_mass) return fail(ctx, SGPU_ERR_ARG, "modal_eigen requires masses on free dofs");
    const double msmall = 1.0e-10;
    double mass_floor = max_mass * 1.0e-12;
    if (mass_floor <= 0.0) mass_floor = 1.0e-12;
    for (int r = 0; r < n; ++r) {
        double mf = 0.0;
        int at = -1;
        for (size_t t = 0; t < mrow[r].size(); ++t)
            if (mrow[r][t].first == r) {
                mf = mrow[r][t].second;
                at = (int)t;
            }
        if (mf == 0.0) mf = kdiag[r] * msmall;
        if (mf <= 0.0) mf = mass_floor;
        if (at < 0) {
            mrow[r].push_back({r, mf});
            std::sort(mrow[r].begin(), mrow[r].end());
        } else {
            mrow[r][at].second = mf;
        }
    }
    // blocks: connected components of the mass pattern
    std::vector<int> parent(n);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    for (int r = 0; r < n; ++r)
        for (auto &pr : mrow[r]) {
            const int a = find(r), b = find(pr.first);
            if (a != b) parent[std::max(a, b)] = std::min(a, b);
        }
    std::vector<std::vector<int>> members(n);
    for (int r = 0; r < n; ++r) members[find(r)].push_back(r);  // ascending within each block
    // per block: the reference's dense Cholesky and inverse over its members
    std::vector<std::vector<std::pair<int, double>>> Lrow(n), Lirow(n);
    for (int b = 0; b < n; ++b) {
        const std::vector<int> &mb = members[b];
        const int s = (int)mb.size();
        if (s == 0) continue;
        std::vector<double> M(s * s, 0.0), L(s * s, 0.0), Li(s * s, 0.0), sol(s);
        for (int x = 0; x < s; ++x)
            for (auto &pr : mrow[mb[x]]) {
                const int y = (int)(std::find(mb.begin(), mb.end(), pr.first) - mb.begin());
                M[x * s + y] = pr.second;
            }
        for (int i = 0; i < s; ++i)  // _cholesky_factor_dense_spd
            for (int j = 0; j <= i; ++j) {
                double value = M[i * s + j];
                for (int k = 0; k < j; ++k) value -= L[i * s + k] * L[j * s + k];
                if (i == j) {
                    if (value <= 1.0e-18) return fail(ctx, SGPU_ERR_ARG, "modal_eigen mass matrix is not positive definite");
                    L[i * s + j] = std::sqrt(value);
                } else {
                    L[i * s + j] = value / L[j * s + j];
                }
            }
        for (int c = 0; c < s; ++c) {  // _solve_lower_triangular per unit column
            for (int i = 0; i < s; ++i) {
                double value = i == c ? 1.0 : 0.0;
                for (int j = 0; j < i; ++j) value -= L[i * s + j] * sol[j];
                sol[i] = value / L[i * s + i];
            }
            for (int r = 0; r < s; ++r) Li[r * s + c] = sol[r];
        }
        for (int x = 0; x < s; ++x)
            for (int y = 0; y < s; ++y) {
                if (L[x * s + y] != 0.0) Lrow[mb[x]].push_back({mb[y], L[x * s + y]});
                if (Li[x * s + y] != 0.0) Lirow[mb[x]].push_back({mb[y], Li[x * s + y]});
            }
    }
    std::vector<int32_t> li_off(n + 1, 0), li_col;
    std::vector<double> li_val;
    for (int r = 0; r < n; ++r) {
        li_off[r] = (int32_t)li_col.size();
        for (auto &pr : Lirow[r]) {
            li_col.push_back(pr.first);
            li_val.push_back(pr.second);
        }
    }
    li_off[n] = (int32_t)li_col.size();
    const size_t li_bytes = sizeof(int32_t) * (n + 1 + li_col.size()) + sizeof(double) * li_val.size();
    if (hipMalloc(&bLi.p, li_bytes + 16) != hipSuccess) return fail(ctx, SGPU_ERR_HIP, "modal: L^-1 upload");
    double *d_li_val = (double *)bLi.p;
    int32_t *d_li_off = (int32_t *)(d_li_val + li_val.size()), *d_li_col = d_li_off + n + 1;
    HIP_OK(ctx, hipMemcpyAsync(d_li_val, li_val.data(), sizeof(double) * li_val.size(), hipMemcpyHostToDevice, st));
    HIP_OK(ctx, hipMemcpyAsync(d_li_off, li_off.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
    HIP_OK(ctx, hipMemcpyAsync(d_li_col, li_col.data(), sizeof(int32_t) * li_col.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(modal_form_a_kernel, blocks_for((int64_t)nn), dim3(256), 0, st, n, d_li_off, d_li_col, d_li_val,
                       K, A);
    hipLaunchKernelGGL(modal_identity_kernel, blocks_for((int64_t)nn), dim3(256), 0, st, n, V);
    HIP_OK(ctx, hipMemsetAsync(state, 0, sizeof(double) * 8, st));
    HIP_OK(ctx, hipGetLastError());
    if (n > 1)
        for (int pass = 0; pass < max_rotations; ++pass) {
            hipLaunchKernelGGL(modal_search_kernel, dim3(kSearchBlocks), dim3(kSearchThreads), 0, st, n, A, state,
                               part_v, part_k);
            hipLaunchKernelGGL(modal_pick_kernel, dim3(1), dim3(1024), 0, st, n, kSearchBlocks, A, part_v, part_k, tol,
                               state);
            hipLaunchKernelGGL(modal_rotate_kernel, blocks_for(n), dim3(256), 0, st, n, A, V, state);
            if ((pass & 63) == 63) {  // stop early once converged
                double done = 0.0;
                HIP_OK(ctx, hipMemcpyAsync(&done, state + 6, sizeof(double), hipMemcpyDeviceToHost, st));
                HIP_OK(ctx, hipStreamSynchronize(st));
                if (done != 0.0) break;
            }
        }
    HIP_OK(ctx, hipGetLastError());
    hipLaunchKernelGGL(modal_diag_kernel, blocks_for(n), dim3(256), 0, st, n, A, dg);
    std::vector<double> eig_raw(n), st8(8);
    HIP_OK(ctx, hipMemcpyAsync(eig_raw.data(), dg, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIP_OK(ctx, hipMemcpyAsync(st8.data(), state, sizeof(double) * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(ctx, hipStreamSynchronize(st));
    if (rotations) *rotations = (int32_t)st8[7];
    // _sorted_indices_ascending, the first num_modes positions
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    for (int i = 0; i < num_modes; ++i) {
        int min_i = i;
        double min_v = eig_raw[order[i]];
        for (int j = i + 1; j < n; ++j)
            if (eig_raw[order[j]] < min_v) {
                min_v = eig_raw[order[j]];
                min_i = j;
            }
        std::swap(order[i], order[min_i]);
    }
    if (num_modes == 0) return SGPU_OK;
    if (hipMalloc(&bCols.p, sizeof(int32_t) * num_modes) != hipSuccess ||
        hipMalloc(&bOut.p, sizeof(double) * (size_t)num_modes * n) != hipSuccess)
        return fail(ctx, SGPU_ERR_HIP, "modal: mode columns");
    HIP_OK(ctx, hipMemcpyAsync(bCols.p, order.data(), sizeof(int32_t) * num_modes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(modal_columns_kernel, blocks_for((int64_t)num_modes * n), dim3(256), 0, st, n, num_modes,
                       (const int32_t *)bCols.p, V, (double *)bOut.p);
    std::vector<double> cols((size_t)num_modes * n);
    HIP_OK(ctx, hipMemcpyAsync(cols.data(), bOut.p, sizeof(double) * cols.size(), hipMemcpyDeviceToHost, st));
    HIP_OK(ctx, hipStreamSynchronize(st));
    std::vector<double> sol(n), full(total);
    for (int m = 0; m < num_modes; ++m) {
        eigenvalues[m] = eig_raw[order[m]];
        const double *v = cols.data() + (size_t)m * n;
        // _solve_upper_from_lower_transpose: x_i = (v_i - sum_{j > i} L_ji x_j) / L_ii,
        // the nonzero L_ji of a block lying in its later members
        for (int i = n - 1; i >= 0; --i) {
            double value = v[i], lii = 1.0;
            const std::vector<int> &mb = members[find(i)];
            for (int j : mb) {
                if (j < i) continue;
                for (auto &pr : Lrow[j])
                    if (pr.first == i) {
                        if (j == i)
                            lii = pr.second;
                        else
                            value -= pr.second * sol[j];
                    }
            }
            sol[i] = value / lii;
        }
        std::fill(full.begin(), full.end(), 0.0);
        for (int i = 0; i < n; ++i) full[h.h_free_list[i]] = sol[i];
        for (int i = 0; i < total; ++i) {  // _enforce_mpc_values
            if (h.h_slave[i]) {
                double value = 0.0;
                for (int t = h.h_ro[i]; t < h.h_ro[i + 1]; ++t) value += h.h_cp[t] * full[h.h_dp[t]];
                full[i] = value;
            } else if (h.h_constrained[i]) {
                full[i] = 0.0;
            }
        }
        std::copy(full.begin(), full.end(), modes + (size_t)m * total);
    }
    return SGPU_OK;
}
