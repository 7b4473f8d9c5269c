// Transformation-handler constraint collapses on gfx950 (rigidDiaphragm,
// equalDOF): K_hat = T^T K T (_collapse_matrix_by_mpc,
// solver/run_case/helpers.mojo:3345-3377), F_hat = T^T F
// (_collapse_vector_by_mpc, :3288-3301) and the slave-DOF update
// (_enforce_mpc_values, :3380-3397), where DOF i expands as
// u_i = sum_t coeff[t] u[dof_pool[t]].
//
// The reference scatters: for every (i, j) with K[i][j] != 0 it adds
// c_i K[i][j] c_j to each (retained i, retained j) pair of their terms.  Here
// every output entry gathers instead: the host builds, for each DOF r, the
// list of (i, c) whose expansion uses r, in ascending i; one thread per
// output (r, s) walks the lists of r and s in that order, so every output
// sees the same additions in the same order as the reference's loops (no
// atomics, the same bits).  HBM-bound: the collapse reads K and writes K_hat.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/strut_gpu.h"
#include "ctx.h"

namespace sgpu {

struct MpcDev {
    int n;
    const int32_t *row_offsets, *dof_pool;
    const double *coeff_pool;
    const int32_t *inv_offsets, *inv_dof;  // per DOF r: the (i, coeff) whose expansion uses r, ascending i
    const double *inv_coeff;
    const int32_t *slave, *constrained;
};

struct MpcHost {
    int n = 0;
    MpcDev dev{};
    double *d_tmp = nullptr;  // slave values of the enforce pass
    // staging of the host-buffer entry points
    double *d_a = nullptr, *d_b = nullptr;
    int64_t cap = 0;
};

void mpc_host_free(MpcHost *h) {
    if (!h) return;
    for (void *p : {(void *)h->d_tmp, (void *)h->d_a, (void *)h->d_b})
        if (p) (void)hipFree(p);
    delete h;
}

__global__ __launch_bounds__(256) void mpc_collapse_matrix_kernel(MpcDev M, const double *__restrict__ K,
                                                                  double *__restrict__ out) {
    const int64_t n = M.n;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const int r = (int)(t / n), s = (int)(t % n);
    double v = 0.0;
    const int a1 = M.inv_offsets[r + 1], b0 = M.inv_offsets[s], b1 = M.inv_offsets[s + 1];
    for (int a = M.inv_offsets[r]; a < a1; ++a) {
        const int64_t i = M.inv_dof[a];
        const double ci = M.inv_coeff[a];
        for (int b = b0; b < b1; ++b) {
            const double value = K[i * n + M.inv_dof[b]];
            if (value == 0.0) continue;  // helpers.mojo:3359-3360
            v += ci * value * M.inv_coeff[b];
        }
    }
    out[t] = v;
}

__global__ __launch_bounds__(256) void mpc_collapse_vector_kernel(MpcDev M, const double *__restrict__ F,
                                                                  double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M.n) return;
    double v = 0.0;
    for (int a = M.inv_offsets[r]; a < M.inv_offsets[r + 1]; ++a) v += M.inv_coeff[a] * F[M.inv_dof[a]];
    out[r] = v;
}

// Slave values as the reference's sequential loop sees them: a retained DOF
// d fixed by a single-point constraint reads 0 when d < i (already zeroed).
__global__ __launch_bounds__(256) void mpc_enforce_slaves_kernel(MpcDev M, const double *__restrict__ u,
                                                                 double *__restrict__ tmp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M.n || !M.slave[i]) return;
    double v = 0.0;
    for (int t = M.row_offsets[i]; t < M.row_offsets[i + 1]; ++t) {
        const int d = M.dof_pool[t];
        const double ud = (M.constrained[d] && !M.slave[d] && d < i) ? 0.0 : u[d];
        v += M.coeff_pool[t] * ud;
    }
    tmp[i] = v;
}

__global__ __launch_bounds__(256) void mpc_enforce_write_kernel(MpcDev M, double *__restrict__ u,
                                                                const double *__restrict__ tmp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M.n) return;
    if (M.slave[i])
        u[i] = tmp[i];
    else if (M.constrained[i])
        u[i] = 0.0;
}

namespace {

MpcHost *mpc_of(sgpu_ctx *ctx) { return ctx->mpc; }

int ensure_stage(sgpu_ctx *ctx, MpcHost &h, int64_t count) {
    if (count <= h.cap) return SGPU_OK;
    for (double **p : {&h.d_a, &h.d_b})
        if (*p) {
            HIP_OK(ctx, hipFree(*p));
            *p = nullptr;
        }
    h.cap = 0;
    HIP_OK(ctx, hipMalloc((void **)&h.d_a, sizeof(double) * count));
    HIP_OK(ctx, hipMalloc((void **)&h.d_b, sizeof(double) * count));
    h.cap = count;
    return SGPU_OK;
}

}  // namespace
}  // namespace sgpu

using namespace sgpu;

int sgpu_mpc_set(sgpu_ctx *ctx, int n, const int32_t *row_offsets, const int32_t *dof_pool, const double *coeff_pool,
                 const int32_t *constrained, const int32_t *slave) {
    if (!ctx || n <= 0 || !row_offsets || !constrained || !slave) return SGPU_ERR_ARG;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const int pool = row_offsets[n];
    if (row_offsets[0] != 0 || pool < 0 || (pool > 0 && (!dof_pool || !coeff_pool)))
        return fail(ctx, SGPU_ERR_ARG, "bad MPC row offsets");
    std::vector<std::vector<std::pair<int32_t, double>>> inv(n);
    for (int i = 0; i < n; ++i) {
        if (row_offsets[i + 1] < row_offsets[i]) return fail(ctx, SGPU_ERR_ARG, "bad MPC row offsets");
        for (int t = row_offsets[i]; t < row_offsets[i + 1]; ++t) {
            const int d = dof_pool[t];
            if (d < 0 || d >= n) return fail(ctx, SGPU_ERR_ARG, "MPC term %d: DOF %d out of range", t, d);
            for (int u = row_offsets[i]; u < t; ++u)
                if (dof_pool[u] == d) return fail(ctx, SGPU_ERR_ARG, "MPC row %d names DOF %d twice", i, d);
            // one pass over the slaves reads only retained values (the
            // loader refuses a retained node that is itself constrained)
            if (slave[i] && slave[d]) return fail(ctx, SGPU_ERR_ARG, "MPC slave DOF %d retains slave DOF %d", i, d);
            inv[d].push_back({i, coeff_pool[t]});
        }
    }
    std::vector<int32_t> io(n + 1, 0), idof;
    std::vector<double> ic;
    for (int r = 0; r < n; ++r) {
        io[r] = (int32_t)idof.size();
        for (auto &e : inv[r]) {
            idof.push_back(e.first);
            ic.push_back(e.second);
        }
    }
    io[n] = (int32_t)idof.size();
    if (!ctx->mpc) ctx->mpc = new MpcHost();
    MpcHost &h = *ctx->mpc;
    int rc;
    MpcDev &M = h.dev;
    M = MpcDev{};
    M.n = n;
    int32_t *p_ro, *p_dp, *p_io, *p_idof, *p_sl, *p_co;
    double *p_cp, *p_ic;
    if ((rc = upload(ctx, &p_ro, std::vector<int32_t>(row_offsets, row_offsets + n + 1)))) return rc;
    if ((rc = upload(ctx, &p_dp, std::vector<int32_t>(dof_pool, dof_pool + pool)))) return rc;
    if ((rc = upload(ctx, &p_cp, std::vector<double>(coeff_pool, coeff_pool + pool)))) return rc;
    if ((rc = upload(ctx, &p_io, io))) return rc;
    if ((rc = upload(ctx, &p_idof, idof))) return rc;
    if ((rc = upload(ctx, &p_ic, ic))) return rc;
    if ((rc = upload(ctx, &p_sl, std::vector<int32_t>(slave, slave + n)))) return rc;
    if ((rc = upload(ctx, &p_co, std::vector<int32_t>(constrained, constrained + n)))) return rc;
    M.row_offsets = p_ro;
    M.dof_pool = p_dp;
    M.coeff_pool = p_cp;
    M.inv_offsets = p_io;
    M.inv_dof = p_idof;
    M.inv_coeff = p_ic;
    M.slave = p_sl;
    M.constrained = p_co;
    if (h.d_tmp) HIP_OK(ctx, hipFree(h.d_tmp));
    h.d_tmp = nullptr;
    HIP_OK(ctx, hipMalloc((void **)&h.d_tmp, sizeof(double) * n));
    h.n = n;
    return SGPU_OK;
}

int sgpu_mpc_collapse_matrix_device(sgpu_ctx *ctx, const double *d_K, double *d_out) {
    if (!ctx || !d_K || !d_out) return SGPU_ERR_ARG;
    MpcHost *h = mpc_of(ctx);
    if (!h) return fail(ctx, SGPU_ERR_STATE, "sgpu_mpc_set has not run");
    const int64_t total = (int64_t)h->n * h->n;
    ProfScope ps(ctx, SGPU_KERNEL_MPC);
    hipLaunchKernelGGL(mpc_collapse_matrix_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream,
                       h->dev, d_K, d_out);
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

int sgpu_mpc_collapse_matrix(sgpu_ctx *ctx, const double *K, double *out) {
    if (!ctx || !K || !out) return SGPU_ERR_ARG;
    MpcHost *h = mpc_of(ctx);
    if (!h) return fail(ctx, SGPU_ERR_STATE, "sgpu_mpc_set has not run");
    const int64_t total = (int64_t)h->n * h->n;
    int rc = ensure_stage(ctx, *h, total);
    if (rc) return rc;
    HIP_OK(ctx, hipMemcpyAsync(h->d_a, K, sizeof(double) * total, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = sgpu_mpc_collapse_matrix_device(ctx, h->d_a, h->d_b))) return rc;
    HIP_OK(ctx, hipMemcpyAsync(out, h->d_b, sizeof(double) * total, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

int sgpu_mpc_collapse_vector(sgpu_ctx *ctx, const double *F, double *out) {
    if (!ctx || !F || !out) return SGPU_ERR_ARG;
    MpcHost *h = mpc_of(ctx);
    if (!h) return fail(ctx, SGPU_ERR_STATE, "sgpu_mpc_set has not run");
    int rc = ensure_stage(ctx, *h, h->n);
    if (rc) return rc;
    HIP_OK(ctx, hipMemcpyAsync(h->d_a, F, sizeof(double) * h->n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(mpc_collapse_vector_kernel, dim3((h->n + 255) / 256), dim3(256), 0, ctx->stream, h->dev, h->d_a,
                       h->d_b);
    HIP_OK(ctx, hipGetLastError());
    HIP_OK(ctx, hipMemcpyAsync(out, h->d_b, sizeof(double) * h->n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

int sgpu_mpc_enforce(sgpu_ctx *ctx, double *u) {
    if (!ctx || !u) return SGPU_ERR_ARG;
    MpcHost *h = mpc_of(ctx);
    if (!h) return fail(ctx, SGPU_ERR_STATE, "sgpu_mpc_set has not run");
    int rc = ensure_stage(ctx, *h, h->n);
    if (rc) return rc;
    HIP_OK(ctx, hipMemcpyAsync(h->d_a, u, sizeof(double) * h->n, hipMemcpyHostToDevice, ctx->stream));
    const dim3 grid((h->n + 255) / 256);
    hipLaunchKernelGGL(mpc_enforce_slaves_kernel, grid, dim3(256), 0, ctx->stream, h->dev, h->d_a, h->d_tmp);
    hipLaunchKernelGGL(mpc_enforce_write_kernel, grid, dim3(256), 0, ctx->stream, h->dev, h->d_a, h->d_tmp);
    HIP_OK(ctx, hipGetLastError());
    HIP_OK(ctx, hipMemcpyAsync(u, h->d_a, sizeof(double) * h->n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}
