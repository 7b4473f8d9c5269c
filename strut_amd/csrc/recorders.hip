// Device-side recorders over a displacement history that stays in HBM: the
// envelope of each recorded DOF (envelope_node_displacement: min, max and
// largest magnitude, the first recorded step seeding all three --
// _update_envelope, solver/run_case/helpers.mojo:3153-3186) and of each
// drift channel, drift = (u_j - u_i) / (coordinate_j - coordinate_i)
// (_drift_value, helpers.mojo:3113-3142).  Large ensembles read back
// 3 values per channel per instance instead of their histories.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/strut_gpu.h"
#include "ctx.h"

namespace sgpu {

// One workgroup per instance, one thread per channel (recorded DOFs first,
// then drift channels); the steps are walked in order, as the reference
// updates its envelope once per converged step.
__global__ __launch_bounds__(256) void envelope_kernel(const double *hist, const int64_t *hist_base,
                                                       const int32_t *steps, int n_record, int n_drift,
                                                       const int32_t *col_i, const int32_t *col_j, const double *dx,
                                                       double *disp_env, double *drift_env) {
    const int inst = blockIdx.x;
    const double *h = hist + hist_base[inst];
    const int n_steps = steps[inst];
    for (int c = threadIdx.x; c < n_record + n_drift; c += blockDim.x) {
        const bool drift = c >= n_record;
        const int d = c - n_record;
        double mn = 0.0, mx = 0.0, ab = 0.0;
        for (int s = 0; s < n_steps; ++s) {
            const double *row = h + (size_t)s * n_record;
            const double v = drift ? (row[col_j[d]] - row[col_i[d]]) / dx[d] : row[c];
            if (s == 0) {
                mn = v;
                mx = v;
                ab = fabs(v);
                continue;
            }
            if (v < mn) mn = v;
            if (v > mx) mx = v;
            if (fabs(v) > ab) ab = fabs(v);
        }
        double *out = drift ? drift_env + ((size_t)inst * n_drift + d) * 3 : disp_env + ((size_t)inst * n_record + c) * 3;
        out[0] = mn;
        out[1] = mx;
        out[2] = ab;
    }
}

// Envelopes of n_inst histories on the context's stream; results to host.
int envelopes_device(sgpu_ctx *ctx, const double *d_hist, const std::vector<int64_t> &bases,
                     const std::vector<int32_t> &steps, int n_record, int n_drift, const int32_t *col_i,
                     const int32_t *col_j, const double *dx, double *disp_env, double *drift_env) {
    const int n_inst = (int)bases.size();
    if (n_inst == 0) return SGPU_OK;
    if (n_record <= 0) return fail(ctx, SGPU_ERR_ARG, "no recorded DOFs");
    if (n_drift < 0 || (n_drift > 0 && (!col_i || !col_j || !dx))) return fail(ctx, SGPU_ERR_ARG, "bad drift channels");
    for (int d = 0; d < n_drift; ++d) {
        if (col_i[d] < 0 || col_i[d] >= n_record || col_j[d] < 0 || col_j[d] >= n_record)
            return fail(ctx, SGPU_ERR_ARG, "drift channel %d: recorded column out of range", d);
        if (dx[d] == 0.0) return fail(ctx, SGPU_ERR_ARG, "drift channel %d: zero denominator", d);
    }
    // small per-call buffers: tables in, envelopes out
    const size_t n_disp = 3 * (size_t)n_inst * n_record, n_dr = 3 * (size_t)n_inst * n_drift;
    const size_t bytes = sizeof(int64_t) * n_inst + sizeof(int32_t) * (n_inst + 2 * (size_t)n_drift) +
                         sizeof(double) * (n_drift + n_disp + n_dr) + 64;
    char *buf = nullptr;
    HIP_OK(ctx, hipMalloc((void **)&buf, bytes));
    double *d_disp = (double *)buf, *d_drift = d_disp + n_disp, *d_dx = d_drift + n_dr;
    int64_t *d_base = (int64_t *)(d_dx + n_drift);
    int32_t *d_steps = (int32_t *)(d_base + n_inst), *d_ci = d_steps + n_inst, *d_cj = d_ci + n_drift;
    int rc = SGPU_OK;
    if (hipMemcpy(d_base, bases.data(), sizeof(int64_t) * n_inst, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_steps, steps.data(), sizeof(int32_t) * n_inst, hipMemcpyHostToDevice) != hipSuccess ||
        (n_drift > 0 && (hipMemcpy(d_ci, col_i, sizeof(int32_t) * n_drift, hipMemcpyHostToDevice) != hipSuccess ||
                         hipMemcpy(d_cj, col_j, sizeof(int32_t) * n_drift, hipMemcpyHostToDevice) != hipSuccess ||
                         hipMemcpy(d_dx, dx, sizeof(double) * n_drift, hipMemcpyHostToDevice) != hipSuccess))) {
        rc = fail(ctx, SGPU_ERR_HIP, "envelope tables upload failed");
    } else {
        hipLaunchKernelGGL(envelope_kernel, dim3(n_inst), dim3(256), 0, ctx->stream, d_hist, d_base, d_steps, n_record,
                           n_drift, d_ci, d_cj, d_dx, d_disp, d_drift);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(disp_env, d_disp, sizeof(double) * n_disp, hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            (n_drift > 0 && hipMemcpyAsync(drift_env, d_drift, sizeof(double) * n_dr, hipMemcpyDeviceToHost,
                                           ctx->stream) != hipSuccess) ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = fail(ctx, SGPU_ERR_HIP, "envelope kernel failed");
    }
    (void)hipFree(buf);
    return rc;
}

}  // namespace sgpu
