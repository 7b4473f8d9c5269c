// Elastic four-node MITC4 shell stiffness on gfx950.
//
// shell4_mindlin_stiffness (shell4.mojo:10-330): membrane, bending and
// assumed transverse shear of an ElasticMembranePlateSection (E, nu, h) at
// 2 x 2 Gauss points plus the drilling term G h, as a 24 x 24 matrix in
// global coordinates (DOFs ux uy uz rx ry rz per node).  The stiffness of an
// elastic shell never changes, so the frame driver forms it once and keeps
// it resident; the reference forms it again at every assembly.
//
// One 64-lane wavefront per shell.  Each lane evaluates the element's
// uniform geometry (basis, local coordinates, tying rotation) itself; per
// Gauss point the lanes split the shape functions, the tied shear rows, the
// per-node strain-displacement blocks B_j, the products B_j^T D dvol and the
// 576 stiffness entries (nine per lane, kept in registers across the four
// points).  Every expression and every sum keeps the reference's operation
// order (the library is built without contraction), so the result equals
// the reference's up to the last bits of atan2 / sin / cos.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/strut_gpu.h"
#include "ctx.h"

namespace sgpu {

namespace {

struct ShellGeo {
    double g1[3], g2[3], g3[3];
    double xl[2][4];
    double Ax, Bx, Cx, Ay, By, Cy;
    double rot[2][2];
};

// :22-61, :126-140; false for a degenerate basis
__device__ bool shell_geometry(const double *p, ShellGeo &g) {
    const double x[4] = {p[0], p[3], p[6], p[9]}, y[4] = {p[1], p[4], p[7], p[10]},
                 z[4] = {p[2], p[5], p[8], p[11]};
    const double v1x = (x[2] + x[1] - x[3] - x[0]) * 0.5, v1y = (y[2] + y[1] - y[3] - y[0]) * 0.5,
                 v1z = (z[2] + z[1] - z[3] - z[0]) * 0.5;
    double v2x = (x[3] + x[2] - x[1] - x[0]) * 0.5, v2y = (y[3] + y[2] - y[1] - y[0]) * 0.5,
           v2z = (z[3] + z[2] - z[1] - z[0]) * 0.5;
    double nrm = sqrt(v1x * v1x + v1y * v1y + v1z * v1z);
    if (nrm == 0.0) return false;
    g.g1[0] = v1x / nrm;
    g.g1[1] = v1y / nrm;
    g.g1[2] = v1z / nrm;
    const double alpha = v2x * g.g1[0] + v2y * g.g1[1] + v2z * g.g1[2];
    v2x -= alpha * g.g1[0];
    v2y -= alpha * g.g1[1];
    v2z -= alpha * g.g1[2];
    nrm = sqrt(v2x * v2x + v2y * v2y + v2z * v2z);
    if (nrm == 0.0) return false;
    g.g2[0] = v2x / nrm;
    g.g2[1] = v2y / nrm;
    g.g2[2] = v2z / nrm;
    g.g3[0] = g.g1[1] * g.g2[2] - g.g1[2] * g.g2[1];
    g.g3[1] = g.g1[2] * g.g2[0] - g.g1[0] * g.g2[2];
    g.g3[2] = g.g1[0] * g.g2[1] - g.g1[1] * g.g2[0];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        g.xl[0][i] = x[i] * g.g1[0] + y[i] * g.g1[1] + z[i] * g.g1[2];
        g.xl[1][i] = x[i] * g.g2[0] + y[i] * g.g2[1] + z[i] * g.g2[2];
    }
    g.Ax = -g.xl[0][0] + g.xl[0][1] + g.xl[0][2] - g.xl[0][3];
    g.Bx = g.xl[0][0] - g.xl[0][1] + g.xl[0][2] - g.xl[0][3];
    g.Cx = -g.xl[0][0] - g.xl[0][1] + g.xl[0][2] + g.xl[0][3];
    g.Ay = -g.xl[1][0] + g.xl[1][1] + g.xl[1][2] - g.xl[1][3];
    g.By = g.xl[1][0] - g.xl[1][1] + g.xl[1][2] - g.xl[1][3];
    g.Cy = -g.xl[1][0] - g.xl[1][1] + g.xl[1][2] + g.xl[1][3];
    const double alph = atan2(g.Ay, g.Ax);
    const double beta = 0.5 * 3.141592653589793 - atan2(g.Cx, g.Cy);
    g.rot[0][0] = sin(beta);
    g.rot[0][1] = -sin(alph);
    g.rot[1][0] = -cos(beta);
    g.rot[1][1] = cos(alph);
    return true;
}

// :96-124: tying matrix entry Gm[k][j]
__device__ double tying(const ShellGeo &g, int k, int j) {
    const double q4 = 0.25;
    const double dx34 = g.xl[0][2] - g.xl[0][3], dy34 = g.xl[1][2] - g.xl[1][3];
    const double dx21 = g.xl[0][1] - g.xl[0][0], dy21 = g.xl[1][1] - g.xl[1][0];
    const double dx32 = g.xl[0][2] - g.xl[0][1], dy32 = g.xl[1][2] - g.xl[1][1];
    const double dx41 = g.xl[0][3] - g.xl[0][0], dy41 = g.xl[1][3] - g.xl[1][0];
    // each tying row touches two nodes: (node, sign) pairs and its edge
    double dx = 0.0, dy = 0.0;
    int na = 0, nb = 0;
    double sa = 0.0, sb = 0.0;
    switch (k) {
        case 0: na = 0; sa = -0.5; nb = 3; sb = 0.5; dx = dx41; dy = dy41; break;
        case 1: na = 0; sa = -0.5; nb = 1; sb = 0.5; dx = dx21; dy = dy21; break;
        case 2: na = 1; sa = -0.5; nb = 2; sb = 0.5; dx = dx32; dy = dy32; break;
        default: na = 2; sa = 0.5; nb = 3; sb = -0.5; dx = dx34; dy = dy34; break;
    }
    const int node = j / 3, c = j % 3;
    if (node != na && node != nb) return 0.0;
    if (c == 0) return node == na ? sa : sb;
    return c == 1 ? -dy * q4 : dx * q4;
}

}  // namespace

// One wavefront per shell: xyz[12 e + 3 i + c] node i coordinate c,
// props[3 e] = (E, nu, h); K[576 e] row major.  status[e] = 1 for a
// degenerate basis, 2 for a zero Jacobian (the reference aborts on both).
__global__ __launch_bounds__(64) void shell4_stiffness_kernel(int n, const double *xyz, const double *props,
                                                              double *K, int32_t *status) {
    __shared__ double s_gm[4][12], s_dd[8][8], s_shp[3][4], s_bs[2][12];
    __shared__ double s_saveB[4][8][6], s_bdr[4][6], s_bjtd[4][6][8], s_bdrj[4][6];
    const int e = blockIdx.x, lane = threadIdx.x;
    if (e >= n) return;
    ShellGeo g;
    if (!shell_geometry(xyz + 12 * (size_t)e, g)) {
        if (lane == 0) status[e] = 1;
        return;
    }
    const double E = props[3 * (size_t)e], nu = props[3 * (size_t)e + 1], h = props[3 * (size_t)e + 2];
    // :63-84 section matrices
    const double factor = E / (1.0 - nu * nu);
    const double bend_factor = E * h * h * h / (12.0 * (1.0 - nu * nu));
    const double G = E / (2.0 * (1.0 + nu));
    const double kappa = 5.0 / 6.0;
    const double Ds = kappa * G * h;
    const double Ktt = G * h;
    {
        const int r = lane / 8, q = lane % 8;
        double v = 0.0;
        if (r < 3 && q < 3) {
            const double dm = (r == 2 || q == 2) ? (r == q ? factor * (1.0 - nu) / 2.0 : 0.0)
                                                 : (r == q ? factor : factor * nu);
            v = dm * h;
        } else if (r >= 3 && r < 6 && q >= 3 && q < 6) {
            const int rr = r - 3, qq = q - 3;
            const double db = (rr == 2 || qq == 2) ? (rr == qq ? bend_factor * (1.0 - nu) / 2.0 : 0.0)
                                                   : (rr == qq ? bend_factor : bend_factor * nu);
            v = -db;
        } else if (r >= 6 && r == q) {
            v = Ds;
        }
        s_dd[r][q] = v;
        if (lane < 48) s_gm[lane / 12][lane % 12] = tying(g, lane / 12, lane % 12);
    }
    const double r3 = 1.0 / sqrt(3.0);
    const double sgv[4] = {-r3, r3, r3, -r3}, tgv[4] = {-r3, -r3, r3, r3};
    const double s4[4] = {-0.5, 0.5, 0.5, -0.5}, t4[4] = {-0.5, -0.5, 0.5, 0.5};
    double acc[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) acc[m] = 0.0;
    for (int gp = 0; gp < 4; ++gp) {
        const double ss = sgv[gp], tt = tgv[gp];
        // :200-229 shape functions, Jacobian, Cartesian derivatives (uniform)
        double shp[3][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            shp[2][i] = (0.5 + s4[i] * ss) * (0.5 + t4[i] * tt);
            shp[0][i] = s4[i] * (0.5 + t4[i] * tt);
            shp[1][i] = t4[i] * (0.5 + s4[i] * ss);
        }
        double xs[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                xs[i][j] = 0.0;
#pragma unroll
                for (int k = 0; k < 4; ++k) xs[i][j] += g.xl[i][k] * shp[j][k];
            }
        const double xsj = xs[0][0] * xs[1][1] - xs[0][1] * xs[1][0];
        if (xsj == 0.0) {
            if (lane == 0) status[e] = 2;
            return;  // uniform: every lane leaves before the next barrier
        }
        const double jinv = 1.0 / xsj;
        const double sx00 = xs[1][1] * jinv, sx11 = xs[0][0] * jinv, sx01 = -xs[0][1] * jinv,
                     sx10 = -xs[1][0] * jinv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double temp = shp[0][i] * sx00 + shp[1][i] * sx10;
            shp[1][i] = shp[0][i] * sx01 + shp[1][i] * sx11;
            shp[0][i] = temp;
        }
        const double dvol = 1.0 * xsj;
        __syncthreads();  // the previous point's readers of s_bs / s_shp are done
        if (lane < 24) {  // :231-251 tied shear row entry Bs[i][j]
            const int i = lane / 12, j = lane % 12;
            const double ms0[4] = {0.0, 1.0 - tt, 0.0, 1.0 + tt}, ms1[4] = {1.0 - ss, 0.0, 1.0 + ss, 0.0};
            double b0 = 0.0, b1 = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) b0 += ms0[k] * s_gm[k][j];
#pragma unroll
            for (int k = 0; k < 4; ++k) b1 += ms1[k] * s_gm[k][j];
            double r1 = g.Cx + ss * g.Bx;
            double r3v = g.Cy + ss * g.By;
            r1 = sqrt(r1 * r1 + r3v * r3v);
            double r2 = g.Ax + tt * g.Bx;
            r3v = g.Ay + tt * g.By;
            r2 = sqrt(r2 * r2 + r3v * r3v);
            b0 = b0 * r1 / (8.0 * xsj);
            b1 = b1 * r2 / (8.0 * xsj);
            s_bs[i][j] = g.rot[i][0] * b0 + g.rot[i][1] * b1;
        } else if (lane < 36) {
            const int c = (lane - 24) / 4, i = (lane - 24) % 4;
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q == i) v = shp[c][q];
            s_shp[c][i] = v;
        }
        __syncthreads();
        if (lane < 32) {  // :253-292 row p of B_j
            const int j = lane / 8, p = lane % 8;
            const double a0 = s_shp[0][j], a1 = s_shp[1][j];
            double row[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (p < 6) {
                // Bmem / Bb row (two entries) times Gmem
                double m0, m1;
                if (p == 0) { m0 = a0; m1 = 0.0; }
                else if (p == 1) { m0 = 0.0; m1 = a1; }
                else if (p == 2) { m0 = a1; m1 = a0; }
                else if (p == 3) { m0 = 0.0; m1 = -a0; }
                else if (p == 4) { m0 = a1; m1 = 0.0; }
                else { m0 = a0; m1 = -a1; }
                const int off = p < 3 ? 0 : 3;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    double sum = 0.0;
                    sum += m0 * g.g1[q];
                    sum += m1 * g.g2[q];
                    row[off + q] = sum;
                }
            } else {
                const int i = p - 6;
                const double b0 = s_bs[i][j * 3], b1 = s_bs[i][j * 3 + 1], b2 = s_bs[i][j * 3 + 2];
                // Gshear rows: (g3 | 0), (0 | g1), (0 | g2)
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const double c0 = q < 3 ? g.g3[q] : 0.0, c1 = q < 3 ? 0.0 : g.g1[q - 3],
                                 c2 = q < 3 ? 0.0 : g.g2[q - 3];
                    double sum = 0.0;
                    sum += b0 * c0;
                    sum += b1 * c1;
                    sum += b2 * c2;
                    row[q] = sum;
                }
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) s_saveB[j][p][q] = row[q];
        } else if (lane < 36) {  // :294-303 drilling row of node j
            const int j = lane - 32;
            const double B1 = -0.5 * s_shp[1][j], B2 = 0.5 * s_shp[0][j], B6 = -s_shp[2][j];
            s_bdr[j][0] = B1 * g.g1[0] + B2 * g.g2[0];
            s_bdr[j][1] = B1 * g.g1[1] + B2 * g.g2[1];
            s_bdr[j][2] = B1 * g.g1[2] + B2 * g.g2[2];
            s_bdr[j][3] = B6 * g.g3[0];
            s_bdr[j][4] = B6 * g.g3[1];
            s_bdr[j][5] = B6 * g.g3[2];
        }
        __syncthreads();
        // :305-318 (B_j^T with the bending block negated) D dvol, and the
        // drilling row times Ktt dvol
        for (int idx = lane; idx < 192; idx += 64) {
            const int j = idx / 48, p = (idx / 8) % 6, q = idx % 8;
            double sum = 0.0;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                double bj = s_saveB[j][r][p];
                if (r >= 3 && r < 6 && p >= 3) bj *= -1.0;
                sum += bj * s_dd[r][q];
            }
            s_bjtd[j][p][q] = sum * dvol;
        }
        if (lane < 24) s_bdrj[lane / 6][lane % 6] = s_bdr[lane / 6][lane % 6] * (Ktt * dvol);
        __syncthreads();
        // :319-330 K[j p][k q] += B_j^T D B_k dvol + drilling
#pragma unroll
        for (int m = 0; m < 9; ++m) {
            const int idx = lane + 64 * m;
            const int row = idx / 24, col = idx % 24;
            const int j = row / 6, p = row % 6, k = col / 6, q = col % 6;
            double sum = 0.0;
#pragma unroll
            for (int r = 0; r < 8; ++r) sum += s_bjtd[j][p][r] * s_saveB[k][r][q];
            acc[m] += sum + s_bdrj[j][p] * s_bdr[k][q];
        }
    }
    double *Ke = K + 576 * (size_t)e;
#pragma unroll
    for (int m = 0; m < 9; ++m) Ke[lane + 64 * m] = acc[m];
    if (lane == 0) status[e] = 0;
}

// n shells on the context's stream (device buffers)
int launch_shell4_stiffness(sgpu_ctx *ctx, int n, const double *d_xyz, const double *d_props, double *d_K,
                            int32_t *d_status) {
    if (n <= 0) return SGPU_OK;
    hipLaunchKernelGGL(shell4_stiffness_kernel, dim3(n), dim3(64), 0, ctx->stream, n, d_xyz, d_props, d_K, d_status);
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

}  // namespace sgpu

int sgpu_shell4_stiffness(sgpu_ctx *ctx, int n, const double *xyz, const double *props, double *K) {
    if (!ctx || n < 0 || (n > 0 && (!xyz || !props || !K))) return SGPU_ERR_ARG;
    if (n == 0) return SGPU_OK;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    for (int e = 0; e < n; ++e)
        if (!(props[3 * (size_t)e + 1] < 1.0 && props[3 * (size_t)e + 1] > -1.0))
            return fail(ctx, SGPU_ERR_ARG, "shell %d: Poisson's ratio %g outside (-1, 1)", e, props[3 * (size_t)e + 1]);
    double *d_xyz = nullptr, *d_props = nullptr, *d_K = nullptr;
    int32_t *d_status = nullptr;
    int rc = SGPU_OK;
    std::vector<int32_t> st(n, 0);
    if (hipMalloc(&d_xyz, sizeof(double) * 12 * (size_t)n) != hipSuccess ||
        hipMalloc(&d_props, sizeof(double) * 3 * (size_t)n) != hipSuccess ||
        hipMalloc(&d_K, sizeof(double) * 576 * (size_t)n) != hipSuccess ||
        hipMalloc(&d_status, sizeof(int32_t) * (size_t)n) != hipSuccess) {
        rc = fail(ctx, SGPU_ERR_HIP, "out of device memory for %d shells", n);
        goto out;
    }
    if (hipMemcpyAsync(d_xyz, xyz, sizeof(double) * 12 * (size_t)n, hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        hipMemcpyAsync(d_props, props, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess) {
        rc = fail(ctx, SGPU_ERR_HIP, "shell upload failed");
        goto out;
    }
    if ((rc = launch_shell4_stiffness(ctx, n, d_xyz, d_props, d_K, d_status))) goto out;
    if (hipMemcpyAsync(K, d_K, sizeof(double) * 576 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(st.data(), d_status, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        rc = fail(ctx, SGPU_ERR_HIP, "shell readback failed");
        goto out;
    }
    for (int e = 0; e < n; ++e)
        if (st[e]) {
            rc = fail(ctx, SGPU_ERR_ARG, "shell %d: %s", e, st[e] == 1 ? "degenerate basis" : "singular Jacobian");
            break;
        }
out:
    (void)hipFree(d_xyz);
    (void)hipFree(d_props);
    (void)hipFree(d_K);
    (void)hipFree(d_status);
    return rc;
}
