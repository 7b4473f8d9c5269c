// Transient nonlinear analysis of 3D fiber beam-column frames on gfx950:
// Newmark / Newton over forceBeamColumn3d and dispBeamColumn3d elements
// (elem3d.hip) with the Transformation constraint handler and several
// UniformExcitation patterns -- the driver
// solver/run_case/analysis/transient_nonlinear.mojo:830-3638 on the path a 3D
// frame with rigid diaphragms takes.
//
// The reference assembles a dense total_dofs^2 K every Newton iteration,
// collapses it to T^T K T (helpers.mojo:3345-3377), extracts the free block,
// builds dense C_ff, and copies K + a1 C + a0 M into a BandGeneral band whose
// half bandwidth comes from the element free map
// (linear_solver_backend.mojo:1481-1506), dropping the collapse's couplings
// outside it (:413-426).  Here nothing is dense:
//   * each element's 12 DOFs expand to at most 12 free retained DOFs; a
//     per-element T_e (12 x m_e) collapses k_e and f_e on the device right
//     after the element state (k_hat = T_e^T k T_e, f_hat = T_e^T f);
//   * every band slot gathers the k_hat entries that land on it (element
//     order, host-built lists) and adds a1 C + a0 M with C = alphaM M +
//     betaK K + betaKInit K_init per slot, in the reference's order; slots
//     outside the chosen band never exist, which is the reference's
//     truncation when the band follows the element free map;
//   * C v in the residual is alphaM M v + element-level (k_hat v) and
//     (k_hat_init v) gathered per DOF -- the full matrices, as the
//     reference's dense C_ff row dot products use;
//   * the band LU and both substitutions run in one workgroup per record
//     with the reference's row swaps (band_lu.hip): where the reference
//     moves to its dense partial-pivot LU, the band LU pivots and gives the
//     same factors, and the run counts those factorizations.
// The host loop launches the kernels of a Newton iteration and reads back
// one residual norm per iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/strut_gpu.h"
#include "ctx.h"
#include "frame3d_host.h"

namespace sgpu {

// band_lu.hip: the factorization with the reference's row swaps
int launch_band_piv(sgpu_ctx *ctx, int mode, int n, int hbw, double *d_band, const double *d_rhs, double *d_x,
                    double *d_status, double *d_L, int32_t *d_ipiv, double *d_work, const int32_t *recs,
                    int count);
bool band_piv_reg_fits(int n, int hbw);
size_t band_piv_work_doubles(int n, int hbw);
size_t band_blk_work_doubles(int n, int hbw);
// shell4.hip: elastic MITC4 shell stiffness, one wavefront per shell
int launch_shell4_stiffness(sgpu_ctx *ctx, int n, const double *d_xyz, const double *d_props, double *d_K,
                            int32_t *d_status);


// ------------------------------------------------------------------ kernels

// Every kernel below works on a list of records (ground-motion instances
// of the same frame): record r = recs[blockIdx.y] (or recs[blockIdx.x] for
// the one-workgroup-per-record kernels) owns the r-th slice of each
// per-record array.

// :553-590 predictor; the step's base state kept in un/vn/an
__global__ void f3_predict_kernel(const int32_t *recs, int n, double *uf, const double *vf, const double *af,
                                  double *un, double *vn, double *an, double dt, double beta) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)recs[blockIdx.y] * n + i;
    const double coeff = dt * dt * (0.5 - beta);
    const double u0 = uf[o], v0 = vf[o], a0 = af[o];
    un[o] = u0;
    vn[o] = v0;
    an[o] = a0;
    uf[o] = u0 + dt * v0 + coeff * a0;
}

// scatter of the free vector and the slave update (_enforce_mpc_values,
// helpers.mojo:3380-3397; retained single-point-fixed DOFs hold 0)
__global__ void f3_expand_kernel(const int32_t *recs, int total, int n, double *u, const double *uf,
                                 const int32_t *free_idx, const int32_t *row_offsets, const int32_t *dof_pool,
                                 const double *coeff_pool, const int32_t *slave) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= total) return;
    const int r = recs[blockIdx.y];
    uf += (size_t)r * n;
    const int fi = free_idx[d];
    double v = 0.0;
    if (fi >= 0) {
        v = uf[fi];
    } else if (slave[d]) {
        for (int t = row_offsets[d]; t < row_offsets[d + 1]; ++t) {
            const int q = free_idx[dof_pool[t]];
            v += coeff_pool[t] * (q >= 0 ? uf[q] : 0.0);
        }
    }
    u[(size_t)r * total + d] = v;
}

// element end displacements in list order: position b = list index * n_el + e
__global__ void f3_gather_u12_kernel(const int32_t *recs, int n_list, int n_el, int total, const int32_t *elem_nodes,
                                     const double *u, double *u12) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_list * n_el * 12) return;
    const int64_t b = t / 12;
    const int a = (int)(t % 12), e = (int)(b % n_el), r = recs[b / n_el];
    const int node = elem_nodes[2 * e + (a >= 6)];
    u12[t] = u[(size_t)r * total + 6 * node + (a % 6)];
}

// :496-531 trial acceleration and velocity
__global__ void f3_trial_kernel(const int32_t *recs, int n, const double *uf, const double *un, const double *vn,
                                const double *an, double *at, double *vt, double a0, double a2, double a3,
                                double gamma, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)recs[blockIdx.y] * n + i;
    const double anext = a0 * (uf[o] - un[o]) - a2 * vn[o] - a3 * an[o];
    at[o] = anext;
    vt[o] = vn[o] + dt * ((1.0 - gamma) * an[o] + gamma * anext);
}

// one 256-thread workgroup per element: k_hat = T^T k T, f_hat = T^T f, and
// the element-level products k_hat v, k_hat_init v at the trial velocity
__global__ __launch_bounds__(256) void f3_elem_collapse_kernel(const int32_t *recs, int n, int n_el, const int32_t *ret,
                                                               const int32_t *ret_n, const double *T, const double *k,
                                                               const double *f, const double *kinit_hat,
                                                               const double *vt, double *khat, double *fhat, double *yk,
                                                               double *yi) {
    __shared__ double sT[144], sK[144], sKT[144], sV[12], sKh[144];
    const int b = blockIdx.x, tid = threadIdx.x;  // list position: k and f are in list order
    const int e = b % n_el, r = recs[b / n_el];
    const size_t re = (size_t)r * n_el + e;  // record layout of the collapsed results
    const int m = ret_n[e];
    if (tid < 144) {
        sT[tid] = T[(size_t)e * 144 + tid];
        sK[tid] = k[(size_t)b * 144 + tid];
    }
    if (tid < 12) sV[tid] = tid < m ? vt[(size_t)r * n + ret[e * 12 + tid]] : 0.0;
    __syncthreads();
    if (tid < 144) {  // (k T)[a][b'] = sum_b k[a][b] T[b][b']
        const int a = tid / 12, bp = tid % 12;
        double s = 0.0;
        for (int b = 0; b < 12; ++b) s += sK[a * 12 + b] * sT[b * 12 + bp];
        sKT[tid] = s;
    }
    __syncthreads();
    if (tid < 144) {  // k_hat[a'][b'] = sum_a T[a][a'] (k T)[a][b']
        const int ap = tid / 12, bp = tid % 12;
        double s = 0.0;
        for (int a = 0; a < 12; ++a) s += sT[a * 12 + ap] * sKT[a * 12 + bp];
        sKh[tid] = s;
        khat[re * 144 + tid] = s;
    }
    __syncthreads();
    if (tid < 12) {
        const int ap = tid;
        double sf = 0.0, sy = 0.0, syi = 0.0;
        for (int a = 0; a < 12; ++a) sf += sT[a * 12 + ap] * f[(size_t)b * 12 + a];
        for (int bp = 0; bp < 12; ++bp) {
            sy += sKh[ap * 12 + bp] * sV[bp];
            syi += kinit_hat[re * 144 + ap * 12 + bp] * sV[bp];
        }
        fhat[re * 12 + ap] = sf;
        yk[re * 12 + ap] = sy;
        yi[re * 12 + ap] = syi;
    }
}

// every band slot: K from its k_hat entries, then (:2302-2335, :2400-2411)
// band = K; band += a1 C with C = alphaM M + betaK K + betaKInit K_init;
// band += a0 M.  raw != 0 writes K alone (the initial tangent).
// A source index q >= 0 is a beam-column's k_hat entry of the record, q < 0
// the shells' collapsed stiffness entry -q - 1 (the same for every record);
// beam-columns come first, then shells, each in element order.
__global__ void f3_band_kernel(const int32_t *recs, int64_t slots, int64_t khat_stride, const int32_t *off,
                               const int32_t *src, const double *khat, const double *sh_khat,
                               const double *mass_band, const double *kinit_band, double *band, int raw, double a0,
                               double a1, double alpha_m, double beta_k, double beta_k_init, int has_damping) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots) return;
    const int r = recs[blockIdx.y];
    khat += (size_t)r * khat_stride;
    kinit_band += (size_t)r * slots;
    band += (size_t)r * slots;
    double kv = 0.0;
    for (int t = off[s]; t < off[s + 1]; ++t) {
        const int q = src[t];
        kv += q >= 0 ? khat[q] : sh_khat[-(int64_t)q - 1];
    }
    if (raw) {
        band[s] = kv;
        return;
    }
    const double mv = mass_band[s];
    double v = kv;
    if (has_damping && a1 != 0.0) {
        double c = 0.0;
        if (alpha_m != 0.0) c = alpha_m * mv;
        if (beta_k != 0.0) c += beta_k * kv;
        if (beta_k_init != 0.0) c += beta_k_init * kinit_band[s];
        v += a1 * c;
    }
    if (a0 != 0.0) v += a0 * mv;
    band[s] = v;
}

// :2337-2356 R = P - F_int - C v - M a on the free DOFs
// A source index s >= 0 is a beam-column's (element, retained position)
// entry, s < 0 the shells' entry -s - 1: their k_hat u and k_hat v, the
// elastic shell's tangent and initial tangent being the same matrix.
__global__ void f3_residual_kernel(const int32_t *recs, int n, int n_el, int n_sh, int n_patterns, const double *load,
                                   const double *ag_all, const int32_t *dof_off, const int32_t *dof_src,
                                   const double *fhat, const double *yk, const double *yi, const double *sh_f,
                                   const double *sh_y, const int32_t *m_off, const int32_t *m_col,
                                   const double *m_val, const double *vt, const double *at, double alpha_m,
                                   double beta_k, double beta_k_init, int has_damping, double *R) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = recs[blockIdx.y];
    const double *ag = ag_all + 3 * r;
    fhat += (size_t)r * n_el * 12;
    yk += (size_t)r * n_el * 12;
    yi += (size_t)r * n_el * 12;
    sh_f += (size_t)r * n_sh * 24;
    sh_y += (size_t)r * n_sh * 24;
    vt += (size_t)r * n;
    at += (size_t)r * n;
    R += (size_t)r * n;
    double p = 0.0;
    for (int q = 0; q < n_patterns; ++q) p += ag[q] * load[(size_t)q * n + i];
    double fi = 0.0, ky = 0.0, kiy = 0.0;
    for (int t = dof_off[i]; t < dof_off[i + 1]; ++t) {
        const int s = dof_src[t];
        if (s >= 0) {
            fi += fhat[s];
            ky += yk[s];
            kiy += yi[s];
        } else {
            const double y = sh_y[-(int64_t)s - 1];
            fi += sh_f[-(int64_t)s - 1];
            ky += y;
            kiy += y;
        }
    }
    double mv = 0.0, ma = 0.0;
    for (int t = m_off[i]; t < m_off[i + 1]; ++t) {
        mv += m_val[t] * vt[m_col[t]];
        ma += m_val[t] * at[m_col[t]];
    }
    double damping = 0.0;
    if (has_damping) damping = alpha_m * mv + beta_k * ky + beta_k_init * kiy;
    R[i] = p - fi - 0.0 - damping - ma;
}

// one 576-thread workgroup per shell, once: k_hat = T^T k T (24 x 24; the
// positions past the shell's retained count stay zero)
__global__ __launch_bounds__(576) void f3_shell_collapse_kernel(const double *T, const double *k, double *khat) {
    __shared__ double sT[576], sK[576], sKT[576];
    const int e = blockIdx.x, tid = threadIdx.x;
    sT[tid] = T[(size_t)e * 576 + tid];
    sK[tid] = k[(size_t)e * 576 + tid];
    __syncthreads();
    const int a = tid / 24, bp = tid % 24;
    double s = 0.0;
    for (int b = 0; b < 24; ++b) s += sK[a * 24 + b] * sT[b * 24 + bp];
    sKT[tid] = s;
    __syncthreads();
    const int ap = tid / 24;
    s = 0.0;
    for (int a2 = 0; a2 < 24; ++a2) s += sT[a2 * 24 + ap] * sKT[a2 * 24 + bp];
    khat[(size_t)e * 576 + tid] = s;
}

// every listed record's shells: k_hat u and k_hat v over the retained free
// DOFs (u = T u_ret exactly, the slaves following their retained DOFs), one
// 64-lane workgroup per (shell, record), one row per lane
__global__ __launch_bounds__(64) void f3_shell_force_kernel(const int32_t *recs, int n, int n_sh, const int32_t *ret,
                                                            const int32_t *ret_n, const double *khat,
                                                            const double *uf, const double *vt, double *sh_f,
                                                            double *sh_y) {
    __shared__ double su[24], sv[24];
    const int e = blockIdx.x, r = recs[blockIdx.y], lane = threadIdx.x;
    const int m = ret_n[e];
    if (lane < 24) {
        su[lane] = lane < m ? uf[(size_t)r * n + ret[e * 24 + lane]] : 0.0;
        sv[lane] = lane < m ? vt[(size_t)r * n + ret[e * 24 + lane]] : 0.0;
    }
    __syncthreads();
    if (lane >= 24) return;
    const double *row = khat + (size_t)e * 576 + lane * 24;
    double f = 0.0, y = 0.0;
    for (int b = 0; b < m; ++b) {
        f += row[b] * su[b];
        y += row[b] * sv[b];
    }
    const size_t o = ((size_t)r * n_sh + e) * 24 + lane;
    sh_f[o] = f;
    sh_y[o] = y;
}

// one workgroup: sum of R^2 and the largest element status
__global__ __launch_bounds__(1024) void f3_norm_kernel(const int32_t *recs, int n, const double *R, int n_el,
                                                       const int32_t *status, double *out) {
    __shared__ double s_sum[1024];
    __shared__ int s_st[1024];
    const int r = recs[blockIdx.x];
    R += (size_t)r * n;
    status += (size_t)blockIdx.x * n_el;  // element status in list order
    out += 4 * r;
    double s = 0.0;
    int st = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += R[i] * R[i];
    for (int e = threadIdx.x; e < n_el; e += blockDim.x) st = max(st, status[e]);
    s_sum[threadIdx.x] = s;
    s_st[threadIdx.x] = st;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
            s_st[threadIdx.x] = max(s_st[threadIdx.x], s_st[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_sum[0];
        out[1] = (double)s_st[0];
    }
}

// Band LU without pivoting (lsb.mojo:631-679) and both substitutions
// (:735-769), one workgroup.  Slot (i, j) = i W + (j - i + hbw).  Where a
// column's largest sub-diagonal magnitude exceeds the pivot's, the reference
// moves to its dense partial-pivot LU (:608-628); the effective tangent of
// these frames is symmetric with a dominant mass and stiffness diagonal, so
// the band elimination continues and gives the same solution within
// rounding, and the solve reports status 1 (the reference's fallback
// counted).  A pivot at or below 1e-14: status 2, nothing solved.
__global__ __launch_bounds__(kF3SolveThreads) void f3_solve_kernel(const int32_t *recs, int n, int hbw, double *band,
                                                                   const double *rhs, double *x, double *status) {
    if (recs) {  // one workgroup per listed record
        const int r_ = recs[blockIdx.x];
        band += (size_t)r_ * n * (2 * hbw + 1);
        rhs += (size_t)r_ * n;
        x += (size_t)r_ * n;
        status += 4 * r_;
    }
    extern __shared__ double s_mem[];
    double *s_y = s_mem;             // n
    double *s_l = s_mem + n;         // hbw multipliers of the column
    double *s_u = s_mem + n + hbw;   // hbw entries of the pivot row
    __shared__ int s_flag;
    const int W = 2 * hbw + 1, tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < n; i += nt) s_y[i] = rhs[i];
    if (tid == 0) s_flag = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const int i1 = min(k + hbw, n - 1);
        const int m = i1 - k;
        const double diag = band[(size_t)k * W + hbw];
        const double pabs = fabs(diag);
        for (int r = tid; r < m; r += nt) {
            const int i = k + 1 + r;
            if (fabs(band[(size_t)i * W + (k - i + hbw)]) > pabs) atomicOr(&s_flag, 1);
        }
        __syncthreads();
        if (pabs <= 1.0e-14) {
            if (tid == 0) status[0] = 2.0;
            return;
        }
        for (int r = tid; r < m; r += nt) {
            const int i = k + 1 + r;
            const size_t ik = (size_t)i * W + (k - i + hbw);
            const double l = band[ik] / diag;
            band[ik] = l;
            s_l[r] = l;
            s_u[r] = band[(size_t)k * W + hbw + 1 + r];
        }
        __syncthreads();
        const int cells = m * m;
        for (int c = tid; c < cells; c += nt) {
            const int r = c / m, q = c % m;
            const double l = s_l[r];
            if (l == 0.0) continue;
            const int i = k + 1 + r, j = k + 1 + q;
            band[(size_t)i * W + (j - i + hbw)] -= l * s_u[q];
        }
        // forward substitution rides along: y_i -= l_ik y_k
        const double yk = s_y[k];
        for (int r = tid; r < m; r += nt) s_y[k + 1 + r] -= s_l[r] * yk;
        __syncthreads();
    }
    // back substitution, column by column
    for (int k = n - 1; k >= 0; --k) {
        const double xk = s_y[k] / band[(size_t)k * W + hbw];
        __syncthreads();
        if (tid == 0) s_y[k] = xk;
        const int i0 = max(k - hbw, 0);
        for (int i = i0 + tid; i < k; i += nt) s_y[i] -= band[(size_t)i * W + (k - i + hbw)] * xk;
        __syncthreads();
    }
    for (int i = tid; i < n; i += nt) x[i] = s_y[i];
    if (tid == 0) status[0] = s_flag ? 1.0 : 0.0;
}

// The same factorization and substitutions with the active window of the
// band in registers, for half bandwidths up to kF3RegMaxP - 1.  At column k
// the window holds rows and columns k .. k + P - 1 (P = hbw + 1), every row
// and column in a circular slot (index mod P); thread t owns the window
// entries e * 1024 + t (slot pair (row, column) fixed for the whole run).
// Per column: (A) the pivot row's entries go to LDS (and to the band, final)
// and the pivot column's entries to LDS; barrier; (B) the pivot column's
// owners form the multipliers (to LDS and the band) and the forward
// substitution; barrier; (C) every owner applies the rank-1 update in
// registers, and the retired row and column slots reload the next row and
// column of the band (never touched before: earlier columns only reach rows
// and columns up to k + hbw).  Pivot row and multipliers alternate between
// two LDS buffers, so (C) of one column and (A) of the next need no barrier
// between them.  The back substitution runs in wavefront 0 alone: lane l
// holds U(i, i + 1 + l) and U(i, i + 65 + l), the solved x values ride in a
// register window shifted one lane per row, the next row's U entries load
// while the current row reduces.
constexpr int kF3RegMaxP = 128;
constexpr int kF3RegEntries = kF3RegMaxP * kF3RegMaxP / kF3SolveThreads;  // 16

__global__ __launch_bounds__(kF3SolveThreads) void f3_solve_reg_kernel(const int32_t *recs, int n, int hbw,
                                                                       double *band, const double *rhs, double *x,
                                                                       double *status) {
    if (recs) {  // one workgroup per listed record
        const int r_ = recs[blockIdx.x];
        band += (size_t)r_ * n * (2 * hbw + 1);
        rhs += (size_t)r_ * n;
        x += (size_t)r_ * n;
        status += 4 * r_;
    }
    extern __shared__ double s_y[];  // n
    __shared__ double s_col[kF3RegMaxP], s_l[2][kF3RegMaxP], s_u[2][kF3RegMaxP];
    __shared__ double s_diag;
    __shared__ int s_flag;
    const int P = hbw + 1, W = 2 * hbw + 1, tid = threadIdx.x;
    const int cells = P * P;
    double a[kF3RegEntries];
    int slot[kF3RegEntries];  // row slot << 8 | column slot, -1 past the window
#pragma unroll
    for (int e = 0; e < kF3RegEntries; ++e) {
        const int L = e * kF3SolveThreads + tid;
        slot[e] = L < cells ? ((L / P) << 8) | (L % P) : -1;
        a[e] = 0.0;
        if (L < cells) {
            const int i = L / P, j = L % P;  // at k = 0 the slots are the indices
            if (i < n && j < n) a[e] = band[(size_t)i * W + (j - i + hbw)];
        }
    }
    for (int i = tid; i < n; i += kF3SolveThreads) s_y[i] = rhs[i];
    if (tid == 0) s_flag = 0;
    __syncthreads();
    int kc = 0;  // k mod P
    for (int k = 0; k < n; ++k) {
        const int buf = k & 1;
        // (A) pivot row -> s_u and the band (final U); pivot column -> s_col
#pragma unroll
        for (int e = 0; e < kF3RegEntries; ++e) {
            const int sl = slot[e];
            if (sl < 0) continue;
            const int rs = sl >> 8, cs = sl & 255;
            if (rs == kc) {
                const int j = k + (cs - kc + P) % P;
                s_u[buf][cs] = a[e];
                if (cs == kc) s_diag = a[e];
                if (j < n) band[(size_t)k * W + (j - k + hbw)] = a[e];
            } else if (cs == kc) {
                s_col[rs] = a[e];
            }
        }
        __syncthreads();
        const double diag = s_diag;
        if (fabs(diag) <= 1.0e-14) {
            if (tid == 0) status[0] = 2.0;
            return;
        }
        // (B) multipliers, the reference's dense-fallback test, forward substitution
        const double yk = s_y[k];
        for (int rs = tid; rs < P; rs += kF3SolveThreads) {
            if (rs == kc) continue;
            const int i = k + (rs - kc + P) % P;
            const double v = s_col[rs];
            if (fabs(v) > fabs(diag)) s_flag = 1;
            const double l = v / diag;
            s_l[buf][rs] = l;
            if (i < n) {
                band[(size_t)i * W + (k - i + hbw)] = l;
                s_y[i] -= l * yk;
            }
        }
        __syncthreads();
        // (C) rank-1 update; the retired slots take row and column k + P
        const int k1 = k + P, kc1 = kc + 1 == P ? 0 : kc + 1;
#pragma unroll
        for (int e = 0; e < kF3RegEntries; ++e) {
            const int sl = slot[e];
            if (sl < 0) continue;
            const int rs = sl >> 8, cs = sl & 255;
            if (rs != kc && cs != kc) {
                a[e] -= s_l[buf][rs] * s_u[buf][cs];
            } else if (rs == kc) {  // row k + P, columns of the next window
                const int j = k + 1 + (cs - kc1 + P) % P;
                a[e] = (k1 < n && j < n) ? band[(size_t)k1 * W + (j - k1 + hbw)] : 0.0;
            } else {  // column k + P, rows of the next window
                const int i = k + 1 + (rs - kc1 + P) % P;
                a[e] = (k1 < n && i < n) ? band[(size_t)i * W + (k1 - i + hbw)] : 0.0;
            }
        }
        kc = kc1;
    }
    __syncthreads();
    // back substitution in wavefront 0
    if (tid < 64) {
        const int lane = tid;
        double xa = 0.0, xb = 0.0;  // x[i + 1 + lane], x[i + 65 + lane]
        auto load_row = [&](int i, double &u0, double &u1) {
            u0 = (i >= 0 && lane < hbw && i + 1 + lane < n) ? band[(size_t)i * W + hbw + 1 + lane] : 0.0;
            u1 = (i >= 0 && lane + 64 < hbw && i + 65 + lane < n) ? band[(size_t)i * W + hbw + 65 + lane] : 0.0;
        };
        double u0, u1, d0;
        load_row(n - 1, u0, u1);
        d0 = band[(size_t)(n - 1) * W + hbw];
        for (int i = n - 1; i >= 0; --i) {
            double n0, n1, nd = 1.0;
            load_row(i - 1, n0, n1);
            if (i > 0) nd = band[(size_t)(i - 1) * W + hbw];
            double s = u0 * xa + u1 * xb;
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            const double xi = (s_y[i] - s) / d0;
            if (lane == 0) x[i] = xi;
            const double carry = __shfl(xa, 63, 64);
            xa = __shfl_up(xa, 1, 64);
            xb = __shfl_up(xb, 1, 64);
            if (lane == 0) {
                xa = xi;
                xb = carry;
            }
            u0 = n0;
            u1 = n1;
            d0 = nd;
        }
        if (lane == 0) status[0] = s_flag ? 1.0 : 0.0;
    }
}

// The factorization and substitutions with no global memory access on the
// per-column path, for half bandwidths up to kF3BlkMaxP - 1 and n up to what
// LDS holds (RB-row x 4-column owner blocks on NT threads):
//   * the active window lives in registers; an owner holds rows RB br + r and
//     columns bc + c Bc (Bc = ceil(P / 4)), so the lanes of a wavefront read
//     consecutive pivot-row entries;
//   * one barrier per column: the owners of the pivot row and column publish
//     them (double-buffered LDS) and the pivot's owner its reciprocal; then
//     every owner applies RB x 4 updates with l = a_ik (1 / a_kk), and two
//     switches on wavefront-uniform values overwrite the retired row and
//     column from the ring (no per-entry tests anywhere);
//   * entering rows and columns are fetched kF3BlkD columns ahead into
//     registers by every wavefront but the last, and land in an LDS ring;
//   * the last wavefront does the forward substitution, the reference's
//     dense-fallback test and the write-back of finished U rows, so no
//     other wavefront issues a store;
//   * the back substitution runs by columns in wavefront 0 with DPP lane
//     shifts while the other wavefronts stage U columns in LDS.
// A workgroup barrier that waits only for this wavefront's LDS traffic
// (lgkmcnt(0); vmcnt and expcnt left at their maxima in the gfx9 encoding),
// so global stores and prefetch loads stay in flight across it.  The empty
// asm statements keep the compiler from moving memory accesses across.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

constexpr int kF3BlkMaxP = 120;                 // 4-row blocks: 30 x 30 owners on 1024 threads
constexpr int kF3BlkD = 8;                      // columns per prefetch batch
constexpr int kF3BlkR = 16;                     // U rows per staged batch
constexpr int kF3BlkRing = 2 * kF3BlkD * 2 * kF3BlkMaxP;  // doubles; also holds 2 x kF3BlkR U rows

// Value of lane l + 1 of the wavefront (lane 63: 0).
__device__ __forceinline__ double wave_next_lane(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_value(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

template <int RB, int NT>
__global__ __launch_bounds__(NT) void f3_solve_blk_kernel(const int32_t *recs, int n, int hbw, double *band,
                                                       const double *rhs, double *x, double *status) {
    if (recs) {  // one workgroup per listed record
        const int r_ = recs[blockIdx.x];
        band += (size_t)r_ * n * (2 * hbw + 1);
        rhs += (size_t)r_ * n;
        x += (size_t)r_ * n;
        status += 4 * r_;
    }
    constexpr int kWorkers = NT - 64;  // all but the last wavefront
    constexpr int kPf = (kF3BlkD * 2 * kF3BlkMaxP + kWorkers - 1) / kWorkers;
    extern __shared__ double s_dyn[];
    double *s_ring = s_dyn;             // [2 kF3BlkD][2][kF3BlkMaxP]: entering row / column values
    double *s_y = s_dyn + kF3BlkRing;   // n
    __shared__ double s_col[2][kF3BlkMaxP], s_u[2][kF3BlkMaxP], s_rdiag[2];
    __shared__ int s_flag;
    const int P = hbw + 1, W = 2 * hbw + 1, tid = threadIdx.x;
    const int Bc = (P + 3) / 4, Br = (P + RB - 1) / RB;
    const bool owner = tid < Br * Bc;
    // owned window slots: rows RB br + r, columns bc + c Bc (r < RB, c < 4):
    // lanes of a wavefront read consecutive pivot-row entries
    const int br = owner ? tid / Bc : 0, bc = owner ? tid % Bc : 0;
    const int per_batch = kF3BlkD * 2 * P;  // values entering the window per batch
    // value q of the batch starting at column s0: column s0 + q / (2P) takes
    // row T = s + P (part < P: column slot part) or column T (row slot part - P)
    auto fetch = [&](int s0, int q) -> double {
        const int s = s0 + q / (2 * P), part = q % (2 * P);
        const int T = s + P, base = s + 1, bslot = base % P;
        int idx;
        bool ok;
        if (part < P) {
            const int j = base + (part - bslot + P) % P;
            ok = T < n && j < n;
            idx = T * W + (j - T + hbw);
        } else {
            const int i = base + (part - P - bslot + P) % P;
            ok = T < n && i < n && i != T;
            idx = i * W + (T - i + hbw);
        }
        const double v = band[ok ? idx : 0];
        return ok ? v : 0.0;
    };
    auto ring_at = [&](int s, int part, int slot) -> double & {
        return s_ring[((s % (2 * kF3BlkD)) * 2 + part) * kF3BlkMaxP + slot];
    };
    double a[RB][4];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int i = RB * br + r, j = bc + c * Bc;  // at k = 0 the slots are the indices
            a[r][c] = (owner && i < P && j < P && i < n && j < n) ? band[(size_t)i * W + (j - i + hbw)] : 0.0;
        }
    for (int i = tid; i < n; i += NT) s_y[i] = rhs[i];
    // slots >= P keep pivot-column and pivot-row entries 0: the fast path
    // updates their (zero, unused) window entries without a test
    for (int i = tid; i < kF3BlkMaxP; i += NT) s_u[0][i] = s_u[1][i] = s_col[0][i] = s_col[1][i] = 0.0;
    if (tid == 0) s_flag = 0;
    // batch 0 straight into the ring; batch 1 into registers
    if (tid < kWorkers)
        for (int q = tid; q < per_batch; q += kWorkers) {
            const int s = q / (2 * P), part = q % (2 * P);
            ring_at(s, part < P ? 0 : 1, part < P ? part : part - P) = fetch(0, q);
        }
    double pf[kPf];
    int pf_so[kPf], pf_part[kPf];  // the thread's batch values: column offset and part, fixed
#pragma unroll
    for (int m = 0; m < kPf; ++m) {
        pf[m] = 0.0;
        const int q = tid + m * kWorkers;
        const bool v = tid < kWorkers && q < per_batch;
        pf_so[m] = v ? q / (2 * P) : -1;
        pf_part[m] = v ? q % (2 * P) : 0;
        if (v) pf[m] = fetch(kF3BlkD, q);
    }
    int land_mod = (kF3BlkD + 1) % P;  // (k + 1) % P at the next landing k, kept without division
    __syncthreads();
    int kc = 0, kq = 0, kr = 0;  // kc = k mod P = kq B + kr
    for (int k = 0; k < n; ++k) {
        const int buf = k & 1;
        if ((k & (kF3BlkD - 1)) == 0 && k > 0) {  // batch k / D lands; fetch the next
#pragma unroll
            for (int m = 0; m < kPf; ++m) {
                if (pf_so[m] < 0) continue;
                const int part = pf_part[m];
                ring_at(k + pf_so[m], part < P ? 0 : 1, part < P ? part : part - P) = pf[m];
            }
            int nm = land_mod + kF3BlkD;  // (k + D + 1) % P: the next batch's first base slot
            while (nm >= P) nm -= P;
            land_mod = nm;
#pragma unroll
            for (int m = 0; m < kPf; ++m) {
                if (pf_so[m] < 0) continue;
                const int so = pf_so[m], part = pf_part[m];
                const int s = k + kF3BlkD + so, T = s + P, base = s + 1;
                int bslot = nm + so;
                while (bslot >= P) bslot -= P;
                int idx;
                bool ok;
                if (part < P) {
                    int d = part - bslot;
                    if (d < 0) d += P;
                    const int j = base + d;
                    ok = T < n && j < n;
                    idx = T * W + (j - T + hbw);
                } else {
                    int d = part - P - bslot;
                    if (d < 0) d += P;
                    const int i = base + d;
                    ok = T < n && i < n && i != T;
                    idx = i * W + (T - i + hbw);
                }
                const double v = band[ok ? idx : 0];
                pf[m] = ok ? v : 0.0;
            }
        }
        // the pivot slot kc = RB rq + rr = kq Bc + kr: its row lies in block
        // row rq (offset rr), its column in column block c = kq of the lanes
        // with bc == kr; rr and kq are the same for every lane
        const int rq = kc / RB, rr = kc % RB;
        const bool in_row = owner && br == rq, in_col = owner && bc == kr;
        // (A) pivot row, pivot column and the pivot's reciprocal to LDS.
        // Slots >= P carry values no valid slot ever reads, so nothing here
        // tests for them.
        switch (rr) {
#define F3_ROW(R)                                                                  \
    case R:                                                                        \
        if constexpr (R < RB) {                                                    \
            if (in_row) {                                                          \
                s_u[buf][bc] = a[R][0];                                            \
                s_u[buf][bc + Bc] = a[R][1];                                       \
                s_u[buf][bc + 2 * Bc] = a[R][2];                                   \
                s_u[buf][bc + 3 * Bc] = a[R][3];                                   \
            }                                                                      \
        }                                                                          \
        break;
            F3_ROW(0) F3_ROW(1) F3_ROW(2) F3_ROW(3) F3_ROW(4) F3_ROW(5) F3_ROW(6) F3_ROW(7)
#undef F3_ROW
        }
        switch (kq) {
#define F3_COL(C)                                                                  \
    case C:                                                                        \
        if (in_col) {                                                              \
            _Pragma("unroll") for (int r = 0; r < RB; ++r) {                       \
                s_col[buf][RB * br + r] = a[r][C];                                 \
                if (in_row && r == rr) s_rdiag[buf] = 1.0 / a[r][C];               \
            }                                                                      \
        }                                                                          \
        break;
            F3_COL(0) F3_COL(1) F3_COL(2) F3_COL(3)
#undef F3_COL
        }
        lds_barrier();
        const double diag = s_u[buf][kc];  // the pivot row's entry at the pivot column
        if (fabs(diag) <= 1.0e-14) {
            if (tid == 0) status[0] = 2.0;
            return;
        }
        const double rdiag = s_rdiag[buf];
        // (C) rank-1 update with the multipliers l = a_ik / a_kk (as a_ik
        // times the reciprocal: within one ulp of the reference's quotient);
        // then the retired row and column slots take row and column k + P
        // (the update wrote junk there)
        if (owner) {
            double l[RB], u[4];
#pragma unroll
            for (int r = 0; r < RB; ++r) l[r] = s_col[buf][RB * br + r] * rdiag;
#pragma unroll
            for (int c = 0; c < 4; ++c) u[c] = s_u[buf][bc + c * Bc];
#pragma unroll
            for (int r = 0; r < RB; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) a[r][c] -= l[r] * u[c];
            switch (rr) {
#define F3_ROWIN(R)                                                                \
    case R:                                                                        \
        if constexpr (R < RB) {                                                    \
            if (in_row) {                                                          \
                a[R][0] = ring_at(k, 0, bc);                                       \
                a[R][1] = ring_at(k, 0, bc + Bc);                                  \
                a[R][2] = ring_at(k, 0, bc + 2 * Bc);                              \
                a[R][3] = ring_at(k, 0, bc + 3 * Bc);                              \
            }                                                                      \
        }                                                                          \
        break;
                F3_ROWIN(0) F3_ROWIN(1) F3_ROWIN(2) F3_ROWIN(3) F3_ROWIN(4) F3_ROWIN(5) F3_ROWIN(6) F3_ROWIN(7)
#undef F3_ROWIN
            }
            switch (kq) {
#define F3_COLIN(C)                                                                \
    case C:                                                                        \
        if (in_col) {                                                              \
            _Pragma("unroll") for (int r = 0; r < RB; ++r)                         \
                if (!(in_row && r == rr)) a[r][C] = ring_at(k, 1, RB * br + r);    \
        }                                                                          \
        break;
                F3_COLIN(0) F3_COLIN(1) F3_COLIN(2) F3_COLIN(3)
#undef F3_COLIN
            }
        } else if (tid >= kWorkers) {
            // wavefront 15: forward substitution (the same multipliers), the
            // reference's dense-fallback test, row k of U back to the band
            const double yk = s_y[k];
            for (int rs = tid - kWorkers; rs < P; rs += 64) {
                const int d = rs >= kc ? rs - kc : rs - kc + P;
                if (d != 0) {
                    const double v = s_col[buf][rs];
                    if (fabs(v) > fabs(diag)) s_flag = 1;
                    if (k + d < n) s_y[k + d] -= (v * rdiag) * yk;
                }
                if (k + d < n) band[(size_t)k * W + (d + hbw)] = s_u[buf][rs];
            }
        }
        if (++kr == Bc) {
            kr = 0;
            ++kq;
        }
        if (++kc == P) {
            kc = 0;
            kq = 0;
            kr = 0;
        }
    }
    if (tid >= kWorkers) __threadfence();  // U rows visible to the staging loads
    __syncthreads();
    // Back substitution by columns in wavefront 0: at row i lane l holds
    // y[i - l] and y[i - 64 - l]; x_i = y_i / U_ii, then every lane takes
    // U(row, i) x_i off its rows and the window moves down one row (DPP lane
    // shift; y[i - 128] enters from LDS, untouched so far because hbw < 128).
    // The other wavefronts stage U columns (diagonal first) kF3BlkR at a time.
    const int R = kF3BlkR, collen = kF3BlkMaxP;
    auto stage = [&](int batch, int half) {
        const int top = n - 1 - batch * R;  // columns top, top - 1, ..., top - R + 1
        for (int q = tid - 64; q < R * P; q += NT - 64) {
            const int rr = q / P, c = q % P, i = top - rr;
            double v = 0.0;
            if (i >= 0 && i - c >= 0) v = band[(size_t)(i - c) * W + (c + hbw)];
            s_ring[(half * R + rr) * collen + c] = v;
        }
    };
    const int n_batches = (n + R - 1) / R;
    if (tid >= 64) stage(0, 0);
    __syncthreads();
    const int lane = tid & 63;
    double ya = 0.0, yb = 0.0;
    if (tid < 64) {
        ya = n - 1 - lane >= 0 ? s_y[n - 1 - lane] : 0.0;
        yb = n - 65 - lane >= 0 ? s_y[n - 65 - lane] : 0.0;
    }
    for (int b = 0; b < n_batches; ++b) {
        const int half = b & 1;
        if (tid >= 64) {
            if (b + 1 < n_batches) stage(b + 1, half ^ 1);
        } else {
            const int top = n - 1 - b * R;
            for (int rr = 0; rr < R; ++rr) {
                const int i = top - rr;
                if (i < 0) break;
                const double *col = s_ring + (half * R + rr) * collen;
                const double ua = lane >= 1 && lane <= hbw ? col[lane] : 0.0;
                const double ub = lane + 64 <= hbw ? col[64 + lane] : 0.0;
                const double yin = i - 128 >= 0 ? s_y[i - 128] : 0.0;
                const double xi = lane_value(ya, 0) / col[0];
                ya -= ua * xi;
                yb -= ub * xi;
                if (lane == 0) s_y[i] = xi;  // y_i is spent: x_i takes its place
                const double yb0 = lane_value(yb, 0);
                ya = wave_next_lane(ya);
                yb = wave_next_lane(yb);
                if (lane == 63) {
                    ya = yb0;
                    yb = yin;
                }
            }
        }
        lds_barrier();
    }
    for (int i = tid; i < n; i += NT) x[i] = s_y[i];
    if (tid == 0) status[0] = s_flag ? 1.0 : 0.0;
}

// recorded DOFs after a converged step: record r's row `row` of its history
__global__ void f3_record_kernel(const int32_t *recs, int n_rec, int total, const int32_t *dofs, const double *u,
                                 double *hist, int64_t rec_stride, const int32_t *row_of) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_rec) return;
    const int r = recs[blockIdx.y];
    hist[(size_t)r * rec_stride + (size_t)row_of[r] * n_rec + c] = u[(size_t)r * total + dofs[c]];
}

__global__ void f3_update_kernel(const int32_t *recs, int n, double *uf, const double *du) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)recs[blockIdx.y] * n + i;
    uf[o] += du[o];
}

// :594-647 converged step: a and v from the final displacement
__global__ void f3_newmark_kernel(const int32_t *recs, int n, const double *uf, const double *un, const double *vn,
                                  const double *an, double *af, double *vf, double a0, double a2, double a3,
                                  double gamma, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)recs[blockIdx.y] * n + i;
    const double anext = a0 * (uf[o] - un[o]) - a2 * vn[o] - a3 * an[o];
    af[o] = anext;
    vf[o] = vn[o] + dt * ((1.0 - gamma) * an[o] + gamma * anext);
}

void frame3d_host_free(Frame3dHost *h) { delete h; }  // device buffers live in ctx->allocations

namespace {

inline dim3 grid_of(int64_t n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

// time_series.mojo:71-96 (Path with dt, start 0, no useLast)
double path_at(const std::vector<double> &v, double dt, double factor, double t) {
    const int n = (int)v.size();
    if (n == 0 || t < 0.0) return 0.0;
    const double incr = t / dt;
    const long i1 = (long)std::floor(incr);
    if (i1 + 1 >= n) return 0.0;
    const double v1 = v[i1], v2 = v[i1 + 1];
    return factor * (v1 + (v2 - v1) * (incr - (double)i1));
}

// Band factor + solve on the stream: the register-window kernel when the
// half bandwidth fits it (variant 0 or 2), else the global-memory one.
// recs == nullptr: one system at the given pointers; else `count` systems,
// record r's band, right-hand side, solution and status at its slices.
// A record's element list: every id an sgpu_elem3d element, none named twice
// within the list or by an earlier record (two workgroups writing one
// element's section states would race).  Claims the ids on success.
int claim_elements(sgpu_ctx *ctx, Frame3dHost &h, const int32_t *ids, int n_el) {
    const int count = elem3d_count(ctx);
    if ((int)h.elem_used.size() != count) h.elem_used.resize(count, 0);
    std::vector<uint8_t> mine(count, 0);
    for (int e = 0; e < n_el; ++e) {
        const int id = ids[e];
        if (id < 0 || id >= count) return fail(ctx, SGPU_ERR_ARG, "element %d: id %d is not an sgpu_elem3d element", e, id);
        if (mine[id] || h.elem_used[id])
            return fail(ctx, SGPU_ERR_ARG, "element %d: id %d is already used by this or another record", e, id);
        mine[id] = 1;
    }
    for (int i = 0; i < count; ++i) h.elem_used[i] |= mine[i];
    return SGPU_OK;
}

int launch_band_solve(sgpu_ctx *ctx, int variant, int n, int hbw, double *d_band, const double *d_rhs, double *d_x,
                      double *d_status, const int32_t *recs = nullptr, int count = 1, double *d_work = nullptr) {
    if (variant == 0 || variant == 5 || variant == 6 || variant == 8)  // the reference's row swaps
        return launch_band_piv(ctx, variant == 0 ? 0 : variant == 8 ? 3 : variant - 4, n, hbw, d_band, d_rhs, d_x,
                               d_status, nullptr, nullptr, d_work, recs, count);
    if (variant == 7) variant = 0;  // no row swaps, kernel picked by size
    const size_t blk_lds = sizeof(double) * ((size_t)kF3BlkRing + n);
    const int P = hbw + 1, Bc = (P + 3) / 4;
    const bool fits8 = ((P + 7) / 8) * Bc <= 512 - 64, fits4 = ((P + 3) / 4) * Bc <= 1024 - 64;
    // dynamic LDS stays under 56 KiB: the kernels' static arrays take the rest of 64 KiB
    const bool lds_ok = P <= kF3BlkMaxP && blk_lds <= 56 * 1024;
    if (variant == 0) variant = lds_ok && fits4 ? 3 : hbw + 1 <= kF3RegMaxP ? 2 : 1;
    if (variant == 3 || variant == 4) {
        if (!lds_ok || (variant == 4 ? !fits8 : !fits4))
            return fail(ctx, SGPU_ERR_ARG, "band %d x %d exceeds the blocked solve (half bandwidth <= %d, LDS)", n,
                        2 * hbw + 1, variant == 4 ? 115 : kF3BlkMaxP - 1);
        if (variant == 4)
            hipLaunchKernelGGL((f3_solve_blk_kernel<8, 512>), dim3(count), dim3(512), blk_lds, ctx->stream, recs, n,
                               hbw, d_band, d_rhs, d_x, d_status);
        else
            hipLaunchKernelGGL((f3_solve_blk_kernel<4, 1024>), dim3(count), dim3(1024), blk_lds, ctx->stream, recs, n,
                               hbw, d_band, d_rhs, d_x, d_status);
        HIP_OK(ctx, hipGetLastError());
        return SGPU_OK;
    }
    const bool reg = variant == 2;
    if (reg && hbw + 1 > kF3RegMaxP)
        return fail(ctx, SGPU_ERR_ARG, "half bandwidth %d exceeds the register-window solve (%d)", hbw,
                    kF3RegMaxP - 1);
    const size_t lds = sizeof(double) * (size_t)(reg ? n : n + 2 * hbw);
    if (lds > 56 * 1024) return fail(ctx, SGPU_ERR_ARG, "band of %d rows: the solve keeps them in LDS", n);
    if (reg)
        hipLaunchKernelGGL(f3_solve_reg_kernel, dim3(count), dim3(kF3SolveThreads), lds, ctx->stream, recs, n, hbw,
                           d_band, d_rhs, d_x, d_status);
    else
        hipLaunchKernelGGL(f3_solve_kernel, dim3(count), dim3(kF3SolveThreads), lds, ctx->stream, recs, n, hbw, d_band,
                           d_rhs, d_x, d_status);
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

}  // namespace
}  // namespace sgpu

using namespace sgpu;

int sgpu_band_solve(sgpu_ctx *ctx, int variant, int n, int hbw, const double *band, const double *rhs, double *x,
                    int32_t *info) {
    if (!ctx || !band || !rhs || !x || !info || n <= 0 || hbw < 0 || hbw > n - 1 || variant < 0 || variant > 8)
        return SGPU_ERR_ARG;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const size_t slots = (size_t)n * (2 * hbw + 1);
    const bool global_piv = variant == 6 || variant == 8 || (variant == 0 && !band_piv_reg_fits(n, hbw));
    const size_t work = global_piv ? band_blk_work_doubles(n, hbw) : 0;
    double *d = nullptr;
    HIP_OK(ctx, hipMalloc((void **)&d, sizeof(double) * (slots + 2 * (size_t)n + 4 + work)));
    double *d_rhs = d + slots, *d_x = d_rhs + n, *d_st = d_x + n, *d_work = global_piv ? d_st + 4 : nullptr;
    int rc = SGPU_OK;
    if (hipMemcpy(d, band, sizeof(double) * slots, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_rhs, rhs, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(ctx, SGPU_ERR_HIP, "band upload failed");
    } else if ((rc = launch_band_solve(ctx, variant, n, hbw, d, d_rhs, d_x, d_st, nullptr, 1, d_work)) == SGPU_OK) {
        double st = 0.0;
        if (hipMemcpyAsync(x, d_x, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(&st, d_st, sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = fail(ctx, SGPU_ERR_HIP, "band solve readback failed");
        *info = (int32_t)st;
    }
    (void)hipFree(d);
    return rc;
}

int sgpu_frame3d_create(sgpu_ctx *ctx, const sgpu_frame3d_desc *d, int32_t *hbw_out, int32_t *free_count_out) {
    if (!ctx || !d) return SGPU_ERR_ARG;
    if (!ctx->finalized) return fail(ctx, SGPU_ERR_STATE, "sgpu_sections_finalize has not run");
    if (ctx->frame3d) return fail(ctx, SGPU_ERR_STATE, "the context already holds a 3D frame");
    if (d->n_nodes <= 0 || d->n_elem <= 0 || !d->sp_constrained || !d->mass || !d->row_offsets || !d->slave ||
        !d->elem_ids || !d->elem_nodes)
        return fail(ctx, SGPU_ERR_ARG, "bad 3D frame description");
    if (d->n_patterns < 0 || d->n_patterns > SGPU_F3_MAX_PATTERNS)
        return fail(ctx, SGPU_ERR_ARG, "at most %d excitation patterns", SGPU_F3_MAX_PATTERNS);
    if (d->beta <= 0.0 || d->dt <= 0.0 || d->max_iters <= 0) return fail(ctx, SGPU_ERR_ARG, "bad Newmark parameters");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const int nn = d->n_nodes, total = 6 * nn, n_el = d->n_elem;
    const int32_t *ro = d->row_offsets, *dp = d->dof_pool;
    const double *cp = d->coeff_pool;
    for (int i = 0; i < total; ++i) {
        if (ro[i + 1] < ro[i]) return fail(ctx, SGPU_ERR_ARG, "bad expansion row offsets");
        for (int t = ro[i]; t < ro[i + 1]; ++t)
            if (dp[t] < 0 || dp[t] >= total) return fail(ctx, SGPU_ERR_ARG, "expansion term %d out of range", t);
    }
    for (int e = 0; e < n_el; ++e)
        for (int a = 0; a < 2; ++a)
            if (d->elem_nodes[2 * e + a] < 0 || d->elem_nodes[2 * e + a] >= nn)
                return fail(ctx, SGPU_ERR_ARG, "element %d: node index out of range", e);
    const int n_sh = d->n_shell;
    if (n_sh < 0 || (n_sh > 0 && (!d->shell_nodes || !d->shell_xyz || !d->shell_props)))
        return fail(ctx, SGPU_ERR_ARG, "bad shell description");
    for (int e = 0; e < n_sh; ++e) {
        for (int a = 0; a < 4; ++a)
            if (d->shell_nodes[4 * e + a] < 0 || d->shell_nodes[4 * e + a] >= nn)
                return fail(ctx, SGPU_ERR_ARG, "shell %d: node index out of range", e);
        const double nu = d->shell_props[3 * e + 1];
        if (!(nu > -1.0 && nu < 1.0)) return fail(ctx, SGPU_ERR_ARG, "shell %d: Poisson's ratio %g", e, nu);
    }
    {  // record 0's element ids, checked before anything is built (claimed below)
        Frame3dHost probe;
        int rc0 = claim_elements(ctx, probe, d->elem_ids, n_el);
        if (rc0) return rc0;
    }
    // loader.mojo:2378-2388, :2443-2450: free DOFs in plain order, slaves out
    std::vector<int32_t> free_idx(total, -1), free_list;
    for (int i = 0; i < total; ++i) {
        const bool con = d->sp_constrained[i] || (d->has_mpc && d->slave[i]);
        if (!con) {
            free_idx[i] = (int32_t)free_list.size();
            free_list.push_back(i);
        }
    }
    const int n = (int)free_list.size();
    if (n == 0) return fail(ctx, SGPU_ERR_ARG, "no free DOFs");
    for (int i = 0; i < total; ++i)  // slaves read only free or fixed retained DOFs
        if (d->has_mpc && d->slave[i])
            for (int t = ro[i]; t < ro[i + 1]; ++t)
                if (d->slave[dp[t]]) return fail(ctx, SGPU_ERR_ARG, "slave DOF %d retains slave DOF %d", i, dp[t]);
    // per element: free retained DOFs and T_e
    std::vector<int32_t> ret((size_t)n_el * 12, 0), ret_n(n_el, 0);
    std::vector<double> T((size_t)n_el * 144, 0.0);
    int hbw = 0;
    for (int e = 0; e < n_el; ++e) {
        int ni = d->elem_nodes[2 * e], nj = d->elem_nodes[2 * e + 1];
        std::vector<int32_t> R;
        int fmap[12];
        for (int a = 0; a < 12; ++a) {
            const int g = a < 6 ? 6 * ni + a : 6 * nj + a - 6;
            fmap[a] = free_idx[g];
            const bool expand = d->has_mpc && d->slave[g];
            const int t0 = expand ? ro[g] : 0, t1 = expand ? ro[g + 1] : (free_idx[g] >= 0 ? 1 : 0);
            for (int t = t0; t < t1; ++t) {
                const int r = expand ? free_idx[dp[t]] : free_idx[g];
                const double c = expand ? cp[t] : 1.0;
                if (r < 0) continue;
                int pos = (int)(std::find(R.begin(), R.end(), r) - R.begin());
                if (pos == (int)R.size()) {
                    if ((int)R.size() == kF3MaxRet)
                        return fail(ctx, SGPU_ERR_ARG, "element %d reaches more than %d free DOFs", e, kF3MaxRet);
                    R.push_back(r);
                }
                T[(size_t)e * 144 + a * 12 + pos] += c;
            }
        }
        ret_n[e] = (int32_t)R.size();
        for (size_t q = 0; q < R.size(); ++q) ret[(size_t)e * 12 + q] = R[q];
        if (d->hbw_rule == 0) {  // lsb.mojo:1481-1506 on the element free map
            for (int a = 0; a < 12; ++a)
                for (int b = a + 1; b < 12; ++b)
                    if (fmap[a] >= 0 && fmap[b] >= 0) hbw = std::max(hbw, std::abs(fmap[a] - fmap[b]));
        } else {
            for (size_t a = 0; a < R.size(); ++a)
                for (size_t b = a + 1; b < R.size(); ++b) hbw = std::max(hbw, std::abs(R[a] - R[b]));
        }
    }
    // per shell: free retained DOFs and T_s (24 x 24), the band from its
    // 24 DOFs as for the beam-columns
    std::vector<int32_t> sh_ret((size_t)n_sh * 24, 0), sh_ret_n(n_sh, 0);
    std::vector<double> sh_T((size_t)n_sh * 576, 0.0);
    for (int e = 0; e < n_sh; ++e) {
        std::vector<int32_t> R;
        int fmap[24];
        for (int a = 0; a < 24; ++a) {
            const int g = 6 * d->shell_nodes[4 * e + a / 6] + a % 6;
            fmap[a] = free_idx[g];
            const bool expand = d->has_mpc && d->slave[g];
            const int t0 = expand ? ro[g] : 0, t1 = expand ? ro[g + 1] : (free_idx[g] >= 0 ? 1 : 0);
            for (int t = t0; t < t1; ++t) {
                const int r = expand ? free_idx[dp[t]] : free_idx[g];
                const double c = expand ? cp[t] : 1.0;
                if (r < 0) continue;
                int pos = (int)(std::find(R.begin(), R.end(), r) - R.begin());
                if (pos == (int)R.size()) {
                    if ((int)R.size() == kF3ShellMaxRet)
                        return fail(ctx, SGPU_ERR_ARG, "shell %d reaches more than %d free DOFs", e, kF3ShellMaxRet);
                    R.push_back(r);
                }
                sh_T[(size_t)e * 576 + a * 24 + pos] += c;
            }
        }
        sh_ret_n[e] = (int32_t)R.size();
        for (size_t q = 0; q < R.size(); ++q) sh_ret[(size_t)e * 24 + q] = R[q];
        if (d->hbw_rule == 0) {
            for (int a = 0; a < 24; ++a)
                for (int b = a + 1; b < 24; ++b)
                    if (fmap[a] >= 0 && fmap[b] >= 0) hbw = std::max(hbw, std::abs(fmap[a] - fmap[b]));
        } else {
            for (size_t a = 0; a < R.size(); ++a)
                for (size_t b = a + 1; b < R.size(); ++b) hbw = std::max(hbw, std::abs(R[a] - R[b]));
        }
    }
    // helpers.mojo:3304-3342: reduced mass, in the reference's addition order
    std::vector<std::vector<std::pair<int, double>>> mrow(n);
    auto madd = [&](int r, int c, double v) {
        for (auto &pr : mrow[r])
            if (pr.first == c) {
                pr.second += v;
                return;
            }
        mrow[r].push_back({c, v});
    };
    for (int i = 0; i < total; ++i) {
        const double mass = d->mass[i];
        if (mass == 0.0) continue;
        const bool expand = d->has_mpc && d->slave[i];
        if (!expand) {
            if (free_idx[i] >= 0) madd(free_idx[i], free_idx[i], mass * 1.0 * 1.0);
            continue;
        }
        for (int ti = ro[i]; ti < ro[i + 1]; ++ti) {
            const int ri = free_idx[dp[ti]];
            if (ri < 0) continue;
            for (int tj = ro[i]; tj < ro[i + 1]; ++tj) {
                const int rj = free_idx[dp[tj]];
                if (rj < 0) continue;
                madd(ri, rj, mass * cp[ti] * cp[tj]);
            }
        }
    }
    if (d->hbw_rule != 0)
        for (int r = 0; r < n; ++r)
            for (auto &pr : mrow[r]) hbw = std::max(hbw, std::abs(pr.first - r));
    hbw = std::min(hbw, n - 1);
    const int W = 2 * hbw + 1;
    const int64_t slots = (int64_t)n * W;
    if (slots > (int64_t)1 << 31) return fail(ctx, SGPU_ERR_ARG, "band of %d x %d is too large", n, W);
    std::vector<int32_t> moff(n + 1, 0), mcol;
    std::vector<double> mval, mass_band(slots, 0.0);
    for (int r = 0; r < n; ++r) {
        std::sort(mrow[r].begin(), mrow[r].end());
        moff[r] = (int32_t)mcol.size();
        for (auto &pr : mrow[r]) {
            mcol.push_back(pr.first);
            mval.push_back(pr.second);
            if (std::abs(pr.first - r) <= hbw) mass_band[(size_t)r * W + (pr.first - r + hbw)] = pr.second;
        }
    }
    moff[n] = (int32_t)mcol.size();
    // band slot gather lists, element order
    std::vector<int32_t> cnt(slots + 1, 0);
    for (int e = 0; e < n_el; ++e)
        for (int a = 0; a < ret_n[e]; ++a)
            for (int b = 0; b < ret_n[e]; ++b) {
                const int r = ret[(size_t)e * 12 + a], s = ret[(size_t)e * 12 + b];
                if (std::abs(r - s) <= hbw) cnt[(size_t)r * W + (s - r + hbw) + 1]++;
            }
    for (int e = 0; e < n_sh; ++e)
        for (int a = 0; a < sh_ret_n[e]; ++a)
            for (int b = 0; b < sh_ret_n[e]; ++b) {
                const int r = sh_ret[(size_t)e * 24 + a], s = sh_ret[(size_t)e * 24 + b];
                if (std::abs(r - s) <= hbw) cnt[(size_t)r * W + (s - r + hbw) + 1]++;
            }
    for (int64_t s = 0; s < slots; ++s) cnt[s + 1] += cnt[s];
    std::vector<int32_t> band_src(cnt[slots]), fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < n_el; ++e)
        for (int a = 0; a < ret_n[e]; ++a)
            for (int b = 0; b < ret_n[e]; ++b) {
                const int r = ret[(size_t)e * 12 + a], s = ret[(size_t)e * 12 + b];
                if (std::abs(r - s) <= hbw) band_src[fill[(size_t)r * W + (s - r + hbw)]++] = e * 144 + a * 12 + b;
            }
    for (int e = 0; e < n_sh; ++e)  // shells after the beam-columns, negative sources
        for (int a = 0; a < sh_ret_n[e]; ++a)
            for (int b = 0; b < sh_ret_n[e]; ++b) {
                const int r = sh_ret[(size_t)e * 24 + a], s = sh_ret[(size_t)e * 24 + b];
                if (std::abs(r - s) <= hbw)
                    band_src[fill[(size_t)r * W + (s - r + hbw)]++] = -(e * 576 + a * 24 + b) - 1;
            }
    // per free DOF gather lists
    std::vector<int32_t> doff(n + 1, 0), dsrc;
    {
        std::vector<std::vector<int32_t>> lists(n);
        for (int e = 0; e < n_el; ++e)
            for (int a = 0; a < ret_n[e]; ++a) lists[ret[(size_t)e * 12 + a]].push_back(e * 12 + a);
        for (int e = 0; e < n_sh; ++e)
            for (int a = 0; a < sh_ret_n[e]; ++a) lists[sh_ret[(size_t)e * 24 + a]].push_back(-(e * 24 + a) - 1);
        for (int r = 0; r < n; ++r) {
            doff[r] = (int32_t)dsrc.size();
            dsrc.insert(dsrc.end(), lists[r].begin(), lists[r].end());
        }
        doff[n] = (int32_t)dsrc.size();
    }
    // collapsed inertia load per pattern: T^T (-M iota_p) on the free DOFs
    std::vector<double> load((size_t)std::max(d->n_patterns, 1) * n, 0.0);
    for (int p = 0; p < d->n_patterns; ++p) {
        const int dir = d->exc_dir[p];
        if (dir < 1 || dir > 3) return fail(ctx, SGPU_ERR_ARG, "UniformExcitation direction %d out of range", dir);
        if (!d->ts_values[p] || d->ts_n[p] <= 0 || d->ts_dt[p] <= 0.0)
            return fail(ctx, SGPU_ERR_ARG, "pattern %d: bad Path series", p);
        for (int i = 0; i < total; ++i) {
            if ((i % 6) + 1 != dir || d->mass[i] == 0.0) continue;
            const double li = -d->mass[i];
            const bool expand = d->has_mpc && d->slave[i];
            if (!expand) {
                if (free_idx[i] >= 0) load[(size_t)p * n + free_idx[i]] += li;
                continue;
            }
            for (int t = ro[i]; t < ro[i + 1]; ++t) {
                const int r = free_idx[dp[t]];
                if (r >= 0) load[(size_t)p * n + r] += cp[t] * li;
            }
        }
    }
    auto *h = new Frame3dHost();
    ctx->frame3d = h;
    h->n_nodes = nn;
    h->total = total;
    h->n = n;
    h->hbw = hbw;
    h->W = W;
    h->n_el = n_el;
    h->dt = d->dt;
    h->gamma = d->gamma;
    h->beta = d->beta;
    h->tol = d->tol;
    h->alpha_m = d->alpha_m;
    h->beta_k = d->beta_k;
    h->beta_k_init = d->beta_k_init;
    h->max_iters = d->max_iters;
    h->n_patterns = d->n_patterns;
    h->n_sh = n_sh;
    h->h_free_list = free_list;
    h->h_constrained.assign(total, 0);
    for (int i = 0; i < total; ++i) h->h_constrained[i] = free_idx[i] < 0 ? 1 : 0;
    h->h_ro.assign(ro, ro + total + 1);
    h->h_dp.assign(dp, dp + ro[total]);
    h->h_cp.assign(cp, cp + ro[total]);
    h->h_slave.assign(total, 0);
    for (int i = 0; i < total; ++i) h->h_slave[i] = d->has_mpc && d->slave[i] ? 1 : 0;
    h->h_moff = moff;
    h->h_mcol = mcol;
    h->h_mval = mval;
    std::vector<int32_t> slave(total), rov(ro, ro + total + 1), dpv(dp, dp + ro[total]), en(d->elem_nodes,
                                                                                           d->elem_nodes + 2 * n_el);
    std::vector<double> cpv(cp, cp + ro[total]);
    for (int i = 0; i < total; ++i) slave[i] = d->has_mpc && d->slave[i] ? 1 : 0;
    int rc;
    if ((rc = upload(ctx, &h->d_free_idx, free_idx)) || (rc = upload(ctx, &h->d_row_offsets, rov)) ||
        (rc = upload(ctx, &h->d_dof_pool, dpv)) || (rc = upload(ctx, &h->d_coeff_pool, cpv)) ||
        (rc = upload(ctx, &h->d_slave, slave)) || (rc = upload(ctx, &h->d_elem_nodes, en)) ||
        (rc = upload(ctx, &h->d_ret, ret)) || (rc = upload(ctx, &h->d_ret_n, ret_n)) ||
        (rc = upload(ctx, &h->d_T, T)) || (rc = upload(ctx, &h->d_band_off, cnt)) ||
        (rc = upload(ctx, &h->d_band_src, band_src)) || (rc = upload(ctx, &h->d_mass_band, mass_band)) ||
        (rc = upload(ctx, &h->d_dof_off, doff)) || (rc = upload(ctx, &h->d_dof_src, dsrc)) ||
        (rc = upload(ctx, &h->d_m_off, moff)) || (rc = upload(ctx, &h->d_m_col, mcol)) ||
        (rc = upload(ctx, &h->d_m_val, mval)) || (rc = upload(ctx, &h->d_load, load)))
        return rc;
    if (n_sh > 0) {  // shell stiffness on the device, once, and its collapse
        std::vector<double> xyz(d->shell_xyz, d->shell_xyz + 12 * (size_t)n_sh),
            props(d->shell_props, d->shell_props + 3 * (size_t)n_sh);
        double *d_xyz = nullptr, *d_props = nullptr, *d_T = nullptr, *d_k = nullptr;
        int32_t *d_st = nullptr;
        if ((rc = upload(ctx, &h->d_sh_ret, sh_ret)) || (rc = upload(ctx, &h->d_sh_ret_n, sh_ret_n)) ||
            (rc = upload(ctx, &d_xyz, xyz)) || (rc = upload(ctx, &d_props, props)) || (rc = upload(ctx, &d_T, sh_T)) ||
            (rc = dalloc(ctx, &d_k, (size_t)n_sh * 576)) || (rc = dalloc(ctx, &h->d_sh_khat, (size_t)n_sh * 576)) ||
            (rc = dalloc(ctx, &d_st, (size_t)n_sh)))
            return rc;
        if ((rc = launch_shell4_stiffness(ctx, n_sh, d_xyz, d_props, d_k, d_st))) return rc;
        std::vector<int32_t> st(n_sh);
        HIP_OK(ctx, hipMemcpyAsync(st.data(), d_st, sizeof(int32_t) * n_sh, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
        for (int e = 0; e < n_sh; ++e)
            if (st[e])
                return fail(ctx, SGPU_ERR_ARG, "shell %d: %s", e, st[e] == 1 ? "degenerate basis" : "singular Jacobian");
        hipLaunchKernelGGL(f3_shell_collapse_kernel, dim3(n_sh), dim3(576), 0, ctx->stream, d_T, d_k, h->d_sh_khat);
        HIP_OK(ctx, hipGetLastError());
    }
    int32_t rec0;
    if ((rc = sgpu_frame3d_add_record(ctx, d->elem_ids, d->ts_values, d->ts_n, d->ts_dt, d->ts_factor, &rec0)))
        return rc;
    if (hbw_out) *hbw_out = hbw;
    if (free_count_out) *free_count_out = n;
    return SGPU_OK;
}

int sgpu_frame3d_add_record(sgpu_ctx *ctx, const int32_t *elem_ids, const double *const *ts_values,
                            const int32_t *ts_n, const double *ts_dt, const double *ts_factor, int32_t *record) {
    if (!ctx || !elem_ids) return SGPU_ERR_ARG;
    Frame3dHost *h = ctx->frame3d;
    if (!h) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame3d_create has not run");
    if (h->allocated) return fail(ctx, SGPU_ERR_STATE, "records are added before the first run");
    if (h->n_rec > 0) return fail(ctx, SGPU_ERR_STATE, "records are added before the recorder is set");
    for (int p = 0; p < h->n_patterns; ++p)
        if (!ts_values || !ts_values[p] || !ts_n || ts_n[p] <= 0 || !ts_dt || ts_dt[p] <= 0.0 || !ts_factor)
            return fail(ctx, SGPU_ERR_ARG, "pattern %d: bad Path series", p);
    int rc = claim_elements(ctx, *h, elem_ids, h->n_el);
    if (rc) return rc;
    h->rec_elem_ids.insert(h->rec_elem_ids.end(), elem_ids, elem_ids + h->n_el);
    for (int p = 0; p < h->n_patterns; ++p) {
        h->rec_ts.emplace_back(ts_values[p], ts_values[p] + ts_n[p]);
        h->rec_ts_dt.push_back(ts_dt[p]);
        h->rec_ts_factor.push_back(ts_factor[p]);
    }
    h->rec_steps.push_back(0);
    h->rec_status.push_back(0);
    if (record) *record = h->R;
    h->R += 1;
    return SGPU_OK;
}

namespace {

// device state of every record, allocated at the first run
int f3_allocate(sgpu_ctx *ctx, Frame3dHost &h) {
    if (h.allocated) return SGPU_OK;
    const size_t R = h.R, n = h.n, n_el = h.n_el, slots = (size_t)h.n * h.W;
    int rc;
    double **vecs[] = {&h.d_uf, &h.d_vf, &h.d_af, &h.d_un, &h.d_vn, &h.d_an, &h.d_vt, &h.d_at, &h.d_R, &h.d_du};
    for (double **p : vecs)
        if ((rc = dalloc(ctx, p, R * n))) return rc;
    if ((rc = dalloc(ctx, &h.d_u, R * h.total)) || (rc = dalloc(ctx, &h.d_u12, R * n_el * 12)) ||
        (rc = dalloc(ctx, &h.d_k, R * n_el * 144)) || (rc = dalloc(ctx, &h.d_f, R * n_el * 12)) ||
        (rc = dalloc(ctx, &h.d_khat, R * n_el * 144)) || (rc = dalloc(ctx, &h.d_kinit_hat, R * n_el * 144)) ||
        (rc = dalloc(ctx, &h.d_fhat, R * n_el * 12)) || (rc = dalloc(ctx, &h.d_yk, R * n_el * 12)) ||
        (rc = dalloc(ctx, &h.d_yi, R * n_el * 12)) || (rc = dalloc(ctx, &h.d_status, R * n_el)) ||
        (rc = dalloc(ctx, &h.d_band, R * slots)) || (rc = dalloc(ctx, &h.d_kinit_band, R * slots)) ||
        (rc = dalloc(ctx, &h.d_scalar, 4 * R)) || (rc = dalloc(ctx, &h.d_list, R)) ||
        (rc = dalloc(ctx, &h.d_list_ids, R * n_el)) || (rc = dalloc(ctx, &h.d_row_of, R)) ||
        (rc = dalloc(ctx, &h.d_ag, 3 * R)))
        return rc;
    if ((rc = dalloc(ctx, &h.d_bwork, R * band_blk_work_doubles(h.n, h.hbw)))) return rc;
    if (h.n_sh > 0 && ((rc = dalloc(ctx, &h.d_sh_f, R * h.n_sh * 24)) || (rc = dalloc(ctx, &h.d_sh_y, R * h.n_sh * 24))))
        return rc;
    for (double **p : vecs) HIP_OK(ctx, hipMemset(*p, 0, sizeof(double) * R * n));
    HIP_OK(ctx, hipMemset(h.d_u, 0, sizeof(double) * R * h.total));
    HIP_OK(ctx, hipMemset(h.d_kinit_hat, 0, sizeof(double) * R * n_el * 144));
    HIP_OK(ctx, hipMemset(h.d_ag, 0, sizeof(double) * 3 * R));
    h.allocated = true;
    return SGPU_OK;
}

// the listed records become the kernels' record list (and their element
// ids, in list order, the element pass's)
int f3_set_list(sgpu_ctx *ctx, Frame3dHost &h, const std::vector<int32_t> &list) {
    if (list == h.list_on_device) return SGPU_OK;  // already the kernels' list
    std::vector<int32_t> ids;
    ids.reserve(list.size() * h.n_el);
    for (int r : list) ids.insert(ids.end(), h.rec_elem_ids.begin() + (size_t)r * h.n_el,
                                  h.rec_elem_ids.begin() + (size_t)(r + 1) * h.n_el);
    HIP_OK(ctx, hipMemcpyAsync(h.d_list, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_OK(ctx, hipMemcpyAsync(h.d_list_ids, ids.data(), sizeof(int32_t) * ids.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));  // the host vectors go out of scope
    h.list_on_device = list;
    return SGPU_OK;
}

inline dim3 grid2(int64_t n, int count) { return dim3((unsigned)((n + 255) / 256), (unsigned)count); }

// one element state determination of the listed records at their current u,
// and its collapse
int f3_elements(sgpu_ctx *ctx, Frame3dHost &h, int count) {
    hipStream_t st = ctx->stream;
    const int64_t m = (int64_t)count * h.n_el;
    hipLaunchKernelGGL(f3_gather_u12_kernel, grid_of(m * 12), dim3(256), 0, st, h.d_list, count, h.n_el, h.total,
                       h.d_elem_nodes, h.d_u, h.d_u12);
    int rc;
    {
        ProfScope ps(ctx, SGPU_KERNEL_F3_ELEM);
        rc = sgpu_elem3d_state_device(ctx, (int)m, h.d_list_ids, h.d_u12, h.d_k, h.d_f, h.d_status);
    }
    if (rc) return rc;
    hipLaunchKernelGGL(f3_elem_collapse_kernel, dim3((unsigned)m), dim3(256), 0, st, h.d_list, h.n, h.n_el, h.d_ret,
                       h.d_ret_n, h.d_T, h.d_k, h.d_f, h.d_kinit_hat, h.d_vt, h.d_khat, h.d_fhat, h.d_yk, h.d_yi);
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

// the initial assembly of every record at u = 0 (:1207), its collapse
// (:1307-1318), the initial tangent (:1329-1337) and the revert (:1452-1453)
int f3_initialize(sgpu_ctx *ctx, Frame3dHost &h) {
    if (h.initialized) return SGPU_OK;
    int rc;
    if ((rc = f3_allocate(ctx, h))) return rc;
    std::vector<int32_t> all(h.R);
    for (int r = 0; r < h.R; ++r) all[r] = r;
    if ((rc = f3_set_list(ctx, h, all))) return rc;
    if ((rc = f3_elements(ctx, h, h.R))) return rc;
    hipStream_t st = ctx->stream;
    HIP_OK(ctx, hipMemcpyAsync(h.d_kinit_hat, h.d_khat, sizeof(double) * (size_t)h.R * h.n_el * 144,
                               hipMemcpyDeviceToDevice, st));
    const int64_t slots = (int64_t)h.n * h.W;
    hipLaunchKernelGGL(f3_band_kernel, grid2(slots, h.R), dim3(256), 0, st, h.d_list, slots, (int64_t)h.n_el * 144,
                       h.d_band_off, h.d_band_src, h.d_khat, h.d_sh_khat, h.d_mass_band, h.d_kinit_band,
                       h.d_kinit_band, 1, 0.0, 0.0, 0.0, 0.0, 0.0, 0);
    HIP_OK(ctx, hipGetLastError());
    if ((rc = sgpu_sections_revert_all(ctx))) return rc;
    h.initialized = true;
    return SGPU_OK;
}

// `steps` more steps of every running record, in lockstep by step; each
// record iterates until it converges, stops or fails.  iters[r * steps +
// step], dense_fallbacks[r] (may be NULL); u_hist (R == 1 only) as
// sgpu_frame3d_run documents.
int f3_run(sgpu_ctx *ctx, Frame3dHost &h, int steps, double *u_hist, int32_t *iters, int32_t *dense_fallbacks) {
    hipStream_t st = ctx->stream;
    const int n = h.n;
    const double dt = h.dt, gamma = h.gamma, beta = h.beta;
    const double a0 = 1.0 / (beta * dt * dt), a1 = gamma / (beta * dt), a2 = 1.0 / (beta * dt);
    const double a3 = 1.0 / (2.0 * beta) - 1.0;
    const int has_damping = h.alpha_m != 0.0 || h.beta_k != 0.0 || h.beta_k_init != 0.0;
    const int64_t slots = (int64_t)n * h.W;
    int rc;
    if ((rc = f3_initialize(ctx, h))) return rc;
    std::vector<double> scal(4 * (size_t)h.R), ag(3 * (size_t)h.R, 0.0);
    for (int step = 0; step < steps; ++step) {
        std::vector<int32_t> live;
        for (int r = 0; r < h.R; ++r)
            if (h.rec_status[r] == 0) live.push_back(r);
        if (live.empty()) break;
        for (int r : live) {
            const double t = (double)(h.rec_steps[r] + 1) * dt;
            for (int p = 0; p < h.n_patterns; ++p) {
                const int q = r * h.n_patterns + p;
                ag[3 * r + p] = path_at(h.rec_ts[q], h.rec_ts_dt[q], h.rec_ts_factor[q], t);
            }
        }
        HIP_OK(ctx, hipMemcpyAsync(h.d_ag, ag.data(), sizeof(double) * ag.size(), hipMemcpyHostToDevice, st));
        if ((rc = f3_set_list(ctx, h, live))) return rc;
        const int nl = (int)live.size();
        hipLaunchKernelGGL(f3_predict_kernel, grid2(n, nl), dim3(256), 0, st, h.d_list, n, h.d_uf, h.d_vf, h.d_af,
                           h.d_un, h.d_vn, h.d_an, dt, beta);
        hipLaunchKernelGGL(f3_expand_kernel, grid2(h.total, nl), dim3(256), 0, st, h.d_list, h.total, n, h.d_u,
                           h.d_uf, h.d_free_idx, h.d_row_offsets, h.d_dof_pool, h.d_coeff_pool, h.d_slave);
        std::vector<int32_t> iter = live, converged;
        for (int it = 0; it < h.max_iters && !iter.empty(); ++it) {
            if (it > 0 && (rc = f3_set_list(ctx, h, iter))) return rc;
            const int ni = (int)iter.size();
            hipLaunchKernelGGL(f3_trial_kernel, grid2(n, ni), dim3(256), 0, st, h.d_list, n, h.d_uf, h.d_un, h.d_vn,
                               h.d_an, h.d_at, h.d_vt, a0, a2, a3, gamma, dt);
            if ((rc = f3_elements(ctx, h, ni))) return rc;
            if (h.n_sh > 0)
                hipLaunchKernelGGL(f3_shell_force_kernel, dim3(h.n_sh, ni), dim3(64), 0, st, h.d_list, n, h.n_sh,
                                   h.d_sh_ret, h.d_sh_ret_n, h.d_sh_khat, h.d_uf, h.d_vt, h.d_sh_f, h.d_sh_y);
            hipLaunchKernelGGL(f3_residual_kernel, grid2(n, ni), dim3(256), 0, st, h.d_list, n, h.n_el, h.n_sh,
                               h.n_patterns, h.d_load, h.d_ag, h.d_dof_off, h.d_dof_src, h.d_fhat, h.d_yk, h.d_yi,
                               h.d_sh_f, h.d_sh_y, h.d_m_off, h.d_m_col, h.d_m_val, h.d_vt, h.d_at, h.alpha_m,
                               h.beta_k, h.beta_k_init, has_damping, h.d_R);
            hipLaunchKernelGGL(f3_norm_kernel, dim3(ni), dim3(1024), 0, st, h.d_list, n, h.d_R, h.n_el, h.d_status,
                               h.d_scalar);
            HIP_OK(ctx, hipGetLastError());
            HIP_OK(ctx, hipMemcpyAsync(scal.data(), h.d_scalar, sizeof(double) * scal.size(), hipMemcpyDeviceToHost,
                                       st));
            HIP_OK(ctx, hipStreamSynchronize(st));
            std::vector<int32_t> next;
            for (int r : iter) {
                iters[(size_t)r * steps + step] = it + 1;
                if (scal[4 * r + 1] != 0.0)  // the reference aborts on an element's local non-convergence
                    h.rec_status[r] = SGPU_F3_ELEMENT_FAILED;
                else if (std::sqrt(scal[4 * r]) <= h.tol)
                    converged.push_back(r);
                else
                    next.push_back(r);
            }
            iter.swap(next);
            if (iter.empty()) break;
            if ((rc = f3_set_list(ctx, h, iter))) return rc;
            const int nu = (int)iter.size();
            hipLaunchKernelGGL(f3_band_kernel, grid2(slots, nu), dim3(256), 0, st, h.d_list, slots,
                               (int64_t)h.n_el * 144, h.d_band_off, h.d_band_src, h.d_khat, h.d_sh_khat,
                               h.d_mass_band, h.d_kinit_band, h.d_band, 0, a0, a1, h.alpha_m, h.beta_k,
                               h.beta_k_init, has_damping);
            {
                ProfScope ps(ctx, SGPU_KERNEL_F3_SOLVE);
                if ((rc = launch_band_solve(ctx, h.solver == 0 ? 0 : h.solver == 1 ? 7 : 8, n, h.hbw, h.d_band, h.d_R,
                                            h.d_du,
                                            h.d_scalar + 2, h.d_list, nu, h.d_bwork)))
                    return rc;
            }
            hipLaunchKernelGGL(f3_update_kernel, grid2(n, nu), dim3(256), 0, st, h.d_list, n, h.d_uf, h.d_du);
            hipLaunchKernelGGL(f3_expand_kernel, grid2(h.total, nu), dim3(256), 0, st, h.d_list, h.total, n, h.d_u,
                               h.d_uf, h.d_free_idx, h.d_row_offsets, h.d_dof_pool, h.d_coeff_pool, h.d_slave);
            HIP_OK(ctx, hipGetLastError());
            HIP_OK(ctx, hipMemcpyAsync(scal.data(), h.d_scalar, sizeof(double) * scal.size(), hipMemcpyDeviceToHost,
                                       st));
            HIP_OK(ctx, hipStreamSynchronize(st));
            next.clear();
            for (int r : iter) {
                if (scal[4 * r + 2] == 1.0 && dense_fallbacks) dense_fallbacks[r]++;
                if (scal[4 * r + 2] == 2.0)
                    h.rec_status[r] = SGPU_F3_SINGULAR;
                else
                    next.push_back(r);
            }
            iter.swap(next);
        }
        for (int r : iter) h.rec_status[r] = h.rec_steps[r] + 1;  // not converged within max_iters
        if (converged.empty()) continue;
        if ((rc = sgpu_sections_commit_all(ctx))) return rc;  // records that stopped are never read again
        if ((rc = f3_set_list(ctx, h, converged))) return rc;
        const int nc = (int)converged.size();
        hipLaunchKernelGGL(f3_newmark_kernel, grid2(n, nc), dim3(256), 0, st, h.d_list, n, h.d_uf, h.d_un, h.d_vn,
                           h.d_an, h.d_af, h.d_vf, a0, a2, a3, gamma, dt);
        HIP_OK(ctx, hipGetLastError());
        if (u_hist && h.R == 1)
            HIP_OK(ctx, hipMemcpyAsync(u_hist + (size_t)step * h.total, h.d_u, sizeof(double) * h.total,
                                       hipMemcpyDeviceToHost, st));
        if (h.n_rec > 0) {
            std::vector<int32_t> row_of(h.R, 0), rec_list;
            for (int r : converged)
                if (h.rec_steps[r] < h.rec_cap) {
                    row_of[r] = h.rec_steps[r];
                    rec_list.push_back(r);
                }
            if (!rec_list.empty()) {
                HIP_OK(ctx, hipMemcpyAsync(h.d_row_of, row_of.data(), sizeof(int32_t) * h.R, hipMemcpyHostToDevice,
                                           st));
                if ((rc = f3_set_list(ctx, h, rec_list))) return rc;
                hipLaunchKernelGGL(f3_record_kernel, grid2(h.n_rec, (int)rec_list.size()), dim3(256), 0, st,
                                   h.d_list, h.n_rec, h.total, h.d_rec_dofs, h.d_u, h.d_rec,
                                   (int64_t)h.rec_cap * h.n_rec, h.d_row_of);
                HIP_OK(ctx, hipStreamSynchronize(st));  // row_of goes out of scope
            }
        }
        for (int r : converged) h.rec_steps[r] += 1;
    }
    HIP_OK(ctx, hipStreamSynchronize(st));
    return SGPU_OK;
}

}  // namespace

namespace sgpu {
int frame3d_initialize(sgpu_ctx *ctx, Frame3dHost &h) { return f3_initialize(ctx, h); }
}  // namespace sgpu

int sgpu_frame3d_run(sgpu_ctx *ctx, int steps, double *u_hist, int32_t *iters, int32_t *steps_done,
                     int32_t *status, int32_t *dense_fallbacks) {
    if (!ctx || steps < 0 || !iters || !steps_done || !status) return SGPU_ERR_ARG;
    Frame3dHost *hp = ctx->frame3d;
    if (!hp) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame3d_create has not run");
    Frame3dHost &h = *hp;
    if (h.R != 1) return fail(ctx, SGPU_ERR_STATE, "the frame holds %d records: use sgpu_frame3d_run_records", h.R);
    HIP_OK(ctx, hipSetDevice(ctx->device));
    if (dense_fallbacks) *dense_fallbacks = 0;
    const int before = h.rec_steps[0];
    for (int s = 0; s < steps; ++s) iters[s] = 0;
    const int rc = h.rec_status[0] == 0 ? f3_run(ctx, h, steps, u_hist, iters, dense_fallbacks) : SGPU_OK;
    *steps_done = h.rec_steps[0] - before;
    *status = h.rec_status[0];
    return rc;
}

int sgpu_frame3d_run_records(sgpu_ctx *ctx, int steps, int32_t *iters, int32_t *steps_done, int32_t *status,
                             int32_t *dense_fallbacks) {
    if (!ctx || steps < 0 || !iters || !steps_done || !status) return SGPU_ERR_ARG;
    Frame3dHost *hp = ctx->frame3d;
    if (!hp) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame3d_create has not run");
    Frame3dHost &h = *hp;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    std::vector<int32_t> before(h.rec_steps);
    for (size_t q = 0; q < (size_t)h.R * steps; ++q) iters[q] = 0;
    if (dense_fallbacks)
        for (int r = 0; r < h.R; ++r) dense_fallbacks[r] = 0;
    const int rc = f3_run(ctx, h, steps, nullptr, iters, dense_fallbacks);
    for (int r = 0; r < h.R; ++r) {
        steps_done[r] = h.rec_steps[r] - before[r];
        status[r] = h.rec_status[r];
    }
    return rc;
}

int sgpu_frame3d_set_band_solver(sgpu_ctx *ctx, int solver) {
    if (!ctx || solver < 0 || solver > 2) return SGPU_ERR_ARG;
    if (!ctx->frame3d) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame3d_create has not run");
    const Frame3dHost &h = *ctx->frame3d;
    if (solver == 1 && (size_t)(h.n + 2 * h.hbw) * sizeof(double) > 56 * 1024)
        return fail(ctx, SGPU_ERR_ARG, "%d free DOFs: the elimination without swaps keeps the right-hand side in LDS "
                    "(at most %zu)", h.n, 56 * 1024 / sizeof(double) - 2 * h.hbw);
    ctx->frame3d->solver = solver;
    return SGPU_OK;
}

int sgpu_frame3d_effective_band(sgpu_ctx *ctx, double *band_out) {
    if (!ctx || !band_out) return SGPU_ERR_ARG;
    Frame3dHost *hp = ctx->frame3d;
    if (!hp) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame3d_create has not run");
    Frame3dHost &h = *hp;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int n = h.n;
    const double dt = h.dt, gamma = h.gamma, beta = h.beta;
    const double a0 = 1.0 / (beta * dt * dt), a1 = gamma / (beta * dt), a2 = 1.0 / (beta * dt);
    const double a3 = 1.0 / (2.0 * beta) - 1.0;
    const int has_damping = h.alpha_m != 0.0 || h.beta_k != 0.0 || h.beta_k_init != 0.0;
    const int64_t slots = (int64_t)n * h.W;
    int rc;
    if ((rc = f3_initialize(ctx, h))) return rc;
    if ((rc = f3_set_list(ctx, h, {0}))) return rc;  // record 0
    hipLaunchKernelGGL(f3_trial_kernel, grid2(n, 1), dim3(256), 0, st, h.d_list, n, h.d_uf, h.d_un, h.d_vn, h.d_an,
                       h.d_at, h.d_vt, a0, a2, a3, gamma, dt);
    if ((rc = f3_elements(ctx, h, 1))) return rc;
    hipLaunchKernelGGL(f3_band_kernel, grid2(slots, 1), dim3(256), 0, st, h.d_list, slots, (int64_t)h.n_el * 144,
                       h.d_band_off, h.d_band_src, h.d_khat, h.d_sh_khat, h.d_mass_band, h.d_kinit_band, h.d_band, 0,
                       a0, a1, h.alpha_m, h.beta_k, h.beta_k_init, has_damping);
    HIP_OK(ctx, hipGetLastError());
    HIP_OK(ctx, hipMemcpyAsync(band_out, h.d_band, sizeof(double) * slots, hipMemcpyDeviceToHost, st));
    HIP_OK(ctx, hipStreamSynchronize(st));
    return SGPU_OK;
}

namespace sgpu {
int envelopes_device(sgpu_ctx *ctx, const double *d_hist, const std::vector<int64_t> &bases,
                     const std::vector<int32_t> &steps, int n_record, int n_drift, const int32_t *col_i,
                     const int32_t *col_j, const double *dx, double *disp_env, double *drift_env);
}

int sgpu_frame3d_set_recorder(sgpu_ctx *ctx, int n_record, const int32_t *dofs, int max_steps) {
    if (!ctx || n_record <= 0 || !dofs || max_steps <= 0) return SGPU_ERR_ARG;
    Frame3dHost *h = ctx->frame3d;
    if (!h) return fail(ctx, SGPU_ERR_STATE, "sgpu_frame3d_create has not run");
    if (h->n_rec > 0) return fail(ctx, SGPU_ERR_STATE, "the recorder is already set");
    for (int c = 0; c < n_record; ++c)
        if (dofs[c] < 0 || dofs[c] >= h->total) return fail(ctx, SGPU_ERR_ARG, "recorded DOF %d out of range", c);
    HIP_OK(ctx, hipSetDevice(ctx->device));
    int rc;
    if ((rc = upload(ctx, &h->d_rec_dofs, std::vector<int32_t>(dofs, dofs + n_record)))) return rc;
    if ((rc = dalloc(ctx, &h->d_rec, (size_t)h->R * n_record * max_steps))) return rc;
    h->n_rec = n_record;
    h->rec_cap = max_steps;
    return SGPU_OK;
}

int sgpu_frame3d_recorded_record(sgpu_ctx *ctx, int32_t record, double *hist, int32_t *steps) {
    if (!ctx || !hist || !steps) return SGPU_ERR_ARG;
    Frame3dHost *h = ctx->frame3d;
    if (!h || h->n_rec == 0) return fail(ctx, SGPU_ERR_STATE, "no recorder is set");
    if (record < 0 || record >= h->R) return fail(ctx, SGPU_ERR_ARG, "record out of range");
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    *steps = std::min(h->rec_steps[record], h->rec_cap);
    HIP_OK(ctx, hipMemcpy(hist, h->d_rec + (size_t)record * h->rec_cap * h->n_rec,
                          sizeof(double) * (size_t)*steps * h->n_rec, hipMemcpyDeviceToHost));
    return SGPU_OK;
}

int sgpu_frame3d_recorded(sgpu_ctx *ctx, double *hist, int32_t *steps) {
    return sgpu_frame3d_recorded_record(ctx, 0, hist, steps);
}

int sgpu_frame3d_envelopes_record(sgpu_ctx *ctx, int32_t record, int n_drift, const int32_t *col_i,
                                  const int32_t *col_j, const double *dx, double *disp_env, double *drift_env) {
    if (!ctx || !disp_env || (n_drift > 0 && !drift_env)) return SGPU_ERR_ARG;
    Frame3dHost *h = ctx->frame3d;
    if (!h || h->n_rec == 0) return fail(ctx, SGPU_ERR_STATE, "no recorder is set");
    if (record < 0 || record >= h->R) return fail(ctx, SGPU_ERR_ARG, "record out of range");
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return envelopes_device(ctx, h->d_rec, {(int64_t)record * h->rec_cap * h->n_rec},
                            {std::min(h->rec_steps[record], h->rec_cap)}, h->n_rec, n_drift, col_i, col_j, dx,
                            disp_env, drift_env);
}

int sgpu_frame3d_envelopes(sgpu_ctx *ctx, int n_drift, const int32_t *col_i, const int32_t *col_j, const double *dx,
                           double *disp_env, double *drift_env) {
    return sgpu_frame3d_envelopes_record(ctx, 0, n_drift, col_i, col_j, dx, disp_env, drift_env);
}
