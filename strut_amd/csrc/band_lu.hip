// BandGeneral factorization and solve with the reference's row-swap
// semantics, on gfx950.
//
// The reference factors the band without pivoting
// (linear_solver_backend.mojo:631-679) and, at the first column whose
// partial-pivot scan picks a row other than k, restarts from the loaded
// values with a dense LU with partial pivoting (:652-657, :175-214).  Both
// do, per entry, factor = a_ik / a_kk and a_ij -= factor * a_kj in
// ascending k; rows outside the band hold zeros, so the dense LU's extra
// work changes nothing.  A band LU with partial pivoting whose U rows are
// 2 hbw wide (a swap brings a row reaching k + 2 hbw) therefore makes the
// reference's pivot choices (the first row of largest magnitude) and its
// values, with no dense matrix: one kernel covers both of the reference's
// paths, and the swap flag says which one the reference took.  The forward
// substitution (rhs[piv[i]] - sum_j L_ij y_j in ascending j, :216-222) is
// the same sum when the swaps are applied as the rows move.  The back
// substitution runs by columns (x_i, then every row above takes U_ri x_i),
// so its sums are in descending order where the reference's are ascending.
//
// Storage.  Input: the reference's band, n x W (W = 2 hbw + 1), slot (i, j)
// at i W + j - i + hbw.  Output, in place: U row k in column-slot order,
// U(k, j) at k W + (j mod W) for j = k .. k + 2 hbw (the same W slots per
// row; every column keeps one slot, so a wavefront stores its columns at
// fixed offsets).  When asked for, L (n x hbw: the multiplier of step k at
// window position i at k hbw + i - k - 1) and piv (n: the position swapped
// with k) for later solves.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/strut_gpu.h"
#include "ctx.h"

namespace sgpu {

constexpr int kBpWaves = 16;
constexpr int kBpThreads = 64 * kBpWaves;
constexpr int kBpMaxP = 128;                                  // window rows: lane l holds slots l, l + 64
constexpr int kBpMaxQ = 2 * kBpMaxP - 1;                      // window columns
constexpr int kBpCols = (kBpMaxQ + kBpWaves - 1) / kBpWaves;  // column slots per wavefront (16)
constexpr int kBpD = 4;                                       // entering rows per prefetch batch (64 values)
constexpr int kBpChunk = 8;                                   // columns per uniform pivot-row batch
constexpr int kBpBack = 8;                                    // U columns per staged back-substitution batch
constexpr int kBpRingRow = kBpWaves * kBpCols;                // 256: one entry per column slot of every wavefront
constexpr int kBpRing = 2 * kBpBack * (kBpMaxQ + 1);          // doubles; >= kBpD * kBpRingRow
constexpr double kBpPivotTol = 1.0e-14;

static_assert(kBpRing >= kBpD * kBpRingRow, "the LDS ring holds kBpD entering rows");

namespace {

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Value of lane l + 1 (lane 63: 0).
__device__ __forceinline__ double next_lane_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Largest value over the wavefront, in every lane: quad swaps, half-row and
// row mirrors, then the four rows.
__device__ __forceinline__ double wave_max_d(double v) {
    v = fmax(v, dpp_d<0xB1>(v));   // quad_perm [1,0,3,2]
    v = fmax(v, dpp_d<0x4E>(v));   // quad_perm [2,3,0,1]
    v = fmax(v, dpp_d<0x141>(v));  // row_half_mirror
    v = fmax(v, dpp_d<0x140>(v));  // row_mirror
    return fmax(fmax(lane_d(v, 0), lane_d(v, 16)), fmax(lane_d(v, 32), lane_d(v, 48)));
}

__device__ __forceinline__ int wave_min_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int first_lane(uint64_t mask) { return (int)__builtin_ctzll(mask); }

__device__ __forceinline__ void bp_lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0) only
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

}  // namespace

// One workgroup per system (recs: one per listed record, at its slices;
// record r's status at status[4 r]).  The active window -- rows at positions
// k .. k + hbw, columns k .. k + 2 hbw -- lives in registers:
//   * column slot c (global column g at slot g mod Q) belongs to wavefront
//     c mod 16, local index c / 16; lane l holds that column's row slots l
//     and l + 64, so a column's pivot search is one wavefront's reduction;
//   * a row keeps its slot for its whole life in the window; a swap only
//     exchanges the positions two slots stand for (pos[], kept in every
//     wavefront), and the pivot row's slot takes the entering row;
//   * per column, one barrier: the pivot column's wavefront picks the pivot
//     (largest magnitude, lowest position among equals), forms the
//     multipliers with the reference's division and publishes them; then
//     every wavefront takes the pivot row's entries of its columns from the
//     pivot lane (readlane), updates its entries, writes its part of U row k
//     and loads its part of the entering row from an LDS ring that it
//     prefetches itself, kBpD rows ahead;
//   * the forward substitution rides along in registers (every wavefront
//     keeps the slots' y), y_k goes to LDS;
//   * after the last column, wavefront 0 runs the back substitution by
//     columns while the others stage U columns in LDS.
// status[0]: 0 no row swap (the reference's band LU), 1 rows swapped (the
// reference's dense fallback, the same values), 2 a column's largest
// magnitude at or below 1e-14 (the reference aborts), nothing solved.
template <bool kStoreL, int kCols>
__global__ __launch_bounds__(kBpThreads) void band_piv_reg_kernel(const int32_t *recs, int n, int hbw,
                                                                  double *band, const double *rhs, double *x,
                                                                  double *status, double *Lmul, int32_t *ipiv) {
    const int P = hbw + 1, Q = 2 * hbw + 1, W = Q;
    const bool has_rhs = rhs != nullptr;  // factor only when null
    if (recs) {
        const int r_ = recs[blockIdx.x];
        band += (size_t)r_ * n * W;
        if (has_rhs) {
            rhs += (size_t)r_ * n;
            x += (size_t)r_ * n;
        }
        status += 4 * r_;
        if (kStoreL) {
            Lmul += (size_t)r_ * n * (hbw > 0 ? hbw : 1);
            ipiv += (size_t)r_ * n;
        }
    }
    __shared__ double s_ring[kBpRing];
    __shared__ double s_l[2][kBpMaxP];
    __shared__ int s_piv[2];
    extern __shared__ double s_y[];  // n
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ncols = w < Q ? (Q - w + kBpWaves - 1) / kBpWaves : 0;  // column slots of this wavefront
    const bool valid0 = lane < P, valid1 = lane + 64 < P;
    double a[2][kCols];
    int pos0 = lane, pos1 = lane + 64;
    // window at k = 0: slot = index
#pragma unroll
    for (int j = 0; j < kCols; ++j) {
        const int g = w + kBpWaves * j;
        a[0][j] = 0.0;
        a[1][j] = 0.0;
        if (j < ncols) {
            if (valid0 && lane < n && g < n && g - lane <= hbw && lane - g <= hbw)
                a[0][j] = band[(size_t)lane * W + (g - lane + hbw)];
            const int r1 = lane + 64;
            if (valid1 && r1 < n && g < n && g - r1 <= hbw && r1 - g <= hbw)
                a[1][j] = band[(size_t)r1 * W + (g - r1 + hbw)];
        }
    }
    double y0 = (has_rhs && valid0 && lane < n) ? rhs[lane] : 0.0;
    double y1 = (has_rhs && valid1 && lane + 64 < n) ? rhs[lane + 64] : 0.0;
    // entering rows, kBpD = 4 steps per batch: lane l carries step
    // s0 + (l >> 4)'s row at local column j = l & 15; the row's value at
    // column slot c = w + 16 j is band[T W + d] with T = s + P and
    // d = (c - s - 1) mod Q, kept incrementally (d falls by kBpD per batch)
    const int pf_j = lane & 15, pf_s = lane >> 4;
    const int pf_c = w + kBpWaves * pf_j;
    const bool pf_on = pf_j < ncols;
    int pf_d = (pf_c - pf_s - 1) % Q;  // for the batch at s0 = 0
    if (pf_d < 0) pf_d += Q;
    // The load is unconditional (its index clamped into the band) and the
    // validity travels beside it, applied when the batch lands: a select
    // on the loaded value would make the compiler wait for the load at once.
    bool pf_ok = false;
    auto fetch = [&](int s0) -> double {  // step s0 + pf_s's row, column slot pf_c
        const int T = s0 + pf_s + P;
        pf_ok = pf_on && T < n && s0 + pf_s + 1 + pf_d < n;
        return band[pf_ok ? (size_t)T * W + pf_d : 0];
    };
    auto next_batch = [&]() {
        pf_d -= kBpD;
        while (pf_d < 0) pf_d += Q;
    };
    // ring row stride kBpRingRow: every column slot 0..255 of every
    // wavefront (unused slots past Q included) has its own entry
    auto ring_at = [&](int s, int j) -> double & { return s_ring[(s % kBpD) * kBpRingRow + w + kBpWaves * j]; };
    // column slots past Q read the ring too: zero it first (each wavefront
    // reads and writes only its own column slots, so no barrier is needed)
#pragma unroll
    for (int j = 0; j < kBpCols; ++j)
        if (lane < kBpD) ring_at(lane, j) = 0.0;
    {
        const double v0 = fetch(0);
        if (pf_on) ring_at(pf_s, pf_j) = pf_ok ? v0 : 0.0;
    }
    next_batch();
    double pf = fetch(kBpD);
    // right-hand sides of the entering rows ride with the batches: lane
    // i < kBpD of rr holds rhs[s0 + i + P] for the current batch, pr the
    // next batch's
    // (a row past n has no right-hand side: its slot's y is never used, so
    // the clamped load's value is harmless there)
    auto fetch_rhs = [&](int s0) -> double {
        const int T = s0 + (lane & (kBpD - 1)) + P;
        return has_rhs ? rhs[T < n ? T : n - 1] : 0.0;
    };
    double rr = fetch_rhs(0), pr = fetch_rhs(kBpD);
    if (tid == 0) s_piv[0] = s_piv[1] = 0;
    int swapped = 0;
    // every load of the initial window has returned before any wavefront
    // overwrites band rows with U rows
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // Pivot of step s from its column's values (c0, c1) in this wavefront:
    // the largest magnitude, the lowest position among equals (the
    // reference's strict scan); multipliers with the reference's division.
    auto pivot_phase = [&](int s, double c0, double c1, int buf) {
        const double m0 = valid0 ? fabs(c0) : -1.0, m1 = valid1 ? fabs(c1) : -1.0;
        const double M = wave_max_d(fmax(m0, m1));
        const uint64_t b0 = __ballot(m0 == M), b1 = __ballot(m1 == M);
        int sp;
        if (__popcll(b0) + __popcll(b1) == 1) {
            sp = b0 ? first_lane(b0) : 64 + first_lane(b1);
        } else {
            const int pmin = wave_min_i(min(m0 == M ? pos0 : 1 << 30, m1 == M ? pos1 : 1 << 30));
            const uint64_t e0 = __ballot(m0 == M && pos0 == pmin), e1 = __ballot(m1 == M && pos1 == pmin);
            sp = e0 ? first_lane(e0) : 64 + first_lane(e1);
        }
        if (!(M > kBpPivotTol)) {
            if (lane == 0) s_piv[buf] = -1;
            return;
        }
        const double pv = sp < 64 ? lane_d(c0, sp) : lane_d(c1, sp - 64);
        const double l0 = (valid0 && lane != sp) ? c0 / pv : 0.0;
        const double l1 = (valid1 && lane + 64 != sp) ? c1 / pv : 0.0;
        s_l[buf][lane] = l0;
        s_l[buf][lane + 64] = l1;
        if (lane == 0) s_piv[buf] = sp;
        if (kStoreL) {
            const int p = sp < 64 ? __builtin_amdgcn_readlane(pos0, sp) : __builtin_amdgcn_readlane(pos1, sp - 64);
            const int q0 = pos0 == s ? p : pos0, q1 = pos1 == s ? p : pos1;
            if (valid0 && lane != sp && q0 < n) Lmul[(size_t)s * hbw + (q0 - s - 1)] = l0;
            if (valid1 && lane + 64 != sp && q1 < n) Lmul[(size_t)s * hbw + (q1 - s - 1)] = l1;
            if (lane == 0) ipiv[s] = p;
        }
    };
    if (w == 0) pivot_phase(0, a[0][0], a[1][0], 0);  // column 0: slot 0 of wavefront 0
    int kc = 0;  // k mod Q
    bool singular = false;
    for (int k = 0; k < n && !singular; ++k) {
        const int buf = k & 1;
        const int wk = kc % kBpWaves, jk = kc / kBpWaves;
        const int kc1 = kc + 1 == Q ? 0 : kc + 1;
        const int wk1 = kc1 % kBpWaves, jk1 = kc1 / kBpWaves;
        bp_lds_barrier();
        // a singular column ends the loop at its head (an exit from the
        // middle of the loop body makes the compiler spill the window); the
        // rest of this column runs on slot 0 and its results are not used
        const int sp_raw = __builtin_amdgcn_readfirstlane(s_piv[buf]);
        singular = sp_raw < 0;
        const int sp = singular ? 0 : sp_raw;
        const double l0 = s_l[buf][lane], l1 = s_l[buf][lane + 64];
        const int Lp = sp & 63, hp = sp >> 6;
        // positions: the slot standing for position k takes the pivot's
        // position p; the pivot slot takes the entering row k + P
        const int p = hp ? __builtin_amdgcn_readlane(pos1, Lp) : __builtin_amdgcn_readlane(pos0, Lp);
        if (p != k) {
            swapped = 1;
            if (pos0 == k) pos0 = p;
            if (pos1 == k) pos1 = p;
        }
        if (hp == 0 && lane == Lp) pos0 = k + P;
        if (hp == 1 && lane == Lp) pos1 = k + P;
        // a batch of entering rows (and their right-hand sides) lands; the
        // next is fetched
        if ((k & (kBpD - 1)) == 0 && k > 0) {
            if (pf_on) ring_at(k + pf_s, pf_j) = pf_ok ? pf : 0.0;
            rr = pr;
            next_batch();
            pf = fetch(k + kBpD);
            pr = fetch_rhs(k + kBpD);
        }
        // forward substitution
        const double rhs_enter = lane_d(rr, k & (kBpD - 1));
        const double yk = hp ? lane_d(y1, Lp) : lane_d(y0, Lp);
        y0 -= l0 * yk;
        y1 -= l1 * yk;
        if (tid == 0) s_y[k] = yk;
        if (hp == 0 && lane == Lp) y0 = rhs_enter;
        if (hp == 1 && lane == Lp) y1 = rhs_enter;
        const double *ring_k = s_ring + (k % kBpD) * kBpRingRow + w;
        // lookahead: the next pivot column first, then its pivot
        if (w == wk1 && k + 1 < n) {
            double c0 = 0.0, c1 = 0.0;
#pragma unroll
            for (int j = 0; j < kCols; ++j)
                if (j == jk1) {
                    c0 = a[0][j];
                    c1 = a[1][j];
                }
            const double uc = lane_d(hp ? c1 : c0, Lp);
            c0 -= l0 * uc;
            c1 -= l1 * uc;
            if (kc1 == kc) {  // a one-column window: the column retires first
                c0 = 0.0;
                c1 = 0.0;
            }
            const double e = ring_k[kBpWaves * jk1];
            if (hp == 0 && lane == Lp) c0 = e;
            if (hp == 1 && lane == Lp) c1 = e;
            pivot_phase(k + 1, c0, c1, buf ^ 1);
        }
        // U row k, from the pivot lane, in column-slot order: U(k, j) at
        // k W + (j mod W), so every local column has a fixed offset
        {
            double *urow = band + (size_t)k * W + w;
            if (lane == Lp) {
                if (hp) {
#pragma unroll
                    for (int j = 0; j < kCols; ++j)
                        if (j < ncols) urow[kBpWaves * j] = a[1][j];
                } else {
#pragma unroll
                    for (int j = 0; j < kCols; ++j)
                        if (j < ncols) urow[kBpWaves * j] = a[0][j];
                }
            }
        }
        // the rank-1 update, in chunks of kBpChunk columns whose pivot-row
        // entries are read from the pivot lane into scalar registers
#pragma unroll
        for (int j0 = 0; j0 < kCols; j0 += kBpChunk) {
            constexpr int C = kBpChunk;
            double u[C];
            if (hp) {
#pragma unroll
                for (int t = 0; t < C; ++t) u[t] = j0 + t < kCols ? lane_d(a[1][j0 + t < kCols ? j0 + t : 0], Lp) : 0.0;
            } else {
#pragma unroll
                for (int t = 0; t < C; ++t) u[t] = j0 + t < kCols ? lane_d(a[0][j0 + t < kCols ? j0 + t : 0], Lp) : 0.0;
            }
#pragma unroll
            for (int t = 0; t < C; ++t) {
                const int j = j0 + t;
                if (j < kCols) {
                    a[0][j] -= l0 * u[t];
                    a[1][j] -= l1 * u[t];
                }
            }
        }
        // the retired column becomes column k + Q, zero in every row present
        if (w == wk) {
#pragma unroll
            for (int j = 0; j < kCols; ++j)
                if (j == jk) {
                    a[0][j] = 0.0;
                    a[1][j] = 0.0;
                }
        }
        // the entering row takes the pivot's slot: loads under a one-lane
        // mask write that lane only (column slots past Q read zeros)
        if (hp == 0 && lane == Lp) {
#pragma unroll
            for (int j = 0; j < kCols; ++j) a[0][j] = ring_k[kBpWaves * j];
        }
        if (hp == 1 && lane == Lp) {
#pragma unroll
            for (int j = 0; j < kCols; ++j) a[1][j] = ring_k[kBpWaves * j];
        }
        kc = kc1;
    }
    // back substitution by columns in wavefront 0: at row i lane l holds
    // y[i - l - 64 m] (m = 0..3); the other wavefronts stage U columns
    // (diagonal first) kBpBack at a time
    if (singular || !has_rhs) {
        if (tid == 0) status[0] = singular ? 2.0 : swapped ? 1.0 : 0.0;
        return;
    }
    __threadfence();
    __syncthreads();
    const int CL = kBpMaxQ + 1;
    auto stage = [&](int batch, int half) {
        const int top = n - 1 - batch * kBpBack;
        for (int q = tid - 64; q < kBpBack * Q; q += kBpThreads - 64) {
            const int rr = q / Q, t = q % Q, i = top - rr;
            double v = 0.0;
            if (i >= 0 && i - t >= 0) v = band[(size_t)(i - t) * W + i % Q];
            s_ring[(half * kBpBack + rr) * CL + t] = v;
        }
    };
    const int n_batches = (n + kBpBack - 1) / kBpBack;
    if (tid >= 64) stage(0, 0);
    __syncthreads();
    double ya[4] = {0.0, 0.0, 0.0, 0.0};
    if (tid < 64) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int r = n - 1 - lane - 64 * m;
            ya[m] = r >= 0 ? s_y[r] : 0.0;
        }
    }
    for (int b = 0; b < n_batches; ++b) {
        const int half = b & 1;
        if (tid >= 64) {
            if (b + 1 < n_batches) stage(b + 1, half ^ 1);
        } else {
            const int top = n - 1 - b * kBpBack;
            for (int rr = 0; rr < kBpBack; ++rr) {
                const int i = top - rr;
                if (i < 0) break;
                const double *col = s_ring + (half * kBpBack + rr) * CL;
                const double xi = lane_d(ya[0], 0) / col[0];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int t = lane + 64 * m;
                    if (t >= 1 && t < Q) ya[m] -= col[t] * xi;
                }
                if (lane == 0) s_y[i] = xi;  // y_i is spent: x_i takes its place
                const int rin = i - 256;
                const double yin = rin >= 0 ? s_y[rin] : 0.0;
                double carry[4];
#pragma unroll
                for (int m = 1; m < 4; ++m) carry[m - 1] = lane_d(ya[m], 0);
                carry[3] = yin;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    ya[m] = next_lane_d(ya[m]);
                    if (lane == 63) ya[m] = carry[m];
                }
            }
        }
        bp_lds_barrier();
    }
    for (int i = tid; i < n; i += kBpThreads) x[i] = s_y[i];
    if (tid == 0) status[0] = swapped ? 1.0 : 0.0;
}

// The same factorization for any half bandwidth, in global memory: `work`
// holds each row at columns i - hbw .. i + 2 hbw (n x (3 hbw + 1)), rows are
// swapped in place over the columns that can hold values (k .. k + 2 hbw),
// earlier multipliers stay where they were formed.  Three barriers per
// column; for bands the register kernel cannot hold.
__global__ __launch_bounds__(1024) void band_piv_global_kernel(const int32_t *recs, int n, int hbw, double *band,
                                                               const double *rhs, double *x, double *status,
                                                               double *Lmul, int32_t *ipiv, double *work) {
    const int W = 2 * hbw + 1, V = 3 * hbw + 1;
    const bool has_rhs = rhs != nullptr;  // factor only when null
    if (recs) {
        const int r_ = recs[blockIdx.x];
        band += (size_t)r_ * n * W;
        if (has_rhs) {
            rhs += (size_t)r_ * n;
            x += (size_t)r_ * n;
        }
        status += 4 * r_;
        work += (size_t)r_ * n * V;
        if (Lmul) Lmul += (size_t)r_ * n * (hbw > 0 ? hbw : 1);
        if (ipiv) ipiv += (size_t)r_ * n;
    }
    __shared__ int s_p, s_swapped;
    const int tid = threadIdx.x, nt = blockDim.x;
    auto ab = [&](int i, int j) -> double & { return work[(size_t)i * V + (j - i + hbw)]; };
    for (size_t q = tid; q < (size_t)n * V; q += nt) {
        const int i = (int)(q / V), o = (int)(q % V);
        work[q] = o < W ? band[(size_t)i * W + o] : 0.0;
    }
    if (has_rhs)
        for (int i = tid; i < n; i += nt) x[i] = rhs[i];
    if (tid == 0) s_swapped = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const int i1 = min(k + hbw, n - 1);
        if (tid < 64) {  // pivot: largest magnitude, first row among equals
            double best = -1.0;
            int bi = 1 << 30;
            for (int i = k + tid; i <= i1; i += 64) {
                const double v = fabs(ab(i, k));
                if (v > best) {
                    best = v;
                    bi = i;
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
            if (tid == 0) s_p = best > kBpPivotTol ? bi : -1;
        }
        __syncthreads();
        const int p = s_p;
        if (p < 0) {
            if (tid == 0) status[0] = 2.0;
            return;
        }
        const int j1 = min(k + 2 * hbw, n - 1);
        if (p != k) {
            for (int j = k + tid; j <= j1; j += nt) {
                const double t = ab(k, j);
                ab(k, j) = ab(p, j);
                ab(p, j) = t;
            }
            if (tid == 0) {
                if (has_rhs) {
                    const double t = x[k];
                    x[k] = x[p];
                    x[p] = t;
                }
                s_swapped = 1;
            }
        }
        if (tid == 0 && ipiv) ipiv[k] = p;
        __syncthreads();
        const double diag = ab(k, k);
        for (int i = k + 1 + tid; i <= i1; i += nt) {
            const double l = ab(i, k) / diag;
            ab(i, k) = l;
            if (Lmul) Lmul[(size_t)k * hbw + (i - k - 1)] = l;
        }
        for (int j = k + tid; j < k + W; j += nt) band[(size_t)k * W + j % W] = j <= j1 ? ab(k, j) : 0.0;
        __syncthreads();
        const double yk = has_rhs ? x[k] : 0.0;
        const int rows = i1 - k, cols = j1 - k;
        for (int q = tid; q < rows * cols; q += nt) {
            const int i = k + 1 + q / cols, j = k + 1 + q % cols;
            ab(i, j) -= ab(i, k) * ab(k, j);
        }
        if (has_rhs)
            for (int i = k + 1 + tid; i <= i1; i += nt) x[i] -= ab(i, k) * yk;
        __syncthreads();
    }
    if (!has_rhs) {
        if (tid == 0) status[0] = s_swapped ? 1.0 : 0.0;
        return;
    }
    // back substitution by columns
    for (int i = n - 1; i >= 0; --i) {
        if (tid == 0) x[i] = x[i] / band[(size_t)i * W + i % W];
        __syncthreads();
        const double xi = x[i];
        for (int r = max(i - 2 * hbw, 0) + tid; r < i; r += nt) x[r] -= band[(size_t)r * W + i % W] * xi;
        __syncthreads();
    }
    if (tid == 0) status[0] = s_swapped ? 1.0 : 0.0;
}

// Solve with a stored factorization (U in the band, L, piv): the swaps and
// multipliers of each step applied in order, then the back substitution by
// columns (linear_solver_backend.mojo:735-769, :216-232).
__global__ __launch_bounds__(1024) void band_lu_solve_kernel(int n, int hbw, const double *band, const double *Lmul,
                                                             const int32_t *ipiv, const double *rhs, double *x) {
    const int W = 2 * hbw + 1, tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < n; i += nt) x[i] = rhs[i];
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const int p = ipiv[k];
        if (tid == 0 && p != k) {
            const double t = x[k];
            x[k] = x[p];
            x[p] = t;
        }
        __syncthreads();
        const double yk = x[k];
        const int i1 = min(k + hbw, n - 1);
        for (int i = k + 1 + tid; i <= i1; i += nt) x[i] -= Lmul[(size_t)k * hbw + (i - k - 1)] * yk;
        __syncthreads();
    }
    for (int i = n - 1; i >= 0; --i) {
        if (tid == 0) x[i] = x[i] / band[(size_t)i * W + i % W];
        __syncthreads();
        const double xi = x[i];
        for (int r = max(i - 2 * hbw, 0) + tid; r < i; r += nt) x[r] -= band[(size_t)r * W + i % W] * xi;
        __syncthreads();
    }
}

// y = alpha y + scale * src over n x W band values (load_reduced_matrix with
// alpha 0, add_reduced_matrix(A, scale) with alpha 1).
__global__ void band_axpy_kernel(int64_t count, double *y, const double *src, double scale, int add) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    y[i] = add ? y[i] + scale * src[i] : (scale == 1.0 ? src[i] : scale * src[i]);
}

// band_blk.hip: the blocked factorization and solve for wide bands
bool band_blk_fits(int n, int hbw);
size_t band_blk_work_doubles(int n, int hbw);
int launch_band_blk(sgpu_ctx *ctx, int n, int hbw, double *d_band, const double *d_rhs, double *d_x,
                    double *d_status, double *d_L, int32_t *d_ipiv, double *d_work, const int32_t *recs, int count);

// LDS the register kernel needs beyond its static arrays (the forward values).
bool band_piv_reg_fits(int n, int hbw) { return hbw + 1 <= kBpMaxP && (size_t)n * sizeof(double) <= 24 * 1024; }

size_t band_piv_work_doubles(int n, int hbw) { return (size_t)n * (3 * (size_t)hbw + 1); }

// Pivoted factor (+ solve when rhs is set) of `count` systems on the stream.
// mode 0: the register kernel when the band fits it, else the global one
// (which needs `work`, band_piv_work_doubles per system); 1: the register
// kernel; 2: the global kernel; 3: the blocked kernels (`work` of
// band_blk_work_doubles per system; mode 0 takes them when the register
// kernel cannot hold the band).
int launch_band_piv(sgpu_ctx *ctx, int mode, int n, int hbw, double *d_band, const double *d_rhs, double *d_x,
                    double *d_status, double *d_L, int32_t *d_ipiv, double *d_work, const int32_t *recs,
                    int count) {
    const bool fits = band_piv_reg_fits(n, hbw);
    if (mode == 0) mode = fits ? 1 : band_blk_fits(n, hbw) ? 3 : 2;
    if (mode == 3) return launch_band_blk(ctx, n, hbw, d_band, d_rhs, d_x, d_status, d_L, d_ipiv, d_work, recs, count);
    if (mode == 1 && !fits)
        return fail(ctx, SGPU_ERR_ARG, "band %d x %d exceeds the register solve (half bandwidth <= %d, n <= 3072)", n,
                    2 * hbw + 1, kBpMaxP - 1);
    if (mode == 2 && !d_work) return fail(ctx, SGPU_ERR_ARG, "the global band solve needs its work array");
    if (mode == 1) {
        const size_t lds = sizeof(double) * (size_t)n;
        const bool c14 = 2 * hbw + 1 <= 14 * kBpWaves;  // at most 14 column slots per wavefront
        auto kern = d_L ? (c14 ? band_piv_reg_kernel<true, 14> : band_piv_reg_kernel<true, 16>)
                        : (c14 ? band_piv_reg_kernel<false, 14> : band_piv_reg_kernel<false, 16>);
        hipLaunchKernelGGL(kern, dim3(count), dim3(kBpThreads), lds, ctx->stream, recs, n, hbw, d_band, d_rhs, d_x,
                           d_status, d_L, d_ipiv);
    } else {
        hipLaunchKernelGGL(band_piv_global_kernel, dim3(count), dim3(1024), 0, ctx->stream, recs, n, hbw, d_band,
                           d_rhs, d_x, d_status, d_L, d_ipiv, d_work);
    }
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

}  // namespace sgpu

// ------------------------------------------------------------------ C ABI:
// a BandGeneral system with the reference's backend calls split
// (initialize, load, factorize_loaded, solve).

struct sgpu_band {
    sgpu_ctx *ctx = nullptr;
    int n = 0, hbw = 0, W = 0;
    bool loaded = false, factored = false;
    double *d_band = nullptr, *d_L = nullptr, *d_work = nullptr, *d_vec = nullptr, *d_status = nullptr;
    int32_t *d_ipiv = nullptr;
    std::vector<double> host_band;  // staging of dense loads
};

using namespace sgpu;

extern "C" {

int sgpu_band_create(sgpu_ctx *ctx, int n, int hbw, sgpu_band **out) {
    if (!ctx || !out) return SGPU_ERR_ARG;
    *out = nullptr;
    if (n <= 0 || hbw < 0 || hbw > n - 1) return fail(ctx, SGPU_ERR_ARG, "band %d rows, half bandwidth %d", n, hbw);
    HIP_OK(ctx, hipSetDevice(ctx->device));
    sgpu_band *b = new sgpu_band();
    b->ctx = ctx;
    b->n = n;
    b->hbw = hbw;
    b->W = 2 * hbw + 1;
    const size_t slots = (size_t)n * b->W;
    hipError_t e = hipMalloc((void **)&b->d_band, sizeof(double) * slots);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_L, sizeof(double) * (size_t)n * (hbw > 0 ? hbw : 1));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_ipiv, sizeof(int32_t) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_vec, sizeof(double) * 2 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_status, sizeof(double) * 4);
    if (e == hipSuccess && !band_piv_reg_fits(n, hbw))
        e = hipMalloc((void **)&b->d_work, sizeof(double) * band_blk_work_doubles(n, hbw));
    if (e != hipSuccess) {
        sgpu_band_destroy(b);
        return fail(ctx, SGPU_ERR_HIP, "band of %d x %d: %s", n, 2 * hbw + 1, hipGetErrorString(e));
    }
    *out = b;
    return SGPU_OK;
}

int sgpu_band_destroy(sgpu_band *b) {
    if (!b) return SGPU_ERR_ARG;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (void *p : {(void *)b->d_band, (void *)b->d_L, (void *)b->d_ipiv, (void *)b->d_vec, (void *)b->d_status,
                    (void *)b->d_work})
        if (p) (void)hipFree(p);
    delete b;
    return SGPU_OK;
}

int sgpu_band_load(sgpu_band *b, int layout, const double *values, double scale, int add) {
    if (!b || !values || (layout != 0 && layout != 1)) return SGPU_ERR_ARG;
    sgpu_ctx *ctx = b->ctx;
    if (add && (!b->loaded || b->factored))
        return fail(ctx, SGPU_ERR_STATE, "add to a band that holds no loaded values");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const int n = b->n, hbw = b->hbw, W = b->W;
    const size_t slots = (size_t)n * W;
    const double *src = values;
    if (layout == 1) {  // dense n x n: the in-band entries (:413-426 drops the rest)
        b->host_band.assign(slots, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = i - hbw < 0 ? 0 : i - hbw; j <= i + hbw && j < n; ++j)
                b->host_band[(size_t)i * W + (j - i + hbw)] = values[(size_t)i * n + j];
        src = b->host_band.data();
    }
    double *d_src = b->d_band;
    double *staging = nullptr;
    if (add) {  // the scaled values go through a staging buffer of the band's size
        if (hipMalloc((void **)&staging, sizeof(double) * slots) != hipSuccess)
            return fail(ctx, SGPU_ERR_HIP, "band staging allocation failed");
        d_src = staging;
    }
    int rc = SGPU_OK;
    if (hipMemcpy(d_src, src, sizeof(double) * slots, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(ctx, SGPU_ERR_HIP, "band upload failed");
    } else if (add || scale != 1.0) {
        hipLaunchKernelGGL(band_axpy_kernel, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, ctx->stream,
                           (int64_t)slots, b->d_band, d_src, scale, add);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = fail(ctx, SGPU_ERR_HIP, "band load kernel failed");
    }
    if (staging) (void)hipFree(staging);
    if (rc == SGPU_OK) {
        b->loaded = true;
        b->factored = false;
    }
    return rc;
}

int sgpu_band_load_device(sgpu_band *b, const double *d_values, double scale, int add) {
    if (!b || !d_values) return SGPU_ERR_ARG;
    sgpu_ctx *ctx = b->ctx;
    if (add && (!b->loaded || b->factored))
        return fail(ctx, SGPU_ERR_STATE, "add to a band that holds no loaded values");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const size_t slots = (size_t)b->n * b->W;
    hipLaunchKernelGGL(band_axpy_kernel, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, ctx->stream,
                       (int64_t)slots, b->d_band, d_values, scale, add);
    HIP_OK(ctx, hipGetLastError());
    b->loaded = true;
    b->factored = false;
    return SGPU_OK;
}

int sgpu_band_factor(sgpu_band *b, int32_t *info) {
    if (!b || !info) return SGPU_ERR_ARG;
    sgpu_ctx *ctx = b->ctx;
    if (!b->loaded || b->factored) return fail(ctx, SGPU_ERR_STATE, "no loaded values to factor");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    int rc = launch_band_piv(ctx, 0, b->n, b->hbw, b->d_band, nullptr, nullptr, b->d_status, b->d_L, b->d_ipiv,
                             b->d_work, nullptr, 1);
    if (rc) return rc;
    double st = 0.0;
    HIP_OK(ctx, hipMemcpyAsync(&st, b->d_status, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    *info = (int32_t)st;
    b->loaded = false;
    b->factored = st != 2.0;
    return SGPU_OK;
}

int sgpu_band_solve_factored(sgpu_band *b, const double *rhs, double *x) {
    if (!b || !rhs || !x) return SGPU_ERR_ARG;
    sgpu_ctx *ctx = b->ctx;
    if (!b->factored) return fail(ctx, SGPU_ERR_STATE, "the band holds no factorization");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const int n = b->n;
    HIP_OK(ctx, hipMemcpyAsync(b->d_vec, rhs, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(band_lu_solve_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, b->hbw, b->d_band, b->d_L,
                       b->d_ipiv, b->d_vec, b->d_vec + n);
    HIP_OK(ctx, hipGetLastError());
    HIP_OK(ctx, hipMemcpyAsync(x, b->d_vec + n, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    return SGPU_OK;
}

int sgpu_band_factors(sgpu_band *b, double *u_rows, double *l_cols, int32_t *piv) {
    if (!b) return SGPU_ERR_ARG;
    sgpu_ctx *ctx = b->ctx;
    if (!b->factored) return fail(ctx, SGPU_ERR_STATE, "the band holds no factorization");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const size_t n = b->n, W = b->W;
    if (u_rows) {  // column-slot order on the device: U(k, k + d) at k W + (k + d) mod W
        std::vector<double> raw(n * W);
        HIP_OK(ctx, hipMemcpy(raw.data(), b->d_band, sizeof(double) * n * W, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k)
            for (size_t d = 0; d < W; ++d) u_rows[k * W + d] = raw[k * W + (k + d) % W];
    }
    if (l_cols && b->hbw > 0)
        HIP_OK(ctx, hipMemcpy(l_cols, b->d_L, sizeof(double) * n * b->hbw, hipMemcpyDeviceToHost));
    if (piv) HIP_OK(ctx, hipMemcpy(piv, b->d_ipiv, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return SGPU_OK;
}

}  // extern "C"
