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
//   * panel: one workgroup per system, one panel row per thread (rows k0 ..
//     k0 + 15 + hbw, sixteen entries each in registers); per column one
//     barrier: before it each wavefront holding rows publishes its first
//     row of largest magnitude (DPP maximum, ballot) with the row's
//     entries, after it every thread picks the pivot from the candidates,
//     the two rows swap from LDS and the rows below update themselves.  The
//     multipliers are stored where they were formed (the reference's band
//     LU keeps them so, and its solves swap the right-hand side step by
//     step); the trailing work takes them as the rows stand after the
//     panel's last swap (rows carry their multipliers when they move),
//     which gives each row the same products in the same order;
//   * the panel's swaps applied to the trailing columns (k0 + 16 .. k0 + 15
//     + 2 hbw) and the right-hand side as one permutation of the rows they
//     touch (each slot loaded from the row whose entry ends there), then the
//     panel's pivot rows finished there (u_kj -= l_kk' u_k'j, k' ascending),
//     in registers, one thread per column, in the same launch;
//   * the trailing update a_ij -= l_ik u_kj over the panel's sixteen k in
//     ascending order, tiles of 64 x 64 entries over as many workgroups as
//     the window holds (the right-hand side is one more column, so the
//     forward substitution rides along);
//   * bands with hbw <= 127 run all of it in one launch, one workgroup
//     per system: the trailing columns one thread each (swaps and pivot rows
//     in registers, the finished rows kept in LDS), then the trailing update
//     over the workgroup's threads; wider bands take one launch for the
//     panel and trailing columns and one for the trailing update per panel;
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
constexpr int kBkMaxRows = kBkThreads;    // panel rows, one per thread: hbw + 16 <= 1024
constexpr int kBkTile = 64;               // trailing update tile (rows and columns)
constexpr int kBkStage = 256;             // trailing columns per LDS stage of the pivot rows
constexpr int kBkFusedMaxHbw = (kBkStage - 2) / 2;  // one-launch path: the window and the rhs in one stage
constexpr int kBkBack = 64;               // back substitution block rows
constexpr double kBkPivotTol = 1.0e-14;

constexpr int kBkTabDoubles = 320;  // per system: the panel's L11 and slot tables for the pivot-rows launch

struct BkSys {
    double *work, *x, *status;
    int32_t *piv;
    double *lt;   // the panel's multipliers with its later swaps applied: 16 columns x (hbw + 16) rows
    double *tab;  // BkTab of the current panel (kBkTabDoubles)
};

__device__ __forceinline__ BkSys bk_sys(const int32_t *recs, int n, int hbw, double *work, double *x,
                                        double *status) {
    const int r = recs ? recs[blockIdx.y] : 0;
    const size_t V = 3 * (size_t)hbw + 1;
    const size_t per =
        (size_t)n * V + ((size_t)n + 1) / 2 + 1 + (size_t)kBkNb * (hbw + kBkNb) + kBkTabDoubles;  // per system
    BkSys s;
    s.work = work + (size_t)r * per;
    s.piv = (int32_t *)(s.work + (size_t)n * V);
    s.lt = s.work + (size_t)n * V + ((size_t)n + 1) / 2 + 1;
    s.tab = s.lt + (size_t)kBkNb * (hbw + kBkNb);
    s.x = x ? x + (size_t)r * n : nullptr;  // null: factor only
    s.status = status + 4 * (size_t)r;
    return s;
}

__device__ __forceinline__ double lane_read(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// LDS of one panel
struct BkPanel {
    double cand_v[2][kBkNb];          // per wavefront: largest magnitude of the column (double buffered)
    int cand_i[2][kBkNb];             //   its first row
    double cand_row[2][kBkNb][kBkNb]; //   and that row's entries
    double krow[2][kBkNb];            // row k's entries (the row a swap sends down)
    double l11[kBkNb][kBkNb];         // the pivot rows' multipliers as the rows finally stand
    double rows[2][kBkNb];            // rows exchanged while the formed multipliers are restored
    int p[kBkNb];                     // row swapped with k0 + q
    int slot_row[2 * kBkNb];          // rows the swaps touch: the pivot rows, then other swapped rows
    int slot_src[2 * kBkNb];          // slot whose row's entries end in slot t after all the swaps
    int nslot;
};

// What the pivot-rows launch needs of a panel (kept in the system's tab
// scratch between the panel launch and it).
struct BkTab {
    double l11[kBkNb][kBkNb];
    int slot_row[2 * kBkNb];
    int slot_src[2 * kBkNb];
    int nslot;
};
static_assert(sizeof(BkTab) <= kBkTabDoubles * sizeof(double), "BkTab fits its scratch");

template <int CTRL>
__device__ __forceinline__ double bk_dpp(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// largest value over the wavefront, in every lane
__device__ __forceinline__ double bk_wave_max(double v) {
    v = fmax(v, bk_dpp<0xB1>(v));   // quad_perm [1,0,3,2]
    v = fmax(v, bk_dpp<0x4E>(v));   // quad_perm [2,3,0,1]
    v = fmax(v, bk_dpp<0x141>(v));  // row_half_mirror
    v = fmax(v, bk_dpp<0x140>(v));  // row_mirror
    return fmax(fmax(lane_read(v, 0), lane_read(v, 16)), fmax(lane_read(v, 32), lane_read(v, 48)));
}

// Factor panel k0 with one panel row per thread (rows k0 .. r_end, at most
// 1024): per column, a block-wide search for the first row of largest
// magnitude (wavefront reductions, then the sixteen candidates in LDS), the
// two rows exchanged through LDS, and every row below updating itself
// against the pivot row.  Rows carry their multipliers when they move, so
// on return r[] holds the panel's multipliers as the rows finally stand
// (and U for the pivot rows).  Writes the U rows, the pivot rows, the slot
// table of the trailing swaps, and with kStoreL the multipliers where they
// were formed (the interleaved-swap convention of the reference's band LU
// and of the later solves).  False (after setting status 2) for a singular
// column.  Every thread of the workgroup calls it.
template <bool kStoreL>
__device__ __forceinline__ bool bk_panel(const BkSys &s, int n, int hbw, int k0, BkPanel &L, double (&r)[kBkNb]) {
    const int V = 3 * hbw + 1;
    auto at = [&](int i, int j) -> double & { return s.work[(size_t)i * V + (j - i + hbw)]; };
    const int tid = threadIdx.x, lane = tid & 63, w = uniform(tid >> 6);
    const int kb = min(kBkNb, n - k0);
    const int r_end = min(k0 + kb - 1 + hbw, n - 1);
    const int row = k0 + tid;
    const bool active = row <= r_end;
#pragma unroll
    for (int c = 0; c < kBkNb; ++c) {
        const int j = k0 + c;
        r[c] = (active && c < kb && row - j <= hbw && j - row <= 2 * hbw) ? at(row, j) : 0.0;
    }
    bool swapped = false;
    const int nw = uniform((r_end - k0 + 64) / 64);  // wavefronts holding panel rows
    // Per column one barrier: before it, every wavefront holding rows
    // publishes its first row of largest magnitude in the column, with the
    // row's entries, and wavefront 0 row k's entries; after it, every thread
    // picks the pivot from the candidates, the two rows swap from LDS, and
    // the rows below update against the pivot row.  jj stays a loop
    // variable: a row's entries keep fixed registers, column jj is picked
    // out by selects.
#pragma unroll
    for (int jj = 0; jj < kBkNb; ++jj) {  // kb is uniform: every thread takes the same path to the barrier
        if (jj >= kb) break;
        const int k = k0 + jj, kmax = min(k + hbw, n - 1), buf = jj & 1;
        if (w < nw) {
            double rj = 0.0;
#pragma unroll
            for (int c = 0; c < kBkNb; ++c) rj = c == jj ? r[c] : rj;
            const double v = (active && row >= k && row <= kmax) ? fabs(rj) : -1.0;
            const double m = bk_wave_max(v);
            const uint64_t hit = __ballot(v == m);
            const int f = (int)__builtin_ctzll(hit);  // rows ascend with the lane: the first of equals
            if (lane == f) {
                L.cand_v[buf][w] = m;
                L.cand_i[buf][w] = row;
#pragma unroll
                for (int c = 0; c < kBkNb; ++c) L.cand_row[buf][w][c] = r[c];
            }
            if (w == 0 && lane == jj)
#pragma unroll
                for (int c = 0; c < kBkNb; ++c) L.krow[buf][c] = r[c];
        }
        __syncthreads();
        // the wavefronts' candidates side by side in lanes 0..nw-1: the
        // largest, and among equals the first wavefront (waves ascend with
        // the rows), as the scan over the rows would pick it
        const double cv = lane < nw ? L.cand_v[buf][lane] : -2.0;
        const double best = bk_wave_max(cv);
        const int wb = uniform((int)__builtin_ctzll(__ballot(cv == best)));
        const int p = uniform(L.cand_i[buf][wb]);
        if (!(best > kBkPivotTol)) {
            if (tid == 0) s.status[0] = 2.0;
            return false;  // uniform: every thread reduced the same candidates
        }
        if (tid == 0) L.p[jj] = p;
        if (w < nw) {
            if (p != k) {
                swapped = true;
                if (row == k)
#pragma unroll
                    for (int c = 0; c < kBkNb; ++c) r[c] = L.cand_row[buf][wb][c];
                if (row == p)
#pragma unroll
                    for (int c = 0; c < kBkNb; ++c) r[c] = L.krow[buf][c];
            }
            if (active && row > k && row <= kmax) {
                double aj = 0.0;
#pragma unroll
                for (int c = 0; c < kBkNb; ++c) aj = c == jj ? r[c] : aj;
                double pj = 0.0;
#pragma unroll
                for (int c = 0; c < kBkNb; ++c) pj = c == jj ? L.cand_row[buf][wb][c] : pj;
                const double l = aj / pj;
#pragma unroll
                for (int c = 0; c < kBkNb; ++c) {
                    const double pc = L.cand_row[buf][wb][c];
                    if (c == jj) r[c] = l;
                    if (c > jj && c < kb) r[c] = r[c] - l * pc;
                }
            }
        }
    }
    // U rows, and the multipliers as the rows stand for the trailing work
    if (active && row < k0 + kb) {
#pragma unroll
        for (int c = 0; c < kBkNb; ++c) {
            const int j = k0 + c;
            if (c < kb && j >= row && j - row <= 2 * hbw) at(row, j) = r[c];
            L.l11[tid][c] = (c < kb && j < row) ? r[c] : 0.0;
        }
    }
    if (tid < kb) s.piv[k0 + tid] = L.p[tid];
    if (tid == 0) {  // the rows the swaps touch, and where their entries end up
        int ns = kb;
        for (int q = 0; q < kb; ++q) L.slot_row[q] = k0 + q;
        int sl_p[kBkNb];
        for (int q = 0; q < kb; ++q) {
            const int pq = L.p[q];
            int sl = pq - k0;
            if (pq >= k0 + kb) {
                sl = -1;
                for (int t = kb; t < ns; ++t)
                    if (L.slot_row[t] == pq) sl = t;
                if (sl < 0) {
                    sl = ns++;
                    L.slot_row[sl] = pq;
                }
            }
            sl_p[q] = sl;
        }
        for (int t = 0; t < ns; ++t) L.slot_src[t] = t;
        for (int q = 0; q < kb; ++q) {
            const int t2 = sl_p[q], tmp = L.slot_src[q];
            L.slot_src[q] = L.slot_src[t2];
            L.slot_src[t2] = tmp;
        }
        L.nslot = ns;
        if (swapped && s.status[0] == 0.0) s.status[0] = 1.0;
    }
    __syncthreads();
    if (kStoreL) {  // later swaps undone in reverse on a copy, then written once
        double b[kBkNb];
#pragma unroll
        for (int c = 0; c < kBkNb; ++c) b[c] = r[c];
        for (int q = kb - 1; q > 0; --q) {
            const int pq = L.p[q], kq = k0 + q;
            if (pq == kq) continue;  // uniform
            if (row == kq)
#pragma unroll
                for (int c = 0; c < kBkNb; ++c) L.rows[0][c] = b[c];
            if (row == pq)
#pragma unroll
                for (int c = 0; c < kBkNb; ++c) L.rows[1][c] = b[c];
            __syncthreads();
            if (row == kq)
#pragma unroll
                for (int c = 0; c < kBkNb; ++c)
                    if (c < q) b[c] = L.rows[1][c];
            if (row == pq)
#pragma unroll
                for (int c = 0; c < kBkNb; ++c)
                    if (c < q) b[c] = L.rows[0][c];
            __syncthreads();
        }
        if (active)
#pragma unroll
            for (int c = 0; c < kBkNb; ++c) {
                const int j = k0 + c;
                if (c < kb && row > j && row - j <= hbw) at(row, j) = b[c];
            }
    }
    return true;
}

// The trailing columns j0 + c (c < ncr; c == nc is the right-hand side),
// kBkStage at a time through LDS: every slot (a row the panel's swaps touch)
// loaded from the row whose entry ends there (the swaps' composition), the
// pivot rows finished per column (u_kj -= l_kk' u_k'j, k' ascending), every
// slot stored back.  The swaps move only entries the band holds or
// structural zeros: a row swapped into a pivot row at a column past that
// pivot row's window is zero there.  On return (with a single chunk)
// stage[q][c] is pivot row k0 + q's finished entry (0 where the band holds
// none).  Every thread of the workgroup calls it.
__device__ __forceinline__ void bk_pivot_rows_staged(const BkSys &s, int hbw, int k0, int kb, int j0, int nc, int ncr,
                                              const BkPanel &L, double (*stage)[kBkStage]) {
    const int V = 3 * hbw + 1, tid = threadIdx.x;
    const int ns = uniform(L.nslot);
    for (int cs = 0; cs < ncr; cs += kBkStage) {
        const int cw = min(kBkStage, ncr - cs);
        for (int e = tid; e < ns * cw; e += kBkThreads) {
            const int t = e / cw, c = e % cw, j = j0 + cs + c;
            const int src = L.slot_row[L.slot_src[t]], o = j - src + hbw;
            double v = 0.0;
            if (cs + c == nc)
                v = s.x[src];
            else if (o >= 0 && o <= 3 * hbw)
                v = s.work[(size_t)src * V + o];
            stage[t][c] = v;
        }
        __syncthreads();
        if (tid < cw) {
            const int c = tid, j = j0 + cs + c;
            const bool rhs = cs + c == nc;
            double u[kBkNb];
#pragma unroll
            for (int q = 0; q < kBkNb; ++q) u[q] = (q < kb && (rhs || j <= k0 + q + 2 * hbw)) ? stage[q][c] : 0.0;
#pragma unroll
            for (int q = 1; q < kBkNb; ++q) {
                if (q < kb && (rhs || j <= k0 + q + 2 * hbw)) {
                    double t = u[q];
#pragma unroll
                    for (int q2 = 0; q2 < q; ++q2)
                        if (rhs || j <= k0 + q2 + 2 * hbw) t = t - L.l11[q][q2] * u[q2];
                    u[q] = t;
                }
            }
#pragma unroll
            for (int q = 0; q < kBkNb; ++q)
                if (q < kb) stage[q][c] = u[q];
        }
        __syncthreads();
        for (int e = tid; e < ns * cw; e += kBkThreads) {
            const int t = e / cw, c = e % cw, j = j0 + cs + c;
            const int dst = L.slot_row[t], o = j - dst + hbw;
            if (cs + c == nc)
                s.x[dst] = stage[t][c];
            else if (o >= 0 && o <= 3 * hbw)
                s.work[(size_t)dst * V + o] = stage[t][c];
        }
        __syncthreads();
    }
}

// The panel's swaps and pivot rows over the trailing columns k0 + kb ..
// jmax (and the right-hand side, column nc; nc < 0 without one): one thread
// per column.  A column's entries in every slot the swaps touch are read
// into registers (slot t takes the entries of slot_src[t]), the pivot rows
// are solved against L11 (the reference's order: u_q -= l_qp u_p, ascending
// p), and every slot is written back.  Columns are independent, so no
// barrier is needed; all of a column's loads issue before its first use.
template <typename Tab>
__device__ __forceinline__ void bk_pivot_rows_cols(const BkSys &s, int hbw, int k0, int kb, int j0, int nc,
                                                   int ncr, const Tab &L, int c_first, int c_step) {
    const int V = 3 * hbw + 1;
    const int ns = uniform(L.nslot);
    // entry (row, o) of the system at 32-bit offset row V + o (a system's
    // work holds n V doubles, under 2^32)
    const double *__restrict__ wk = s.work;
    for (int c = c_first; c < ncr; c += c_step) {
        // keeps the L11 reads in the loop (hoisted, 120 values would take
        // the registers the column needs)
        asm volatile("" ::: "memory");
        const int j = j0 + c;
        const bool rhs = c == nc;
        double v[2 * kBkNb];
#pragma unroll
        for (int t = 0; t < 2 * kBkNb; ++t) {
            v[t] = 0.0;
            if (t < ns) {
                const int src = L.slot_row[L.slot_src[t]], o = j - src + hbw;
                if (rhs)
                    v[t] = s.x[src];
                else if (o >= 0 && o <= 3 * hbw)
                    v[t] = wk[(uint32_t)src * (uint32_t)V + (uint32_t)o];
            }
        }
#pragma unroll
        for (int q = 0; q < kBkNb; ++q)
            if (q < kb && !(rhs || j <= k0 + q + 2 * hbw)) v[q] = 0.0;
#pragma unroll
        for (int q = 1; q < kBkNb; ++q) {
            if (q < kb && (rhs || j <= k0 + q + 2 * hbw)) {
                double t = v[q];
#pragma unroll
                for (int q2 = 0; q2 < q; ++q2)
                    if (rhs || j <= k0 + q2 + 2 * hbw) t = t - L.l11[q][q2] * v[q2];
                v[q] = t;
            }
        }
#pragma unroll
        for (int t = 0; t < 2 * kBkNb; ++t) {
            if (t < ns) {
                const int dst = L.slot_row[t], o = j - dst + hbw;
                if (rhs)
                    s.x[dst] = v[t];
                else if (o >= 0 && o <= 3 * hbw)
                    s.work[(uint32_t)dst * (uint32_t)V + (uint32_t)o] = v[t];
            }
        }
    }
}

// x_i = (y_i - sum_{j > i} u_ij x_j) / u_ii by blocks of 64 rows from the
// bottom; U row i at work columns i .. i + 2 hbw.  Every thread calls it.
// s_tri (LDS, kBkBack x (kBkBack + 1)): the block's triangle, staged by every
// thread while the sums run, so the triangle's sequential steps read LDS
// instead of waiting on a global load each.
__device__ __forceinline__ void bk_back(const BkSys &s, int n, int hbw, double *s_s,
                                        double (*s_tri)[kBkBack + 1]) {
    const int V = 3 * hbw + 1;
    auto at = [&](int i, int j) -> double { return s.work[(size_t)i * V + (j - i + hbw)]; };
    const int tid = threadIdx.x, lane = tid & 63, w = uniform(tid >> 6);
    for (int r1 = n; r1 > 0; r1 -= kBkBack) {
        const int r0 = max(r1 - kBkBack, 0);
        for (int e = tid; e < kBkBack * kBkBack; e += kBkThreads) {
            const int a = e / kBkBack, b = e % kBkBack, q = r0 + a, i = r0 + b;
            s_tri[a][b] = (i < r1 && q <= i && i <= q + 2 * hbw) ? at(q, i) : 0.0;
        }
        // the block's sums over the solved unknowns: a wavefront's rows
        // i = r0 + w + 16 m side by side, so their loads are in flight together
        {
            constexpr int kRows = kBkBack / kBkNb;
            double part[kRows];
            int jend[kRows];
#pragma unroll
            for (int m = 0; m < kRows; ++m) {
                const int i = r0 + w + kBkNb * m;
                part[m] = 0.0;
                jend[m] = i < r1 ? min(i + 2 * hbw, n - 1) : -1;
            }
            const int jmax = min(r0 + w + kBkNb * (kRows - 1) + 2 * hbw, n - 1);
#pragma unroll 4
            for (int j = r1 + lane; j <= jmax; j += 64) {
                const double xj = s.x[j];
#pragma unroll
                for (int m = 0; m < kRows; ++m)
                    if (j <= jend[m]) part[m] += at(r0 + w + kBkNb * m, j) * xj;
            }
#pragma unroll
            for (int m = 0; m < kRows; ++m) {
                double p = part[m];
                for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
                if (lane == 0 && jend[m] >= 0) s_s[w + kBkNb * m] = p;
            }
        }
        __syncthreads();
        if (w == 0) {  // the block's triangle, by columns
            const int q = r0 + lane;
            double t = q < r1 ? s.x[q] - s_s[lane] : 0.0;
            for (int i = r1 - 1; i >= r0; --i) {
                double xi = 0.0;
                if (lane == i - r0) xi = t / s_tri[i - r0][i - r0];
                xi = lane_read(xi, i - r0);
                if (lane == i - r0) s.x[i] = xi;
                if (q < i && i <= q + 2 * hbw) t = t - s_tri[lane][i - r0] * xi;
            }
        }
        __syncthreads();
    }
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

// Panel k0 of every listed system; the multipliers as the rows finally
// stand go to the system's scratch for bk_update_kernel, and L11 with the
// slot tables for bk_pivot_rows_kernel (the trailing columns' swaps and
// pivot rows, launched next).
template <bool kStoreL>
__global__ __launch_bounds__(kBkThreads) void bk_panel_kernel(const int32_t *recs, int n, int hbw, int k0,
                                                              double *work, double *x, double *status) {
    const BkSys s = bk_sys(recs, n, hbw, work, x, status);
    if (s.status[0] == 2.0) return;  // singular earlier: uniform exit
    __shared__ BkPanel L;
    double r[kBkNb];
    if (!bk_panel<kStoreL>(s, n, hbw, k0, L, r)) return;
    const int tid = threadIdx.x;
    const int kb = min(kBkNb, n - k0), r_end = min(k0 + kb - 1 + hbw, n - 1);
    if (tid >= kb && k0 + tid <= r_end)
#pragma unroll
        for (int c = 0; c < kBkNb; ++c)
            if (c < kb) s.lt[(size_t)c * (hbw + kBkNb) + tid] = r[c];
    // the tables the pivot-rows launch needs (bk_panel ended with a barrier
    // after they were formed)
    BkTab *tab = reinterpret_cast<BkTab *>(s.tab);
    for (int e = tid; e < kBkNb * kBkNb; e += kBkThreads) tab->l11[e / kBkNb][e % kBkNb] = L.l11[e / kBkNb][e % kBkNb];
    if (tid < 2 * kBkNb) {
        tab->slot_row[tid] = L.slot_row[tid];
        tab->slot_src[tid] = L.slot_src[tid];
    }
    if (tid == 0) tab->nslot = L.nslot;
}

// The panel's swaps and pivot rows over its trailing columns (see
// bk_pivot_rows_cols), 256 columns per workgroup, from the tables the panel
// launch left in the system's scratch.
__global__ __launch_bounds__(256) void bk_pivot_rows_kernel(const int32_t *recs, int n, int hbw, int k0,
                                                            double *work, double *x, double *status) {
    const BkSys s = bk_sys(recs, n, hbw, work, x, status);
    if (s.status[0] == 2.0) return;
    __shared__ BkTab T;
    const BkTab *tab = reinterpret_cast<const BkTab *>(s.tab);
    const int tid = threadIdx.x;
    for (int e = tid; e < kBkNb * kBkNb; e += 256) T.l11[e / kBkNb][e % kBkNb] = tab->l11[e / kBkNb][e % kBkNb];
    if (tid < 2 * kBkNb) {
        T.slot_row[tid] = tab->slot_row[tid];
        T.slot_src[tid] = tab->slot_src[tid];
    }
    if (tid == 0) T.nslot = tab->nslot;
    __syncthreads();
    const int kb = min(kBkNb, n - k0);
    const int j0 = k0 + kb, jmax = min(k0 + kb - 1 + 2 * hbw, n - 1);
    const int nc = jmax >= j0 ? jmax - j0 + 1 : 0;
    const int ncr = nc + (s.x ? 1 : 0);  // the right-hand side is column nc
    const int c = blockIdx.x * 256 + tid;
    if (c < ncr) bk_pivot_rows_cols(s, hbw, k0, kb, j0, s.x ? nc : -1, c + 1, T, c, 256);
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
    // 4 x 4 entries per thread, kept in registers over the panel's columns
    // q: each q reads four multipliers and four pivot-row entries from LDS
    // for sixteen updates.  Every entry still sees t - l_iq u_qj for
    // q = 0 .. kb - 1 in order.
    const int tr = (tid / 16) * 4, tc = (tid % 16) * 4;
    double t[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int i = i0 + ti * kBkTile + tr + a;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int cc = tj * kBkTile + tc + b;
            t[a][b] = 0.0;
            if (i <= r_end && cc < nc) t[a][b] = (has_rhs && cc == nc - 1) ? s.x[i] : at(i, j0 + cc);
        }
    }
#pragma unroll
    for (int q = 0; q < kBkNb; ++q) {
        if (q < kb) {
            double lt[4], ut[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) lt[a] = s_lt[tr + a][q];
#pragma unroll
            for (int b = 0; b < 4; ++b) ut[b] = s_ut[q][tc + b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) t[a][b] = t[a][b] - lt[a] * ut[b];
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int i = i0 + ti * kBkTile + tr + a;
        if (i > r_end) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int cc = tj * kBkTile + tc + b;
            if (cc >= nc) continue;
            if (has_rhs && cc == nc - 1)
                s.x[i] = t[a][b];
            else
                at(i, j0 + cc) = t[a][b];
        }
    }
}

__global__ __launch_bounds__(kBkThreads) void bk_back_kernel(const int32_t *recs, int n, int hbw, double *work,
                                                             double *x, double *status) {
    const BkSys s = bk_sys(recs, n, hbw, work, x, status);
    if (s.status[0] == 2.0) return;
    __shared__ double s_s[kBkBack];
    __shared__ double s_tri[kBkBack][kBkBack + 1];
    bk_back(s, n, hbw, s_s, s_tri);
}

// Narrow bands in one launch, one workgroup per system: every panel, its
// trailing columns (one thread per column: swaps and pivot rows, the
// results kept in LDS) and its trailing update (the workgroup's threads
// over the window's entries), then the back substitution.
template <int M, bool kStoreL>
__global__ __launch_bounds__(kBkThreads) void bk_fused_kernel(const int32_t *recs, int n, int hbw, double *work,
                                                              double *x, double *status) {
    const BkSys s = bk_sys(recs, n, hbw, work, x, status);
    if (s.status[0] == 2.0) return;
    const int V = 3 * hbw + 1;
    auto at = [&](int i, int j) -> double & { return s.work[(size_t)i * V + (j - i + hbw)]; };
    __shared__ BkPanel L;
    __shared__ double s_lt[64 * M][kBkNb + 1];   // multipliers as the rows finally stand
    __shared__ double stage[2 * kBkNb][kBkStage];  // the window's slots; then its finished pivot rows
    __shared__ double s_s[kBkBack];
    const int tid = threadIdx.x;
    const bool has_rhs = s.x != nullptr;
    double r[kBkNb];
    for (int k0 = 0; k0 < n; k0 += kBkNb) {
        if (!bk_panel<kStoreL>(s, n, hbw, k0, L, r)) return;
        const int kb = min(kBkNb, n - k0), r_end = min(k0 + kb - 1 + hbw, n - 1);
        if (tid >= kb && k0 + tid <= r_end)
#pragma unroll
            for (int c = 0; c < kBkNb; ++c) s_lt[tid][c] = c < kb ? r[c] : 0.0;
        const int j0 = k0 + kb, jmax = min(k0 + kb - 1 + 2 * hbw, n - 1);
        const int nc = jmax >= j0 ? jmax - j0 + 1 : 0;
        const int ncr = nc + (has_rhs ? 1 : 0);  // <= kBkStage: one stage
        __syncthreads();  // s_lt complete
        bk_pivot_rows_staged(s, hbw, k0, kb, j0, has_rhs ? nc : -1, ncr, L, stage);
        const int i0 = k0 + kb, nr = r_end - i0 + 1;
        for (int e = tid; e < nr * ncr; e += kBkThreads) {
            const int rr = e / ncr, c = e % ncr, i = i0 + rr;
            double *tp = (has_rhs && c == nc) ? s.x + i : &at(i, j0 + c);
            double t = *tp;
#pragma unroll
            for (int q = 0; q < kBkNb; ++q)
                if (q < kb) t = t - s_lt[i - k0][q] * stage[q][c];
            *tp = t;
        }
        __syncthreads();
    }
    // the back substitution's triangle takes the stage's LDS (no longer used)
    static_assert(sizeof(stage) >= sizeof(double) * kBkBack * (kBkBack + 1), "triangle fits the stage");
    if (has_rhs) bk_back(s, n, hbw, s_s, reinterpret_cast<double (*)[kBkBack + 1]>(&stage[0][0]));
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

bool band_blk_fits(int n, int hbw) { return n > 0 && hbw >= 0 && hbw + kBkNb <= kBkMaxRows; }

size_t band_blk_work_doubles(int n, int hbw) {
    return (size_t)n * (3 * (size_t)hbw + 1) + ((size_t)n + 1) / 2 + 1 + (size_t)kBkNb * (hbw + kBkNb) +
           kBkTabDoubles;
}

// Factor and solve `count` systems (recs: their record indices; nullptr for
// one system at the given pointers) on the context's stream.
// d_rhs == nullptr: factor only; d_L: keep the factors (U in the band, L,
// ipiv) for later solves.
int launch_band_blk(sgpu_ctx *ctx, int n, int hbw, double *d_band, const double *d_rhs, double *d_x,
                    double *d_status, double *d_L, int32_t *d_ipiv, double *d_work, const int32_t *recs, int count) {
    if (!band_blk_fits(n, hbw))
        return fail(ctx, SGPU_ERR_ARG, "half bandwidth %d exceeds the blocked band solve (%d)", hbw,
                    kBkMaxRows - kBkNb);
    if (!d_work || (d_rhs && !d_x)) return fail(ctx, SGPU_ERR_ARG, "the blocked band solve needs work (and x)");
    if (!d_rhs) d_x = nullptr;
    hipStream_t st = ctx->stream;
    const int64_t cells = (int64_t)n * (3 * hbw + 1);
    hipLaunchKernelGGL(bk_init_kernel, dim3((unsigned)((cells + 255) / 256), count), dim3(256), 0, st, recs, n, hbw,
                       d_band, d_rhs, d_work, d_x, d_status);
    const int M = (hbw + kBkNb + 63) / 64;
    const bool keep = d_L != nullptr;  // the multipliers where they were formed, for later solves
    if (hbw <= kBkFusedMaxHbw) {  // one launch: panels, trailing work and back substitution in one workgroup
        auto fused = keep ? (M <= 1 ? bk_fused_kernel<1, true> : M == 2 ? bk_fused_kernel<2, true>
                                                                       : bk_fused_kernel<3, true>)
                          : (M <= 1 ? bk_fused_kernel<1, false> : M == 2 ? bk_fused_kernel<2, false>
                                                                        : bk_fused_kernel<3, false>);
        hipLaunchKernelGGL(fused, dim3(1, count), dim3(kBkThreads), 0, st, recs, n, hbw, d_work, d_x, d_status);
    } else {
        auto panel = keep ? bk_panel_kernel<true> : bk_panel_kernel<false>;
        for (int k0 = 0; k0 < n; k0 += kBkNb) {
            hipLaunchKernelGGL(panel, dim3(1, count), dim3(kBkThreads), 0, st, recs, n, hbw, k0, d_work, d_x,
                               d_status);
            const int kb = std::min(kBkNb, n - k0);
            {
                const int jmax = std::min(k0 + kb - 1 + 2 * hbw, n - 1);
                const int ncr = (jmax >= k0 + kb ? jmax - (k0 + kb) + 1 : 0) + (d_x ? 1 : 0);
                if (ncr > 0)
                    hipLaunchKernelGGL(bk_pivot_rows_kernel, dim3((ncr + 255) / 256, count), dim3(256), 0, st, recs,
                                       n, hbw, k0, d_work, d_x, d_status);
            }
            const int nr = std::min(k0 + kb - 1 + hbw, n - 1) - (k0 + kb) + 1;
            if (nr > 0) {
                const int jmax = std::min(k0 + kb - 1 + 2 * hbw, n - 1);
                const int nc = (jmax >= k0 + kb ? jmax - (k0 + kb) + 1 : 0) + 1;
                const int tiles = ((nr + kBkTile - 1) / kBkTile) * ((nc + kBkTile - 1) / kBkTile);
                hipLaunchKernelGGL(bk_update_kernel, dim3(tiles, count), dim3(256), 0, st, recs, n, hbw, k0, d_work,
                                   d_x, d_status);
            }
        }
        if (d_x)
            hipLaunchKernelGGL(bk_back_kernel, dim3(1, count), dim3(kBkThreads), 0, st, recs, n, hbw, d_work, d_x,
                               d_status);
    }
    if (d_L || !d_x) {
        const int64_t slots = (int64_t)n * (2 * hbw + 1);
        hipLaunchKernelGGL(bk_store_kernel, dim3((unsigned)((slots + 255) / 256), count), dim3(256), 0, st, recs, n,
                           hbw, d_work, d_status, d_band, d_L, d_ipiv);
    }
    HIP_OK(ctx, hipGetLastError());
    return SGPU_OK;
}

}  // namespace sgpu
