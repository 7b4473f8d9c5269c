// 3D fiber beam-column elements on gfx950: forceBeamColumn3d
// (force_beam_column3d.mojo:1203-1607, :2549-2990) and dispBeamColumn3d
// (beam_column3d_nonlinear.mojo:357-575) for fiber sections without element
// loads, Linear or PDelta transformation, elastic torsion GJ/L.
//
// One block per element, one wavefront per integration point.  A section
// evaluation runs on all integration points at once (wave_section_3d on
// wavefront w for point w); the element-level arithmetic (basic forces,
// 5 x 5 flexibility, local iteration, subdivision) runs uniformly on every
// thread with scalars in registers and per-point data in LDS.  Every
// expression keeps the reference's operation order, including terms whose
// factors are zero, so that the element tangents and forces are the same bits
// as the CPU path's whenever the section results are.
//
// Trial fiber states are stored by every section evaluation; the last
// evaluation of each point is the converged one, so the commit of an element
// is the copy of its sections' trial states (fiber3d.mojo:873-926).
#include <hip/hip_runtime.h>

#include "elem3d_types.h"
#include "material_laws.h"
#include "section_device.h"

namespace sgpu {

constexpr int kElem3dThreads = 64 * kMaxNip3d;
constexpr int kMats3dInLds = 16;  // material tables up to this size are staged in LDS
constexpr int kMaxTries3d = 4096;  // forceBeamColumn3d: tries of one state determination

struct Sec3dOut {
    double N, My, Mz, k11, k12, k13, k22, k23, k33;
};

// fiber_section3d_set_trial_from_offset (fiber3d.mojo:769-836) by one wave:
// elastic W-chunks folded by the halving tree, then every other slot as a
// single chunk in input order; lanes 0..8 add the chunk partials one after
// another.  The uniform-modulus rule zeroes k12 and k13 after the elastic
// chunks (:245-254).  The trial state of every material slot is stored.
// The definition needs at most 63 batches.
__device__ Sec3dOut wave_section_3d(const SecDefDev &d, const SlotTables &tb, const MatDev *mats,
                                    const StateArrays &c, const StateArrays &t, int inst, double eps0, double ky,
                                    double kz, double *red) {
    const int lane = threadIdx.x & 63;
    const int n_batches = d.n_batches, n_el = d.n_elastic_chunks, n_slots = d.n_slots;
    const int64_t inst_state = d.state_base + (int64_t)inst * n_slots;
    const int64_t inst_aux = d.aux_base + (int64_t)inst * d.n_aux;
    const int bcb_lane = lane <= n_batches ? tb.batch_chunk_begin[d.batch_base + lane] : 0;
    const bool zero_lane = d.zero_k12 && (lane == 4 || lane == 5);
    double acc = 0.0;  // lane q < 9: running sum of resultant q
    for (int b = 0; b < n_batches; ++b) {
        const int s = b * kBatch + lane;
        double fs = 0.0, ks = 0.0, y = 0.0, z = 0.0;
        int code = -1;
        if (s < n_slots) {
            const int64_t gs = d.slot_base + s;
            const int mi = tb.mat[gs];
            code = tb.chunk[gs];
            y = tb.y[gs];
            z = tb.z[gs];
            if (mi >= 0) {
                const int aux = tb.aux[gs];
                const double strain = eps0 + z * ky - y * kz;
                eval_slot(mats, mi, strain, tb.area[gs], c, t, inst_state + s, inst_aux + (aux < 0 ? 0 : aux), fs,
                          ks);
            }
        }
        double term[9] = {fs, fs * z, -fs * y, ks, ks * z, -ks * y, ks * z * z, -ks * z * y, ks * y * y};
        const bool w8 = code >= 0 && (code & 1);
        if (__builtin_amdgcn_ballot_w64(w8) != 0) {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const double r = tree8_dpp(term[q]);
                if (w8) term[q] = r;
            }
        }
        if (code >= 0) {
#pragma unroll
            for (int q = 0; q < 9; ++q) red[q * kWaveRedStride + (code >> 1)] = term[q];
        }
        wave_sync();
        const int c0 = __builtin_amdgcn_readlane(bcb_lane, b), c1 = __builtin_amdgcn_readlane(bcb_lane, b + 1);
        if (lane < 9) {
            const double *row = red + lane * kWaveRedStride - c0;
            int cc = c0;
            if (zero_lane && n_el >= c0 && n_el < c1) {
                for (; cc < n_el; ++cc) acc += row[cc];
                acc = 0.0;  // uniform elastic modulus: k12, k13 restart with the nonlinear fibers
            }
            for (; cc < c1; ++cc) acc += row[cc];
        }
        wave_sync();
    }
    if (zero_lane && __builtin_amdgcn_readlane(bcb_lane, n_batches) <= n_el) acc = 0.0;
    Sec3dOut o;
    o.N = readlane_d(acc, 0);
    o.My = readlane_d(acc, 1);
    o.Mz = readlane_d(acc, 2);
    o.k11 = readlane_d(acc, 3);
    o.k12 = readlane_d(acc, 4);
    o.k13 = readlane_d(acc, 5);
    o.k22 = readlane_d(acc, 6);
    o.k23 = readlane_d(acc, 7);
    o.k33 = readlane_d(acc, 8);
    return o;
}

// _invert_3x3_values (force_beam_column3d.mojo:679-727) as
// _fiber_section3d_response_flexibility uses it (:1093-1117):
// f = {f11, f12, f13, f22, f23, f33}.
__device__ __forceinline__ bool section_flex3(const Sec3dOut &r, double *f) {
    const double a = r.k11, b = r.k12, c = r.k13, d = r.k12, e = r.k22, ff = r.k23, g = r.k13, h = r.k23,
                 i = r.k33;
    const double co00 = e * i - ff * h;
    const double co01 = -(d * i - ff * g);
    const double co02 = d * h - e * g;
    const double co10 = -(b * i - c * h);
    const double co11 = a * i - c * g;
    const double co20 = b * ff - c * e;
    const double co21 = -(a * ff - c * d);
    const double co22 = a * e - b * d;
    const double det = a * co00 + b * co01 + c * co02;
    if (fabs(det) <= 1.0e-40) return false;
    const double inv_det = 1.0 / det;
    f[0] = co00 * inv_det;
    f[1] = co10 * inv_det;
    f[2] = co20 * inv_det;
    f[3] = co11 * inv_det;
    f[4] = co21 * inv_det;
    f[5] = co22 * inv_det;
    return true;
}

// _invert_5x5_flat_into (force_beam_column3d.mojo:788-833): Gauss-Jordan
// with partial pivoting; a is overwritten.
__device__ __forceinline__ bool invert5(double *a, double *inv) {
#pragma unroll
    for (int k = 0; k < 25; ++k) inv[k] = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) inv[i * 5 + i] = 1.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        int pivot = i;
        double max_val = fabs(a[i * 5 + i]);
#pragma unroll
        for (int r = i + 1; r < 5; ++r) {
            const double v = fabs(a[r * 5 + i]);
            if (v > max_val) {
                max_val = v;
                pivot = r;
            }
        }
        if (max_val <= 1.0e-40) return false;
        if (pivot != i) {
#pragma unroll
            for (int r = i + 1; r < 5; ++r) {
                if (r != pivot) continue;
#pragma unroll
                for (int cc = 0; cc < 5; ++cc) {
                    double tmp = a[i * 5 + cc];
                    a[i * 5 + cc] = a[r * 5 + cc];
                    a[r * 5 + cc] = tmp;
                    tmp = inv[i * 5 + cc];
                    inv[i * 5 + cc] = inv[r * 5 + cc];
                    inv[r * 5 + cc] = tmp;
                }
            }
        }
        const double piv = a[i * 5 + i];
#pragma unroll
        for (int cc = 0; cc < 5; ++cc) {
            a[i * 5 + cc] /= piv;
            inv[i * 5 + cc] /= piv;
        }
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            if (r == i) continue;
            const double factor = a[r * 5 + i];
            if (factor == 0.0) continue;
#pragma unroll
            for (int cc = 0; cc < 5; ++cc) {
                a[r * 5 + cc] -= factor * a[i * 5 + cc];
                inv[r * 5 + cc] -= factor * inv[i * 5 + cc];
            }
        }
    }
    return true;
}

// force_beam_column3d.mojo:419-463 (rows 0..4 basic forces, row 5 torsion)
__device__ __forceinline__ double basic_coeff3(int row, int col, double inv_L) {
    switch (row) {
    case 0: return col == 0 ? -1.0 : (col == 6 ? 1.0 : 0.0);
    case 1: return col == 1 ? inv_L : (col == 5 ? 1.0 : (col == 7 ? -inv_L : 0.0));
    case 2: return col == 1 ? inv_L : (col == 7 ? -inv_L : (col == 11 ? 1.0 : 0.0));
    case 3: return col == 2 ? -inv_L : (col == 4 ? 1.0 : (col == 8 ? inv_L : 0.0));
    case 4: return col == 2 ? -inv_L : (col == 8 ? inv_L : (col == 10 ? 1.0 : 0.0));
    default: return col == 3 ? -1.0 : (col == 9 ? 1.0 : 0.0);
    }
}

// (N, My, Mz) components of row a of b(xi) (:1490-1505)
__device__ __forceinline__ void basic_b3(int a, double b_mi, double b_mj, double &bn, double &bmy, double &bmz) {
    bn = 0.0;
    bmy = 0.0;
    bmz = 0.0;
    if (a == 0)
        bn = 1.0;
    else if (a == 1)
        bmz = b_mi;
    else if (a == 2)
        bmz = b_mj;
    else if (a == 3)
        bmy = b_mi;
    else
        bmy = b_mj;
}

struct Elem3dLds {
    double res[kMaxNip3d][10];  // per point: N, My, Mz, then f11..f33 (force-based) or k11..k33, valid
    double qs[5 + 3 * kMaxNip3d];
    double vs_e[kMaxNip3d], vs_ky[kMaxNip3d], vs_kz[kMaxNip3d];
    double sub_e[kMaxNip3d], sub_ky[kMaxNip3d], sub_kz[kMaxNip3d];
    double sub_n[kMaxNip3d], sub_my[kMaxNip3d], sub_mz[kMaxNip3d];
    double sub_f[kMaxNip3d][6], init_f[kMaxNip3d][6];
    double pe[kMaxNip3d], pky[kMaxNip3d], pkz[kMaxNip3d];
    double k_basic[25];
    double kl[144], work[144];
    double red[kMaxNip3d][9 * kWaveRedStride];
    MatDev smats[kMats3dInLds];
};

struct Elem3dCtx {
    const Elem3dArgs *A;
    const Elem3dDev *el;
    Elem3dLds *S;
    const MatDev *mats;
};

// Section evaluation of every integration point at (pe, pky, pkz), called by
// every thread; flex: store the point flexibility (else the tangent).
// Returns whether every point is valid (non-singular).
__device__ bool eval3(const Elem3dCtx &E, bool flex) {
    const Elem3dArgs &A = *E.A;
    const Elem3dDev &el = *E.el;
    Elem3dLds &S = *E.S;
    const int w = threadIdx.x >> 6;
    __syncthreads();  // inputs written, previous results read
    if (w < el.nip) {
        const Sec3dOut r =
            wave_section_3d(A.defs[el.def], A.tb, E.mats, A.c, A.t, el.inst0 + w, S.pe[w], S.pky[w], S.pkz[w], S.red[w]);
        if ((threadIdx.x & 63) == 0) {
            double *o = S.res[w];
            o[0] = r.N;
            o[1] = r.My;
            o[2] = r.Mz;
            if (flex) {
                o[9] = section_flex3(r, o + 3) ? 1.0 : 0.0;
            } else {
                o[3] = r.k11;
                o[4] = r.k12;
                o[5] = r.k13;
                o[6] = r.k22;
                o[7] = r.k23;
                o[8] = r.k33;
                o[9] = 1.0;
            }
        }
    }
    __syncthreads();
    bool ok = true;
    for (int ip = 0; ip < el.nip; ++ip)
        if (S.res[ip][9] == 0.0) ok = false;
    return ok;
}

// _force_beam_column3d_try_increment (force_beam_column3d.mojo:1203-1607).
// q-state in S.qs; returns true with the converged basic forces in qout.
__device__ bool fbc3_try(const Elem3dCtx &E, double L, const double *base_v, const double *dv, int scheme,
                         double *qout) {
    const Elem3dDev &el = *E.el;
    Elem3dLds &S = *E.S;
    const int nip = el.nip;
    double q0 = S.qs[0], q1 = S.qs[1], q2 = S.qs[2], q3 = S.qs[3], q4 = S.qs[4];
    double target[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) target[k] = base_v[k] + dv[k];
    // predictor sections at the current deformations (:1263-1303)
    for (int ip = 0; ip < nip; ++ip) {
        const double e0 = S.qs[5 + ip], ky = S.qs[5 + nip + ip], kz = S.qs[5 + 2 * nip + ip];
        S.vs_e[ip] = e0;
        S.vs_ky[ip] = ky;
        S.vs_kz[ip] = kz;
        S.pe[ip] = e0;
        S.pky[ip] = ky;
        S.pkz[ip] = kz;
    }
    if (!eval3(E, true)) return false;
    for (int ip = 0; ip < nip; ++ip) {
        S.sub_e[ip] = S.pe[ip];
        S.sub_ky[ip] = S.pky[ip];
        S.sub_kz[ip] = S.pkz[ip];
        S.sub_n[ip] = S.res[ip][0];
        S.sub_my[ip] = S.res[ip][1];
        S.sub_mz[ip] = S.res[ip][2];
        for (int k = 0; k < 6; ++k) S.sub_f[ip][k] = S.res[ip][3 + k];
    }
    // initial flexibility from the committed state at zero deformation (:1305-1322)
    __syncthreads();  // every wave has read this evaluation's deformations
    for (int ip = 0; ip < nip; ++ip) {
        S.pe[ip] = 0.0;
        S.pky[ip] = 0.0;
        S.pkz[ip] = 0.0;
    }
    if (!eval3(E, true)) return false;
    for (int ip = 0; ip < nip; ++ip)
        for (int k = 0; k < 6; ++k) S.init_f[ip][k] = S.res[ip][3 + k];
    {  // predictor stiffness and force update (:1323-1393)
        double pf[25];
#pragma unroll
        for (int k = 0; k < 25; ++k) pf[k] = 0.0;
        for (int ip = 0; ip < nip; ++ip) {
            const double xi = el.xi[ip], wL = el.w[ip] * L, b_mi = xi - 1.0, b_mj = xi;
            const double *f = S.sub_f[ip];
            const double f11 = f[0], f12 = f[1], f13 = f[2], f22 = f[3], f23 = f[4], f33 = f[5];
            pf[0] += wL * f11;
            pf[1] += wL * f13 * b_mi;
            pf[2] += wL * f13 * b_mj;
            pf[3] += wL * f12 * b_mi;
            pf[4] += wL * f12 * b_mj;
            pf[5] += wL * b_mi * f13;
            pf[6] += wL * b_mi * f33 * b_mi;
            pf[7] += wL * b_mi * f33 * b_mj;
            pf[8] += wL * b_mi * f23 * b_mi;
            pf[9] += wL * b_mi * f23 * b_mj;
            pf[10] += wL * b_mj * f13;
            pf[11] += wL * b_mj * f33 * b_mi;
            pf[12] += wL * b_mj * f33 * b_mj;
            pf[13] += wL * b_mj * f23 * b_mi;
            pf[14] += wL * b_mj * f23 * b_mj;
            pf[15] += wL * b_mi * f12;
            pf[16] += wL * b_mi * f23 * b_mi;
            pf[17] += wL * b_mi * f23 * b_mj;
            pf[18] += wL * b_mi * f22 * b_mi;
            pf[19] += wL * b_mi * f22 * b_mj;
            pf[20] += wL * b_mj * f12;
            pf[21] += wL * b_mj * f23 * b_mi;
            pf[22] += wL * b_mj * f23 * b_mj;
            pf[23] += wL * b_mj * f22 * b_mi;
            pf[24] += wL * b_mj * f22 * b_mj;
        }
        double k[25];
        const bool inv_ok = invert5(pf, k);
        // the reference's k_basic holds this inverse (also a partial one) until the next inversion
#pragma unroll
        for (int j = 0; j < 25; ++j) S.k_basic[j] = k[j];
        if (inv_ok) {
            q0 += k[0] * dv[0] + k[1] * dv[1] + k[2] * dv[2] + k[3] * dv[3] + k[4] * dv[4];
            q1 += k[5] * dv[0] + k[6] * dv[1] + k[7] * dv[2] + k[8] * dv[3] + k[9] * dv[4];
            q2 += k[10] * dv[0] + k[11] * dv[1] + k[12] * dv[2] + k[13] * dv[3] + k[14] * dv[4];
            q3 += k[15] * dv[0] + k[16] * dv[1] + k[17] * dv[2] + k[18] * dv[3] + k[19] * dv[4];
            q4 += k[20] * dv[0] + k[21] * dv[1] + k[22] * dv[2] + k[23] * dv[3] + k[24] * dv[4];
        }
    }
    const int n_iters = scheme == 1 ? 100 : 10;
    for (int it = 0; it < n_iters; ++it) {  // :1395-1605
        const bool use_init = scheme == 1 || (scheme == 2 && it == 0);
        __syncthreads();  // every wave has read the previous evaluation's deformations
        for (int ip = 0; ip < nip; ++ip) {
            const double xi = el.xi[ip], b_mi = xi - 1.0, b_mj = xi;
            const double axial_target = q0 + 0.0;
            const double mz_target = b_mi * q1 + b_mj * q2 + 0.0;
            const double my_target = b_mi * q3 + b_mj * q4 + 0.0;
            const double dss0 = axial_target - S.sub_n[ip];
            const double dss1 = my_target - S.sub_my[ip];
            const double dss2 = mz_target - S.sub_mz[ip];
            const double *sf = use_init ? S.init_f[ip] : S.sub_f[ip];
            const double s11 = sf[0], s12 = sf[1], s13 = sf[2], s22 = sf[3], s23 = sf[4], s33 = sf[5];
            const double dvs0 = s11 * dss0 + s12 * dss1 + s13 * dss2;
            const double dvs1 = s12 * dss0 + s22 * dss1 + s23 * dss2;
            const double dvs2 = s13 * dss0 + s23 * dss1 + s33 * dss2;
            S.pe[ip] = S.sub_e[ip] + dvs0;
            S.pky[ip] = S.sub_ky[ip] + dvs1;
            S.pkz[ip] = S.sub_kz[ip] + dvs2;
        }
        if (!eval3(E, true)) return false;
        double fb[25], vf[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 25; ++k) fb[k] = 0.0;
        for (int ip = 0; ip < nip; ++ip) {
            const double xi = el.xi[ip], wL = el.w[ip] * L, b_mi = xi - 1.0, b_mj = xi;
            const double axial_target = q0 + 0.0;
            const double mz_target = b_mi * q1 + b_mj * q2 + 0.0;
            const double my_target = b_mi * q3 + b_mj * q4 + 0.0;
            const double v0 = S.pe[ip], v1 = S.pky[ip], v2 = S.pkz[ip];
            const double *r = S.res[ip];
            const double f6[6] = {r[3], r[4], r[5], r[6], r[7], r[8]};
            S.sub_e[ip] = v0;
            S.sub_ky[ip] = v1;
            S.sub_kz[ip] = v2;
            S.sub_n[ip] = r[0];
            S.sub_my[ip] = r[1];
            S.sub_mz[ip] = r[2];
            for (int k = 0; k < 6; ++k) S.sub_f[ip][k] = f6[k];
            const double dr0 = axial_target - r[0];
            const double dr1 = my_target - r[1];
            const double dr2 = mz_target - r[2];
            const double e0 = f6[0] * dr0 + f6[1] * dr1 + f6[2] * dr2;
            const double e1 = f6[1] * dr0 + f6[3] * dr1 + f6[4] * dr2;
            const double e2 = f6[2] * dr0 + f6[4] * dr1 + f6[5] * dr2;
            vf[0] += wL * (v0 + e0);
            vf[1] += wL * b_mi * (v2 + e2);
            vf[2] += wL * b_mj * (v2 + e2);
            vf[3] += wL * b_mi * (v1 + e1);
            vf[4] += wL * b_mj * (v1 + e1);
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                double an, amy, amz;
                basic_b3(a, b_mi, b_mj, an, amy, amz);
#pragma unroll
                for (int b = 0; b < 5; ++b) {
                    double bn, bmy, bmz;
                    basic_b3(b, b_mi, b_mj, bn, bmy, bmz);
                    const double dEps = f6[0] * bn + f6[1] * bmy + f6[2] * bmz;
                    const double dKy = f6[1] * bn + f6[3] * bmy + f6[4] * bmz;
                    const double dKz = f6[2] * bn + f6[4] * bmy + f6[5] * bmz;
                    fb[a * 5 + b] += wL * (an * dEps + amy * dKy + amz * dKz);
                }
            }
        }
        double k[25];
        const bool inv_ok = invert5(fb, k);
#pragma unroll
        for (int j = 0; j < 25; ++j) S.k_basic[j] = k[j];
        if (!inv_ok) return false;
        double res[5], dq[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) res[j] = target[j] - vf[j];
#pragma unroll
        for (int r = 0; r < 5; ++r)
            dq[r] = k[r * 5] * res[0] + k[r * 5 + 1] * res[1] + k[r * 5 + 2] * res[2] + k[r * 5 + 3] * res[3] +
                    k[r * 5 + 4] * res[4];
        const double work = fabs(res[0] * dq[0] + res[1] * dq[1] + res[2] * dq[2] + res[3] * dq[3] + res[4] * dq[4]);
        q0 += dq[0];
        q1 += dq[1];
        q2 += dq[2];
        q3 += dq[3];
        q4 += dq[4];
        if (work < 1.0e-12) {
            S.qs[0] = q0;
            S.qs[1] = q1;
            S.qs[2] = q2;
            S.qs[3] = q3;
            S.qs[4] = q4;
            for (int ip = 0; ip < nip; ++ip) {
                S.qs[5 + ip] = S.sub_e[ip];
                S.qs[5 + nip + ip] = S.sub_ky[ip];
                S.qs[5 + 2 * nip + ip] = S.sub_kz[ip];
                S.vs_e[ip] = S.sub_e[ip];
                S.vs_ky[ip] = S.sub_ky[ip];
                S.vs_kz[ip] = S.sub_kz[ip];
            }
            qout[0] = q0;
            qout[1] = q1;
            qout[2] = q2;
            qout[3] = q3;
            qout[4] = q4;
            return true;
        }
    }
    return false;
}

// :1149-1173
__device__ __forceinline__ void fbc3_restore(Elem3dLds &S, int nip, const double *qa) {
#pragma unroll
    for (int k = 0; k < 5; ++k) S.qs[k] = qa[k];
    for (int ip = 0; ip < nip; ++ip) {
        S.qs[5 + ip] = S.vs_e[ip];
        S.qs[5 + nip + ip] = S.vs_ky[ip];
        S.qs[5 + 2 * nip + ip] = S.vs_kz[ip];
    }
}

// Element outputs: k_global = R^T-blocks (k_local R-blocks)
// (force_beam_column3d.mojo:836-867) and f_global (:371-385); S.kl holds
// k_local, fl the local forces (uniform registers).  Every thread.
__device__ void fbc3_outputs(const Elem3dCtx &E, const double *fl, double *k_out, double *f_out) {
    const Elem3dDev &el = *E.el;
    Elem3dLds &S = *E.S;
    const double *R = el.R;
    __syncthreads();
    for (int t = threadIdx.x; t < 144; t += blockDim.x) {
        const int row = t / 12, col = t % 12, j = col % 3, base = row * 12 + (col - j);
        const double s0 = S.kl[base], s1 = S.kl[base + 1], s2 = S.kl[base + 2];
        S.work[t] = s0 * R[j] + s1 * R[3 + j] + s2 * R[6 + j];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 144; t += blockDim.x) {
        const int row = t / 12, col = t % 12, i = row % 3, rb = row - i;
        k_out[t] = R[i] * S.work[rb * 12 + col] + R[3 + i] * S.work[(rb + 1) * 12 + col] +
                   R[6 + i] * S.work[(rb + 2) * 12 + col];
    }
    if (threadIdx.x < 12) {
        const int t = threadIdx.x, i = t % 3, o = t - i;
        f_out[t] = R[i] * fl[o] + R[3 + i] * fl[o + 1] + R[6 + i] * fl[o + 2];
    }
}

// force_beam_column3d_fiber_global_tangent_and_internal
// (force_beam_column3d.mojo:2549-2990).  Returns 0, or 1 when the element
// compatibility does not converge (the reference aborts there, :2846).
__device__ int fbc3d_element(const Elem3dCtx &E, const double *u, double *q_glob, double *k_out, double *f_out) {
    const Elem3dDev &el = *E.el;
    Elem3dLds &S = *E.S;
    const int nip = el.nip, nq = 5 + 3 * nip;
    const double L = el.L;
    const double *R = el.R;
    for (int k = threadIdx.x; k < nq; k += blockDim.x) S.qs[k] = q_glob[k];
    __syncthreads();
    double ul[12];  // :340-355
#pragma unroll
    for (int blk = 0; blk < 4; ++blk) {
        const int o = 3 * blk;
        const double g0 = u[o], g1 = u[o + 1], g2 = u[o + 2];
#pragma unroll
        for (int i = 0; i < 3; ++i) ul[o + i] = R[3 * i] * g0 + R[3 * i + 1] * g1 + R[3 * i + 2] * g2;
    }
    // :1610-1622
    const double chord_z = (ul[7] - ul[1]) / L, chord_y = (ul[8] - ul[2]) / L;
    const double vb[5] = {ul[6] - ul[0], ul[5] - chord_z, ul[11] - chord_z, ul[4] + chord_y, ul[10] + chord_y};
    const double torsion_basic = ul[9] - ul[3];
    // accepted basic deformation from the section deformations (:1176-1200)
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int ip = 0; ip < nip; ++ip) {
        const double xi = el.xi[ip], wL = el.w[ip] * L, b_mi = xi - 1.0, b_mj = xi;
        acc[0] += wL * S.qs[5 + ip];
        acc[1] += wL * b_mi * S.qs[5 + 2 * nip + ip];
        acc[2] += wL * b_mj * S.qs[5 + 2 * nip + ip];
        acc[3] += wL * b_mi * S.qs[5 + nip + ip];
        acc[4] += wL * b_mj * S.qs[5 + nip + ip];
    }
    double qa[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) qa[k] = S.qs[k];
    for (int ip = 0; ip < nip; ++ip) {
        S.vs_e[ip] = S.qs[5 + ip];
        S.vs_ky[ip] = S.qs[5 + nip + ip];
        S.vs_kz[ip] = S.qs[5 + 2 * nip + ip];
    }
    for (int j = 0; j < 25; ++j) S.k_basic[j] = 0.0;
    // subdivision, cutback and schemes (:2733-2831)
    bool converged = false;
    double att[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) att[k] = vb[k] - acc[k];
    int num_subdivide = 1;
    // Each accepted increment restarts the subdivision count, as in the
    // reference; the device bounds the total number of tries, so that a
    // compatibility loop that keeps accepting tiny increments ends with the
    // non-converged status instead of running on.
    int tries = 0;
    while (num_subdivide <= 4 && tries < kMaxTries3d) {
        double target[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) target[k] = acc[k] + att[k];
        bool success = false;
        for (int scheme = 0; scheme < 3; ++scheme) {
            ++tries;
            fbc3_restore(S, nip, qa);
            double qs[5];
            if (!fbc3_try(E, L, acc, att, scheme, qs)) continue;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                acc[k] = target[k];
                qa[k] = qs[k];
            }
            for (int ip = 0; ip < nip; ++ip) {
                S.vs_e[ip] = S.qs[5 + ip];
                S.vs_ky[ip] = S.qs[5 + nip + ip];
                S.vs_kz[ip] = S.qs[5 + 2 * nip + ip];
            }
            double rem[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) rem[k] = vb[k] - acc[k];
            const double m01 = fmax(fabs(rem[0]), fabs(rem[1]));
            const double m23 = fmax(fabs(rem[2]), fabs(rem[3]));
            const double rn = fmax(m01, fmax(m23, fabs(rem[4])));
            if (rn <= 1.0e-12) {
                converged = true;
            } else {
#pragma unroll
                for (int k = 0; k < 5; ++k) att[k] = rem[k];
                num_subdivide = 1;
            }
            success = true;
            break;
        }
        if (converged) break;
        if (success) continue;
#pragma unroll
        for (int k = 0; k < 5; ++k) att[k] /= 10.0;
        num_subdivide += 1;
    }
    if (!converged) fbc3_restore(S, nip, qa);
    __syncthreads();
    for (int k = threadIdx.x; k < nq; k += blockDim.x) q_glob[k] = S.qs[k];
    if (!converged) return 1;
    const double q0 = qa[0], q1 = qa[1], q2 = qa[2], q3 = qa[3], q4 = qa[4];
    const double inv_L = 1.0 / L;
    const bool torsion = el.G > 0.0 && el.J > 0.0;
    const double torsion_k = torsion ? el.G * el.J / L : 0.0;
    // k_local (:2867-2922), one entry per thread
    for (int t = threadIdx.x; t < 144; t += blockDim.x) {
        const int a = t / 12, b = t % 12;
        double v = 0.0;
        for (int i = 0; i < 5; ++i) {
            const double ai = basic_coeff3(i, a, inv_L);
            for (int j = 0; j < 5; ++j) v += ai * S.k_basic[i * 5 + j] * basic_coeff3(j, b, inv_L);
        }
        if (torsion) v += basic_coeff3(5, a, inv_L) * torsion_k * basic_coeff3(5, b, inv_L);
        if (el.transf == 1) {
            if ((a == 1 && b == 1) || (a == 7 && b == 7) || (a == 2 && b == 2) || (a == 8 && b == 8)) v += q0 / L;
            if ((a == 1 && b == 7) || (a == 7 && b == 1) || (a == 2 && b == 8) || (a == 8 && b == 2)) v -= q0 / L;
        }
        S.kl[t] = v;
    }
    // f_local (:2924-2953)
    const double q5 = torsion ? (el.G * el.J / L) * torsion_basic : 0.0;
    double fl[12];
#pragma unroll
    for (int a = 0; a < 12; ++a) {
        fl[a] = 0.0;
        fl[a] += basic_coeff3(0, a, inv_L) * q0;
        fl[a] += basic_coeff3(1, a, inv_L) * q1;
        fl[a] += basic_coeff3(2, a, inv_L) * q2;
        fl[a] += basic_coeff3(3, a, inv_L) * q3;
        fl[a] += basic_coeff3(4, a, inv_L) * q4;
        fl[a] += basic_coeff3(5, a, inv_L) * q5;
    }
    fl[0] += 0.0;  // fixed-end reactions of element loads (none)
    fl[1] += 0.0;
    fl[2] += 0.0;
    fl[7] += 0.0;
    fl[8] += 0.0;
    fbc3_outputs(E, fl, k_out, f_out);
    return 0;
}

// beam_column3d_fiber_global_tangent_and_internal
// (beam_column3d_nonlinear.mojo:357-575): T = diag(R, R, R, R),
// u_local = T u, section kinematics (:154-171), k_local / f_local sums over
// the points, torsion, P-Delta, then k_global = T^T (k_local T) by
// linalg.mojo:20-36 (terms with a zero left factor skipped) and f_global = T^T f.
__device__ int dbc3d_element(const Elem3dCtx &E, const double *u, double *k_out, double *f_out) {
    const Elem3dDev &el = *E.el;
    Elem3dLds &S = *E.S;
    const int nip = el.nip;
    const double L = el.L;
    const double *R = el.R;
    auto T = [&](int i, int j) -> double {  // 12 x 12 transformation
        return (i / 3 == j / 3) ? R[3 * (i % 3) + (j % 3)] : 0.0;
    };
    double ul[12];  // :92-102
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        double sum = 0.0;
#pragma unroll
        for (int j = 0; j < 12; ++j) sum += T(i, j) * u[j];
        ul[i] = sum;
    }
    const int w = threadIdx.x >> 6;
    if (w < nip) {
        const double xi = el.xi[w], inv_L = 1.0 / L;
        S.pe[w] = (-inv_L) * ul[0] + inv_L * ul[6];
        S.pky[w] = ((-6.0 + 12.0 * xi) / (L * L)) * ul[2] + ((4.0 - 6.0 * xi) / L) * ul[4] +
                   ((6.0 - 12.0 * xi) / (L * L)) * ul[8] + ((2.0 - 6.0 * xi) / L) * ul[10];
        S.pkz[w] = ((-6.0 + 12.0 * xi) / (L * L)) * ul[1] + ((-4.0 + 6.0 * xi) / L) * ul[5] +
                   ((6.0 - 12.0 * xi) / (L * L)) * ul[7] + ((-2.0 + 6.0 * xi) / L) * ul[11];
    }
    eval3(E, false);
    // B rows of point ip at column a
    auto b_eps = [&](int a) -> double { return a == 0 ? -1.0 / L : (a == 6 ? 1.0 / L : 0.0); };
    auto b_ky = [&](int a, double xi) -> double {
        return a == 2 ? (-6.0 + 12.0 * xi) / (L * L)
                      : (a == 4 ? (4.0 - 6.0 * xi) / L
                                : (a == 8 ? (6.0 - 12.0 * xi) / (L * L) : (a == 10 ? (2.0 - 6.0 * xi) / L : 0.0)));
    };
    auto b_kz = [&](int a, double xi) -> double {
        return a == 1 ? (-6.0 + 12.0 * xi) / (L * L)
                      : (a == 5 ? (-4.0 + 6.0 * xi) / L
                                : (a == 7 ? (6.0 - 12.0 * xi) / (L * L) : (a == 11 ? (-2.0 + 6.0 * xi) / L : 0.0)));
    };
    const bool torsion = el.G > 0.0 && el.J > 0.0;
    const double torsion_k = torsion ? el.G * el.J / L : 0.0;
    double axial_force_avg = 0.0;
    for (int ip = 0; ip < nip; ++ip) axial_force_avg += el.w[ip] * S.res[ip][0];
    for (int t = threadIdx.x; t < 144; t += blockDim.x) {  // :437-487
        const int a = t / 12, b = t % 12;
        double v = 0.0;
        for (int ip = 0; ip < nip; ++ip) {
            const double xi = el.xi[ip], wL = el.w[ip] * L;
            const double *r = S.res[ip];
            const double k11 = r[3], k12 = r[4], k13 = r[5], k22 = r[6], k23 = r[7], k33 = r[8];
            const double ba_n = b_eps(a), ba_my = b_ky(a, xi), ba_mz = b_kz(a, xi);
            const double bb_n = b_eps(b), bb_my = b_ky(b, xi), bb_mz = b_kz(b, xi);
            const double dN = k11 * bb_n + k12 * bb_my + k13 * bb_mz;
            const double dMy = k12 * bb_n + k22 * bb_my + k23 * bb_mz;
            const double dMz = k13 * bb_n + k23 * bb_my + k33 * bb_mz;
            v += wL * (ba_n * dN + ba_my * dMy + ba_mz * dMz);
        }
        if (torsion) {  // :489-496
            if ((a == 3 && b == 3) || (a == 9 && b == 9)) v += torsion_k;
            if ((a == 3 && b == 9) || (a == 9 && b == 3)) v -= torsion_k;
        }
        if (el.transf == 1) {  // :118-131
            const double N_over_L = axial_force_avg / L;
            if ((a == 1 && b == 1) || (a == 7 && b == 7) || (a == 2 && b == 2) || (a == 8 && b == 8)) v += N_over_L;
            if ((a == 1 && b == 7) || (a == 7 && b == 1) || (a == 2 && b == 8) || (a == 8 && b == 2)) v -= N_over_L;
        }
        S.kl[t] = v;
    }
    double fl[12];
    const double q5 = torsion ? torsion_k * (ul[9] - ul[3]) : 0.0;
#pragma unroll
    for (int a = 0; a < 12; ++a) {
        double v = 0.0;
        for (int ip = 0; ip < nip; ++ip) {
            const double xi = el.xi[ip], wL = el.w[ip] * L;
            const double *r = S.res[ip];
            v += wL * (b_eps(a) * r[0] + b_ky(a, xi) * r[1] + b_kz(a, xi) * r[2]);
        }
        v += (a == 3 ? -1.0 : (a == 9 ? 1.0 : 0.0)) * q5;  // :503-526
        fl[a] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 144; t += blockDim.x) {  // work = k_local T
        const int i = t / 12, j = t % 12;
        double v = 0.0;
        for (int k = 0; k < 12; ++k) {
            const double aik = S.kl[i * 12 + k];
            if (aik == 0.0) continue;
            v += aik * T(k, j);
        }
        S.work[t] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 144; t += blockDim.x) {  // k_global = T^T work
        const int i = t / 12, j = t % 12;
        double v = 0.0;
        for (int k = 0; k < 12; ++k) {
            const double aik = T(k, i);
            if (aik == 0.0) continue;
            v += aik * S.work[k * 12 + j];
        }
        k_out[t] = v;
    }
    if (threadIdx.x < 12) {  // :105-115
        const int i = threadIdx.x;
        double sum = 0.0;
        for (int j = 0; j < 12; ++j) sum += T(j, i) * fl[j];
        f_out[i] = sum;
    }
    return 0;
}

// Listed elements, one block each.
__global__ __launch_bounds__(kElem3dThreads) void elem3d_kernel(Elem3dArgs A) {
    __shared__ Elem3dLds S;
    const int p = blockIdx.x;
    if (p >= A.n) return;
    const int id = A.ids[p];
    if (id < 0 || id >= A.n_elems) {
        if (threadIdx.x == 0) A.status[p] = -1;
        return;
    }
    const Elem3dDev el = A.elems[id];
    const MatDev *mats = A.mats;
    if (A.n_mats <= kMats3dInLds) {
        const double *src = reinterpret_cast<const double *>(A.mats);
        double *dst = reinterpret_cast<double *>(S.smats);
        const int nd = A.n_mats * (int)(sizeof(MatDev) / sizeof(double));
        for (int k = threadIdx.x; k < nd; k += blockDim.x) dst[k] = src[k];
        __syncthreads();
        mats = S.smats;
    }
    Elem3dCtx E;
    E.A = &A;
    E.el = &el;
    E.S = &S;
    E.mats = mats;
    double u[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) u[k] = A.u[12 * (int64_t)p + k];
    double *k_out = A.k + 144 * (int64_t)p, *f_out = A.f + 12 * (int64_t)p;
    const int st = el.type == kForceBeamColumn3d ? fbc3d_element(E, u, A.q_state + el.q_off, k_out, f_out)
                                                 : dbc3d_element(E, u, k_out, f_out);
    if (threadIdx.x == 0) A.status[p] = st;
}

void launch_elem3d(hipStream_t stream, const Elem3dArgs &args) {
    if (args.n <= 0) return;
    hipLaunchKernelGGL(elem3d_kernel, dim3(args.n), dim3(kElem3dThreads), 0, stream, args);
}

}  // namespace sgpu
