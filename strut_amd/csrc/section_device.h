// Per-fiber device routines shared by the fiber-section kernels: slot
// history load/store, the material-law dispatch, the W = 8 halving tree and
// a wave-wide 2D section evaluation used inside the element kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "material_laws.h"
#include "strut_gpu_types.h"

namespace sgpu {

constexpr int kBatch = 64;

__device__ __forceinline__ double tree8(double v) {
    double w = v + __shfl_xor(v, 4, 64);
    double x = w + __shfl_xor(w, 2, 64);
    return x + __shfl_xor(x, 1, 64);
}

// Value of lane l + N (N = 1, 2, 4) within the lane's 16-lane DPP row
// (row_shl); lanes whose source falls outside the row read 0.
template <int N>
__device__ __forceinline__ double dpp_shl(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x100 | N, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x100 | N, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// The W = 8 halving tree ((v0+v4)+(v2+v6))+((v1+v5)+(v3+v7)) of the eight
// lanes 8k..8k+7, left in lane 8k (other lanes hold partial garbage).  Same
// pairs as tree8, formed with DPP row shifts instead of LDS permutes.
__device__ __forceinline__ double tree8_dpp(double v) {
    const double w = v + dpp_shl<4>(v);
    const double x = w + dpp_shl<2>(w);
    return x + dpp_shl<1>(x);
}

__device__ __forceinline__ void load_hist(const StateArrays &s, int64_t gi, int64_t ai, int type,
                                          FiberHist &h) {
    h.a0 = h.a1 = h.a2 = h.a3 = h.a4 = 0.0;
    h.kon = 0;
    h.eps = s.eps[gi];
    h.sig = s.sig[gi];
    h.h0 = s.h0[gi];
    h.h1 = s.h1[gi];
    if (type == kSteel02) {
        h.tan = 0.0;  // Steel02 never reads its committed tangent
        h.a0 = s.a0[ai];
        h.a1 = s.a1[ai];
        h.a2 = s.a2[ai];
        h.a3 = s.a3[ai];
        h.a4 = s.a4[ai];
        h.kon = s.kon[ai];
    } else {
        h.tan = s.tan[gi];
        if (type == kConcrete01) h.a0 = s.a0[ai];
    }
}

__device__ __forceinline__ void store_hist(const StateArrays &s, int64_t gi, int64_t ai, int type,
                                           const FiberHist &h) {
    s.eps[gi] = h.eps;
    s.sig[gi] = h.sig;
    s.tan[gi] = h.tan;
    s.h0[gi] = h.h0;
    s.h1[gi] = h.h1;
    if (type == kSteel02) {
        s.a0[ai] = h.a0;
        s.a1[ai] = h.a1;
        s.a2[ai] = h.a2;
        s.a3[ai] = h.a3;
        s.a4[ai] = h.a4;
        s.kon[ai] = h.kon;
    } else if (type == kConcrete01) {
        s.a0[ai] = h.a0;
    }
}

// Evaluate one slot: returns (fs, ks) = (sigma * A, E_t * A) for material
// fibers and (E A strain, E A) for elastic fibers (fiber2d.mojo:3480-3484).
__device__ __forceinline__ void eval_slot(const MatDev *__restrict__ mats, int mi, double strain,
                                          double area, const StateArrays &c, const StateArrays &t,
                                          int64_t gi, int64_t ai, double &fs, double &ks) {
    const MatDev &m = mats[mi];
    const int type = m.type;
    if (type == kElastic) {
        ks = m.p[0] * area;
        fs = ks * strain;
        return;
    }
    FiberHist hc, ht;
    load_hist(c, gi, ai, type, hc);
    ht = hc;
    switch (type) {
    case kConcrete02: concrete02_trial(m, hc, strain, ht); break;
    case kSteel02: steel02_trial(m, hc, strain, ht); break;
    case kSteel01: steel01_trial(m, hc, strain, ht); break;
    default: concrete01_trial(m, hc, strain, ht); break;
    }
    store_hist(t, gi, ai, type, ht);
    fs = ht.sig * area;
    ks = ht.tan * area;
}

// Same for the laws without aux state (Elastic, Concrete02, Steel01); the
// aux-state laws (Steel02, Concrete01) are evaluated by the aux pre-pass of
// the section-point kernel, which keeps the heavy Steel02 code (and its pow)
// out of the register budget of the main pass.
__device__ __forceinline__ void eval_slot_common(const MatDev *__restrict__ mats, int mi, double strain,
                                                 double area, const StateArrays &c, const StateArrays &t,
                                                 int64_t gi, double &fs, double &ks) {
    const MatDev &m = mats[mi];
    const int type = m.type;
    if (type == kElastic) {
        ks = m.p[0] * area;
        fs = ks * strain;
        return;
    }
    FiberHist hc, ht;
    load_hist(c, gi, 0, kConcrete02, hc);
    ht = hc;
    if (type == kConcrete02)
        concrete02_trial(m, hc, strain, ht);
    else
        steel01_trial(m, hc, strain, ht);
    store_hist(t, gi, 0, kConcrete02, ht);
    fs = ht.sig * area;
    ks = ht.tan * area;
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

constexpr int kWaveRedStride = kBatch + 1;  // LDS row stride of wave_section_2d (odd: rows in different banks)

// Orders the LDS traffic of one wavefront: writes before the barrier are
// seen by reads after it (LDS requests of one wave complete in order; this
// only keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// What one lane reads for its slot of a batch before the law runs: the
// table entry and the committed common history (read for every slot: the
// address does not depend on the table, so the loads of a batch issue
// together).
struct WaveSlotIn {
    int mi, code, aux;
    double y, area;
    double eps, sig, tan, h0, h1;
};

__device__ __forceinline__ void wave_slot_load(const SecDefDev &d, const SlotTables &tb, const StateArrays &c,
                                               int64_t inst_state, int s, WaveSlotIn &in) {
    in.mi = -1;
    in.code = -1;
    in.aux = -1;
    in.y = in.area = 0.0;
    in.eps = in.sig = in.tan = in.h0 = in.h1 = 0.0;
    if (s >= d.n_slots) return;
    const int64_t gs = d.slot_base + s;
    const int64_t gi = inst_state + s;
    in.mi = tb.mat[gs];
    in.code = tb.chunk[gs];
    in.aux = tb.aux[gs];
    in.y = tb.y[gs];
    in.area = tb.area[gs];
    in.eps = c.eps[gi];
    in.sig = c.sig[gi];
    in.tan = c.tan[gi];
    in.h0 = c.h0[gi];
    in.h1 = c.h1[gi];
}

// Section resultants of a 2D fiber section evaluated by one whole wave
// (fiber2d.mojo:3982-4024).  Each batch leaves its chunk partials in the
// wave's LDS rows `red` (5 x kWaveRedStride doubles); lanes 0..4 add them
// one after another onto their running sums, so the sums see the chunks in
// the reference's order; every lane returns the same five sums.
//
// kTrial = false evaluates without storing the trial fiber states: a trial
// state is a function of the committed state and the strain only, so a
// caller that records the section's last deformation can rebuild it at
// commit (commit_section_2d) and keeps the Newton loop free of stores.
// The definition needs at most 63 batches (chunk offsets live in one lane
// register each).
struct Sec2dOut {
    double N, Mz, k11, k12, k22;
};

template <bool kTrial>
__device__ __forceinline__ Sec2dOut wave_section_2d(const SecDefDev &d, const SlotTables &tb,
                                                    const MatDev *__restrict__ mats, const StateArrays &c,
                                                    const StateArrays &t, int inst, double eps0, double kappa,
                                                    double *red) {
    const int lane = threadIdx.x & 63;
    const int n_batches = d.n_batches, n_el = d.n_elastic_chunks;
    const int64_t inst_state = d.state_base + (int64_t)inst * d.n_slots;
    const int64_t inst_aux = d.aux_base + (int64_t)inst * d.n_aux;
    // chunk offset of batch `lane`, read back with readlane
    const int bcb_lane = lane <= n_batches ? tb.batch_chunk_begin[d.batch_base + lane] : 0;
    const bool zero_lane = d.zero_k12 && lane == 3;
    double acc = 0.0;  // lanes 0..4: running sum of N, Mz, k11, k12, k22
    WaveSlotIn cur;
    wave_slot_load(d, tb, c, inst_state, lane, cur);
    for (int b = 0; b < n_batches; ++b) {
        // aux history of this batch's Steel02 / Concrete01 lanes, issued before the prefetch
        const int mt = cur.mi >= 0 ? mats[cur.mi].type : kDead;
        const int64_t ai = inst_aux + (cur.aux < 0 ? 0 : cur.aux);
        FiberHist hc;
        hc.eps = cur.eps;
        hc.sig = cur.sig;
        hc.tan = cur.tan;
        hc.h0 = cur.h0;
        hc.h1 = cur.h1;
        hc.a0 = hc.a1 = hc.a2 = hc.a3 = hc.a4 = 0.0;
        hc.kon = 0;
        if (mt == kSteel02) {
            hc.tan = 0.0;  // Steel02 never reads its committed tangent
            hc.a0 = c.a0[ai];
            hc.a1 = c.a1[ai];
            hc.a2 = c.a2[ai];
            hc.a3 = c.a3[ai];
            hc.a4 = c.a4[ai];
            hc.kon = c.kon[ai];
        } else if (mt == kConcrete01) {
            hc.a0 = c.a0[ai];
        }
        WaveSlotIn nxt;
        if (b + 1 < n_batches) wave_slot_load(d, tb, c, inst_state, (b + 1) * kBatch + lane, nxt);
        double fs = 0.0, ks = 0.0;
        if (cur.mi >= 0) {
            const MatDev &m = mats[cur.mi];
            const double strain = eps0 - cur.y * kappa;
            if (mt == kElastic) {
                ks = m.p[0] * cur.area;
                fs = ks * strain;
            } else {
                FiberHist ht = hc;
                switch (mt) {
                case kConcrete02: concrete02_trial(m, hc, strain, ht); break;
                case kSteel02: steel02_trial(m, hc, strain, ht); break;
                case kSteel01: steel01_trial(m, hc, strain, ht); break;
                default: concrete01_trial(m, hc, strain, ht); break;
                }
                if (kTrial) store_hist(t, inst_state + b * kBatch + lane, ai, mt, ht);
                fs = ht.sig * cur.area;
                ks = ht.tan * cur.area;
            }
        }
        const double y = cur.y;
        const double t0 = fs, t1 = -fs * y, t2 = ks, t3 = -ks * y, t4 = ks * y * y;
        // chunk partial in the chunk's first lane: W-wide tree or the lane's own term
        const bool w8 = cur.code >= 0 && (cur.code & 1);
        double p0 = t0, p1 = t1, p2 = t2, p3 = t3, p4 = t4;
        if (__builtin_amdgcn_ballot_w64(w8) != 0) {
            const double r0 = tree8_dpp(t0), r1 = tree8_dpp(t1), r2 = tree8_dpp(t2), r3 = tree8_dpp(t3),
                         r4 = tree8_dpp(t4);
            if (w8) {
                p0 = r0;
                p1 = r1;
                p2 = r2;
                p3 = r3;
                p4 = r4;
            }
        }
        if (cur.code >= 0) {
            const int dst = cur.code >> 1;
            red[dst] = p0;
            red[kWaveRedStride + dst] = p1;
            red[2 * kWaveRedStride + dst] = p2;
            red[3 * kWaveRedStride + dst] = p3;
            red[4 * kWaveRedStride + dst] = p4;
        }
        wave_sync();
        const int c0 = __builtin_amdgcn_readlane(bcb_lane, b), c1 = __builtin_amdgcn_readlane(bcb_lane, b + 1);
        if (lane < 5) {
            const double *row = red + lane * kWaveRedStride - c0;
            int cc = c0;
            if (zero_lane && n_el >= c0 && n_el < c1) {
                for (; cc < n_el; ++cc) acc += row[cc];
                acc = 0.0;  // uniform elastic modulus: k12 restarts with the nonlinear families
            }
#pragma unroll 8
            for (; cc < c1; ++cc) acc += row[cc];
        }
        wave_sync();
        cur = nxt;
    }
    if (zero_lane && __builtin_amdgcn_readlane(bcb_lane, n_batches) <= n_el) acc = 0.0;
    Sec2dOut o;
    o.N = readlane_d(acc, 0);
    o.Mz = readlane_d(acc, 1);
    o.k11 = readlane_d(acc, 2);
    o.k12 = readlane_d(acc, 3);
    o.k22 = readlane_d(acc, 4);
    return o;
}

// Commit one 2D section instance by one wave: every fiber's law evaluated
// from its committed state at the section's last deformation (eps0, kappa),
// written back as the committed state.  Identical to copying the trial state
// of that last evaluation (fiber2d.mojo:4059-4092).
__device__ __forceinline__ void commit_section_2d(const SecDefDev &d, const SlotTables &tb,
                                                  const MatDev *__restrict__ mats, const StateArrays &c, int inst,
                                                  double eps0, double kappa) {
    const int lane = threadIdx.x & 63;
    const int64_t inst_state = d.state_base + (int64_t)inst * d.n_slots;
    const int64_t inst_aux = d.aux_base + (int64_t)inst * d.n_aux;
    for (int s = lane; s < d.n_slots; s += kBatch) {
        const int64_t gs = d.slot_base + s;
        const int mi = tb.mat[gs];
        if (mi < 0) continue;
        const MatDev &m = mats[mi];
        if (m.type == kElastic) continue;
        const int32_t aux = tb.aux[gs];
        const int64_t gi = inst_state + s, ai = inst_aux + (aux < 0 ? 0 : aux);
        FiberHist hc, ht;
        load_hist(c, gi, ai, m.type, hc);
        ht = hc;
        const double strain = eps0 - tb.y[gs] * kappa;
        switch (m.type) {
        case kConcrete02: concrete02_trial(m, hc, strain, ht); break;
        case kSteel02: steel02_trial(m, hc, strain, ht); break;
        case kSteel01: steel01_trial(m, hc, strain, ht); break;
        default: concrete01_trial(m, hc, strain, ht); break;
        }
        store_hist(c, gi, ai, m.type, ht);
    }
}

}  // namespace sgpu
