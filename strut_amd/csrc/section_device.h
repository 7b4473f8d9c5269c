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

// Section resultants of a 2D fiber section evaluated by one whole wave
// (fiber2d.mojo:3982-4024): every lane returns the same five sums, formed
// in the reference's order (chunk trees, then chunks one after another).
struct Sec2dOut {
    double N, Mz, k11, k12, k22;
};

__device__ __forceinline__ Sec2dOut wave_section_2d(const SecDefDev &d, const SlotTables &tb,
                                                    const MatDev *__restrict__ mats, const StateArrays &c,
                                                    const StateArrays &t, int inst, double eps0, double kappa) {
    const int lane = threadIdx.x & 63;
    const int64_t inst_state = d.state_base + (int64_t)inst * d.n_slots;
    const int64_t inst_aux = d.aux_base + (int64_t)inst * d.n_aux;
    const int32_t *bcb = tb.batch_chunk_begin + d.batch_base;
    const int32_t *cslot = tb.chunk_slot + d.chunk_base;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0, acc4 = 0.0;
    for (int b = 0; b < d.n_batches; ++b) {
        const int s = b * kBatch + lane;
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
        int code = -1;
        if (s < d.n_slots) {
            const int64_t gs = d.slot_base + s;
            const int mi = tb.mat[gs];
            code = tb.chunk[gs];
            if (mi >= 0) {
                const double y = tb.y[gs];
                const double area = tb.area[gs];
                const double strain = eps0 - y * kappa;
                const int32_t aux = tb.aux[gs];
                double fs, ks;
                eval_slot(mats, mi, strain, area, c, t, inst_state + s, inst_aux + (aux < 0 ? 0 : aux), fs, ks);
                t0 = fs;
                t1 = -fs * y;
                t2 = ks;
                t3 = -ks * y;
                t4 = ks * y * y;
            }
        }
        // chunk partial in the chunk's first lane: W-wide tree or the lane's own term
        const double r0 = tree8(t0), r1 = tree8(t1), r2 = tree8(t2), r3 = tree8(t3), r4 = tree8(t4);
        const bool w8 = code >= 0 && (code & 1);
        const double p0 = w8 ? r0 : t0, p1 = w8 ? r1 : t1, p2 = w8 ? r2 : t2, p3 = w8 ? r3 : t3,
                     p4 = w8 ? r4 : t4;
        const int c0 = bcb[b], c1 = bcb[b + 1];
        for (int cc = c0; cc < c1; ++cc) {
            const int src = cslot[cc] & (kBatch - 1);
            if (d.zero_k12 && cc == d.n_elastic_chunks) acc3 = 0.0;
            acc0 += readlane_d(p0, src);
            acc1 += readlane_d(p1, src);
            acc2 += readlane_d(p2, src);
            acc3 += readlane_d(p3, src);
            acc4 += readlane_d(p4, src);
        }
    }
    if (d.zero_k12 && bcb[d.n_batches] <= d.n_elastic_chunks) acc3 = 0.0;
    Sec2dOut o;
    o.N = acc0;
    o.Mz = acc1;
    o.k11 = acc2;
    o.k12 = acc3;
    o.k22 = acc4;
    return o;
}

}  // namespace sgpu
