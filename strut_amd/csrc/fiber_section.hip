// Fiber-section state determination kernels for gfx950.
//
// One 64-lane wavefront evaluates one section point (definition, instance,
// deformation).  Lanes walk the section's slots 64 at a time: each lane
// loads its fiber's committed history (structure-of-arrays, so a wave reads
// 512 contiguous bytes per field), applies the material law, stores the
// trial history and forms the section terms.  The terms are then reduced in
// exactly the order Strut's CPU path uses (fiber2d.mojo:2065-2082,
// :1584-1601, :3459-3501, :3994-4022; fiber3d.mojo:189-256, :804-822):
// W = 8 lanes are folded by the halving tree (xor 4, 2, 1), and the chunk
// partials are added one after another in chunk order by one lane per
// resultant.  Floating-point addition is commutative, so the xor butterfly
// gives the same bits as the reference's reduce_add, and the section
// resultants match the CPU restatement bit for bit whenever the material
// laws do.
#include <hip/hip_runtime.h>

#include "../../include/strut_gpu.h"
#include "material_laws.h"
#include "section_device.h"
#include "strut_gpu_types.h"

namespace sgpu {

constexpr int kPointsPerBlock = 4;         // one wavefront per section point
constexpr int kRedStride = kBatch + 1;     // LDS row stride (odd, so the NQ rows sit in different banks)
constexpr int kLdsMaxMats = 128;           // material tables up to this size are staged in LDS
constexpr int kLdsMaxAux = 128;            // aux results of a point staged in LDS up to this many slots

// Section term q of one slot from its fiber force fs = sigma A and stiffness
// ks = E_t A (fiber2d.mojo:2065-2082, fiber3d.mojo:804-822).
template <int DIM>
__device__ __forceinline__ double slot_term(int q, double fs, double ks, double y, double z) {
    if (DIM == 2) {
        switch (q) {
        case 0: return fs;
        case 1: return -fs * y;
        case 2: return ks;
        case 3: return -ks * y;
        default: return ks * y * y;
        }
    }
    switch (q) {
    case 0: return fs;
    case 1: return fs * z;
    case 2: return -fs * y;
    case 3: return ks;
    case 4: return ks * z;
    case 5: return -ks * y;
    case 6: return ks * z * z;
    case 7: return -ks * z * y;
    default: return ks * y * y;
    }
}

// Everything one lane reads for its slot of a batch before the law runs:
// the slot's table entry and the committed history of the common fields.
struct SlotIn {
    int mi, code, aux;
    double y, z, area;
    double eps, sig, tan, h0, h1;
};

// gs = table index of the slot, gi = its state index.  Every field is loaded
// (z only in 3D): a lane past the section's slots reads slot 0 of its
// instance and is masked by slot_in_ready, so no register of a set is left
// with a load in flight when the set is loaded again.
template <int DIM>
__device__ __forceinline__ void load_slot_in(const SlotTables &tb, const StateArrays &c, int64_t gs, int64_t gi,
                                             SlotIn &in) {
    in.mi = tb.mat[gs];
    in.code = tb.chunk[gs];
    in.aux = tb.aux[gs];
    in.y = tb.y[gs];
    in.area = tb.area[gs];
    in.z = DIM == 3 ? tb.z[gs] : 0.0;
    in.eps = c.eps[gi];
    in.sig = c.sig[gi];
    in.tan = c.tan[gi];
    in.h0 = c.h0[gi];
    in.h1 = c.h1[gi];
}

// Wait for every load of a set before it is used (an empty asm that reads
// the registers), then blank the lanes past the section's slots.
__device__ __forceinline__ void slot_in_ready(SlotIn &in, bool valid) {
    asm volatile("" ::"v"(in.mi), "v"(in.code), "v"(in.aux), "v"(in.y), "v"(in.z), "v"(in.area), "v"(in.eps),
                 "v"(in.sig), "v"(in.tan), "v"(in.h0), "v"(in.h1)
                 : "memory");
    if (!valid) {
        in.mi = -1;
        in.code = -1;
        in.aux = -1;
        in.y = in.z = in.area = 0.0;
    }
}

// (fs, ks) = (sigma A, E_t A) of one slot of a batch, from its loaded
// committed history; the trial history of a material slot is stored at gi.
// Aux slots (Steel02, Concrete01) take the aux pre-pass result.
template <int DIM, typename MatPtr, typename AuxPtr>
__device__ __forceinline__ void slot_eval(const SlotIn &cur, MatPtr mats, const StateArrays &t, int64_t gi,
                                          double eps0, double kz, double ky, AuxPtr fsks, double &fs, double &ks) {
    fs = 0.0;
    ks = 0.0;
    if (cur.mi < 0) return;
    if (cur.aux >= 0) {
        fs = fsks[2 * cur.aux];
        ks = fsks[2 * cur.aux + 1];
        return;
    }
    const MatDev &m = mats[cur.mi];
    const double strain = DIM == 2 ? eps0 - cur.y * kz : eps0 + cur.z * ky - cur.y * kz;
    if (m.type == kElastic) {
        ks = m.p[0] * cur.area;
        fs = ks * strain;
        return;
    }
    FiberHist hc, ht;
    hc.eps = cur.eps;
    hc.sig = cur.sig;
    hc.tan = cur.tan;
    hc.h0 = cur.h0;
    hc.h1 = cur.h1;
    hc.a0 = hc.a1 = hc.a2 = hc.a3 = hc.a4 = 0.0;
    hc.kon = 0;
    ht = hc;
    if (m.type == kConcrete02)
        concrete02_trial(m, hc, strain, ht);
    else
        steel01_trial(m, hc, strain, ht);
    t.eps[gi] = ht.eps;
    t.sig[gi] = ht.sig;
    t.tan[gi] = ht.tan;
    t.h0[gi] = ht.h0;
    t.h1[gi] = ht.h1;
    fs = ht.sig * cur.area;
    ks = ht.tan * cur.area;
}

// The chunk partials of one batch: W-chunks folded by the halving tree,
// single-slot chunks as they are.  Chunk k of the batch (its code >> 1)
// lands in red[q * stride + cbase + k]; the whole wave must be active.
template <int DIM>
__device__ __forceinline__ void batch_partials(const SlotIn &cur, double fs, double ks, double *red, int stride,
                                               int cbase) {
    constexpr int NQ = DIM == 2 ? 5 : 9;
    const int dst = cur.code >= 0 ? cbase + (cur.code >> 1) : -1;
    const bool w8 = cur.code >= 0 && (cur.code & 1);
    if (__builtin_amdgcn_ballot_w64(w8) != 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double term = slot_term<DIM>(q, fs, ks, cur.y, cur.z);
            const double tr = tree8_dpp(term);
            if (dst >= 0) red[q * stride + dst] = w8 ? tr : term;
        }
    } else {  // every chunk of the batch is a single slot
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            if (dst >= 0) red[q * stride + dst] = slot_term<DIM>(q, fs, ks, cur.y, cur.z);
    }
}

// Section resultants, tangent and flexibility of one point from its reduced
// sums fin[0..NQ) (fiber2d.mojo:3074-3085; force_beam_column3d.mojo:679-727
// applied as :1093-1117).
template <int DIM>
__device__ __forceinline__ void finish_point(const double *fin, sgpu_section_result *out) {
    sgpu_section_result r;
    for (int q = 0; q < 3; ++q) r.s[q] = 0.0;
    for (int q = 0; q < 6; ++q) {
        r.k[q] = 0.0;
        r.f[q] = 0.0;
    }
    r.pad_ = 0;
    if (DIM == 2) {
        // fiber2d.mojo:3074-3085
        const double k11 = fin[2], k12 = fin[3], k22 = fin[4];
        r.s[0] = fin[0];
        r.s[1] = fin[1];
        r.k[0] = k11;
        r.k[1] = k12;
        r.k[2] = k22;
        const double det = k11 * k22 - k12 * k12;
        if (fabs(det) <= 1.0e-40) {
            r.valid = 0;
        } else {
            const double inv_det = 1.0 / det;
            r.f[0] = k22 * inv_det;
            r.f[1] = -k12 * inv_det;
            r.f[2] = k11 * inv_det;
            r.valid = 1;
        }
    } else {
        // force_beam_column3d.mojo:679-727 applied as :1093-1117
        r.s[0] = fin[0];
        r.s[1] = fin[1];
        r.s[2] = fin[2];
        const double a = fin[3], bb = fin[4], cq = fin[5], e = fin[6], f = fin[7], i = fin[8];
        const double dd = bb, g = cq, h = f;
        r.k[0] = a;
        r.k[1] = bb;
        r.k[2] = cq;
        r.k[3] = e;
        r.k[4] = f;
        r.k[5] = i;
        const double co00 = e * i - f * h;
        const double co01 = -(dd * i - f * g);
        const double co02 = dd * h - e * g;
        const double co10 = -(bb * i - cq * h);
        const double co11 = a * i - cq * g;
        const double co20 = bb * f - cq * e;
        const double co21 = -(a * f - cq * dd);
        const double co22 = a * e - bb * dd;
        const double det = a * co00 + bb * co01 + cq * co02;
        if (fabs(det) <= 1.0e-40) {
            r.valid = 0;
        } else {
            const double inv_det = 1.0 / det;
            r.f[0] = co00 * inv_det;
            r.f[1] = co10 * inv_det;
            r.f[2] = co20 * inv_det;
            r.f[3] = co11 * inv_det;
            r.f[4] = co21 * inv_det;
            r.f[5] = co22 * inv_det;
            r.valid = 1;
        }
    }
    *out = r;
}

// Batch b of one section point: evaluate the lanes' slots, fold the chunk
// partials into the wave's LDS rows, and let lanes 0..NQ-1 add them one after
// another onto their running sums acc (chunks [cbase, c1) of the section).
template <int DIM, typename MatPtr, typename AuxPtr>
__device__ __forceinline__ void section_batch(SlotIn cur, int b, int cbase, int c1, int n_slots, int ne,
                                              MatPtr mats, const StateArrays &t, int64_t inst_state, double eps0,
                                              double kz, double ky, AuxPtr fsks, double *red, bool zero_lane,
                                              double &acc) {
    constexpr int NQ = DIM == 2 ? 5 : 9;
    const int lane = threadIdx.x & (kBatch - 1);
    slot_in_ready(cur, b * kBatch + lane < n_slots);
    double fs, ks;
    slot_eval<DIM>(cur, mats, t, inst_state + b * kBatch + lane, eps0, kz, ky, fsks, fs, ks);
    batch_partials<DIM>(cur, fs, ks, red, kRedStride, 0);
    wave_sync();
    if (lane < NQ) {
        const double *row = red + lane * kRedStride - cbase;
        int cc = cbase;
        if (zero_lane && ne >= cbase && ne < c1) {
            for (; cc < ne; ++cc) acc += row[cc];
            acc = 0.0;  // uniform elastic modulus: k12 (k13) restarts with the nonlinear families
        }
#pragma unroll 8
        for (; cc < c1; ++cc) acc += row[cc];
    }
    wave_sync();
}

// One section point, evaluated by one wavefront.  Lanes walk the slots 64 at
// a time; each batch leaves its chunk partials in the wave's LDS rows, and
// lanes 0..NQ-1 add them one after another onto their running sums, so the
// sums see the chunks in exactly the sequential order of the reference.
// Two register sets A and B alternate: while batch b is evaluated the loads
// of batch b + 1 are in flight, and batch b + 2's loads issue as soon as b
// is done, with no copy between the sets (a copy would wait for the loads).
// Batches run in pages of 62 whose chunk ranges sit in one register.
// Every other read inside the loop comes from registers or LDS (material
// table, aux results, chunk ranges): vector-memory reads return in issue
// order, so waiting for one there would wait for the batches in flight too.
template <int DIM, typename MatPtr, typename AuxPtr>
__device__ __forceinline__ void section_point(const SecDefDev &d, const SlotTables &tb, MatPtr mats,
                                              const StateArrays &t, int inst, double eps0, double kz, double ky,
                                              AuxPtr fsks, double *red, const StateArrays &c,
                                              sgpu_section_result *out) {
    constexpr int NQ = DIM == 2 ? 5 : 9;
    const int lane = threadIdx.x & (kBatch - 1);
    const int n_slots = d.n_slots, n_batches = d.n_batches;
    const int64_t slot_base = d.slot_base;
    const int64_t inst_state = d.state_base + (int64_t)inst * n_slots;
    const int32_t *bcb = tb.batch_chunk_begin + d.batch_base;
    const int ne = d.n_elastic_chunks;
    const bool zero_lane = d.zero_k12 && (DIM == 2 ? lane == 3 : (lane == 4 || lane == 5));
    // slot index of a lane in batch b, or 0 past the section's slots
    auto slot_of = [&](int b) { const int sl = b * kBatch + lane; return sl < n_slots ? sl : 0; };
    double acc = 0.0;
    int n_chunks = 0;
    // pages of up to kPage batches; a page's chunk range starts are held one
    // per lane (bcb[pb + l] in lane l), read once before its batch loads issue
    constexpr int kPage = kBatch - 2;
    for (int pb = 0; pb < n_batches; pb += kPage) {
        const int pn = n_batches - pb < kPage ? n_batches - pb : kPage;
        const int lane_bcb = lane <= pn ? bcb[pb + lane] : 0;
        asm volatile("" ::"v"(lane_bcb) : "memory");
        n_chunks = __builtin_amdgcn_readlane(lane_bcb, pn);
        // The set loads are unconditional (past the section a lane reads slot
        // 0), so that every path issues the same loads and the wait for a set
        // leaves the other set's loads in flight.
        SlotIn A, B;
        load_slot_in<DIM>(tb, c, slot_base + slot_of(pb), inst_state + slot_of(pb), A);
        load_slot_in<DIM>(tb, c, slot_base + slot_of(pb + 1), inst_state + slot_of(pb + 1), B);
        for (int k = 0; k < pn; k += 2) {
            section_batch<DIM>(A, pb + k, __builtin_amdgcn_readlane(lane_bcb, k),
                               __builtin_amdgcn_readlane(lane_bcb, k + 1), n_slots, ne, mats, t, inst_state, eps0, kz,
                               ky, fsks, red, zero_lane, acc);
            load_slot_in<DIM>(tb, c, slot_base + slot_of(pb + k + 2), inst_state + slot_of(pb + k + 2), A);
            if (k + 1 < pn)
                section_batch<DIM>(B, pb + k + 1, __builtin_amdgcn_readlane(lane_bcb, k + 1),
                                   __builtin_amdgcn_readlane(lane_bcb, k + 2), n_slots, ne, mats, t, inst_state, eps0,
                                   kz, ky, fsks, red, zero_lane, acc);
            load_slot_in<DIM>(tb, c, slot_base + slot_of(pb + k + 3), inst_state + slot_of(pb + k + 3), B);
        }
    }
    if (zero_lane && n_chunks <= ne) acc = 0.0;
    double fin[9];
#pragma unroll
    for (int q = 0; q < NQ; ++q) fin[q] = readlane_d(acc, q);
    if (lane != 0) return;
    finish_point<DIM>(fin, out);
}

__device__ __forceinline__ bool point_ok(const SecDefDev *__restrict__ defs, int n_defs,
                                         const sgpu_section_point &pt) {
    return pt.def_id >= 0 && pt.def_id < n_defs && pt.instance >= 0 && pt.instance < defs[pt.def_id].n_inst;
}

// Aux pre-pass: the Steel02 and Concrete01 slots of every point (few per
// section) evaluated from their committed history; the trial history is
// stored and (sigma A, E_t A) left in fsks[point][aux][2] for the main pass.
// Lanes are packed densely: lane l of a block serves point
// blockIdx.x * (blockDim.x / per) + l / per and aux index l % per (plus
// multiples of per below the definition's n_aux), per = min(max_aux, 256).
__global__ __launch_bounds__(256) void aux_eval_kernel(const SecDefDev *__restrict__ defs, int n_defs,
                                                       SlotTables tb, const MatDev *__restrict__ mats,
                                                       StateArrays c, StateArrays t,
                                                       const sgpu_section_point *__restrict__ pts, int n,
                                                       double *__restrict__ fsks, int max_aux, int per) {
    const int ppb = blockDim.x / per;
    const int k = threadIdx.x / per;
    if (k >= ppb) return;
    const int p = blockIdx.x * ppb + k;
    if (p >= n) return;
    const sgpu_section_point pt = pts[p];
    if (!point_ok(defs, n_defs, pt)) return;
    const SecDefDev &d = defs[pt.def_id];
    const int64_t inst_state = d.state_base + (int64_t)pt.instance * d.n_slots;
    const int64_t inst_aux = d.aux_base + (int64_t)pt.instance * d.n_aux;
    for (int j = threadIdx.x - k * per; j < d.n_aux; j += per) {
        const int64_t ja = d.aux_slot_base + j;
        const int s = tb.aux_slot[ja];
        const int mi = tb.aux_mat[ja];
        const double y = tb.aux_y[ja];
        const double area = tb.aux_area[ja];
        const double strain = d.dim == 2 ? pt.eps0 - y * pt.kappa_z : pt.eps0 + tb.aux_z[ja] * pt.kappa_y - y * pt.kappa_z;
        double fs, ks;
        eval_slot(mats, mi, strain, area, c, t, inst_state + s, inst_aux + j, fs, ks);
        double *o = fsks + ((int64_t)p * max_aux + j) * 2;
        o[0] = fs;
        o[1] = ks;
    }
}

// Main pass: every point's slots without aux state, and the section
// reduction.  Wave w of a block evaluates point blockIdx.x * kPointsPerBlock + w.
// LDS variant (kLds): the material table and each point's aux results are
// staged in LDS first (dynamic LDS: n_mats MatDev, then max_aux (fs, ks)
// pairs per wave).
template <bool kLds>
__global__ __launch_bounds__(kPointsPerBlock * kBatch) void section_eval_kernel(
    const SecDefDev *__restrict__ defs, int n_defs, SlotTables tb, const MatDev *__restrict__ mats, int n_mats,
    StateArrays c, StateArrays t, const sgpu_section_point *__restrict__ pts, sgpu_section_result *__restrict__ res,
    int n, const double *__restrict__ fsks, int max_aux) {
    __shared__ double red[kPointsPerBlock][9 * kRedStride];
    extern __shared__ double lds_dyn[];
    MatDev *smat = reinterpret_cast<MatDev *>(lds_dyn);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & (kBatch - 1);
    if (kLds) {
        constexpr int per = sizeof(MatDev) / sizeof(double);
        const double *src = reinterpret_cast<const double *>(mats);
        for (int i = threadIdx.x; i < n_mats * per; i += blockDim.x) lds_dyn[i] = src[i];
        __syncthreads();
    }
    // wave-uniform point index: point and definition loads go through the scalar unit
    const int p = blockIdx.x * kPointsPerBlock + wave;
    if (p >= n) return;
    const sgpu_section_point pt = pts[p];
    if (!point_ok(defs, n_defs, pt)) {
        if (lane == 0) {
            sgpu_section_result r = {};
            r.valid = -1;
            res[p] = r;
        }
        return;
    }
    const SecDefDev &d = defs[pt.def_id];
    const double *pf = fsks + (int64_t)p * max_aux * 2;
    if (kLds) {
        double *saux = lds_dyn + (size_t)n_mats * (sizeof(MatDev) / sizeof(double)) + (size_t)wave * max_aux * 2;
        for (int j = lane; j < 2 * d.n_aux; j += kBatch) saux[j] = pf[j];
        wave_sync();
        const MatDev *ms = smat;
        const double *as = saux;
        if (d.dim == 2)
            section_point<2>(d, tb, ms, t, pt.instance, pt.eps0, pt.kappa_z, 0.0, as, red[wave], c, res + p);
        else
            section_point<3>(d, tb, ms, t, pt.instance, pt.eps0, pt.kappa_z, pt.kappa_y, as, red[wave], c, res + p);
    } else {
        if (d.dim == 2)
            section_point<2>(d, tb, mats, t, pt.instance, pt.eps0, pt.kappa_z, 0.0, pf, red[wave], c, res + p);
        else
            section_point<3>(d, tb, mats, t, pt.instance, pt.eps0, pt.kappa_z, pt.kappa_y, pf, red[wave], c, res + p);
    }
}

// Plain copies between the committed and trial arrays (commit: src = trial).
__global__ void copy_state_kernel(StateArrays dst, StateArrays src, int64_t n_slots, int64_t n_aux) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_slots; i += stride) {
        dst.eps[i] = src.eps[i];
        dst.sig[i] = src.sig[i];
        dst.tan[i] = src.tan[i];
        dst.h0[i] = src.h0[i];
        dst.h1[i] = src.h1[i];
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_aux; i += stride) {
        dst.a0[i] = src.a0[i];
        dst.a1[i] = src.a1[i];
        dst.a2[i] = src.a2[i];
        dst.a3[i] = src.a3[i];
        dst.a4[i] = src.a4[i];
        dst.kon[i] = src.kon[i];
    }
}

// Copy the slots of listed (definition, instance) pairs.
__global__ __launch_bounds__(256) void copy_instances_kernel(const SecDefDev *__restrict__ defs,
                                                             const int32_t *__restrict__ def_ids,
                                                             const int32_t *__restrict__ insts, int n,
                                                             StateArrays dst, StateArrays src) {
    const int e = blockIdx.x;
    if (e >= n) return;
    const SecDefDev d = defs[def_ids[e]];
    const int64_t sb = d.state_base + (int64_t)insts[e] * d.n_slots;
    const int64_t ab = d.aux_base + (int64_t)insts[e] * d.n_aux;
    for (int s = threadIdx.x; s < d.n_slots; s += blockDim.x) {
        dst.eps[sb + s] = src.eps[sb + s];
        dst.sig[sb + s] = src.sig[sb + s];
        dst.tan[sb + s] = src.tan[sb + s];
        dst.h0[sb + s] = src.h0[sb + s];
        dst.h1[sb + s] = src.h1[sb + s];
    }
    for (int s = threadIdx.x; s < d.n_aux; s += blockDim.x) {
        dst.a0[ab + s] = src.a0[ab + s];
        dst.a1[ab + s] = src.a1[ab + s];
        dst.a2[ab + s] = src.a2[ab + s];
        dst.a3[ab + s] = src.a3[ab + s];
        dst.a4[ab + s] = src.a4[ab + s];
        dst.kon[ab + s] = src.kon[ab + s];
    }
}

void launch_aux_eval(hipStream_t stream, const SecDefDev *defs, int n_defs, const SlotTables &tb,
                     const MatDev *mats, const StateArrays &c, const StateArrays &t,
                     const sgpu_section_point *pts, int n, double *fsks, int max_aux) {
    if (n <= 0 || max_aux <= 0) return;
    const int per = max_aux < 256 ? max_aux : 256;
    const int ppb = 256 / per;
    hipLaunchKernelGGL(aux_eval_kernel, dim3((n + ppb - 1) / ppb), dim3(256), 0, stream, defs, n_defs, tb, mats, c,
                       t, pts, n, fsks, max_aux, per);
}

void launch_section_eval(hipStream_t stream, const SecDefDev *defs, int n_defs, const SlotTables &tb,
                         const MatDev *mats, int n_mats, const StateArrays &c, const StateArrays &t,
                         const sgpu_section_point *pts, sgpu_section_result *res, int n, const double *fsks,
                         int max_aux) {
    if (n <= 0) return;
    const int blocks = (n + kPointsPerBlock - 1) / kPointsPerBlock;
    if (n_mats <= kLdsMaxMats && max_aux <= kLdsMaxAux) {
        const size_t lds = sizeof(MatDev) * (size_t)n_mats + sizeof(double) * 2 * (size_t)max_aux * kPointsPerBlock;
        hipLaunchKernelGGL(section_eval_kernel<true>, dim3(blocks), dim3(kPointsPerBlock * kBatch), lds, stream,
                           defs, n_defs, tb, mats, n_mats, c, t, pts, res, n, fsks, max_aux);
    } else {
        hipLaunchKernelGGL(section_eval_kernel<false>, dim3(blocks), dim3(kPointsPerBlock * kBatch), 0, stream,
                           defs, n_defs, tb, mats, n_mats, c, t, pts, res, n, fsks, max_aux);
    }
}

void launch_copy_state(hipStream_t stream, const StateArrays &dst, const StateArrays &src, int64_t n_slots,
                       int64_t n_aux) {
    const int64_t n = n_slots > n_aux ? n_slots : n_aux;
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(copy_state_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, dst, src, n_slots, n_aux);
}

void launch_copy_instances(hipStream_t stream, const SecDefDev *defs, const int32_t *def_ids,
                           const int32_t *insts, int n, const StateArrays &dst, const StateArrays &src) {
    if (n <= 0) return;
    hipLaunchKernelGGL(copy_instances_kernel, dim3(n), dim3(256), 0, stream, defs, def_ids, insts, n, dst, src);
}

}  // namespace sgpu
