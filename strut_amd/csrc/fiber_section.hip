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

template <int DIM>
__device__ void section_point(const SecDefDev &d, const SlotTables &tb, const MatDev *__restrict__ mats,
                              const StateArrays &c, const StateArrays &t, int inst, double eps0,
                              double kz, double ky, sgpu_section_result *out) {
    constexpr int NQ = DIM == 2 ? 5 : 9;
    __shared__ double red[9][kBatch + 1];
    __shared__ double fin[9];
    const int lane = threadIdx.x;
    const int64_t inst_state = d.state_base + (int64_t)inst * d.n_slots;
    const int64_t inst_aux = d.aux_base + (int64_t)inst * d.n_aux;
    const int32_t *bcb = tb.batch_chunk_begin + d.batch_base;
    const int n_chunks_total = bcb[d.n_batches];
    double acc = 0.0;
    for (int b = 0; b < d.n_batches; ++b) {
        const int s = b * kBatch + lane;
        double term[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) term[q] = 0.0;
        int code = -1;
        if (s < d.n_slots) {
            const int64_t gs = d.slot_base + s;
            const int mi = tb.mat[gs];
            code = tb.chunk[gs];
            if (mi >= 0) {
                const double y = tb.y[gs];
                const double area = tb.area[gs];
                double strain, z = 0.0;
                if (DIM == 2) {
                    strain = eps0 - y * kz;
                } else {
                    z = tb.z[gs];
                    strain = eps0 + z * ky - y * kz;
                }
                const int32_t aux = tb.aux[gs];
                double fs, ks;
                eval_slot(mats, mi, strain, area, c, t, inst_state + s, inst_aux + (aux < 0 ? 0 : aux), fs,
                          ks);
                if (DIM == 2) {
                    term[0] = fs;
                    term[1] = -fs * y;
                    term[2] = ks;
                    term[3] = -ks * y;
                    term[4] = ks * y * y;
                } else {
                    term[0] = fs;
                    term[1] = fs * z;
                    term[2] = -fs * y;
                    term[3] = ks;
                    term[4] = ks * z;
                    term[5] = -ks * y;
                    term[6] = ks * z * z;
                    term[7] = -ks * z * y;
                    term[8] = ks * y * y;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double tr = tree8(term[q]);
            if (code >= 0) red[q][code >> 1] = (code & 1) ? tr : term[q];
        }
        __syncthreads();
        if (lane < NQ) {
            const int c0 = bcb[b], c1 = bcb[b + 1];
            const bool zero_lane = d.zero_k12 && (DIM == 2 ? lane == 3 : (lane == 4 || lane == 5));
            for (int cc = c0; cc < c1; ++cc) {
                if (zero_lane && cc == d.n_elastic_chunks) acc = 0.0;
                acc += red[lane][cc - c0];
            }
        }
        __syncthreads();
    }
    if (lane < NQ) {
        const bool zero_lane = d.zero_k12 && (DIM == 2 ? lane == 3 : (lane == 4 || lane == 5));
        if (zero_lane && n_chunks_total <= d.n_elastic_chunks) acc = 0.0;
        fin[lane] = acc;
    }
    __syncthreads();
    if (lane == 0) {
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
}

__global__ __launch_bounds__(64) void section_eval_kernel(const SecDefDev *__restrict__ defs, int n_defs,
                                                          SlotTables tb, const MatDev *__restrict__ mats,
                                                          StateArrays c, StateArrays t,
                                                          const sgpu_section_point *__restrict__ pts,
                                                          sgpu_section_result *__restrict__ res, int n) {
    const int p = blockIdx.x;
    if (p >= n) return;
    const sgpu_section_point pt = pts[p];
    if (pt.def_id < 0 || pt.def_id >= n_defs || pt.instance < 0 || pt.instance >= defs[pt.def_id].n_inst) {
        if (threadIdx.x == 0) {
            sgpu_section_result r = {};
            r.valid = -1;
            res[p] = r;
        }
        return;
    }
    const SecDefDev d = defs[pt.def_id];
    if (d.dim == 2)
        section_point<2>(d, tb, mats, c, t, pt.instance, pt.eps0, pt.kappa_z, 0.0, res + p);
    else
        section_point<3>(d, tb, mats, c, t, pt.instance, pt.eps0, pt.kappa_z, pt.kappa_y, res + p);
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

void launch_section_eval(hipStream_t stream, const SecDefDev *defs, int n_defs, const SlotTables &tb,
                         const MatDev *mats, const StateArrays &c, const StateArrays &t,
                         const sgpu_section_point *pts, sgpu_section_result *res, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(section_eval_kernel, dim3(n), dim3(kBatch), 0, stream, defs, n_defs, tb, mats, c, t,
                       pts, res, n);
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
