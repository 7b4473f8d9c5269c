// Device-side tables shared by the fiber-section kernels and the host code
// of libstrut_gpu.  Plain structs only; no HIP types so the host side can
// build them with ordinary C++.
#pragma once

#include <stdint.h>

namespace sgpu {

enum MatType : int32_t {
    kElastic = 0,
    kSteel01 = 1,
    kConcrete01 = 2,
    kSteel02 = 3,
    kConcrete02 = 4,
    kDead = -1,  // layout filler slot, never evaluated
};

// One uniaxial material definition (UniMaterialDef, core.mojo:48-118) plus
// constants the reference recomputes inside its loops with the same
// expressions (fiber2d.mojo:3230-3308):
//   Concrete02: pre = {ec0, ft/ec0, ft*(1/Ets+1/ec0), epsr, ec0*epsr}
//   Steel02:    pre = {b*E0, Fy/E0, sigini/E0}
struct MatDev {
    int32_t type;
    int32_t pad_;
    double p[11];
    double pre[5];
};
static_assert(sizeof(MatDev) % sizeof(double) == 0, "MatDev is copied as doubles");

// One fiber-section definition in slot layout.  Slots are the section's
// fibers placed in the reference's reduction order, 64 to a batch:
//   2D: elastic W-chunks | elastic tail singles | Steel02 groups (padded to
//       W) | Concrete02 groups (padded to W) | Steel01 | Concrete01
//   3D: elastic W-chunks | elastic tail singles | nonlinear fibers in order
// Every W-chunk starts on a multiple of W so a chunk never straddles a batch.
struct SecDefDev {
    int32_t dim;               // 2 or 3
    int32_t n_slots;
    int32_t n_fibers;          // real fibers (pads and filler excluded)
    int32_t n_aux;             // aux slots per instance
    int32_t n_batches;         // ceil(n_slots / 64)
    int32_t n_elastic_chunks;  // chunks [0, n) are elastic
    int32_t zero_k12;          // uniform elastic modulus: zero k12 (and k13 in 3D)
    int32_t n_inst;
    int64_t slot_base;         // into the per-slot definition tables
    int64_t batch_base;        // into batch_chunk_begin (n_batches + 1 entries)
    int64_t chunk_base;        // into chunk_slot (one entry per chunk)
    int64_t state_base;        // first slot-state index of instance 0
    int64_t aux_base;          // first aux-state index of instance 0
    int64_t aux_slot_base;     // into aux_slot (n_aux entries)
    double init_flex[6];       // 2D: {f00, f01, f11}; valid flag in init_flex_valid
    int32_t init_flex_valid;
    int32_t pad_;
};

// Per-slot definition tables (length = total slots over all definitions).
// chunk code: -1 = not the first slot of a reduction chunk; otherwise
// (index of the chunk within its 64-slot batch) * 2 + (chunk is W wide).
struct SlotTables {
    const double *y;
    const double *z;
    const double *area;
    const int32_t *mat;     // material index, or -1 for filler slots
    const int32_t *aux;     // aux index within the instance, or -1
    const int32_t *chunk;   // chunk code (see above)
    const int32_t *batch_chunk_begin;
    const int32_t *chunk_slot;  // first slot of each chunk
    const int32_t *aux_slot;    // slot of each aux index
    // the slot tables again, indexed like aux_slot (aux_slot_base + j), so
    // that the aux pre-pass reads them without waiting for aux_slot
    const int32_t *aux_mat;
    const double *aux_y, *aux_z, *aux_area;
};

// Fiber state, structure of arrays over all instances of all definitions.
struct StateArrays {
    double *eps, *sig, *tan, *h0, *h1;          // per slot
    double *a0, *a1, *a2, *a3, *a4;             // per aux slot
    int32_t *kon;                               // per aux slot
};

}  // namespace sgpu
