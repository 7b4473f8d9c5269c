// Device records of the 3D fiber beam-column elements (forceBeamColumn3d,
// dispBeamColumn3d) shared by elem3d.hip and elem3d_api.cpp.  Plain structs.
#pragma once

#include <stdint.h>

#include "strut_gpu_types.h"

namespace sgpu {

constexpr int kMaxNip3d = 5;  // integration points per 3D element on the device path

enum Elem3dType : int32_t { kForceBeamColumn3d = 0, kDispBeamColumn3d = 1 };

// One element: its section definition and first section instance (points
// ip = 0..nip-1 use instances inst0 + ip), the geometry the reference caches
// per element (length and rotation R, rows = local x, y, z axes,
// force_beam_column3d.mojo:895-953), elastic torsion and the integration
// rule (xi on [0, 1], weights halved).
struct Elem3dDev {
    int32_t type, def, nip, transf;  // transf: 0 Linear, 1 PDelta
    int32_t inst0, pad_;
    int64_t q_off;  // force-based: [q0..q4 | eps0[nip] | kappa_y[nip] | kappa_z[nip]] at q_state + q_off
    double L;
    double R[9];
    double G, J;
    double xi[kMaxNip3d], w[kMaxNip3d];
};

struct Elem3dArgs {
    const SecDefDev *defs;
    SlotTables tb;
    const MatDev *mats;
    int n_mats;
    StateArrays c, t;
    const Elem3dDev *elems;
    int n_elems;
    double *q_state;
    // one call: n elements listed by id, their global displacements (12 per
    // element) and outputs (144 + 12 per element, status per element)
    int n;
    const int32_t *ids;
    const double *u;
    double *k;
    double *f;
    int32_t *status;
};

}  // namespace sgpu
