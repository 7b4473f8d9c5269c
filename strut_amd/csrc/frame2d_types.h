// Device tables of the 2D fiber-frame transient (frame2d.hip).
#pragma once

#include <stdint.h>

#include "mlp_device.h"
#include "strut_gpu_types.h"

namespace sgpu {

constexpr int kMaxNipDev = 5;  // integration points per element on the device path

// One forceBeamColumn2d or dispBeamColumn2d element (Linear transformation,
// no element loads).
struct ElemDev {
    int32_t nip;
    int32_t def;        // fiber-section definition of its integration points
    int32_t type;       // 0 forceBeamColumn2d, 1 dispBeamColumn2d
    int32_t all_elastic;  // every fiber elastic: forceBeamColumn2d's exact state (force_beam_column2d.mojo:608-732)
    int32_t fmap[6];    // free-DOF index of the element DOFs, -1 if constrained
    double L, inv_L, c, s;  // hypot length, 1/L, and dx/L as dx*(1/L) (force_beam_column2d.mojo:383-408)
                            // or, for dispBeamColumn2d, as dx/L (disp_beam_column2d.mojo:137-140)
    double L2, c2, s2;      // sqrt length and dx/L2 of the committed sync (helpers.mojo:1855-1866)
    double xi[kMaxNipDev], w[kMaxNipDev];
    double if00, if01, if11;  // initial section flexibility (fiber2d.mojo:3448-3454)
    int32_t if_valid;
    int32_t q_off;      // element state offset within the instance's state block
};

// One frame (model) shared by all its ground-motion instances.
struct FrameDev {
    int32_t n_elem, n_free, hbw, bw;
    int32_t n_record, max_iters, exc_dir, has_damping;
    int32_t q_len;      // element state doubles per instance
    int32_t pad_;
    double dt, gamma, beta, tol;
    double a0, a1, a2, a3, coeff, omg;
    double alpha_m, beta_k, beta_k_init;
    int64_t elem_base;  // into the ElemDev table
    int64_t free_base;  // into mass_ff / exc_ff
    int64_t kg_off_base, fg_off_base;  // into the gather offset tables
    int64_t rec_base;   // into rec_fidx
};

// One (frame, ground motion) instance.
struct InstDev {
    int32_t frame, n_steps;
    int32_t n_ranges, pad_;
    int64_t ag_base;        // ground acceleration at t = (step + 1) dt
    int64_t elem_inst_base; // section instance of integration point 0, per element
    int64_t q_base, vec_base, kinit_base, scratch_base, hist_base, iters_base;
    // vec: 10 n = u v a un vn an P, then the last three displacement increments
    int64_t sample_base;    // training samples (features then targets per step) into samples, or -1
    int64_t dense_base;     // n*n dense fallback matrix followed by the n*bw band copy
    int64_t piv_base;       // n row permutation of the dense fallback
    int64_t last_base;      // per element and integration point: last evaluated (eps0, kappa)
    int64_t slot_begin[4], slot_count[4], aux_begin[4], aux_count[4];  // fiber state ranges
};

struct InstState {
    int32_t step;       // next step to run
    int32_t status;     // 0 running / done; > 0: step that did not converge; < 0: failure code
    int32_t setup_done;
    int32_t total_iters;
    int32_t dense_fallbacks;  // BandGeneral factorizations that switched to the dense LU
    int32_t iter;       // assemblies so far in the current step
    int32_t proposed;   // the current iterate is the learned predictor's proposal (not yet judged)
    int32_t proposals;  // steps started from a learned proposal
    int32_t fallbacks;  // proposals replaced by the Newmark predictor
    int32_t commit_pending;  // the last converged step's commit is done by the next round's element tasks
};

enum InstStatus : int32_t {
    kInstOk = 0,
    kInstBandPivot = -10,    // BandGeneral would switch to its dense LU (lsb.mojo:652-657)
    kInstSingular = -11,     // singular band pivot (lsb.mojo:646-647, :661-662)
    kInstElementFail = -12,  // forceBeamColumn2d tangent recovery failed (force_beam_column2d.mojo:1825)
};

// Everything the frame kernel reads or writes, as device pointers.
struct Frame2dArgs {
    const FrameDev *frames;
    const InstDev *insts;
    InstState *state;
    const ElemDev *elems;
    const double *mass_ff;
    const double *exc_ff;
    const int32_t *exc_flag;
    const int32_t *rec_fidx;
    const int32_t *kg_off, *kg;
    const int32_t *fg_off, *fg;
    const int32_t *elem_inst;
    const int32_t *pair_inst, *pair_elem;  // element-kernel block -> (instance, element)
    int32_t *eflag;                        // per instance: an element state failed this round
    const double *ag;
    double *q, *vec, *kinit, *scratch, *hist, *dense, *last;
    int32_t *iters, *piv;
    const SecDefDev *defs;
    SlotTables tb;
    const MatDev *mats;
    int32_t n_mats;
    StateArrays c, t;
    const MlpDev *preds;  // per frame: learned predictor (n_in == 0: the plain Newmark predictor)
    float *samples;       // training-sample recorder, or null
    unsigned long long *sec_evals;  // per instance: section point evaluations so far
};

// Work-queue schedule: a ring of tasks, (instance << 16) | (element + 1),
// element + 1 == 0 being the instance's system step (sys_round).
struct QueueDev {
    unsigned long long *head, *tail;  // claimed / published task counts
    int32_t *slots;                   // ring of cap_mask + 1 entries, -1 when empty
    int32_t *pending;                 // per instance: element tasks of the current round not yet done
    int32_t *live;                    // instances still running in this launch
    int32_t *step0;                   // per instance: step at the launch
    int32_t *abort;                   // set when a wait exceeds its bound (the launch then fails)
    unsigned long long *start;        // realtime counter at the launch
    unsigned long long budget;        // 100 MHz ticks after which instances park (0: none)
    int32_t cap_mask;
    int32_t n_inst;
};

}  // namespace sgpu
