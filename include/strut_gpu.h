/*
 * strut_gpu.h -- C ABI of libstrut_gpu, the MI355X (gfx950) implementation
 * of Strut's nonlinear fiber-section state determination.
 *
 * Plain C: integers, doubles, pointers and sizes only.  One opaque context
 * owns every device buffer and one HIP stream; calls on one context are not
 * thread-safe; several contexts may live in one process.  No call aborts the
 * process: every entry point returns an sgpu_status, and sgpu_last_error()
 * holds the message of the last failure on that context.
 *
 * Each declaration names the reference interface it replaces (paths are
 * relative to the Strut repository root, src/mojo/...).
 */
#ifndef STRUT_GPU_H
#define STRUT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGPU_ABI_VERSION 1

typedef struct sgpu_ctx sgpu_ctx;

/* Status codes.  Negative values are API or runtime errors; positive values
 * are numerical outcomes the host maps onto the reference's messages. */
typedef enum {
    SGPU_OK = 0,
    SGPU_ERR_ARG = -1,       /* bad argument (size, index, null pointer)   */
    SGPU_ERR_HIP = -2,       /* HIP runtime error                          */
    SGPU_ERR_NODEV = -3,     /* no usable gfx950 device                    */
    SGPU_ERR_STATE = -4,     /* call out of order (e.g. before finalize)   */
    SGPU_SINGULAR_SECTION = 1, /* section |det k| <= 1e-40 (fiber2d.mojo:3074-3077) */
    SGPU_ERR_LOCAL_NONCONV = 2 /* forceBeamColumn3d element compatibility did not
                                  converge (force_beam_column3d.mojo:2846 aborts) */
} sgpu_status;

/* Material type tags: tag_types.mojo:136-142 (UniMaterialTypeTag). */
enum {
    SGPU_ELASTIC = 0,
    SGPU_STEEL01 = 1,
    SGPU_CONCRETE01 = 2,
    SGPU_STEEL02 = 3,
    SGPU_CONCRETE02 = 4
};

int sgpu_abi_version(void);

/* Create a context on HIP device `device`.  Replaces nothing in the
 * reference (the reference has no device); owns all state the reference
 * keeps in RunCaseState (solver/run_case/loader.mojo:68). */
int sgpu_ctx_create(int device, sgpu_ctx **out);
int sgpu_ctx_destroy(sgpu_ctx *ctx);
const char *sgpu_last_error(const sgpu_ctx *ctx);
/* The context's hipStream_t, for callers that order their own work on it. */
void *sgpu_stream(sgpu_ctx *ctx);
int sgpu_synchronize(sgpu_ctx *ctx);

/* Upload the uniaxial material table: n definitions, tags in `types`,
 * 11 parameters each (p0..p10) already normalised as
 * solver/run_case/loader.mojo:963-1096 does.  Replaces the
 * List[UniMaterialDef] of core.mojo:48-118. */
int sgpu_material_defs_upload(sgpu_ctx *ctx, int n, const int32_t *types, const double *params);

/* Create a fiber-section definition from its discretised fibers (input
 * order, as the patch/layer generators of fiber2d.mojo:3504-3878 and
 * fiber3d.mojo:258-470 emit them).  dim = 2 ignores z.  Replaces
 * _build_fiber_section2d_def (fiber2d.mojo:3323-3455) and
 * _build_fiber_section3d_def (fiber3d.mojo:157-186). */
int sgpu_section_def_create(sgpu_ctx *ctx, int dim, int n_fibers, const double *y, const double *z,
                            const double *area, const int32_t *mat_index, int32_t *def_id);

/* Allocate n runtime instances of a definition (one per element
 * integration point); returns the first instance index.  Replaces
 * fiber_section2d_runtime_alloc_instances (fiber2d.mojo:961-1006) and the
 * per-element state allocation of loader.mojo:1760-1869. */
int sgpu_section_instances_alloc(sgpu_ctx *ctx, int32_t def_id, int n, int32_t *first_instance);

/* Freeze definitions and instances, upload the slot tables and the initial
 * material states (UniMaterialState(mat_def), core.mojo:203-263).  No
 * definition or instance may be added afterwards. */
int sgpu_sections_finalize(sgpu_ctx *ctx);

/* Static facts of a definition: out[0] fiber count, out[1] slot count,
 * out[2] y centroid, out[3] z centroid, out[4] initial flexibility valid,
 * out[5..7] initial 2D flexibility {f00, f01, f11}
 * (fiber2d.mojo:3448-3454). */
int sgpu_section_def_info(sgpu_ctx *ctx, int32_t def_id, double *out8);

/* One section evaluation point: a (definition, instance) pair and its
 * section deformation.  2D uses {eps0, kappa_z}; 3D uses
 * {eps0, kappa_z, kappa_y}.  Mirrors FiberSection2dBatchPoint
 * (fiber2d.mojo:621-676). */
typedef struct {
    int32_t def_id;
    int32_t instance;
    double eps0;
    double kappa_z;
    double kappa_y;
} sgpu_section_point;

/* Section response.  2D: s = {N, Mz, 0}, k = {k11, k12, k22, 0, 0, 0},
 * f = {f00, f01, f11, 0, 0, 0}.  3D: s = {N, My, Mz},
 * k = {k11, k12, k13, k22, k23, k33}, f = the same entries of k^-1.
 * valid = 0 when |det k| <= 1e-40.  Mirrors FiberSection2dResponse and
 * FiberSection2dBatchResult (fiber2d.mojo:586-612, :679-705) and
 * FiberSection3dResponse (fiber3d.mojo:98-143). */
typedef struct {
    double s[3];
    double k[6];
    double f[6];
    int32_t valid;
    int32_t pad_;
} sgpu_section_result;

/* Evaluate n points: trial strains, fiber trial states, section resultants,
 * tangents and flexibilities.  Fiber trial states are always computed from
 * the committed states, so points must name distinct instances within one
 * call.  Replaces fiber_section2d_set_trial_from_offset
 * (fiber2d.mojo:3982-4024), fiber_section2d_batch_eval_same_def
 * (:3042-3086) and fiber_section3d_set_trial_from_offset
 * (fiber3d.mojo:769-836).  Host buffers; synchronous. */
int sgpu_section_eval(sgpu_ctx *ctx, int n, const sgpu_section_point *points,
                      sgpu_section_result *results);

/* Same with device-resident points/results; asynchronous on the context's
 * stream.  Used by hosts that keep the Newton loop on the device. */
int sgpu_section_eval_device(sgpu_ctx *ctx, int n, const sgpu_section_point *d_points,
                             sgpu_section_result *d_results);

/* sgpu_section_eval_device for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  The same kernels and results; the aux
 * pre-pass results go to a fixed device array of 2^20 doubles (8 MiB) that
 * each GPU holds for the life of the process, so n * (most aux slots of any
 * definition) * 2 must not exceed 2^20 (else SGPU_ERR_ARG, nothing
 * launched).  Calls no hipMalloc/hipFree and records no profiling events.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (two streams or two contexts at once) race; calls in order on one stream
 * do not (every entry the main pass reads is written by the same call's aux
 * pass first). */
int sgpu_section_eval_device_noalloc(sgpu_ctx *ctx, int n, const sgpu_section_point *d_points,
                                     sgpu_section_result *d_results);

/* Commit (trial -> committed) or revert (committed -> trial) every fiber of
 * every instance.  Replaces fiber_section2d_commit_runtime_all /
 * revert_trial_runtime_all (fiber2d.mojo:4191-4252) and
 * fiber_section3d_commit_all / revert_trial_all (fiber3d.mojo:905-926).
 * Asynchronous on the context's stream. */
int sgpu_sections_commit_all(sgpu_ctx *ctx);
int sgpu_sections_revert_all(sgpu_ctx *ctx);

/* Commit or revert listed instances only.  Replaces
 * fiber_section2d_commit_from_offset / revert_trial_from_offset
 * (fiber2d.mojo:4059-4128) and fiber_section3d_commit / revert_trial
 * (fiber3d.mojo:873-902).  Host lists; synchronous. */
int sgpu_sections_commit(sgpu_ctx *ctx, int n, const int32_t *def_ids, const int32_t *instances);
int sgpu_sections_revert(sgpu_ctx *ctx, int n, const int32_t *def_ids, const int32_t *instances);

/* Read the per-fiber state of one instance in input fiber order:
 * out[3*i..3*i+2] = {eps, sig, tangent}; which = 0 trial, 1 committed.
 * Elastic fibers report zeros (the 2D reference keeps no state for them). */
int sgpu_fiber_state_read(sgpu_ctx *ctx, int32_t def_id, int32_t instance, int which, double *out);

/* Total real fibers over all instances (the unit of the fiber-update
 * count, benchmarks_micro.mojo:250) and total slots. */
int sgpu_fiber_totals(sgpu_ctx *ctx, int64_t *fibers, int64_t *slots);

/* ------------------------------------------------------------------
 * 3D fiber beam-column elements: forceBeamColumn3d and dispBeamColumn3d
 * on 3D fiber sections, Linear or PDelta transformation, elastic torsion,
 * no element loads.  The element state (the force-based basic forces and
 * section deformations, the fibers' trial states) is device-resident.
 * ------------------------------------------------------------------ */

#define SGPU_MAX_NIP3D 5
enum { SGPU_ELEM3D_FORCE = 0, SGPU_ELEM3D_DISP = 1 };
enum { SGPU_TRANSF_LINEAR = 0, SGPU_TRANSF_PDELTA = 1 };

/* One element as the loader hands it to the element routines: type, 3D
 * fiber-section definition, integration rule (points on [0, 1], weights
 * halved: beam_integration.mojo:224-241), geomTransf, end coordinates,
 * optional vecxz (utils.mojo:96-139) and the elastic torsion G, J
 * (applied when both are > 0). */
typedef struct {
    int32_t type;
    int32_t def_id;
    int32_t nip;
    int32_t transf;
    double xyz1[3], xyz2[3];
    int32_t has_vecxz;
    int32_t pad_;
    double vecxz[3];
    double G, J;
    double xi[SGPU_MAX_NIP3D], w[SGPU_MAX_NIP3D];
} sgpu_elem3d_desc;

/* Register an element before sgpu_sections_finalize: allocates its nip
 * section instances of def_id (loader.mojo:1821-1850; the force-based
 * basic state [q0..q4 | eps0[nip] | kappa_y[nip] | kappa_z[nip]] starts at
 * zero, :1826-1841) and caches its length and rotation
 * (force_beam_column3d.mojo:895-953). */
int sgpu_elem3d_add(sgpu_ctx *ctx, const sgpu_elem3d_desc *desc, int32_t *elem_id);

/* State determination of n listed elements at their global end
 * displacements u12[12 n]: global tangents k144[144 n] (row major) and
 * resisting forces f12[12 n].  Replaces
 * force_beam_column3d_fiber_global_tangent_and_internal
 * (force_beam_column3d.mojo:2549-2990) and
 * beam_column3d_fiber_global_tangent_and_internal
 * (beam_column3d_nonlinear.mojo:357-575) in batch.  The force-based
 * elements keep their trial basic state across calls (transient Newton
 * iterations).  status[i] = SGPU_ERR_LOCAL_NONCONV where the reference
 * aborts; the call then returns that code.  Host buffers; synchronous. */
int sgpu_elem3d_state(sgpu_ctx *ctx, int n, const int32_t *elem_ids, const double *u12, double *k144, double *f12,
                      int32_t *status);

/* Same with device buffers; asynchronous on the context's stream, status
 * per element left in d_status. */
int sgpu_elem3d_state_device(sgpu_ctx *ctx, int n, const int32_t *d_elem_ids, const double *d_u12, double *d_k144,
                             double *d_f12, int32_t *d_status);

/* Commit the listed elements' sections (trial -> committed,
 * fiber3d.mojo:873-926, as transient_nonlinear.mojo:3081-3101 does after a
 * converged step). */
int sgpu_elem3d_commit(sgpu_ctx *ctx, int n, const int32_t *elem_ids);

/* The force-based element's basic state (5 + 3 nip values). */
int sgpu_elem3d_basic_state(sgpu_ctx *ctx, int32_t elem_id, double *q);

/* ------------------------------------------------------------------
 * Elastic four-node shells (ShellMITC4 with an ElasticMembranePlateSection)
 * ------------------------------------------------------------------ */

/* Global stiffness of n elastic MITC4 shells: xyz[12 e + 3 i + c] is
 * coordinate c of node i of shell e, props[3 e] = (E, nu, h); K[576 e] is
 * the 24 x 24 matrix, row major, DOFs ux uy uz rx ry rz per node in the
 * given node order.  Replaces shell4_mindlin_stiffness
 * (elements/shell4.mojo:10-330), which _assemble_surface_soa_indices calls
 * for every shell at every assembly
 * (solver/assembly/stiffness_internal_surface.mojo:238-280).  A degenerate
 * basis or a zero Jacobian is SGPU_ERR_ARG (the reference aborts). */
int sgpu_shell4_stiffness(sgpu_ctx *ctx, int n, const double *xyz, const double *props, double *K);

/* ------------------------------------------------------------------
 * 2D fiber beam-column elements outside any frame: the element calls of
 * the 2D frame assembly (stiffness_internal_frame2d.mojo:336-634) for a
 * host that scatters them itself.  The device routines are the 2D frame
 * path's (forceBeamColumn2d with its local iteration, subdivision, initial-
 * tangent fallback and exact all-elastic state; dispBeamColumn2d), Linear
 * transformation, no element loads.
 * ------------------------------------------------------------------ */
#define SGPU_MAX_NIP2D 5
enum { SGPU_ELEM2D_FORCE = 0, SGPU_ELEM2D_DISP = 1 };

/* type, 2D fiber-section definition, integration rule (points on [0, 1],
 * weights summing to 1: beam_integration.mojo:224-241), end coordinates. */
typedef struct {
    int32_t type;
    int32_t def_id;
    int32_t nip;
    int32_t pad_;
    double xy1[2], xy2[2];
    double xi[SGPU_MAX_NIP2D], w[SGPU_MAX_NIP2D];
} sgpu_elem2d_desc;

/* Register an element before sgpu_sections_finalize: allocates its nip
 * section instances of def_id and a zero force-based basic state
 * [q0..q2 | eps0[nip] | kappa[nip] | v0..v2] (loader.mojo:1760-1869). */
int sgpu_elem2d_add(sgpu_ctx *ctx, const sgpu_elem2d_desc *desc, int32_t *elem_id);

/* State determination of n distinct listed elements at their global end
 * displacements u6[6 n]: global tangents k36[36 n] (row major) and
 * resisting forces f6[6 n].  Replaces
 * force_beam_column2d_fiber_global_tangent_and_internal
 * (force_beam_column2d.mojo:1418-1919) and the dispBeamColumn2d state
 * (disp_beam_column2d.mojo:114-254) in batch.  Force-based elements keep
 * their trial basic state across calls (Newton iterations).  status[i] =
 * SGPU_ERR_LOCAL_NONCONV where the reference aborts (no tangent could be
 * recovered); the call then returns that code.  Host buffers, synchronous. */
int sgpu_elem2d_state(sgpu_ctx *ctx, int n, const int32_t *elem_ids, const double *u6, double *k36, double *f6,
                      int32_t *status);

/* Commit the listed elements' sections at their last evaluated
 * deformations (transient_nonlinear.mojo:3081-3101); with u6 (the
 * converged end displacements, or NULL) a force-based element's basic
 * deformation is re-synced from them (helpers.mojo:1832-1883). */
int sgpu_elem2d_commit(sgpu_ctx *ctx, int n, const int32_t *elem_ids, const double *u6);

/* The forceBeamColumn2d element's basic state (3 + 2 nip + 3 values). */
int sgpu_elem2d_basic_state(sgpu_ctx *ctx, int32_t elem_id, double *q);

/* ------------------------------------------------------------------
 * Multi-point constraints of the Transformation handler (rigidDiaphragm,
 * equalDOF).  DOF i of n expands as u_i = sum_t coeff_pool[t] u[dof_pool[t]]
 * over t in [row_offsets[i], row_offsets[i + 1]) -- the tables
 * solver/run_case/loader.mojo:2230-2372 builds (identity rows for ordinary
 * DOFs, the retained DOFs for a constrained one, none for an SP-fixed one).
 * ------------------------------------------------------------------ */

/* Upload the tables (replaces earlier ones); constrained / slave: 0 or 1
 * per DOF (loader.mojo:2245-2345).  A row may not name a DOF twice. */
int sgpu_mpc_set(sgpu_ctx *ctx, int n, const int32_t *row_offsets, const int32_t *dof_pool, const double *coeff_pool,
                 const int32_t *constrained, const int32_t *slave);

/* out = T^T K T on dense n x n row-major matrices: _collapse_matrix_by_mpc
 * (solver/run_case/helpers.mojo:3345-3377), the same additions in the same
 * order (entries of K equal to 0 contribute nothing).  Host buffers,
 * synchronous; the _device form takes device buffers and is asynchronous. */
int sgpu_mpc_collapse_matrix(sgpu_ctx *ctx, const double *K, double *out);
int sgpu_mpc_collapse_matrix_device(sgpu_ctx *ctx, const double *d_K, double *d_out);

/* out = T^T F: _collapse_vector_by_mpc (helpers.mojo:3288-3301). */
int sgpu_mpc_collapse_vector(sgpu_ctx *ctx, const double *F, double *out);

/* In place: slave DOFs from their retained DOFs, SP-fixed DOFs to 0 --
 * _enforce_mpc_values (helpers.mojo:3380-3397). */
int sgpu_mpc_enforce(sgpu_ctx *ctx, double *u);

/* ------------------------------------------------------------------
 * Transient nonlinear analysis of one 3D fiber beam-column frame
 * (transient_nonlinear.mojo:830-3638 with ndf 6): the sgpu_elem3d elements
 * of the context, the Transformation handler's expansion tables
 * (loader.mojo:2230-2372; identity rows when has_mpc is 0), lumped masses,
 * Rayleigh alphaM / betaK / betaKInit, Newmark, Newton with the
 * NormUnbalance test and a BandGeneral band; one or more UniformExcitation
 * patterns whose collapsed inertia loads are summed (one is Strut's schema,
 * two is the bidirectional extension).
 * ------------------------------------------------------------------ */
#define SGPU_F3_MAX_PATTERNS 3

/* run status beyond 0 (all steps done) and k > 0 (step k did not converge
 * within max_iters, the reference's abort) */
#define SGPU_F3_ELEMENT_FAILED (-1)   /* forceBeamColumn3d compatibility failed (the reference aborts) */
#define SGPU_F3_SINGULAR (-11)        /* a band pivot at or below 1e-14 */

typedef struct {
    int n_nodes;                     /* 6 DOFs per node, plain numbering */
    const int32_t *sp_constrained;   /* 6 n_nodes single-point flags */
    const double *mass;              /* 6 n_nodes lumped masses */
    int has_mpc;                     /* Transformation handler with constraints */
    const int32_t *row_offsets;      /* 6 n_nodes + 1 */
    const int32_t *dof_pool;
    const double *coeff_pool;
    const int32_t *slave;            /* 6 n_nodes */
    int n_elem;
    const int32_t *elem_ids;         /* sgpu_elem3d ids in element order */
    const int32_t *elem_nodes;       /* 2 node indices per element */
    double dt, gamma, beta, tol;
    int max_iters;
    double alpha_m, beta_k, beta_k_init;
    int n_patterns;
    int exc_dir[SGPU_F3_MAX_PATTERNS];             /* 1-based direction */
    int ts_n[SGPU_F3_MAX_PATTERNS];                /* Path values (start 0) */
    const double *ts_values[SGPU_F3_MAX_PATTERNS];
    double ts_dt[SGPU_F3_MAX_PATTERNS], ts_factor[SGPU_F3_MAX_PATTERNS];
    int hbw_rule;   /* 0: half bandwidth from the element free map, as the
                       reference (linear_solver_backend.mojo:1481-1506),
                       which drops the collapse's couplings outside it;
                       1: every coupling of the collapsed system */
    /* elastic MITC4 shells (ElasticMembranePlateSection), assembled after
     * the beam-columns (stiffness_internal_surface.mojo:238-280); their
     * stiffness is formed once on the device (sgpu_shell4_stiffness's
     * kernel) and collapsed once by the Transformation handler */
    int n_shell;
    const int32_t *shell_nodes;      /* 4 node indices per shell */
    const double *shell_xyz;         /* 12 per shell: node i coordinate c at 3 i + c */
    const double *shell_props;       /* 3 per shell: E, nu, h */
} sgpu_frame3d_desc;

/* Tables of the frame (after sgpu_sections_finalize; one frame per
 * context).  Replaces the set-up half of run_transient_nonlinear
 * (:830-1700) for this path. */
int sgpu_frame3d_create(sgpu_ctx *ctx, const sgpu_frame3d_desc *desc, int32_t *hbw, int32_t *free_count);

/* Run `steps` more steps (the step loop :1698-3110): u_hist[step * 6
 * n_nodes] the full displacement after each converged step (slaves
 * enforced) or NULL, iters[step] the Newton iterations, *status as above.
 * The band LU makes the reference's row swaps (its dense partial-pivot
 * fallback, linear_solver_backend.mojo:652-657, :175-214, gives the same
 * factors as a band LU with partial pivoting); *dense_fallbacks (may be
 * NULL) counts the factorizations that swapped rows, where the reference
 * moved to its dense LU. */
int sgpu_frame3d_run(sgpu_ctx *ctx, int steps, double *u_hist, int32_t *iters, int32_t *steps_done, int32_t *status,
                     int32_t *dense_fallbacks);

/* BandGeneral factor and solve of one system on the device
 * (linear_solver_backend.mojo:631-679, :735-769): band n x (2 hbw + 1),
 * row major as the reference stores it (slot (i, j) at i (2 hbw + 1) + j - i
 * + hbw).  Variants 0, 5 and 6 make the reference's row swaps (its band LU,
 * and at the first needed swap its dense partial-pivot LU, :652-657,
 * :175-214: the same factors as a band LU with partial pivoting whose U is
 * 2 hbw wide); *info: 0 solved without a swap; 1 solved with swaps (where
 * the reference took its dense fallback); 2 a column's largest magnitude
 * at or below 1e-14 (the reference aborts), nothing solved.  0 picks the
 * register kernel (half bandwidth <= 127, n <= 3072), else the blocked
 * kernels (half bandwidth <= 1008: sixteen-column panels, trailing updates
 * over many workgroups), else the global-memory one; 5 the register kernel,
 * 6 the global-memory one, 8 the blocked kernels.
 * Variants 1-4 and 7 eliminate without row swaps (*info 1: the reference
 * would have swapped; the solution is the unpivoted one): 1 global memory,
 * 2 register window, 3 blocked, 4 blocked with 8-row blocks on 512
 * threads, 7 picked by size.  Host buffers; allocates. */
int sgpu_band_solve(sgpu_ctx *ctx, int variant, int n, int hbw, const double *band, const double *rhs, double *x,
                    int32_t *info);

/* Recorder of the 3D frame: the full displacement of each listed DOF
 * (6 node + dof - 1) after every converged step, up to max_steps, kept on
 * the device (node_displacement recorders, transient_nonlinear.mojo:3228-
 * 3272).  Set before sgpu_frame3d_run. */
int sgpu_frame3d_set_recorder(sgpu_ctx *ctx, int n_record, const int32_t *dofs, int max_steps);
/* The recorded rows (steps x n_record) and their number. */
int sgpu_frame3d_recorded(sgpu_ctx *ctx, double *hist, int32_t *steps);
/* Envelopes over the recorded steps, computed on the device:
 * disp_env[3 c + {0,1,2}] = min, max and largest magnitude of recorded
 * column c (envelope_node_displacement, _update_envelope,
 * helpers.mojo:3153-3186, the first step seeding all three);
 * drift_env[3 d + ...] the same for drift channel d,
 * (u[col_j[d]] - u[col_i[d]]) / dx[d] (_drift_value, helpers.mojo:3113-3142;
 * dx = coordinate_j - coordinate_i along perp_dirn). */
int sgpu_frame3d_envelopes(sgpu_ctx *ctx, int n_drift, const int32_t *col_i, const int32_t *col_j, const double *dx,
                           double *disp_env, double *drift_env);

/* Modal analysis of the frame at rest: run_modal_eigen
 * (solver/run_case/analysis/modal_eigen.mojo:143-362) with
 * jacobi_symmetric_eigen (solver/eigen.mojo:15-79).  K_ff from the
 * beam-columns' initial tangents and the shells' stiffness, collapsed; M_ff
 * the reduced mass with the reference's floors; A = L^-1 K L^-T (M = L L^T);
 * classical Jacobi, one rotation per pass on the first largest off-diagonal
 * magnitude in row-major order, at most max_rotations passes (<= 0: the
 * reference's 100), stopping at or below tol (the reference's 1e-12).
 * eigenvalues[num_modes] ascending (the reference's selection sort),
 * modes[num_modes x 6 n_nodes] = L^-T v expanded to every DOF
 * (_enforce_mpc_values), rotations = passes that rotated.  Needs 24 n^2
 * bytes of device memory for the call (n free DOFs). */
int sgpu_frame3d_modal(sgpu_ctx *ctx, int num_modes, int max_rotations, double tol, double *eigenvalues,
                       double *modes, int32_t *rotations);

/* Another record (ground-motion instance) of the same frame, before the
 * first run and before the recorder: its own elements (sgpu_elem3d_add, so
 * their own section states; element i of this list plays element i of the
 * frame) and its own series for each of the frame's patterns.  The frame's
 * description is record 0.  Records run together: every kernel of a Newton
 * iteration covers all records still iterating, one band solve workgroup
 * per record. */
int sgpu_frame3d_add_record(sgpu_ctx *ctx, const int32_t *elem_ids, const double *const *ts_values,
                            const int32_t *ts_n, const double *ts_dt, const double *ts_factor, int32_t *record);
/* `steps` more steps of every running record, in lockstep by step, each
 * record converging, stopping or failing on its own: iters[r * steps +
 * step], steps_done[r], status[r] (as sgpu_frame3d_run), dense_fallbacks[r]
 * (may be NULL).  sgpu_frame3d_run is the one-record form. */
int sgpu_frame3d_run_records(sgpu_ctx *ctx, int steps, int32_t *iters, int32_t *steps_done, int32_t *status,
                             int32_t *dense_fallbacks);
/* The recorder's rows and envelopes of one record (the forms without a
 * record read record 0). */
int sgpu_frame3d_recorded_record(sgpu_ctx *ctx, int32_t record, double *hist, int32_t *steps);
int sgpu_frame3d_envelopes_record(sgpu_ctx *ctx, int32_t record, int n_drift, const int32_t *col_i,
                                  const int32_t *col_j, const double *dx, double *disp_env, double *drift_env);

/* The effective tangent K + a1 C + a0 M in the band (n_free x (2 hbw + 1),
 * row major, slot (i, j) at i (2 hbw + 1) + j - i + hbw), as the reference
 * loads it before each factorization (transient_nonlinear.mojo:2400-2411),
 * with record 0's elements evaluated at its current displacement (a trial
 * state; before the first step: at rest). */
int sgpu_frame3d_effective_band(sgpu_ctx *ctx, double *band);

/* The band solver of the 3D driver: 0 (default) the reference's row swaps
 * (the register kernel when the band fits it, else the blocked kernels);
 * 1 the band elimination without swaps (an opt-in for comparison; it counts
 * where the reference would have swapped, and its solution differs from the
 * reference's there by rounding, or more on an indefinite tangent); 2 the
 * reference's row swaps by the blocked kernels at any band width they
 * hold. */
int sgpu_frame3d_set_band_solver(sgpu_ctx *ctx, int solver);

/* ------------------------------------------------------------------
 * A BandGeneral system with the reference's backend calls split
 * (LinearSolverBackend, linear_solver_backend.mojo:11-92): create
 * (initialize_structure / initialize_symbolic, :1664-1685, with the half
 * bandwidth given), load (load_reduced_matrix :1877-1913 and
 * add_reduced_matrix :1916-1981), factor (factorize_loaded :2174 ->
 * :631-679 with the dense fallback :652-657) and solve (:2277-2306 ->
 * :735-769, :216-232), so one factorization serves several right-hand
 * sides.  Device-resident; calls on one band are not thread-safe.
 * ------------------------------------------------------------------ */
typedef struct sgpu_band sgpu_band;
int sgpu_band_create(sgpu_ctx *ctx, int n, int hbw, sgpu_band **out);
/* Destroy a band before its context (the band uses the context's stream). */
int sgpu_band_destroy(sgpu_band *b);
/* layout 0: n x (2 hbw + 1) band values; 1: a dense n x n row-major matrix
 * whose entries outside the band are dropped (:413-426).  add 0: the band
 * becomes scale * values; add 1: scale * values is added to the loaded
 * values.  Host buffer, synchronous. */
int sgpu_band_load(sgpu_band *b, int layout, const double *values, double scale, int add);
/* The same from device band values (layout 0), asynchronous. */
int sgpu_band_load_device(sgpu_band *b, const double *d_values, double scale, int add);
/* *info as sgpu_band_solve; the loaded values become the factors. */
int sgpu_band_factor(sgpu_band *b, int32_t *info);
/* x = A^-1 rhs with the last factorization; host buffers, synchronous. */
int sgpu_band_solve_factored(sgpu_band *b, const double *rhs, double *x);
/* The last factorization, each buffer may be NULL: u_rows n x (2 hbw + 1),
 * U(k, k + d) at k (2 hbw + 1) + d (the reference's U rows, which reach
 * 2 hbw past the diagonal after its dense fallback); l_cols n x hbw, the
 * multiplier of step k for the row at position i (after step k's swap) at
 * k hbw + i - k - 1; piv[k] the position swapped with k at step k. */
int sgpu_band_factors(sgpu_band *b, double *u_rows, double *l_cols, int32_t *piv);

/* ------------------------------------------------------------------
 * Transient nonlinear analysis of 2D forceBeamColumn fiber frames.
 * One device workgroup runs one (frame, ground motion) instance's whole
 * Newmark / Newton loop; many instances run in one launch.
 * ------------------------------------------------------------------ */

/* A 2D frame after loading (solver/run_case/loader.mojo:758-2460): nodes,
 * plain DOF numbering (3 per node), constraint flags, lumped masses
 * (M_total), forceBeamColumn2d / dispBeamColumn2d elements (elem_type 0 / 1;
 * NULL: all forceBeamColumn2d) with a fiber-section definition
 * and integration rule each (10 slots per element in elem_xi / elem_w,
 * points on [0, 1], weights halved: beam_integration.mojo:224-241), and the
 * transient_nonlinear analysis record: Newmark gamma / beta, NormUnbalance
 * tolerance, Newton iteration limit, Rayleigh alphaM / betaK / betaKInit and
 * the UniformExcitation direction (input_types.mojo:1821-2290).  Recorded
 * DOFs are global DOF indices. */
typedef struct {
    int32_t n_nodes;
    const double *x;
    const double *y;
    const int32_t *constrained;
    const double *mass;
    int32_t n_elem;
    const int32_t *elem_nodes;
    const int32_t *elem_def;
    const int32_t *elem_nip;
    const double *elem_xi;
    const double *elem_w;
    double dt, gamma, beta, tol;
    int32_t max_iters;
    double alpha_m, beta_k, beta_k_init;
    int32_t exc_dir;
    int32_t n_record;
    const int32_t *record_dofs;
    const int32_t *elem_type;
} sgpu_frame2d_desc;

/* Path time series with a constant dt (time_series.mojo:71-96). */
typedef struct {
    int32_t n;
    const double *values;
    double dt, factor, start_time;
    int32_t use_last;
} sgpu_path_series;

/* Register a frame.  Replaces the set-up half of run_transient_nonlinear
 * (transient_nonlinear.mojo:830-1700): Newmark constants, the BandGeneral
 * structure from the element free maps (linear_solver_backend.mojo:
 * 1481-1506, :1664-1676) and the element-to-band gather in element order
 * (stiffness_internal_frame2d.mojo:594-613). */
int sgpu_frame2d_create(sgpu_ctx *ctx, const sgpu_frame2d_desc *desc, int32_t *frame_id);

/* Add one instance of a frame driven by one ground-acceleration record for
 * n_steps steps; allocates its fiber-section instances.  The acceleration is
 * evaluated at t = (step + 1) dt as transient_nonlinear.mojo:1800-1852 does. */
int sgpu_frame2d_add_instance(sgpu_ctx *ctx, int32_t frame_id, int32_t n_steps, const sgpu_path_series *accel,
                              int32_t *instance_id);

/* Freeze frames and instances (finalizes the sections too) and upload. */
int sgpu_frame2d_finalize(sgpu_ctx *ctx);

/* Advance every unfinished instance by `rounds` rounds (asynchronous): one
 * round is one Newton iteration of every active instance, whatever step it is
 * on (the first round of an instance is its set-up assembly); instances that
 * have finished or stopped are masked out.  The rounds are captured once in a
 * HIP graph and replayed.  sgpu_frame2d_run repeats this until every instance
 * has finished (synchronous).  Replaces the step loop of
 * run_transient_nonlinear (transient_nonlinear.mojo:1698-3110) with its
 * Newton loop (:1959-2846). */
int sgpu_frame2d_launch(sgpu_ctx *ctx, int32_t rounds);
int sgpu_frame2d_run(sgpu_ctx *ctx, int32_t rounds);

/* The persistent schedule: one workgroup per instance runs its own Newton
 * loop for up to max_steps steps without waiting for any other instance
 * (asynchronous).  Same per-instance state and arithmetic as the rounds, so
 * the two schedules can alternate within one run with identical results. */
int sgpu_frame2d_launch_persistent(sgpu_ctx *ctx, int32_t max_steps);

/* Work-queue schedule: persistent worker blocks (as many as fit on the
 * device) take element and system tasks from a device ring, so each
 * instance runs its own Newton loop with its elements in parallel and no
 * instance waits for another; every unfinished instance advances up to
 * max_steps steps.  Same per-instance arithmetic as the other schedules. */
int sgpu_frame2d_launch_queue(sgpu_ctx *ctx, int32_t max_steps);

/* Schedule used by sgpu_frame2d_run.  AUTO (the default) times a chunk of
 * the rounds and of the work queue on the live ensemble, keeps the faster
 * and measures again whenever the number of active instances has halved and
 * every 64 chunks.
 * sgpu_frame2d_schedule_info: out5 = {schedule in use, last measured
 * GM-steps/s of the rounds, of the persistent schedule, tunings, of the
 * work-queue schedule}. */
enum { SGPU_SCHEDULE_AUTO = 0, SGPU_SCHEDULE_ROUNDS = 1, SGPU_SCHEDULE_PERSISTENT = 2, SGPU_SCHEDULE_QUEUE = 3 };
int sgpu_frame2d_set_schedule(sgpu_ctx *ctx, int schedule);
int sgpu_frame2d_schedule_info(sgpu_ctx *ctx, double *out5);

/* Per-step recorded displacements (n_steps x n_record), per-step Newton
 * iteration counts (assemblies, including the converged one), status
 * (0 = ran to the end; k > 0 = step k did not converge, where the reference
 * aborts; -10 = BandGeneral needed a row swap, where the reference switches
 * to its dense LU; -11 = singular band pivot; -12 = element tangent
 * recovery failed) and steps completed.  Any pointer may be NULL. */
int sgpu_frame2d_result(sgpu_ctx *ctx, int32_t instance, double *u_hist, int32_t *iters, int32_t *status,
                        int32_t *steps_done);

/* Envelopes of every instance of frame_id over its completed steps,
 * computed on the device from the recorded history (no history download):
 * per instance, in instance order, disp_env[(i n_record + c) 3 + {0,1,2}]
 * the min, max and largest magnitude of recorded column c, and
 * drift_env[(i n_drift + d) 3 + ...] those of drift channel d,
 * (u[col_j[d]] - u[col_i[d]]) / dx[d] (_update_envelope and _drift_value,
 * helpers.mojo:3113-3186). */
int sgpu_frame2d_envelopes(sgpu_ctx *ctx, int32_t frame_id, int n_drift, const int32_t *col_i, const int32_t *col_j,
                           const double *dx, double *disp_env, double *drift_env);

/* out4 = {free DOFs, band half-width, recorded DOFs, total DOFs}. */
int sgpu_frame2d_info(sgpu_ctx *ctx, int32_t frame_id, int32_t *out4);

/* ------------------------------------------------------------------
 * Learned Newton predictor.  No counterpart in the reference: its Newmark
 * predictor (solver/run_case/analysis/transient_nonlinear.mojo:553-590,
 * :1757-1766) is the plain method and stays the fallback.
 *
 * The network: 4 hidden layers of 256 SiLU units and a linear output,
 * bf16 weights and activations on the gfx950 matrix cores, FP32
 * accumulation, FP32 master weights trained with Adam.  For a frame with n
 * free DOFs, the features of step k are the last three displacement
 * increments, v_k and a_k at every free DOF and the ground acceleration at
 * t_k and t_{k+1} (n_in = 5 n + 2); the output is u_{k+1} minus the Newmark
 * predictor (n_out = n).  Features and targets are normalised per column
 * (x_n = (x - x_mean) * x_scale; y = y_n * y_scale + y_mean).
 *
 * Device pointers (x, y below) are [count][n_in] / [count][n_out] FP32
 * arrays in the context's device memory.  A network belongs to its
 * context and is destroyed before it.
 * ------------------------------------------------------------------ */
typedef struct sgpu_mlp sgpu_mlp;
int sgpu_mlp_create(sgpu_ctx *ctx, int32_t n_in, int32_t n_out, uint64_t seed, sgpu_mlp **out);
void sgpu_mlp_destroy(sgpu_mlp *m);
int64_t sgpu_mlp_param_count(const sgpu_mlp *m);
/* out8 = {n_in, n_out, padded n_in, padded n_out, hidden width, layers, parameters, Adam steps}. */
int sgpu_mlp_layout(const sgpu_mlp *m, int32_t *out8);
/* FP32 master weights, flat: per layer W [N][K] (row-major, padded) then b [N]. */
int sgpu_mlp_get_params(sgpu_mlp *m, float *out);
int sgpu_mlp_set_params(sgpu_mlp *m, const float *in);
int sgpu_mlp_set_normalization(sgpu_mlp *m, const float *x_mean, const float *x_scale, const float *y_mean,
                               const float *y_scale);
/* out: x_mean[n_in], x_scale[n_in], y_mean[n_out], y_scale[n_out]. */
int sgpu_mlp_get_normalization(sgpu_mlp *m, float *out);
/* Per-column mean and 1/std (features) or std (targets) of device samples. */
int sgpu_mlp_fit_normalization(sgpu_mlp *m, const float *x, const float *y, int64_t count);
/* Forward and backward pass over `rows` device samples without an update:
 * gradients of the mean squared error of the normalised outputs, in the
 * layout of sgpu_mlp_get_params (host buffer), and that loss. */
int sgpu_mlp_gradients(sgpu_mlp *m, const float *x, const float *y, int32_t rows, float *grad_out, double *loss);
/* Adam over `epochs` passes of shuffled minibatches (batch a multiple of
 * 16; the remainder of each pass is skipped); learning rate lr with a
 * linear warm-up and cosine decay.  epoch_loss (host, may be NULL) receives
 * each pass's mean squared error. */
int sgpu_mlp_train(sgpu_mlp *m, const float *x, const float *y, int64_t count, int32_t batch, int32_t epochs,
                   float lr, uint64_t seed, double *epoch_loss);
/* Batched inference: de-normalised outputs of `count` device samples. */
int sgpu_mlp_predict(sgpu_mlp *m, const float *x, int64_t count, float *y);

/* Record every instance's (features, target) pair of each converged step
 * on the device.  Chosen before sgpu_frame2d_finalize. */
int sgpu_frame2d_record_samples(sgpu_ctx *ctx, int32_t on);
/* Copy the recorded pairs of a frame's instances into device arrays x
 * ([capacity][5 n + 2]) and y ([capacity][n]); *count receives the number
 * of pairs (x or y NULL: count only). */
int sgpu_frame2d_samples(sgpu_ctx *ctx, int32_t frame_id, float *x, float *y, int64_t capacity, int64_t *count);
/* Use network m (a snapshot of its current weights; NULL: none) as the
 * predictor of every instance of a frame.  Each step then starts from the
 * network's proposal; the unchanged FP64 residual test decides convergence,
 * and the step restarts from the Newmark predictor when the proposal's
 * residual exceeds the first-order estimate of the residual at the Newmark
 * predictor, R(u_p) + K_eff (u_p - u_nm), or an element state fails there. */
int sgpu_frame2d_set_predictor(sgpu_ctx *ctx, int32_t frame_id, const sgpu_mlp *m);
/* out4 = {Newton iterations, steps started from a proposal, proposals
 * replaced by the Newmark predictor, dense-LU fallbacks} of an instance. */
int sgpu_frame2d_inst_stats(sgpu_ctx *ctx, int32_t instance, int32_t *out4);
/* Section point evaluations of instances [first, first + count) so far,
 * counted where the reference's runtime metrics count them (a batch of an
 * element's integration points in the predictor and each local iteration,
 * force_beam_column2d.mojo:910 and :1082; the points of a dispBeamColumn2d
 * state determination) plus the re-evaluation at the committed
 * deformations after a failed local iteration. */
int sgpu_frame2d_section_evals(sgpu_ctx *ctx, int32_t first, int32_t count, int64_t *out);

/* ------------------------------------------------------------------
 * Per-kernel timing.  Mirrors the phase timers of Strut's --profile output
 * (phase_times_us.json, solver/run_case/helpers.mojo), per device kernel:
 * while enabled, every launch on the context's stream is bracketed by HIP
 * events; sgpu_profile_read synchronizes the stream and returns the summed
 * milliseconds and the launch count of one kernel since the last enable.
 * ------------------------------------------------------------------ */
enum {
    SGPU_KERNEL_AUX_EVAL = 0,     /* aux_eval_kernel: Steel02 / Concrete01 slots   */
    SGPU_KERNEL_SECTION_EVAL = 1, /* section_eval_kernel: other slots + reduction  */
    SGPU_KERNEL_COMMIT = 2,       /* commit / revert copies                        */
    SGPU_KERNEL_FRAME2D = 3,      /* frame2d_kernel                                */
    SGPU_KERNEL_MPC = 4,          /* mpc_collapse_matrix_kernel                    */
    SGPU_KERNEL_F3_ELEM = 5,      /* 3D frame driver: element state launches       */
    SGPU_KERNEL_F3_SOLVE = 6,     /* 3D frame driver: band LU and substitutions    */
    SGPU_KERNEL_COUNT = 7
};
int sgpu_profile_enable(sgpu_ctx *ctx, int on);
int sgpu_profile_read(sgpu_ctx *ctx, int kernel, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* STRUT_GPU_H */
