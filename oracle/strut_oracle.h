/*
 * strut_oracle.h -- CPU restatement of Strut's fiber-section state
 * determination path, used ONLY as the parity checker and the timed CPU
 * baseline.  TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library
 * (libstrut_gpu.so) never links, calls or falls back to anything here.
 *
 * Every function restates the reference Mojo source named in its comment
 * (paths relative to the Strut repository root).  Arithmetic order follows
 * the reference expression by expression; the library is compiled with
 * -ffp-contract=off so no fused multiply-add changes a rounding.
 *
 * Choices the reference leaves to its host (parity unpinned, see the test
 * suite for the invariants that cover them):
 *   - SIMD width of the fiber family loops: W = 8 (simd_width_of[float64]
 *     on an AVX-512 host; src/mojo/solver/simd_contiguous.mojo:6).
 *   - SIMD.reduce_add order: halving tree ((v0+v4)+(v2+v6))+((v1+v5)+(v3+v7)).
 *   - vectorize tail: one scalar call per remaining element.
 *   - pow: the C library pow (the reference's `**` on Float64).
 */
#ifndef STRUT_ORACLE_H
#define STRUT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_SIMD_W 8

/* tag_types.mojo:136-142 */
enum { OR_ELASTIC = 0, OR_STEEL01 = 1, OR_CONCRETE01 = 2, OR_STEEL02 = 3, OR_CONCRETE02 = 4 };

/* materials/uniaxial/core.mojo:48-118 (UniMaterialDef: tag + p0..p10) */
typedef struct {
    int type;
    double p[11];
} or_mat_def;

/* materials/uniaxial/core.mojo:121-263 (UniMaterialState, AoS) */
typedef struct {
    double eps_c, sig_c, tangent_c, eps_p_c, alpha_c;
    double eps_t, sig_t, tangent_t, eps_p_t, alpha_t;
    double min_strain_c, end_strain_c, unload_slope_c;
    double min_strain_t, end_strain_t, unload_slope_t;
    double s2_epsmin_c, s2_epsmax_c, s2_epspl_c, s2_epss0_c, s2_sigs0_c, s2_epsr_c, s2_sigr_c;
    long s2_kon_c;
    double s2_epsmin_t, s2_epsmax_t, s2_epspl_t, s2_epss0_t, s2_sigs0_t, s2_epsr_t, s2_sigr_t;
    long s2_kon_t;
    double c2_ecmin_c, c2_dept_c, c2_ecmin_t, c2_dept_t;
} or_mat_state;

/* ---- uniaxial materials (AoS path) ---- */
double or_mat_initial_tangent(const or_mat_def *d);
void or_mat_state_init(const or_mat_def *d, or_mat_state *s);
void or_uniaxial_set_trial(const or_mat_def *d, or_mat_state *s, double eps);
void or_uniaxial_commit(or_mat_state *s);
void or_uniaxial_revert(or_mat_state *s);
/* law cores without the AoS eps == eps_t early return (2D SoA runtime) */
void or_steel01_core(const or_mat_def *d, or_mat_state *s, double eps);
void or_concrete01_core(const or_mat_def *d, or_mat_state *s, double eps);
void or_steel02_core(const or_mat_def *d, or_mat_state *s, double eps);
void or_concrete02_core(const or_mat_def *d, or_mat_state *s, double eps);
/* flat helpers for ctypes: state as 38 doubles (see materials.c) */
int or_mat_state_size(void);
void or_material_path(int type, const double *params, int n, const double *strains,
                      int commit_each, double *out_sig, double *out_tan);

/* ---- FiberSection2d (SoA runtime, family-grouped, W-chunked) ---- */
typedef struct or_sec2d or_sec2d;
or_sec2d *or_sec2d_create(int n_fibers, const double *y, const double *area, const int *mat_index,
                          int n_mats, const int *mat_type, const double *mat_params);
void or_sec2d_free(or_sec2d *s);
int or_sec2d_alloc_instances(or_sec2d *s, int n);
double or_sec2d_ybar(const or_sec2d *s);
int or_sec2d_set_trial(or_sec2d *s, int inst, double eps0, double kappa, double *out5);
int or_sec2d_batch_eval(or_sec2d *s, int n, const int *inst, const double *eps0, const double *kappa,
                        double *out8, int *valid);
void or_sec2d_commit(or_sec2d *s, int inst);
void or_sec2d_revert(or_sec2d *s, int inst);
void or_sec2d_commit_all(or_sec2d *s);
void or_sec2d_revert_all(or_sec2d *s);
int or_sec2d_initial_flex(const or_sec2d *s, double *f3);
int or_sec2d_fiber_trial(const or_sec2d *s, int inst, int fiber, double *out3);
int or_sec2d_fiber_count(const or_sec2d *s);
int or_sec2d_nonlinear_only(const or_sec2d *s);

/* ---- FiberSection3d (AoS runtime, scalar per fiber) ---- */
typedef struct or_sec3d or_sec3d;
or_sec3d *or_sec3d_create(int n_fibers, const double *y, const double *z, const double *area,
                          const int *mat_index, int n_mats, const int *mat_type,
                          const double *mat_params);
void or_sec3d_free(or_sec3d *s);
int or_sec3d_alloc_instances(or_sec3d *s, int n);
int or_sec3d_set_trial(or_sec3d *s, int inst, double eps0, double kappa_y, double kappa_z,
                       double *out9);
void or_sec3d_commit(or_sec3d *s, int inst);
void or_sec3d_revert(or_sec3d *s, int inst);
int or_sec3d_fiber_trial(const or_sec3d *s, int inst, int fiber, double *out3);
void or_sec3d_centroid(const or_sec3d *s, double *yz);
int or_sec3d_flexibility(const double *k6, double *f6);
/* timing helper: set_trial on n instances, then commit them if commit != 0 */
void or_sec3d_eval_many(or_sec3d *s, int n, const int *inst, const double *eps0, const double *ky,
                        const double *kz, int commit);

#ifdef __cplusplus
}
#endif
#endif
