/* Internal types of the oracle's element and transient layers (test
 * infrastructure only). */
#ifndef STRUT_ORACLE_INTERNAL_H
#define STRUT_ORACLE_INTERNAL_H

#include "strut_oracle.h"

#define OR_MAX_NIP 10

typedef struct {
    long section_evals;   /* section point evaluations */
    long local_iters;     /* forceBeamColumn local iterations */
    long subdivisions;    /* subdivision cutbacks */
    long unconverged;     /* element states left unconverged */
    long fallbacks;       /* initial-tangent fallbacks */
} or_stats;

typedef struct {
    or_sec2d *sec;
    int first_inst;       /* section instance of integration point 0 */
    int nip;
    double xi[OR_MAX_NIP], w[OR_MAX_NIP]; /* on [0,1], weights halved */
    double L, inv_L, c, s; /* cached geometry (force_beam_column2d.mojo:383-408) */
    int type;             /* 0 forceBeamColumn2d, 1 dispBeamColumn2d */
    double dc, ds;        /* dx / L, dy / L (disp_beam_column2d.mojo:137-140) */
    or_stats *stats;
} or_elem2d;

/* 3D fiber beam-column element (forceBeamColumn3d / dispBeamColumn3d) */
enum { OR_TRANSF_LINEAR = 0, OR_TRANSF_PDELTA = 1 };

typedef struct {
    or_sec3d *sec;
    int first_inst;       /* section instance of integration point 0 */
    int nip;
    double xi[OR_MAX_NIP], w[OR_MAX_NIP]; /* on [0,1], weights halved */
    double x1, y1, z1, x2, y2, z2;
    int has_vecxz;
    double vecxz[3];
    double G, J;          /* elastic torsion GJ (both > 0 to apply) */
    int transf;           /* OR_TRANSF_* */
    or_stats *stats;      /* may be NULL */
} or_elem3d;

int or_beam3d_rotation(double x1, double y1, double z1, double x2, double y2, double z2, int has_vecxz,
                       const double *vecxz, double R[9]);
int or_invert5(const double a[25], double inv[25]);
int or_dbc3d_state(const or_elem3d *el, const double u_global[12], double k_out[144], double f_out[12]);
int or_fbc3d_state(const or_elem3d *el, double *q_state, const double u_global[12], double k_out[144],
                   double f_out[12]);

/* Elastic MITC4 shell stiffness (shell4.c: shell4.mojo:10-330), 24 x 24
 * row major, DOFs ux uy uz rx ry rz per node */
int or_shell4_stiffness(double E, double nu, double h, const double x[4], const double y[4], const double z[4],
                        double K[576]);

/* Transformation-handler collapses (mpc.c) */
void or_mpc_collapse_matrix(int n, const double *K, const int *row_offsets, const int *dof_pool,
                            const double *coeff_pool, double *out);
void or_mpc_collapse_vector(int n, const double *F, const int *row_offsets, const int *dof_pool,
                            const double *coeff_pool, double *out);
void or_mpc_enforce(int n, double *u, const int *constrained, const int *slave, const int *row_offsets,
                    const int *dof_pool, const double *coeff_pool);

/* BandGeneral storage and no-pivot LU with the dense fallback
 * (transient2d.c: lsb.mojo:175-240, :377-381, :631-679, :735-769) and the
 * chunked dot product (simd_contiguous.mojo:197-225), shared by the
 * transient drivers */
typedef struct {
    int n, hbw, w;
    double *band;       /* n x (2 hbw + 1) */
    double *dense;      /* n x n dense fallback storage */
    int *piv;
    double *work;
    int dense_fallback;
    long dense_fallbacks;
} band_t;
int band_slot(int hbw, int row, int col);
int band_factor(band_t *b);
void band_solve(band_t *b, const double *rhs, double *x);
double or_tree8(const double *v);
double or_dot_w(const double *a, const double *b, int n);

int or_sec2d_all_elastic(const or_sec2d *s);
int or_dbc2d_state(const or_elem2d *el, const double u_global[6], double k_out[36], double f_out[6]);
int or_fbc2d_state(const or_elem2d *el, double *q_state, const double u_global[6], double k_out[36],
                   double f_out[6]);

#endif
