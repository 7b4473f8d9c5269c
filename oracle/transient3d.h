/* Case and result records of the oracle's 3D transient driver (test
 * infrastructure only): 3D fiber beam-column frames (6 DOFs per node) with
 * the Transformation constraint handler and one or more UniformExcitation
 * patterns (one pattern is Strut's schema; several, summed in order, is the
 * bidirectional extension). */
#ifndef STRUT_ORACLE_TRANSIENT3D_H
#define STRUT_ORACLE_TRANSIENT3D_H

#define OR_TR3D_MAX_PATTERNS 3

typedef struct {
    int n_nodes;
    const int *sp_constrained;   /* n_nodes * 6 single-point flags */
    const double *mass;          /* n_nodes * 6 lumped masses (M_total) */
    /* expansion tables (loader.mojo:2230-2372): identity rows, retained
     * terms for slaves, empty rows for single-point-fixed DOFs */
    const int *row_offsets, *dof_pool;
    const double *coeff_pool;
    const int *slave;
    int has_mpc;                 /* Transformation handler with constraints */
    int n_elem;
    const int *elem_nodes;       /* 2 per element (node indices) */
    void *const *elem;           /* or_elem3d* per element */
    const int *elem_kind;        /* 0 forceBeamColumn3d, 1 dispBeamColumn3d */
    double *const *elem_q;       /* force-based basic state per element */
    /* elastic MITC4 shells (ElasticMembranePlateSection), assembled after
     * the beam-columns as _assemble_surface_soa_indices does */
    int n_shell;
    const int *shell_nodes;      /* 4 per shell (node indices) */
    const double *shell_xyz;     /* 12 per shell: node i coordinate c at 3 i + c */
    const double *shell_props;   /* 3 per shell: E, nu, h */
    double dt;
    int steps;
    double gamma, beta, tol;
    int max_iters;
    double alpha_m, beta_k, beta_k_init;
    int n_patterns;
    int exc_dir[OR_TR3D_MAX_PATTERNS];          /* 1-based direction */
    int ts_n[OR_TR3D_MAX_PATTERNS];
    const double *ts_values[OR_TR3D_MAX_PATTERNS];
    double ts_dt[OR_TR3D_MAX_PATTERNS], ts_factor[OR_TR3D_MAX_PATTERNS];
    int hbw_rule;                /* 0: element free map (the reference), 1: collapsed couplings */
} or_tr3d_case;

typedef struct {
    double *u_hist;  /* steps * 6 n_nodes, or NULL */
    int *iters;      /* Newton iterations (assemblies) per step */
    int steps_done;
    int hbw, free_count;
    long dense_fallbacks;
    double *band_first; /* n x (2 hbw + 1): K_eff of the first Newton solve, or NULL */
} or_tr3d_out;

int or_tr3d_run(const or_tr3d_case *c, or_tr3d_out *out);
/* run_modal_eigen (modal_eigen.mojo:143-362) on the same model: num_modes
 * eigenvalues and mode shapes (num_modes x 6 n_nodes); the case's analysis
 * fields are not read */
int or_tr3d_modal(const or_tr3d_case *c, int num_modes, int max_rotations, double tol, double *eig, double *modes,
                  int *rotations);

#endif
