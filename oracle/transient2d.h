/* Case and result records of the oracle's 2D transient driver (test
 * infrastructure only).  Field meanings follow Strut's JSON case schema
 * (docs/harness.md) after loading (solver/run_case/loader.mojo). */
#ifndef STRUT_ORACLE_TRANSIENT2D_H
#define STRUT_ORACLE_TRANSIENT2D_H

typedef struct {
    int n_nodes;
    const double *x, *y;
    const int *constrained;  /* n_nodes * 3 flags, plain DOF numbering */
    const double *mass;      /* n_nodes * 3 lumped masses (M_total) */
    int n_elem;
    const int *elem_nodes;   /* 2 per element (node indices) */
    void *const *elem_sec;   /* or_sec2d* per element */
    const int *elem_first_inst;
    const int *elem_nip;
    const double *elem_xi, *elem_w; /* OR_MAX_NIP per element, on [0,1] */
    const int *q_offset;     /* element state offset in q_state */
    double *q_state;         /* 3 + 2 nip + 3 doubles per element, zeroed */
    double dt;
    int steps;
    double gamma, beta, tol;
    int max_iters;
    double alpha_m, beta_k, beta_k_init;
    int exc_dir;             /* UniformExcitation direction (1-based), 0 = none */
    int ts_n;
    const double *ts_values;
    double ts_dt, ts_factor, ts_start;
    int ts_use_last;
    const int *elem_type;    /* 0 forceBeamColumn2d, 1 dispBeamColumn2d per element; NULL: all 0 */
    /* Not in the reference (0 restates it): every step's Newton start, the
     * Newmark predictor, scaled by 1 + start_eps (-1)^i on free DOF i -- a
     * probe of how far the converged solution depends on where Newton
     * starts, as a learned proposal moves it. */
    double start_eps;
} or_tr2d_case;

typedef struct {
    double *u_hist;  /* steps * 3 n_nodes, or NULL */
    int *iters;      /* Newton iterations (assemblies) per step */
    int steps_done;
    int hbw, free_count;
    long section_evals, local_iters, subdivisions, fallbacks, dense_fallbacks;
} or_tr2d_out;

double or_path_eval(const or_tr2d_case *c, double t);
double or_path_eval_raw(int n, const double *values, double dt, double factor, double start_time, int use_last,
                        double t);
int or_band_solve(int n, int hbw, const double *band, const double *rhs, double *x, int *dense_fallback);
int or_band_factor_rows(int n, int hbw, const double *band, double *u_rows, int *dense_fallback);
int or_tr2d_run(const or_tr2d_case *c, or_tr2d_out *out);

#endif
