/*
 * Transient nonlinear analysis of 3D fiber beam-column frames in the oracle
 * (test infrastructure only).  Restates the Newmark / Newton driver
 * src/mojo/solver/run_case/analysis/transient_nonlinear.mojo:830-3638 on the
 * path a 3D frame with rigid diaphragms takes: the dense global assembly
 * (stiffness_internal_frame3d.mojo:336-876), the Transformation handler's
 * collapses of K, F_int and the inertia loads (helpers.mojo:3288-3377, driver
 * :1307-1318, :1844-1850, :2093-2110), the reduced mass
 * (_build_reduced_diagonal_matrix_by_mpc, helpers.mojo:3304-3342, driver
 * :1026-1041), the slave update after the predictor and every correction
 * (:1717-1790, :2543-2560), Rayleigh alphaM / betaK / betaKInit on the
 * reduced matrices (:2302-2335), the residual and NormUnbalance test
 * (:2337-2371) and the BandGeneral backend, whose half bandwidth comes from
 * the element free map (linear_solver_backend.mojo:1481-1506, :1664-1676)
 * and truncates the collapsed matrix to that band (:413-426).
 *
 * The free DOFs are numbered in plain DOF order without slaves
 * (loader.mojo:2378-2453 with the Transformation handler).  Several
 * UniformExcitation patterns add their collapsed inertia loads in order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "strut_oracle.h"
#include "oracle_internal.h"
#include "transient3d.h"

/* time_series.mojo:71-96 (Path with dt, start 0, no useLast) */
static double path_at(int n, const double *values, double dt, double factor, double t) {
    if (n == 0 || t < 0.0) return 0.0;
    double incr = t / dt;
    long i1 = (long)floor(incr);
    long i2 = i1 + 1;
    if (i2 >= n) return 0.0;
    double v1 = values[i1], v2 = values[i2];
    return factor * (v1 + (v2 - v1) * (incr - (double)i1));
}

/* stiffness_internal_frame3d.mojo:336-876: dense K and F_int of the
 * beam-columns, then the shells (stiffness_internal_surface.mojo:238-280:
 * every row of the shell's K scattered and dotted with the element's
 * displacements in W-wide chunks).  shell_k holds each shell's
 * shell4_mindlin_stiffness, formed once (it depends on the geometry
 * alone). */
static int assemble(const or_tr3d_case *c, const double *shell_k, const double *u, double *K, double *F,
                    int total) {
    memset(K, 0, sizeof(double) * (size_t)total * total);
    memset(F, 0, sizeof(double) * (size_t)total);
    for (int e = 0; e < c->n_elem; ++e) {
        int dof[12];
        int ni = c->elem_nodes[2 * e], nj = c->elem_nodes[2 * e + 1];
        for (int k = 0; k < 6; ++k) {
            dof[k] = 6 * ni + k;
            dof[6 + k] = 6 * nj + k;
        }
        double ue[12], ke[144], fe[12];
        for (int k = 0; k < 12; ++k) ue[k] = u[dof[k]];
        const or_elem3d *el = (const or_elem3d *)c->elem[e];
        int rc = c->elem_kind[e] == 0 ? or_fbc3d_state(el, c->elem_q[e], ue, ke, fe) : or_dbc3d_state(el, ue, ke, fe);
        if (rc < 0) return -1;
        for (int a = 0; a < 12; ++a)
            for (int b = 0; b < 12; ++b) K[(size_t)dof[a] * total + dof[b]] += ke[12 * a + b];
        for (int a = 0; a < 12; ++a) F[dof[a]] += fe[a];
    }
    for (int e = 0; e < c->n_shell; ++e) {
        int dof[24];
        double ue[24];
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 6; ++k) dof[6 * i + k] = 6 * c->shell_nodes[4 * e + i] + k;
        for (int a = 0; a < 24; ++a) ue[a] = u[dof[a]];
        const double *ke = shell_k + (size_t)576 * e;
        for (int a = 0; a < 24; ++a) {
            for (int b = 0; b < 24; ++b) K[(size_t)dof[a] * total + dof[b]] += ke[24 * a + b];
            F[dof[a]] += or_dot_w(ke + 24 * a, ue, 24);
        }
    }
    return 0;
}

static void sections_apply(const or_tr3d_case *c, int commit) {
    for (int e = 0; e < c->n_elem; ++e) {
        const or_elem3d *el = (const or_elem3d *)c->elem[e];
        for (int ip = 0; ip < el->nip; ++ip) {
            if (commit)
                or_sec3d_commit(el->sec, el->first_inst + ip);
            else
                or_sec3d_revert(el->sec, el->first_inst + ip);
        }
    }
}

static void enforce(const or_tr3d_case *c, double *u, const int *constrained, int total) {
    if (c->has_mpc) or_mpc_enforce(total, u, constrained, c->slave, c->row_offsets, c->dof_pool, c->coeff_pool);
}

int or_tr3d_run(const or_tr3d_case *c, or_tr3d_out *out) {
    const int total = 6 * c->n_nodes;
    int *constrained = malloc(sizeof(int) * total);
    int *free_idx = malloc(sizeof(int) * total);
    int *free_list = malloc(sizeof(int) * total);
    int n = 0;
    for (int i = 0; i < total; ++i) { /* loader.mojo:2378-2388, :2443-2450 */
        constrained[i] = c->sp_constrained[i] || (c->has_mpc && c->slave[i]);
        free_idx[i] = -1;
        if (!constrained[i]) {
            free_idx[i] = n;
            free_list[n++] = i;
        }
    }
    double *shell_k = calloc((size_t)576 * (c->n_shell > 0 ? c->n_shell : 1), sizeof(double));
    for (int e = 0; e < c->n_shell; ++e) {
        const double *p = c->shell_xyz + (size_t)12 * e;
        const double xs[4] = {p[0], p[3], p[6], p[9]}, ys[4] = {p[1], p[4], p[7], p[10]},
                     zs[4] = {p[2], p[5], p[8], p[11]};
        if (or_shell4_stiffness(c->shell_props[3 * e], c->shell_props[3 * e + 1], c->shell_props[3 * e + 2], xs, ys,
                                zs, shell_k + (size_t)576 * e) != 0) {
            free(shell_k);
            free(constrained); free(free_idx); free(free_list);
            return -3;
        }
    }
    /* half bandwidth */
    int hbw = 0;
    for (int e = 0; e < c->n_elem; ++e) {
        int ni = c->elem_nodes[2 * e], nj = c->elem_nodes[2 * e + 1];
        int fd[12 * 3], m = 0;
        for (int k = 0; k < 12; ++k) {
            int g = k < 6 ? 6 * ni + k : 6 * nj + k - 6;
            if (c->hbw_rule == 0) { /* lsb.mojo:1481-1506 on the element free map */
                fd[m++] = free_idx[g];
            } else { /* the couplings the collapse makes */
                for (int t = c->row_offsets[g]; t < c->row_offsets[g + 1]; ++t) fd[m++] = free_idx[c->dof_pool[t]];
            }
        }
        for (int a = 0; a < m; ++a)
            for (int b = a + 1; b < m; ++b)
                if (fd[a] >= 0 && fd[b] >= 0 && abs(fd[a] - fd[b]) > hbw) hbw = abs(fd[a] - fd[b]);
    }
    for (int e = 0; e < c->n_shell; ++e) {
        int fd[24 * 3], m = 0;
        for (int k = 0; k < 24; ++k) {
            int g = 6 * c->shell_nodes[4 * e + k / 6] + k % 6;
            if (c->hbw_rule == 0) {
                fd[m++] = free_idx[g];
            } else {
                for (int t = c->row_offsets[g]; t < c->row_offsets[g + 1]; ++t) fd[m++] = free_idx[c->dof_pool[t]];
            }
        }
        for (int a = 0; a < m; ++a)
            for (int b = a + 1; b < m; ++b)
                if (fd[a] >= 0 && fd[b] >= 0 && abs(fd[a] - fd[b]) > hbw) hbw = abs(fd[a] - fd[b]);
    }
    if (c->hbw_rule == 1) /* the reduced mass couples a slave's retained DOFs */
        for (int i = 0; i < total; ++i) {
            if (c->mass[i] == 0.0) continue;
            for (int a = c->row_offsets[i]; a < c->row_offsets[i + 1]; ++a)
                for (int b = a + 1; b < c->row_offsets[i + 1]; ++b) {
                    int fa = free_idx[c->dof_pool[a]], fb = free_idx[c->dof_pool[b]];
                    if (fa >= 0 && fb >= 0 && abs(fa - fb) > hbw) hbw = abs(fa - fb);
                }
        }
    if (hbw > n - 1) hbw = n - 1;
    out->hbw = hbw;
    out->free_count = n;
    const double dt = c->dt, gamma = c->gamma, beta = c->beta;
    const double a0 = 1.0 / (beta * dt * dt), a1 = gamma / (beta * dt), a2 = 1.0 / (beta * dt);
    const double a3 = 1.0 / (2.0 * beta) - 1.0;
    int has_damping = c->alpha_m != 0.0 || c->beta_k != 0.0 || c->beta_k_init != 0.0;
    double *K = calloc((size_t)total * total, sizeof(double));
    double *Kc = calloc((size_t)total * total, sizeof(double));
    double *F = calloc(total, sizeof(double)), *Fc = calloc(total, sizeof(double));
    double *Kff = calloc((size_t)n * n, sizeof(double));
    double *Kinit = calloc((size_t)n * n, sizeof(double));
    double *Cff = calloc((size_t)n * n, sizeof(double));
    double *Mff = calloc((size_t)n * n, sizeof(double));
    double *u = calloc(total, sizeof(double));
    double *uf = calloc(n, sizeof(double)), *vf = calloc(n, sizeof(double)), *af = calloc(n, sizeof(double));
    double *un = calloc(n, sizeof(double)), *vn = calloc(n, sizeof(double)), *an = calloc(n, sizeof(double));
    double *vt = calloc(n, sizeof(double)), *at = calloc(n, sizeof(double));
    double *P = calloc(n, sizeof(double)), *R = calloc(n, sizeof(double)), *du = calloc(n, sizeof(double));
    double *Fext = calloc(total, sizeof(double)), *load = calloc(total, sizeof(double));
    double *loadc = calloc(total, sizeof(double));
    band_t bd;
    memset(&bd, 0, sizeof(bd));
    bd.n = n;
    bd.hbw = hbw;
    bd.w = 2 * hbw + 1;
    bd.band = calloc((size_t)n * (2 * hbw + 1), sizeof(double));
    bd.dense = calloc((size_t)n * n, sizeof(double));
    bd.piv = calloc(n, sizeof(int));
    bd.work = calloc(n, sizeof(double));
    /* helpers.mojo:3304-3342: M_ff from the lumped masses through the
     * expansion (identity rows without constraints) */
    for (int i = 0; i < total; ++i) {
        double mass = c->mass[i];
        if (mass == 0.0) continue;
        for (int ti = c->row_offsets[i]; ti < c->row_offsets[i + 1]; ++ti) {
            int ri = free_idx[c->dof_pool[ti]];
            if (ri < 0) continue;
            double ci = c->coeff_pool[ti];
            for (int tj = c->row_offsets[i]; tj < c->row_offsets[i + 1]; ++tj) {
                int rj = free_idx[c->dof_pool[tj]];
                if (rj < 0) continue;
                Mff[(size_t)ri * n + rj] += mass * ci * c->coeff_pool[tj];
            }
        }
    }
    const double *Kuse = K, *Fuse = F;
    if (c->has_mpc) {
        Kuse = Kc;
        Fuse = Fc;
    }
    int status = 0;
    /* initial assembly at u (:1207), collapse (:1307-1318), initial tangent
     * (:1329-1337), revert (:1452-1453) */
    if (assemble(c, shell_k, u, K, F, total) < 0) { status = -1; goto done; }
    if (c->has_mpc) or_mpc_collapse_matrix(total, K, c->row_offsets, c->dof_pool, c->coeff_pool, Kc);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Kinit[(size_t)i * n + j] = Kuse[(size_t)free_list[i] * total + free_list[j]];
    sections_apply(c, 0);
    for (int step = 0; step < c->steps; ++step) {
        enforce(c, u, constrained, total); /* :1717-1750 */
        memcpy(un, uf, sizeof(double) * n);
        memcpy(vn, vf, sizeof(double) * n);
        memcpy(an, af, sizeof(double) * n);
        double coeff = dt * dt * (0.5 - beta); /* :553-590 */
        for (int i = 0; i < n; ++i) uf[i] = un[i] + dt * vn[i] + coeff * an[i];
        for (int i = 0; i < n; ++i) u[free_list[i]] = uf[i];
        enforce(c, u, constrained, total); /* :1768-1790 */
        double t = (double)(step + 1) * dt;
        memset(Fext, 0, sizeof(double) * total);
        for (int p = 0; p < c->n_patterns; ++p) { /* :1823-1852 */
            double ag = path_at(c->ts_n[p], c->ts_values[p], c->ts_dt[p], c->ts_factor[p], t);
            memset(load, 0, sizeof(double) * total);
            for (int i = 0; i < total; ++i)
                if ((i % 6) + 1 == c->exc_dir[p]) load[i] += -c->mass[i] * ag;
            const double *lp = load;
            if (c->has_mpc) {
                or_mpc_collapse_vector(total, load, c->row_offsets, c->dof_pool, c->coeff_pool, loadc);
                lp = loadc;
            }
            for (int i = 0; i < total; ++i) Fext[i] += lp[i];
        }
        for (int i = 0; i < n; ++i) P[i] = Fext[free_list[i]];
        int converged = 0, iters = 0;
        for (int it = 0; it < c->max_iters; ++it) {
            iters++;
            if (assemble(c, shell_k, u, K, F, total) < 0) { status = -1; goto done; }
            if (c->has_mpc) { /* :2093-2110 */
                or_mpc_collapse_matrix(total, K, c->row_offsets, c->dof_pool, c->coeff_pool, Kc);
                or_mpc_collapse_vector(total, F, c->row_offsets, c->dof_pool, c->coeff_pool, Fc);
            }
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) Kff[(size_t)i * n + j] = Kuse[(size_t)free_list[i] * total + free_list[j]];
            if (has_damping) { /* :2302-2335 */
                for (size_t k = 0; k < (size_t)n * n; ++k) Cff[k] = 0.0;
                if (c->alpha_m != 0.0)
                    for (size_t k = 0; k < (size_t)n * n; ++k) Cff[k] = c->alpha_m * Mff[k];
                if (c->beta_k != 0.0)
                    for (size_t k = 0; k < (size_t)n * n; ++k) Cff[k] += c->beta_k * Kff[k];
                if (c->beta_k_init != 0.0)
                    for (size_t k = 0; k < (size_t)n * n; ++k) Cff[k] += c->beta_k_init * Kinit[k];
            }
            double omg = 1.0 - gamma; /* :496-531 */
            for (int i = 0; i < n; ++i) {
                double anext = a0 * (uf[i] - un[i]) - a2 * vn[i] - a3 * an[i];
                at[i] = anext;
                vt[i] = vn[i] + dt * (omg * an[i] + gamma * anext);
            }
            for (int i = 0; i < n; ++i) { /* :2337-2356 */
                double damping_force = 0.0;
                if (has_damping) damping_force = or_dot_w(Cff + (size_t)i * n, vt, n);
                double inertia_force = or_dot_w(Mff + (size_t)i * n, at, n);
                R[i] = P[i] - Fuse[free_list[i]] - 0.0 - damping_force - inertia_force;
            }
            double norm = sqrt(or_dot_w(R, R, n)); /* :2357-2371 */
            if (norm <= c->tol) {
                converged = 1;
                break;
            }
            /* :2400-2411: K_eff = K + a1 C + a0 M, each truncated to the band */
            for (int row = 0; row < n; ++row) {
                int c0 = row - hbw < 0 ? 0 : row - hbw;
                int c1 = row + hbw > n - 1 ? n - 1 : row + hbw;
                for (int col = 0; col < 2 * hbw + 1; ++col) bd.band[(size_t)row * (2 * hbw + 1) + col] = 0.0;
                for (int col = c0; col <= c1; ++col) bd.band[band_slot(hbw, row, col)] = Kff[(size_t)row * n + col];
            }
            if (has_damping && a1 != 0.0)
                for (int row = 0; row < n; ++row) {
                    int c0 = row - hbw < 0 ? 0 : row - hbw;
                    int c1 = row + hbw > n - 1 ? n - 1 : row + hbw;
                    for (int col = c0; col <= c1; ++col)
                        bd.band[band_slot(hbw, row, col)] += a1 * Cff[(size_t)row * n + col];
                }
            if (a0 != 0.0)
                for (int row = 0; row < n; ++row) {
                    int c0 = row - hbw < 0 ? 0 : row - hbw;
                    int c1 = row + hbw > n - 1 ? n - 1 : row + hbw;
                    for (int col = c0; col <= c1; ++col)
                        bd.band[band_slot(hbw, row, col)] += a0 * Mff[(size_t)row * n + col];
                }
            if (out->band_first && step == 0 && it == 0)
                memcpy(out->band_first, bd.band, sizeof(double) * (size_t)n * (2 * hbw + 1));
            if (!band_factor(&bd)) { status = -2; goto done; }
            band_solve(&bd, R, du);
            for (int i = 0; i < n; ++i) uf[i] += du[i];
            for (int i = 0; i < n; ++i) u[free_list[i]] = uf[i];
            enforce(c, u, constrained, total); /* :2543-2560 */
        }
        out->iters[step] = iters;
        if (!converged) {
            status = step + 1;
            goto done;
        }
        sections_apply(c, 1); /* :3081-3101 */
        double omg = 1.0 - gamma; /* :594-647 */
        for (int i = 0; i < n; ++i) {
            double anext = a0 * (uf[i] - un[i]) - a2 * vn[i] - a3 * an[i];
            af[i] = anext;
            vf[i] = vn[i] + dt * (omg * an[i] + gamma * anext);
        }
        if (out->u_hist) memcpy(out->u_hist + (size_t)step * total, u, sizeof(double) * total);
        out->steps_done = step + 1;
    }
done:
    out->dense_fallbacks = bd.dense_fallbacks;
    free(constrained); free(free_idx); free(free_list); free(K); free(Kc); free(F); free(Fc); free(Kff);
    free(Kinit); free(Cff); free(Mff); free(u); free(uf); free(vf); free(af); free(un); free(vn); free(an);
    free(vt); free(at); free(P); free(R); free(du); free(Fext); free(load); free(loadc); free(bd.band);
    free(bd.dense); free(bd.piv); free(bd.work); free(shell_k);
    return status;
}

/* ---------------------------------------------------------------------
 * Modal analysis of the same model: run_modal_eigen
 * (solver/run_case/analysis/modal_eigen.mojo:143-362) with
 * jacobi_symmetric_eigen (solver/eigen.mojo:15-79).  K at rest (beam-columns
 * and shells, the dense assembly), collapsed by the Transformation handler
 * when constrained, restricted to the free DOFs; M_ff the reduced mass with
 * the reference's floors (a zero diagonal gets K_ii 1e-10, a non-positive
 * one max mass 1e-12); dense Cholesky M = L L^T, A = L^-1 K L^-T by the
 * reference's two dense products, classical Jacobi (one rotation per pass,
 * the first largest off-diagonal entry in row-major order, at most
 * max_rotations passes, stop at or below tol), eigenvalues sorted by the
 * reference's selection sort, mode m = L^-T v expanded to every DOF
 * (_enforce_mpc_values).  Returns 0, -1 an element failed, -2 num_modes
 * exceeds the free count, -3 no mass, -4 M not positive definite, -5 a
 * degenerate shell. */
int or_tr3d_modal(const or_tr3d_case *c, int num_modes, int max_rotations, double tol, double *eig, double *modes,
                  int *rotations) {
    const int total = 6 * c->n_nodes;
    int *constrained = malloc(sizeof(int) * total), *free_idx = malloc(sizeof(int) * total);
    int *free_list = malloc(sizeof(int) * total);
    int n = 0, status = 0;
    for (int i = 0; i < total; ++i) {
        constrained[i] = c->sp_constrained[i] || (c->has_mpc && c->slave[i]);
        free_idx[i] = -1;
        if (!constrained[i]) {
            free_idx[i] = n;
            free_list[n++] = i;
        }
    }
    double *shell_k = calloc((size_t)576 * (c->n_shell > 0 ? c->n_shell : 1), sizeof(double));
    double *K = calloc((size_t)total * total, sizeof(double)), *Kc = calloc((size_t)total * total, sizeof(double));
    double *F = calloc(total, sizeof(double)), *u = calloc(total, sizeof(double));
    double *Kff = calloc((size_t)n * n, sizeof(double)), *Mff = calloc((size_t)n * n, sizeof(double));
    double *L = calloc((size_t)n * n, sizeof(double)), *Li = calloc((size_t)n * n, sizeof(double));
    double *P = calloc((size_t)n * n, sizeof(double)), *A = calloc((size_t)n * n, sizeof(double));
    double *V = calloc((size_t)n * n, sizeof(double)), *col = calloc(n, sizeof(double));
    double *sol = calloc(n, sizeof(double)), *full = calloc(total, sizeof(double));
    int *order = malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int e = 0; e < c->n_shell; ++e) {
        const double *p = c->shell_xyz + (size_t)12 * e;
        const double xs[4] = {p[0], p[3], p[6], p[9]}, ys[4] = {p[1], p[4], p[7], p[10]},
                     zs[4] = {p[2], p[5], p[8], p[11]};
        if (or_shell4_stiffness(c->shell_props[3 * e], c->shell_props[3 * e + 1], c->shell_props[3 * e + 2], xs, ys,
                                zs, shell_k + (size_t)576 * e) != 0) { status = -5; goto done; }
    }
    if (assemble(c, shell_k, u, K, F, total) < 0) { status = -1; goto done; }
    sections_apply(c, 0);
    const double *Kuse = K;
    if (c->has_mpc) {
        or_mpc_collapse_matrix(total, K, c->row_offsets, c->dof_pool, c->coeff_pool, Kc);
        Kuse = Kc;
    }
    if (num_modes > n) { status = -2; goto done; }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Kff[(size_t)i * n + j] = Kuse[(size_t)free_list[i] * total + free_list[j]];
    for (int i = 0; i < total; ++i) { /* _build_reduced_diagonal_matrix_by_mpc */
        double mass = c->mass[i];
        if (mass == 0.0) continue;
        for (int ti = c->row_offsets[i]; ti < c->row_offsets[i + 1]; ++ti) {
            int ri = free_idx[c->dof_pool[ti]];
            if (ri < 0) continue;
            double ci = c->coeff_pool[ti];
            for (int tj = c->row_offsets[i]; tj < c->row_offsets[i + 1]; ++tj) {
                int rj = free_idx[c->dof_pool[tj]];
                if (rj < 0) continue;
                Mff[(size_t)ri * n + rj] += mass * ci * c->coeff_pool[tj];
            }
        }
    }
    {
        double max_mass = 0.0;
        int has_mass = 0;
        for (int i = 0; i < n; ++i) {
            if (Mff[(size_t)i * n + i] > max_mass) max_mass = Mff[(size_t)i * n + i];
            for (int j = 0; j < n; ++j)
                if (Mff[(size_t)i * n + j] != 0.0) has_mass = 1;
        }
        if (!has_mass) { status = -3; goto done; }
        double msmall = 1.0e-10, mass_floor = max_mass * 1.0e-12;
        if (mass_floor <= 0.0) mass_floor = 1.0e-12;
        for (int i = 0; i < n; ++i) {
            double mf = Mff[(size_t)i * n + i];
            if (mf == 0.0) mf = Kff[(size_t)i * n + i] * msmall;
            if (mf <= 0.0) mf = mass_floor;
            Mff[(size_t)i * n + i] = mf;
        }
    }
    for (int i = 0; i < n; ++i) /* _cholesky_factor_dense_spd */
        for (int j = 0; j <= i; ++j) {
            double value = Mff[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) value -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            if (i == j) {
                if (value <= 1.0e-18) { status = -4; goto done; }
                L[(size_t)i * n + j] = sqrt(value);
            } else {
                L[(size_t)i * n + j] = value / L[(size_t)j * n + j];
            }
        }
    for (int cidx = 0; cidx < n; ++cidx) { /* L^-1 column by column (_solve_lower_triangular) */
        for (int i = 0; i < n; ++i) {
            double value = i == cidx ? 1.0 : 0.0;
            for (int j = 0; j < i; ++j) value -= L[(size_t)i * n + j] * sol[j];
            sol[i] = value / L[(size_t)i * n + i];
        }
        for (int r = 0; r < n; ++r) Li[(size_t)r * n + cidx] = sol[r];
    }
    for (int i = 0; i < n; ++i) /* _multiply_dense_matrices(L^-1, K) */
        for (int k = 0; k < n; ++k) {
            double lik = Li[(size_t)i * n + k];
            if (lik == 0.0) continue;
            for (int j = 0; j < n; ++j) P[(size_t)i * n + j] += lik * Kff[(size_t)k * n + j];
        }
    for (int i = 0; i < n; ++i) /* _multiply_dense_by_transpose(P, L^-1) */
        for (int j = 0; j < n; ++j) {
            double value = 0.0;
            for (int k = 0; k < n; ++k) value += P[(size_t)i * n + k] * Li[(size_t)j * n + k];
            A[(size_t)i * n + j] = value;
        }
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    int done_rot = 0;
    if (n > 1)
        for (int sweep = 0; sweep < max_rotations; ++sweep) { /* eigen.mojo:34-74 */
            int p = 0, q = 1;
            double max_off = fabs(A[1]);
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j) {
                    double v = fabs(A[(size_t)i * n + j]);
                    if (v > max_off) {
                        max_off = v;
                        p = i;
                        q = j;
                    }
                }
            if (max_off <= tol) break;
            double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q], apq = A[(size_t)p * n + q];
            double phi = 0.5 * atan2(2.0 * apq, aqq - app);
            double cs = cos(phi), sn = sin(phi);
            for (int i = 0; i < n; ++i) {
                if (i == p || i == q) continue;
                double aip = A[(size_t)i * n + p], aiq = A[(size_t)i * n + q];
                A[(size_t)i * n + p] = cs * aip - sn * aiq;
                A[(size_t)p * n + i] = A[(size_t)i * n + p];
                A[(size_t)i * n + q] = sn * aip + cs * aiq;
                A[(size_t)q * n + i] = A[(size_t)i * n + q];
            }
            A[(size_t)p * n + p] = cs * cs * app - 2.0 * sn * cs * apq + sn * sn * aqq;
            A[(size_t)q * n + q] = sn * sn * app + 2.0 * sn * cs * apq + cs * cs * aqq;
            A[(size_t)p * n + q] = 0.0;
            A[(size_t)q * n + p] = 0.0;
            for (int i = 0; i < n; ++i) {
                double vip = V[(size_t)i * n + p], viq = V[(size_t)i * n + q];
                V[(size_t)i * n + p] = cs * vip - sn * viq;
                V[(size_t)i * n + q] = sn * vip + cs * viq;
            }
            done_rot++;
        }
    if (rotations) *rotations = done_rot;
    for (int i = 0; i < n; ++i) order[i] = i; /* _sorted_indices_ascending */
    for (int i = 0; i < n; ++i) {
        int min_i = i;
        double min_v = A[(size_t)order[i] * n + order[i]];
        for (int j = i + 1; j < n; ++j) {
            double v = A[(size_t)order[j] * n + order[j]];
            if (v < min_v) {
                min_v = v;
                min_i = j;
            }
        }
        if (min_i != i) {
            int tmp = order[i];
            order[i] = order[min_i];
            order[min_i] = tmp;
        }
    }
    for (int m = 0; m < num_modes; ++m) {
        int cm = order[m];
        eig[m] = A[(size_t)cm * n + cm];
        for (int i = 0; i < n; ++i) col[i] = V[(size_t)i * n + cm];
        for (int i = n - 1; i >= 0; --i) { /* _solve_upper_from_lower_transpose */
            double value = col[i];
            for (int j = i + 1; j < n; ++j) value -= L[(size_t)j * n + i] * sol[j];
            sol[i] = value / L[(size_t)i * n + i];
        }
        for (int i = 0; i < total; ++i) full[i] = 0.0;
        for (int i = 0; i < n; ++i) full[free_list[i]] = sol[i];
        for (int i = 0; i < total; ++i) { /* _enforce_mpc_values */
            if (c->has_mpc && c->slave[i]) {
                double value = 0.0;
                for (int t = c->row_offsets[i]; t < c->row_offsets[i + 1]; ++t)
                    value += c->coeff_pool[t] * full[c->dof_pool[t]];
                full[i] = value;
            } else if (constrained[i]) {
                full[i] = 0.0;
            }
        }
        memcpy(modes + (size_t)m * total, full, sizeof(double) * total);
    }
done:
    free(constrained); free(free_idx); free(free_list); free(shell_k); free(K); free(Kc); free(F); free(u);
    free(Kff); free(Mff); free(L); free(Li); free(P); free(A); free(V); free(col); free(sol); free(full); free(order);
    return status;
}
