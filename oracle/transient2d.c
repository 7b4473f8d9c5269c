/*
 * Transient nonlinear analysis of 2D forceBeamColumn fiber frames in the
 * oracle (test infrastructure only).  Restates the Newmark / Newton driver
 * src/mojo/solver/run_case/analysis/transient_nonlinear.mojo:830-3638 on the
 * path the fiber-frame cases take (Plain constraints, UniformExcitation,
 * Newton with the NormUnbalance test, BandGeneral, Rayleigh alphaM / betaK /
 * betaKInit), the dense assembly of solver/assembly/
 * stiffness_internal_frame2d.mojo:336-634, the BandGeneral backend of
 * solver/run_case/linear_solver_backend.mojo (:377-426, :631-679, :735-769,
 * :1481-1506, :1877-1981, dense fallback :175-240), the Path time series of
 * solver/time_series.mojo:71-118 and the committed basic-deformation sync of
 * solver/run_case/helpers.mojo:1832-1883.
 *
 * Dot products and the residual norm follow the reference's vectorize loops
 * (solver/simd_contiguous.mojo:197-270): W-chunks reduced by the halving
 * tree, scalar tail.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "strut_oracle.h"
#include "oracle_internal.h"
#include "transient2d.h"

double or_tree8(const double *v) {
    double w0 = v[0] + v[4], w1 = v[1] + v[5], w2 = v[2] + v[6], w3 = v[3] + v[7];
    double x0 = w0 + w2, x1 = w1 + w3;
    return x0 + x1;
}

/* simd_contiguous.mojo:197-225 with factor 1.0 */
double or_dot_w(const double *a, const double *b, int n) {
    double sum = 0.0;
    int i = 0;
    for (; i + OR_SIMD_W <= n; i += OR_SIMD_W) {
        double v[OR_SIMD_W];
        for (int l = 0; l < OR_SIMD_W; ++l) v[l] = (a[i + l] * b[i + l]) * 1.0;
        sum += or_tree8(v);
    }
    for (; i < n; ++i) sum += (a[i] * b[i]) * 1.0;
    return sum;
}

/* time_series.mojo:71-96 (Path with dt) */
double or_path_eval(const or_tr2d_case *c, double t) {
    int n = c->ts_n;
    if (n == 0) return 0.0;
    if (t < c->ts_start) return 0.0;
    double incr = (t - c->ts_start) / c->ts_dt;
    long incr1 = (long)floor(incr);
    long incr2 = incr1 + 1;
    if (incr2 >= n) {
        if (c->ts_use_last) return c->ts_factor * c->ts_values[n - 1];
        return 0.0;
    }
    double v1 = c->ts_values[incr1], v2 = c->ts_values[incr2];
    return c->ts_factor * (v1 + (v2 - v1) * (incr - (double)incr1));
}

int band_slot(int hbw, int row, int col) { /* lsb.mojo:377-381 */
    int off = col - row + hbw;
    if (off < 0 || off > 2 * hbw) return -1;
    return row * (2 * hbw + 1) + off;
}

/* lsb.mojo:175-214 */
static int dense_factor(band_t *b) {
    int n = b->n;
    double *A = b->dense;
    for (int i = 0; i < n; ++i) b->piv[i] = i;
    const double eps = 1.0e-14;
    for (int col = 0; col < n; ++col) {
        int prow = col;
        double pabs = fabs(A[col * n + col]);
        for (int row = col + 1; row < n; ++row) {
            double cand = fabs(A[row * n + col]);
            if (cand > pabs) {
                pabs = cand;
                prow = row;
            }
        }
        if (pabs <= eps) return 0;
        if (prow != col) {
            for (int j = 0; j < n; ++j) {
                double tmp = A[col * n + j];
                A[col * n + j] = A[prow * n + j];
                A[prow * n + j] = tmp;
            }
            int t = b->piv[col];
            b->piv[col] = b->piv[prow];
            b->piv[prow] = t;
        }
        double diag = A[col * n + col];
        if (fabs(diag) <= eps) return 0;
        for (int row = col + 1; row < n; ++row) {
            double factor = A[row * n + col] / diag;
            A[row * n + col] = factor;
            if (factor == 0.0) continue;
            for (int j = col + 1; j < n; ++j) A[row * n + j] -= factor * A[col * n + j];
        }
    }
    return 1;
}

/* lsb.mojo:631-679 */
int band_factor(band_t *b) {
    int n = b->n, hbw = b->hbw;
    size_t bsz = (size_t)n * (2 * hbw + 1);
    double *orig = malloc(sizeof(double) * bsz);
    memcpy(orig, b->band, sizeof(double) * bsz);
    const double pivot_tol = 1.0e-14;
    int rc = 1;
    for (int k = 0; k < n; ++k) {
        int i1 = k + hbw;
        if (i1 > n - 1) i1 = n - 1;
        int prow = k;
        double pabs = 0.0;
        for (int row = k; row <= i1; ++row) {
            int sl = band_slot(hbw, row, k);
            if (sl < 0) continue;
            double cand = fabs(b->band[sl]);
            if (cand > pabs) {
                pabs = cand;
                prow = row;
            }
        }
        if (pabs <= pivot_tol) { rc = 0; break; }
        if (prow != k) {
            /* lsb.mojo:608-628: band values -> dense, then dense LU */
            memset(b->dense, 0, sizeof(double) * (size_t)n * n);
            for (int row = 0; row < n; ++row) {
                int c0 = row - hbw < 0 ? 0 : row - hbw;
                int c1 = row + hbw > n - 1 ? n - 1 : row + hbw;
                for (int col = c0; col <= c1; ++col) b->dense[row * n + col] = orig[band_slot(hbw, row, col)];
            }
            rc = dense_factor(b) ? 1 : 0;
            b->dense_fallback = 1;
            b->dense_fallbacks++;
            free(orig);
            return rc;
        }
        double diag = b->band[band_slot(hbw, k, k)];
        if (fabs(diag) <= pivot_tol) { rc = 0; break; }
        int j1 = k + hbw;
        if (j1 > n - 1) j1 = n - 1;
        for (int i = k + 1; i <= i1; ++i) {
            int ik = band_slot(hbw, i, k);
            if (ik < 0) continue;
            double factor = b->band[ik] / diag;
            b->band[ik] = factor;
            if (factor == 0.0) continue;
            for (int j = k + 1; j <= j1; ++j) {
                int kj = band_slot(hbw, k, j), ij = band_slot(hbw, i, j);
                if (ij < 0 || kj < 0) continue;
                b->band[ij] -= factor * b->band[kj];
            }
        }
    }
    b->dense_fallback = 0;
    free(orig);
    return rc;
}

/* lsb.mojo:735-769 and :216-232 */
void band_solve(band_t *b, const double *rhs, double *x) {
    int n = b->n, hbw = b->hbw;
    if (b->dense_fallback) {
        double *A = b->dense;
        for (int i = 0; i < n; ++i) {
            double s = rhs[b->piv[i]];
            for (int j = 0; j < i; ++j) s -= A[i * n + j] * b->work[j];
            b->work[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = b->work[i];
            for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * x[j];
            x[i] = s / A[i * n + i];
        }
        return;
    }
    for (int i = 0; i < n; ++i) {
        double sum = rhs[i];
        int j0 = i - hbw < 0 ? 0 : i - hbw;
        for (int j = j0; j < i; ++j) {
            int sl = band_slot(hbw, i, j);
            if (sl < 0) continue;
            sum -= b->band[sl] * b->work[j];
        }
        b->work[i] = sum;
    }
    for (int i = n - 1; i >= 0; --i) {
        double sum = b->work[i];
        int j1 = i + hbw > n - 1 ? n - 1 : i + hbw;
        for (int j = i + 1; j <= j1; ++j) {
            int sl = band_slot(hbw, i, j);
            if (sl < 0) continue;
            sum -= b->band[sl] * x[j];
        }
        x[i] = sum / b->band[band_slot(hbw, i, i)];
    }
}

/* Test entry: factor and solve one BandGeneral system (n x (2 hbw + 1)
 * values, row-major as lsb.mojo:377-381).  Returns 1 on success, 0 when a
 * pivot is below the tolerance; *dense_fallback = 1 when the factorization
 * switched to the dense partial-pivot LU. */
int or_band_solve(int n, int hbw, const double *band, const double *rhs, double *x, int *dense_fallback) {
    band_t b;
    memset(&b, 0, sizeof(b));
    b.n = n;
    b.hbw = hbw;
    b.w = 2 * hbw + 1;
    b.band = malloc(sizeof(double) * (size_t)n * b.w);
    b.dense = calloc((size_t)n * n, sizeof(double));
    b.piv = calloc((size_t)n, sizeof(int));
    b.work = calloc((size_t)n, sizeof(double));
    memcpy(b.band, band, sizeof(double) * (size_t)n * b.w);
    int ok = band_factor(&b);
    if (ok) band_solve(&b, rhs, x);
    if (dense_fallback) *dense_fallback = b.dense_fallback;
    free(b.band);
    free(b.dense);
    free(b.piv);
    free(b.work);
    return ok;
}

/* Test entry: the factorization's U rows, n x (2 hbw + 1) with the diagonal
 * first (U(k, k + d) at k (2 hbw + 1) + d): from the band when the band LU
 * ran to the end (d <= hbw), from the dense LU's rows when it fell back
 * (lsb.mojo:652-657, :175-214; d <= 2 hbw).  Returns 1 on success, 0 on a
 * pivot below the tolerance. */
int or_band_factor_rows(int n, int hbw, const double *band, double *u_rows, int *dense_fallback) {
    band_t b;
    memset(&b, 0, sizeof(b));
    b.n = n;
    b.hbw = hbw;
    b.w = 2 * hbw + 1;
    b.band = malloc(sizeof(double) * (size_t)n * b.w);
    b.dense = calloc((size_t)n * n, sizeof(double));
    b.piv = calloc((size_t)n, sizeof(int));
    b.work = calloc((size_t)n, sizeof(double));
    memcpy(b.band, band, sizeof(double) * (size_t)n * b.w);
    int ok = band_factor(&b);
    memset(u_rows, 0, sizeof(double) * (size_t)n * b.w);
    if (ok)
        for (int k = 0; k < n; ++k)
            for (int d = 0; d < b.w && k + d < n; ++d) {
                if (b.dense_fallback)
                    u_rows[(size_t)k * b.w + d] = b.dense[(size_t)k * n + k + d];
                else if (d <= hbw)
                    u_rows[(size_t)k * b.w + d] = b.band[band_slot(hbw, k, k + d)];
            }
    if (dense_fallback) *dense_fallback = b.dense_fallback;
    free(b.band);
    free(b.dense);
    free(b.piv);
    free(b.work);
    return ok;
}

/* Test entry: the Path series of a case record evaluated at t. */
double or_path_eval_raw(int n, const double *values, double dt, double factor, double start_time, int use_last,
                        double t) {
    or_tr2d_case c;
    memset(&c, 0, sizeof(c));
    c.ts_n = n;
    c.ts_values = values;
    c.ts_dt = dt;
    c.ts_factor = factor;
    c.ts_start = start_time;
    c.ts_use_last = use_last;
    return or_path_eval(&c, t);
}

/* stiffness_internal_frame2d.mojo:336-634 for forceBeamColumn2d elements */
static int assemble(const or_tr2d_case *c, or_elem2d *els, double *q, const double *u, double *K, double *F,
                    int total) {
    memset(K, 0, sizeof(double) * (size_t)total * total);
    memset(F, 0, sizeof(double) * (size_t)total);
    int unconverged = 0;
    for (int e = 0; e < c->n_elem; ++e) {
        int dof[6];
        int ni = c->elem_nodes[2 * e], nj = c->elem_nodes[2 * e + 1];
        for (int k = 0; k < 3; ++k) {
            dof[k] = 3 * ni + k;
            dof[3 + k] = 3 * nj + k;
        }
        double ue[6], ke[36], fe[6];
        for (int k = 0; k < 6; ++k) ue[k] = u[dof[k]];
        int rc = els[e].type == 1 ? or_dbc2d_state(&els[e], ue, ke, fe)
                                  : or_fbc2d_state(&els[e], q + (size_t)c->q_offset[e], ue, ke, fe);
        if (rc < 0) return -1;
        unconverged += rc;
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) K[(size_t)dof[a] * total + dof[b]] += ke[6 * a + b];
        for (int a = 0; a < 6; ++a) F[dof[a]] += fe[a];
    }
    return unconverged;
}

/* helpers.mojo:1832-1883 */
static void sync_basic(const or_tr2d_case *c, double *q, const double *u) {
    for (int e = 0; e < c->n_elem; ++e) {
        if (c->elem_type && c->elem_type[e] != 0) continue; /* forceBeamColumn2d only (:1844-1845) */
        int nip = c->elem_nip[e];
        int ni = c->elem_nodes[2 * e], nj = c->elem_nodes[2 * e + 1];
        double dx = c->x[nj] - c->x[ni], dy = c->y[nj] - c->y[ni];
        double L = sqrt(dx * dx + dy * dy);
        double cs = dx / L, sn = dy / L;
        const double *u1 = u + 3 * ni, *u2 = u + 3 * nj;
        double ul[6];
        ul[0] = cs * u1[0] + sn * u1[1];
        ul[1] = -sn * u1[0] + cs * u1[1];
        ul[2] = u1[2];
        ul[3] = cs * u2[0] + sn * u2[1];
        ul[4] = -sn * u2[0] + cs * u2[1];
        ul[5] = u2[2];
        double chord = (ul[4] - ul[1]) / L;
        double *b = q + c->q_offset[e] + 3 + 2 * nip;
        b[0] = ul[3] - ul[0];
        b[1] = ul[2] - chord;
        b[2] = ul[5] - chord;
    }
}

int or_tr2d_run(const or_tr2d_case *c, or_tr2d_out *out) {
    const int total = 3 * c->n_nodes;
    int *free_idx = malloc(sizeof(int) * total);
    int *free_list = malloc(sizeof(int) * total);
    int n = 0;
    for (int i = 0; i < total; ++i) {
        free_idx[i] = -1;
        if (!c->constrained[i]) {
            free_idx[i] = n;
            free_list[n++] = i;
        }
    }
    or_stats stats;
    memset(&stats, 0, sizeof(stats));
    or_elem2d *els = calloc(c->n_elem, sizeof(or_elem2d));
    int hbw = 0;
    for (int e = 0; e < c->n_elem; ++e) {
        or_elem2d *el = &els[e];
        el->sec = (or_sec2d *)c->elem_sec[e];
        el->first_inst = c->elem_first_inst[e];
        el->nip = c->elem_nip[e];
        for (int ip = 0; ip < el->nip; ++ip) {
            el->xi[ip] = c->elem_xi[OR_MAX_NIP * e + ip];
            el->w[ip] = c->elem_w[OR_MAX_NIP * e + ip];
        }
        int ni = c->elem_nodes[2 * e], nj = c->elem_nodes[2 * e + 1];
        double dx = c->x[nj] - c->x[ni], dy = c->y[nj] - c->y[ni];
        el->L = hypot(dx, dy); /* force_beam_column2d.mojo:383-395 */
        el->inv_L = 1.0 / el->L;
        el->c = dx * el->inv_L;
        el->s = dy * el->inv_L;
        el->type = c->elem_type ? c->elem_type[e] : 0;
        el->dc = dx / el->L;
        el->ds = dy / el->L;
        el->stats = &stats;
        int fd[6];
        for (int k = 0; k < 3; ++k) {
            fd[k] = free_idx[3 * ni + k];
            fd[3 + k] = free_idx[3 * nj + k];
        }
        for (int a = 0; a < 6; ++a) /* lsb.mojo:1481-1506 */
            for (int b = a + 1; b < 6; ++b)
                if (fd[a] >= 0 && fd[b] >= 0 && abs(fd[a] - fd[b]) > hbw) hbw = abs(fd[a] - fd[b]);
    }
    out->hbw = hbw;
    out->free_count = n;
    const double dt = c->dt, gamma = c->gamma, beta = c->beta;
    const double a0 = 1.0 / (beta * dt * dt), a1 = gamma / (beta * dt), a2 = 1.0 / (beta * dt);
    const double a3 = 1.0 / (2.0 * beta) - 1.0;
    int has_damping = c->alpha_m != 0.0 || c->beta_k != 0.0 || c->beta_k_init != 0.0;
    double *K = calloc((size_t)total * total, sizeof(double));
    double *F = calloc(total, sizeof(double));
    double *Kff = calloc((size_t)n * n, sizeof(double));
    double *Kinit = calloc((size_t)n * n, sizeof(double));
    double *Cff = calloc((size_t)n * n, sizeof(double));
    double *Mff = calloc((size_t)n * n, sizeof(double));
    double *u = calloc(total, sizeof(double));
    double *uf = calloc(n, sizeof(double)), *vf = calloc(n, sizeof(double)), *af = calloc(n, sizeof(double));
    double *un = calloc(n, sizeof(double)), *vn = calloc(n, sizeof(double)), *an = calloc(n, sizeof(double));
    double *vt = calloc(n, sizeof(double)), *at = calloc(n, sizeof(double));
    double *P = calloc(n, sizeof(double)), *R = calloc(n, sizeof(double)), *du = calloc(n, sizeof(double));
    double *Fext = calloc(total, sizeof(double));
    band_t bd;
    bd.n = n;
    bd.hbw = hbw;
    bd.band = calloc((size_t)n * (2 * hbw + 1), sizeof(double));
    bd.dense = calloc((size_t)n * n, sizeof(double));
    bd.piv = calloc(n, sizeof(int));
    bd.work = calloc(n, sizeof(double));
    bd.dense_fallback = 0;
    bd.dense_fallbacks = 0;
    double *q = c->q_state;
    for (int i = 0; i < n; ++i) Mff[(size_t)i * n + i] = c->mass[free_list[i]];
    int status = 0;
    /* initial assembly at u (:1207), initial tangent copy (:1329-1337), revert (:1452-1453) */
    if (assemble(c, els, q, u, K, F, total) < 0) { status = -1; goto done; }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Kinit[(size_t)i * n + j] = K[(size_t)free_list[i] * total + free_list[j]];
    for (int e = 0; e < c->n_elem; ++e)
        for (int ip = 0; ip < els[e].nip; ++ip) or_sec2d_revert(els[e].sec, els[e].first_inst + ip);
    for (int step = 0; step < c->steps; ++step) {
        memcpy(un, uf, sizeof(double) * n);
        memcpy(vn, vf, sizeof(double) * n);
        memcpy(an, af, sizeof(double) * n);
        double coeff = dt * dt * (0.5 - beta); /* :553-590 */
        for (int i = 0; i < n; ++i) uf[i] = un[i] + dt * vn[i] + coeff * an[i];
        if (c->start_eps != 0.0)
            for (int i = 0; i < n; ++i) uf[i] *= 1.0 + ((i & 1) ? -c->start_eps : c->start_eps);
        for (int i = 0; i < n; ++i) u[free_list[i]] = uf[i];
        double t = (double)(step + 1) * dt;
        memset(Fext, 0, sizeof(double) * total);
        if (c->exc_dir > 0) { /* :1823-1852 */
            double ag = or_path_eval(c, t);
            for (int i = 0; i < total; ++i) {
                double load = 0.0;
                if ((i % 3) + 1 == c->exc_dir) load += -c->mass[i] * ag;
                Fext[i] += load;
            }
        }
        for (int i = 0; i < n; ++i) P[i] = Fext[free_list[i]];
        int converged = 0, iters = 0;
        for (int it = 0; it < c->max_iters; ++it) {
            iters++;
            int unc = assemble(c, els, q, u, K, F, total);
            if (unc < 0) { status = -1; goto done; }
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) Kff[(size_t)i * n + j] = K[(size_t)free_list[i] * total + free_list[j]];
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
                R[i] = P[i] - F[free_list[i]] - 0.0 - damping_force - inertia_force;
            }
            double norm = sqrt(or_dot_w(R, R, n)); /* :2357-2371 */
            if (norm <= c->tol) {
                converged = 1;
                break;
            }
            /* :2400-2411: K_eff = K + a1 C + a0 M restricted to the band */
            for (int row = 0; row < n; ++row) {
                int c0 = row - hbw < 0 ? 0 : row - hbw;
                int c1 = row + hbw > n - 1 ? n - 1 : row + hbw;
                for (int col = 0; col < 2 * hbw + 1; ++col) bd.band[(size_t)row * (2 * hbw + 1) + col] = 0.0;
                for (int col = c0; col <= c1; ++col)
                    bd.band[band_slot(hbw, row, col)] = Kff[(size_t)row * n + col];
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
            if (!band_factor(&bd)) { status = -2; goto done; }
            band_solve(&bd, R, du);
            for (int i = 0; i < n; ++i) uf[i] += du[i];
            for (int i = 0; i < n; ++i) u[free_list[i]] = uf[i];
        }
        out->iters[step] = iters;
        if (!converged) {
            status = step + 1;
            goto done;
        }
        for (int e = 0; e < c->n_elem; ++e)
            for (int ip = 0; ip < els[e].nip; ++ip) or_sec2d_commit(els[e].sec, els[e].first_inst + ip);
        sync_basic(c, q, u);
        double omg = 1.0 - gamma; /* :594-647 */
        for (int i = 0; i < n; ++i) {
            double anext = a0 * (uf[i] - un[i]) - a2 * vn[i] - a3 * an[i];
            af[i] = anext;
            vf[i] = vn[i] + dt * (omg * an[i] + gamma * anext);
        }
        for (int i = 0; i < n; ++i) u[free_list[i]] = uf[i];
        if (out->u_hist) memcpy(out->u_hist + (size_t)step * total, u, sizeof(double) * total);
        out->steps_done = step + 1;
    }
done:
    out->section_evals = stats.section_evals;
    out->local_iters = stats.local_iters;
    out->subdivisions = stats.subdivisions;
    out->fallbacks = stats.fallbacks;
    out->dense_fallbacks = bd.dense_fallbacks;
    free(free_idx); free(free_list); free(els); free(K); free(F); free(Kff); free(Kinit); free(Cff); free(Mff);
    free(u); free(uf); free(vf); free(af); free(un); free(vn); free(an); free(vt); free(at); free(P); free(R);
    free(du); free(Fext); free(bd.band); free(bd.dense); free(bd.piv); free(bd.work);
    return status;
}
