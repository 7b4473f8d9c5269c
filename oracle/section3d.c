/*
 * FiberSection3d of the oracle (test infrastructure only).
 * Restates src/mojo/sections/fiber3d.mojo: definition build (:157-186), the
 * W-chunked elastic response with the uniform-modulus k12/k13 rule
 * (:189-256), the AoS per-fiber nonlinear loop through the dispatching
 * uniaxial_set_trial_strain (:769-836) and commit/revert (:873-926).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "strut_oracle.h"

struct or_sec3d {
    int fiber_count;
    double y_bar, z_bar;
    or_mat_def *mats;
    int ne, nn;
    double *el_y, *el_z, *el_a, *el_E;
    double *nl_y, *nl_z, *nl_a;
    int *nl_def;
    int *fiber_kind, *fiber_sub;
    int n_inst;
    or_mat_state *rt; /* instances * fiber_count, elastic first then nonlinear */
};

static double tree8(const double *v) {
    double w0 = v[0] + v[4], w1 = v[1] + v[5], w2 = v[2] + v[6], w3 = v[3] + v[7];
    double x0 = w0 + w2, x1 = w1 + w3;
    return x0 + x1;
}

/* fiber3d.mojo:157-186; centroid sums as in :418-470 */
or_sec3d *or_sec3d_create(int n_fibers, const double *y, const double *z, const double *area,
                          const int *mat_index, int n_mats, const int *mat_type,
                          const double *mat_params) {
    if (n_fibers <= 0) return NULL;
    or_sec3d *s = (or_sec3d *)calloc(1, sizeof(or_sec3d));
    s->fiber_count = n_fibers;
    s->mats = (or_mat_def *)calloc(n_mats, sizeof(or_mat_def));
    for (int m = 0; m < n_mats; ++m) {
        s->mats[m].type = mat_type[m];
        for (int k = 0; k < 11; ++k) s->mats[m].p[k] = mat_params[m * 11 + k];
    }
    double area_sum = 0.0, qy_sum = 0.0, qz_sum = 0.0;
    for (int i = 0; i < n_fibers; ++i) {
        area_sum += area[i];
        qz_sum += area[i] * y[i];
        qy_sum += area[i] * z[i];
    }
    s->y_bar = qz_sum / area_sum;
    s->z_bar = qy_sum / area_sum;
    size_t n = (size_t)n_fibers;
    s->el_y = malloc(sizeof(double) * n); s->el_z = malloc(sizeof(double) * n);
    s->el_a = malloc(sizeof(double) * n); s->el_E = malloc(sizeof(double) * n);
    s->nl_y = malloc(sizeof(double) * n); s->nl_z = malloc(sizeof(double) * n);
    s->nl_a = malloc(sizeof(double) * n); s->nl_def = malloc(sizeof(int) * n);
    s->fiber_kind = malloc(sizeof(int) * n); s->fiber_sub = malloc(sizeof(int) * n);
    for (int i = 0; i < n_fibers; ++i) {
        const or_mat_def *md = &s->mats[mat_index[i]];
        double y_rel = y[i] - s->y_bar, z_rel = z[i] - s->z_bar;
        if (md->type == OR_ELASTIC) {
            s->fiber_kind[i] = 0; s->fiber_sub[i] = s->ne;
            s->el_y[s->ne] = y_rel; s->el_z[s->ne] = z_rel; s->el_a[s->ne] = area[i];
            s->el_E[s->ne] = or_mat_initial_tangent(md);
            s->ne++;
        } else {
            s->fiber_kind[i] = 1; s->fiber_sub[i] = s->nn;
            s->nl_y[s->nn] = y_rel; s->nl_z[s->nn] = z_rel; s->nl_a[s->nn] = area[i];
            s->nl_def[s->nn] = mat_index[i];
            s->nn++;
        }
    }
    return s;
}

void or_sec3d_free(or_sec3d *s) {
    if (!s) return;
    free(s->mats); free(s->el_y); free(s->el_z); free(s->el_a); free(s->el_E);
    free(s->nl_y); free(s->nl_z); free(s->nl_a); free(s->nl_def);
    free(s->fiber_kind); free(s->fiber_sub); free(s->rt);
    free(s);
}

void or_sec3d_centroid(const or_sec3d *s, double *yz) {
    yz[0] = s->y_bar;
    yz[1] = s->z_bar;
}

/* fiber3d.mojo:715-766 (init_states): elastic states then nonlinear, each
 * UniMaterialState(mat_def). */
int or_sec3d_alloc_instances(or_sec3d *s, int n) {
    int old = s->n_inst, tot = old + n;
    s->rt = realloc(s->rt, sizeof(or_mat_state) * (size_t)tot * s->fiber_count);
    for (int inst = old; inst < tot; ++inst) {
        or_mat_state *base = s->rt + (size_t)inst * s->fiber_count;
        int k = 0;
        for (int i = 0; i < s->fiber_count; ++i)
            if (s->fiber_kind[i] == 0) {
                /* elastic fiber state (unused by the response) */
                or_mat_def e = {OR_ELASTIC, {s->el_E[k], 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
                or_mat_state_init(&e, &base[k]);
                k++;
            }
        for (int j = 0; j < s->nn; ++j) or_mat_state_init(&s->mats[s->nl_def[j]], &base[s->ne + j]);
    }
    s->n_inst = tot;
    return old;
}

/* fiber3d.mojo:769-836 */
int or_sec3d_set_trial(or_sec3d *s, int inst, double eps0, double kappa_y, double kappa_z,
                       double *out9) {
    if (inst < 0 || inst >= s->n_inst) return -1;
    double r[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int i = 0;
    for (; i + OR_SIMD_W <= s->ne; i += OR_SIMD_W) {
        double v[9][OR_SIMD_W];
        for (int l = 0; l < OR_SIMD_W; ++l) {
            double yv = s->el_y[i + l], zv = s->el_z[i + l];
            double ks = s->el_E[i + l] * s->el_a[i + l];
            double fs = ks * (eps0 + zv * kappa_y - yv * kappa_z);
            v[0][l] = fs; v[1][l] = fs * zv; v[2][l] = -fs * yv;
            v[3][l] = ks; v[4][l] = ks * zv; v[5][l] = -ks * yv;
            v[6][l] = ks * zv * zv; v[7][l] = -ks * zv * yv; v[8][l] = ks * yv * yv;
        }
        for (int q = 0; q < 9; ++q) r[q] += tree8(v[q]);
    }
    for (; i < s->ne; ++i) {
        double yv = s->el_y[i], zv = s->el_z[i];
        double ks = s->el_E[i] * s->el_a[i];
        double fs = ks * (eps0 + zv * kappa_y - yv * kappa_z);
        r[0] += fs; r[1] += fs * zv; r[2] += -fs * yv;
        r[3] += ks; r[4] += ks * zv; r[5] += -ks * yv;
        r[6] += ks * zv * zv; r[7] += -ks * zv * yv; r[8] += ks * yv * yv;
    }
    if (s->ne > 0) {
        int uniform = 1;
        for (int j = 1; j < s->ne; ++j)
            if (s->el_E[j] != s->el_E[0]) { uniform = 0; break; }
        if (uniform) { r[4] = 0.0; r[5] = 0.0; }
    }
    or_mat_state *base = s->rt + (size_t)inst * s->fiber_count + s->ne;
    for (int j = 0; j < s->nn; ++j) {
        double yv = s->nl_y[j], zv = s->nl_z[j];
        double eps = eps0 + zv * kappa_y - yv * kappa_z;
        or_mat_state *st = &base[j];
        or_uniaxial_set_trial(&s->mats[s->nl_def[j]], st, eps);
        double a = s->nl_a[j];
        double fs = st->sig_t * a, ks = st->tangent_t * a;
        r[0] += fs; r[1] += fs * zv; r[2] += -fs * yv;
        r[3] += ks; r[4] += ks * zv; r[5] += -ks * yv;
        r[6] += ks * zv * zv; r[7] += -ks * zv * yv; r[8] += ks * yv * yv;
    }
    for (int q = 0; q < 9; ++q) out9[q] = r[q];
    return 0;
}

/* fiber3d.mojo:873-926 */
void or_sec3d_commit(or_sec3d *s, int inst) {
    for (int i = 0; i < s->fiber_count; ++i) or_uniaxial_commit(&s->rt[(size_t)inst * s->fiber_count + i]);
}

void or_sec3d_revert(or_sec3d *s, int inst) {
    for (int i = 0; i < s->fiber_count; ++i) or_uniaxial_revert(&s->rt[(size_t)inst * s->fiber_count + i]);
}

int or_sec3d_fiber_trial(const or_sec3d *s, int inst, int fiber, double *out3) {
    if (inst < 0 || inst >= s->n_inst || fiber < 0 || fiber >= s->fiber_count) return -1;
    if (s->fiber_kind[fiber] == 0) return 1;
    const or_mat_state *st = &s->rt[(size_t)inst * s->fiber_count + s->ne + s->fiber_sub[fiber]];
    out3[0] = st->eps_t;
    out3[1] = st->sig_t;
    out3[2] = st->tangent_t;
    return 0;
}

/* force_beam_column3d.mojo:679-727 (_invert_3x3_values) applied as in
 * _fiber_section3d_response_flexibility (:1093-1117): returns 1 and
 * f = {f11, f12, f13, f22, f23, f33} when |det| > 1e-40. */
int or_sec3d_flexibility(const double *k6, double *f6) {
    double a = k6[0], b = k6[1], c = k6[2], d = k6[1], e = k6[3], f = k6[4];
    double g = k6[2], h = k6[4], i = k6[5];
    double co00 = e * i - f * h;
    double co01 = -(d * i - f * g);
    double co02 = d * h - e * g;
    double co10 = -(b * i - c * h);
    double co11 = a * i - c * g;
    double co20 = b * f - c * e;
    double co21 = -(a * f - c * d);
    double co22 = a * e - b * d;
    double det = a * co00 + b * co01 + c * co02;
    if (fabs(det) <= 1.0e-40) {
        for (int q = 0; q < 6; ++q) f6[q] = 0.0;
        return 0;
    }
    double inv_det = 1.0 / det;
    f6[0] = co00 * inv_det;
    f6[1] = co10 * inv_det;
    f6[2] = co20 * inv_det;
    f6[3] = co11 * inv_det;
    f6[4] = co21 * inv_det;
    f6[5] = co22 * inv_det;
    return 1;
}

/* Timing helper for the CPU baseline: evaluate n points (instances inst[i]
 * at the given deformations) in order, then commit those instances. */
/* Timing helper: set_trial on n instances, then commit them when commit != 0. */
void or_sec3d_eval_many(or_sec3d *s, int n, const int *inst, const double *eps0, const double *ky,
                        const double *kz, int commit) {
    double out[9];
    for (int i = 0; i < n; ++i) or_sec3d_set_trial(s, inst[i], eps0[i], ky[i], kz[i], out);
    if (commit)
        for (int i = 0; i < n; ++i) or_sec3d_commit(s, inst[i]);
}
