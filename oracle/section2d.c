/*
 * FiberSection2d of the oracle (test infrastructure only).
 * Restates src/mojo/sections/fiber2d.mojo: definition build with family
 * grouping and padding (:3167-3455), SoA runtime allocation (:961-1006),
 * set_trial_from_offset (:3982-4024) with the W-wide family loops
 * (Steel02 :2301-2516, Concrete02 :1785-1972, elastic :3459-3501, scalar
 * Steel01/Concrete01 :1371-1431), batch evaluation with the 2x2 flexibility
 * (:3042-3086) and commit/revert (:4059-4252).
 *
 * Each padded family slot keeps a full or_mat_state; the laws only read the
 * fields the reference's SoA arrays hold, so the storage shape does not
 * change any value.  Padding slots start all-zero, as the reference's
 * List.resize(…, 0.0) leaves them (:868-911).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "strut_oracle.h"

typedef struct {
    int count;          /* real fibers in the family */
    int padded;         /* padded slot count (instance stride) */
    double *y, *area;   /* per padded slot; pads carry y = area = 0 */
    int *def;           /* per padded slot material def index */
    int ngroups;
    int *goff, *gpad, *gdef;
    int *pos_by_nl;     /* nonlinear index -> family slot, -1 otherwise */
    or_mat_state *rt;   /* instances * padded */
} family2d;

struct or_sec2d {
    int fiber_count;
    double y_bar;
    int n_mats;
    or_mat_def *mats;
    int ne;
    double *el_y, *el_a, *el_E;
    int nn;
    double *nl_y, *nl_a;
    int *nl_def;
    int n_s01, *s01_idx;
    int n_c01, *c01_idx;
    family2d s2, c2;
    int init_flex_valid;
    double if00, if01, if11;
    int n_inst;
    or_mat_state *rt; /* instances * fiber_count: slot elastic_count + i for nonlinear i */
    int *fiber_kind;  /* input fiber -> 0 elastic, 1 nonlinear */
    int *fiber_sub;   /* input fiber -> elastic or nonlinear index */
};

/* fiber2d.mojo:31-35 */
static int padded_count(int count) {
    if (count <= 0) return 0;
    return ((count + OR_SIMD_W - 1) / OR_SIMD_W) * OR_SIMD_W;
}

/* halving-tree SIMD reduce_add of a W = 8 vector */
static double tree8(const double *v) {
    double w0 = v[0] + v[4], w1 = v[1] + v[5], w2 = v[2] + v[6], w3 = v[3] + v[7];
    double x0 = w0 + w2, x1 = w1 + w3;
    return x0 + x1;
}

/* fiber2d.mojo:3146-3228 (_fiber_section2d_rebuild_grouped_family): regroup a
 * family by material definition in order of first appearance; pad each
 * group to a multiple of W. */
static void build_family(family2d *f, int fam_type, const or_sec2d *s) {
    int n = 0;
    for (int i = 0; i < s->nn; ++i)
        if (s->mats[s->nl_def[i]].type == fam_type) n++;
    f->count = n;
    f->pos_by_nl = (int *)malloc(sizeof(int) * (s->nn > 0 ? s->nn : 1));
    for (int i = 0; i < s->nn; ++i) f->pos_by_nl[i] = -1;
    int *uniq = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int nu = 0;
    for (int i = 0; i < s->nn; ++i) {
        if (s->mats[s->nl_def[i]].type != fam_type) continue;
        int d = s->nl_def[i], seen = 0;
        for (int j = 0; j < nu; ++j)
            if (uniq[j] == d) { seen = 1; break; }
        if (!seen) uniq[nu++] = d;
    }
    int cap = n + OR_SIMD_W * (nu + 1);
    f->y = (double *)calloc(cap, sizeof(double));
    f->area = (double *)calloc(cap, sizeof(double));
    f->def = (int *)calloc(cap, sizeof(int));
    f->goff = (int *)malloc(sizeof(int) * (nu + 1));
    f->gpad = (int *)malloc(sizeof(int) * (nu + 1));
    f->gdef = (int *)malloc(sizeof(int) * (nu + 1));
    f->ngroups = 0;
    int len = 0;
    for (int g = 0; g < nu; ++g) {
        int d = uniq[g], off = len, cnt = 0;
        for (int i = 0; i < s->nn; ++i) {
            if (s->mats[s->nl_def[i]].type != fam_type || s->nl_def[i] != d) continue;
            f->y[len] = s->nl_y[i];
            f->area[len] = s->nl_a[i];
            f->def[len] = d;
            f->pos_by_nl[i] = len;
            len++;
            cnt++;
        }
        if (cnt <= 0) continue;
        int pad = padded_count(cnt);
        f->goff[f->ngroups] = off;
        f->gpad[f->ngroups] = pad;
        f->gdef[f->ngroups] = d;
        f->ngroups++;
        for (int k = cnt; k < pad; ++k) {
            f->y[len] = 0.0;
            f->area[len] = 0.0;
            f->def[len] = d;
            len++;
        }
    }
    f->padded = len;
    f->rt = NULL;
    free(uniq);
}

/* fiber2d.mojo:3323-3455 (_build_fiber_section2d_def); y_bar follows
 * append_fiber_section2d_from_input (:3881-3943): running area and A*y sums
 * in fiber order. */
or_sec2d *or_sec2d_create(int n_fibers, const double *y, const double *area, const int *mat_index,
                          int n_mats, const int *mat_type, const double *mat_params) {
    if (n_fibers <= 0) return NULL;
    or_sec2d *s = (or_sec2d *)calloc(1, sizeof(or_sec2d));
    s->fiber_count = n_fibers;
    s->n_mats = n_mats;
    s->mats = (or_mat_def *)calloc(n_mats, sizeof(or_mat_def));
    for (int m = 0; m < n_mats; ++m) {
        s->mats[m].type = mat_type[m];
        for (int k = 0; k < 11; ++k) s->mats[m].p[k] = mat_params[m * 11 + k];
    }
    double area_sum = 0.0, qz_sum = 0.0;
    for (int i = 0; i < n_fibers; ++i) {
        area_sum += area[i];
        qz_sum += area[i] * y[i];
    }
    s->y_bar = qz_sum / area_sum;
    s->el_y = (double *)malloc(sizeof(double) * n_fibers);
    s->el_a = (double *)malloc(sizeof(double) * n_fibers);
    s->el_E = (double *)malloc(sizeof(double) * n_fibers);
    s->nl_y = (double *)malloc(sizeof(double) * n_fibers);
    s->nl_a = (double *)malloc(sizeof(double) * n_fibers);
    s->nl_def = (int *)malloc(sizeof(int) * n_fibers);
    s->s01_idx = (int *)malloc(sizeof(int) * n_fibers);
    s->c01_idx = (int *)malloc(sizeof(int) * n_fibers);
    s->fiber_kind = (int *)malloc(sizeof(int) * n_fibers);
    s->fiber_sub = (int *)malloc(sizeof(int) * n_fibers);
    double k11 = 0.0, k12 = 0.0, k22 = 0.0;
    for (int i = 0; i < n_fibers; ++i) {
        const or_mat_def *md = &s->mats[mat_index[i]];
        double y_rel = y[i] - s->y_bar;
        double ks = or_mat_initial_tangent(md) * area[i];
        k11 += ks;
        k12 += -ks * y_rel;
        k22 += ks * y_rel * y_rel;
        if (md->type == OR_ELASTIC) {
            s->fiber_kind[i] = 0;
            s->fiber_sub[i] = s->ne;
            s->el_y[s->ne] = y_rel;
            s->el_a[s->ne] = area[i];
            s->el_E[s->ne] = or_mat_initial_tangent(md);
            s->ne++;
        } else {
            int nl = s->nn++;
            s->fiber_kind[i] = 1;
            s->fiber_sub[i] = nl;
            s->nl_y[nl] = y_rel;
            s->nl_a[nl] = area[i];
            s->nl_def[nl] = mat_index[i];
            if (md->type == OR_STEEL01) s->s01_idx[s->n_s01++] = nl;
            else if (md->type == OR_CONCRETE01) s->c01_idx[s->n_c01++] = nl;
        }
    }
    build_family(&s->s2, OR_STEEL02, s);
    build_family(&s->c2, OR_CONCRETE02, s);
    double det = k11 * k22 - k12 * k12;
    if (fabs(det) > 1.0e-40) {
        double inv_det = 1.0 / det;
        s->init_flex_valid = 1;
        s->if00 = k22 * inv_det;
        s->if01 = -k12 * inv_det;
        s->if11 = k11 * inv_det;
    }
    return s;
}

static void free_family(family2d *f) {
    free(f->y); free(f->area); free(f->def); free(f->goff); free(f->gpad); free(f->gdef);
    free(f->pos_by_nl); free(f->rt);
}

void or_sec2d_free(or_sec2d *s) {
    if (!s) return;
    free(s->mats); free(s->el_y); free(s->el_a); free(s->el_E); free(s->nl_y); free(s->nl_a);
    free(s->nl_def); free(s->s01_idx); free(s->c01_idx); free(s->fiber_kind); free(s->fiber_sub);
    free_family(&s->s2); free_family(&s->c2); free(s->rt);
    free(s);
}

/* fiber2d.mojo:961-1006 */
int or_sec2d_alloc_instances(or_sec2d *s, int n) {
    int old = s->n_inst, tot = old + n;
    s->rt = (or_mat_state *)realloc(s->rt, sizeof(or_mat_state) * (size_t)tot * s->fiber_count);
    memset(s->rt + (size_t)old * s->fiber_count, 0, sizeof(or_mat_state) * (size_t)n * s->fiber_count);
    family2d *fams[2] = {&s->s2, &s->c2};
    for (int k = 0; k < 2; ++k) {
        family2d *f = fams[k];
        if (f->padded == 0) continue;
        f->rt = (or_mat_state *)realloc(f->rt, sizeof(or_mat_state) * (size_t)tot * f->padded);
        memset(f->rt + (size_t)old * f->padded, 0, sizeof(or_mat_state) * (size_t)n * f->padded);
    }
    for (int inst = old; inst < tot; ++inst) {
        for (int i = 0; i < s->nn; ++i) {
            const or_mat_def *md = &s->mats[s->nl_def[i]];
            if (md->type == OR_STEEL02)
                or_mat_state_init(md, &s->s2.rt[(size_t)inst * s->s2.padded + s->s2.pos_by_nl[i]]);
            else if (md->type == OR_CONCRETE02)
                or_mat_state_init(md, &s->c2.rt[(size_t)inst * s->c2.padded + s->c2.pos_by_nl[i]]);
            else
                or_mat_state_init(md, &s->rt[(size_t)inst * s->fiber_count + s->ne + i]);
        }
    }
    s->n_inst = tot;
    return old;
}

double or_sec2d_ybar(const or_sec2d *s) { return s->y_bar; }
int or_sec2d_fiber_count(const or_sec2d *s) { return s->fiber_count; }
int or_sec2d_nonlinear_only(const or_sec2d *s) { return s->ne == 0; }
/* force_beam_column2d.mojo:327-331 (_fiber_section2d_all_materials_elastic) */
int or_sec2d_all_elastic(const or_sec2d *s) { return s->nn == 0; }

int or_sec2d_initial_flex(const or_sec2d *s, double *f3) {
    f3[0] = s->if00;
    f3[1] = s->if01;
    f3[2] = s->if11;
    return s->init_flex_valid;
}

/* accumulate one W-chunk of per-lane (sig, tangent) into the five resultants
 * (fiber2d.mojo:2065-2082, :1584-1601) */
static void accumulate_chunk(const double *y, const double *a, const double *sig, const double *tan,
                             double *r) {
    double fs[OR_SIMD_W], mz[OR_SIMD_W], ks[OR_SIMD_W], k12[OR_SIMD_W], k22[OR_SIMD_W];
    for (int l = 0; l < OR_SIMD_W; ++l) {
        fs[l] = sig[l] * a[l];
        ks[l] = tan[l] * a[l];
        mz[l] = -fs[l] * y[l];
        k12[l] = -ks[l] * y[l];
        k22[l] = ks[l] * y[l] * y[l];
    }
    r[0] += tree8(fs);
    r[1] += tree8(mz);
    r[2] += tree8(ks);
    r[3] += tree8(k12);
    r[4] += tree8(k22);
}

/* fiber2d.mojo:2301-2516 / :1785-1972: one homogeneous group, W lanes at a
 * time; each lane follows the scalar law (the masked selects reproduce it
 * lane by lane). */
static void apply_family(const or_sec2d *s, family2d *f, int inst, int is_steel, double eps0,
                         double kappa, double *r) {
    for (int g = 0; g < f->ngroups; ++g) {
        const or_mat_def *md = &s->mats[f->gdef[g]];
        for (int i = 0; i < f->gpad[g]; i += OR_SIMD_W) {
            int base = f->goff[g] + i;
            double sig[OR_SIMD_W], tan[OR_SIMD_W];
            for (int l = 0; l < OR_SIMD_W; ++l) {
                double strain = eps0 - f->y[base + l] * kappa;
                or_mat_state *st = &f->rt[(size_t)inst * f->padded + base + l];
                if (is_steel) or_steel02_core(md, st, strain);
                else or_concrete02_core(md, st, strain);
                sig[l] = st->sig_t;
                tan[l] = st->tangent_t;
            }
            accumulate_chunk(&f->y[base], &f->area[base], sig, tan, r);
        }
    }
}

/* fiber2d.mojo:3982-4024 */
int or_sec2d_set_trial(or_sec2d *s, int inst, double eps0, double kappa, double *out5) {
    if (inst < 0 || inst >= s->n_inst) return -1;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    /* elastic part, :3459-3501 */
    int i = 0;
    for (; i + OR_SIMD_W <= s->ne; i += OR_SIMD_W) {
        double fs[OR_SIMD_W], mz[OR_SIMD_W], ks[OR_SIMD_W], k12[OR_SIMD_W], k22[OR_SIMD_W];
        for (int l = 0; l < OR_SIMD_W; ++l) {
            double yv = s->el_y[i + l];
            ks[l] = s->el_E[i + l] * s->el_a[i + l];
            fs[l] = ks[l] * (eps0 - yv * kappa);
            mz[l] = -fs[l] * yv;
            k12[l] = -ks[l] * yv;
            k22[l] = ks[l] * yv * yv;
        }
        r[0] += tree8(fs); r[1] += tree8(mz); r[2] += tree8(ks); r[3] += tree8(k12); r[4] += tree8(k22);
    }
    for (; i < s->ne; ++i) {
        double yv = s->el_y[i];
        double ks = s->el_E[i] * s->el_a[i];
        double fs = ks * (eps0 - yv * kappa);
        r[0] += fs; r[1] += -fs * yv; r[2] += ks; r[3] += -ks * yv; r[4] += ks * yv * yv;
    }
    if (s->ne > 0) {
        int uniform = 1;
        for (int j = 1; j < s->ne; ++j)
            if (s->el_E[j] != s->el_E[0]) { uniform = 0; break; }
        if (uniform) r[3] = 0.0;
    }
    apply_family(s, &s->s2, inst, 1, eps0, kappa, r);
    apply_family(s, &s->c2, inst, 0, eps0, kappa, r);
    /* :1371-1431 scalar Steel01 then Concrete01 */
    int lists[2] = {s->n_s01, s->n_c01};
    const int *idx[2] = {s->s01_idx, s->c01_idx};
    for (int k = 0; k < 2; ++k) {
        for (int j = 0; j < lists[k]; ++j) {
            int nl = idx[k][j];
            double yv = s->nl_y[nl], a = s->nl_a[nl];
            or_mat_state *st = &s->rt[(size_t)inst * s->fiber_count + s->ne + nl];
            const or_mat_def *md = &s->mats[s->nl_def[nl]];
            double strain = eps0 - yv * kappa;
            if (k == 0) or_steel01_core(md, st, strain);
            else or_concrete01_core(md, st, strain);
            double fs = st->sig_t * a, ks = st->tangent_t * a;
            r[0] += fs; r[1] += -fs * yv; r[2] += ks; r[3] += -ks * yv; r[4] += ks * yv * yv;
        }
    }
    for (int q = 0; q < 5; ++q) out5[q] = r[q];
    return 0;
}

/* fiber2d.mojo:3042-3086: evaluates points in order and stops at the first
 * singular section flexibility (|det| <= 1e-40). */
int or_sec2d_batch_eval(or_sec2d *s, int n, const int *inst, const double *eps0, const double *kappa,
                        double *out8, int *valid) {
    for (int p = 0; p < n; ++p) valid[p] = 0;
    for (int p = 0; p < n; ++p) {
        double r[5];
        if (or_sec2d_set_trial(s, inst[p], eps0[p], kappa[p], r) != 0) return -1;
        double *o = out8 + 8 * p;
        for (int q = 0; q < 5; ++q) o[q] = r[q];
        double det = r[2] * r[4] - r[3] * r[3];
        if (fabs(det) <= 1.0e-40) {
            o[5] = o[6] = o[7] = 0.0;
            return 1;
        }
        double inv_det = 1.0 / det;
        o[5] = r[4] * inv_det;
        o[6] = -r[3] * inv_det;
        o[7] = r[2] * inv_det;
        valid[p] = 1;
    }
    return 0;
}

/* fiber2d.mojo:4059-4092 / :4095-4128 */
static void commit_or_revert(or_sec2d *s, int inst, int commit) {
    family2d *fams[2] = {&s->s2, &s->c2};
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < fams[k]->padded; ++j) {
            or_mat_state *st = &fams[k]->rt[(size_t)inst * fams[k]->padded + j];
            if (commit) or_uniaxial_commit(st); else or_uniaxial_revert(st);
        }
    for (int i = 0; i < s->nn; ++i) {
        int t = s->mats[s->nl_def[i]].type;
        if (t == OR_STEEL02 || t == OR_CONCRETE02) continue;
        or_mat_state *st = &s->rt[(size_t)inst * s->fiber_count + s->ne + i];
        if (commit) or_uniaxial_commit(st); else or_uniaxial_revert(st);
    }
}

void or_sec2d_commit(or_sec2d *s, int inst) { commit_or_revert(s, inst, 1); }
void or_sec2d_revert(or_sec2d *s, int inst) { commit_or_revert(s, inst, 0); }
void or_sec2d_commit_all(or_sec2d *s) { for (int i = 0; i < s->n_inst; ++i) commit_or_revert(s, i, 1); }
void or_sec2d_revert_all(or_sec2d *s) { for (int i = 0; i < s->n_inst; ++i) commit_or_revert(s, i, 0); }

/* trial (eps, sig, tangent) of input-order fiber `fiber`.  Elastic fibers
 * keep no runtime state in the reference's 2D SoA path; they return 1. */
int or_sec2d_fiber_trial(const or_sec2d *s, int inst, int fiber, double *out3) {
    if (inst < 0 || inst >= s->n_inst || fiber < 0 || fiber >= s->fiber_count) return -1;
    if (s->fiber_kind[fiber] == 0) return 1;
    int nl = s->fiber_sub[fiber];
    int t = s->mats[s->nl_def[nl]].type;
    const or_mat_state *st;
    if (t == OR_STEEL02) st = &s->s2.rt[(size_t)inst * s->s2.padded + s->s2.pos_by_nl[nl]];
    else if (t == OR_CONCRETE02) st = &s->c2.rt[(size_t)inst * s->c2.padded + s->c2.pos_by_nl[nl]];
    else st = &s->rt[(size_t)inst * s->fiber_count + s->ne + nl];
    out3[0] = st->eps_t;
    out3[1] = st->sig_t;
    out3[2] = st->tangent_t;
    return 0;
}
