/*
 * Uniaxial material laws of the oracle (test infrastructure only).
 * Restates src/mojo/materials/uniaxial/{core,ops,steel02,concrete02,concrete01}.mojo.
 */
#include <math.h>
#include <string.h>
#include "strut_oracle.h"

static double or_abs(double x) { return x < 0.0 ? -x : x; }          /* core.mojo:6-9 */
static double or_sign(double x) { return x < 0.0 ? -1.0 : 1.0; }     /* core.mojo:12-15 */
static double or_pow_abs(double x, double p) { return pow(or_abs(x), p); } /* core.mojo:18-19 */

/* core.mojo:270-282 */
double or_mat_initial_tangent(const or_mat_def *d) {
    switch (d->type) {
    case OR_ELASTIC: return d->p[0];
    case OR_STEEL01: return d->p[1];
    case OR_CONCRETE01: return (2.0 * d->p[0]) / d->p[1];
    case OR_STEEL02: return d->p[1];
    case OR_CONCRETE02: return (2.0 * d->p[0]) / d->p[1];
    default: return 0.0;
    }
}

/* core.mojo:203-263 (UniMaterialState(mat_def)) */
void or_mat_state_init(const or_mat_def *d, or_mat_state *s) {
    memset(s, 0, sizeof(*s));
    double tangent = or_mat_initial_tangent(d);
    s->tangent_c = tangent;
    s->tangent_t = tangent;
    if (d->type == OR_CONCRETE01) {
        s->unload_slope_c = tangent;
        s->unload_slope_t = tangent;
    }
    if (d->type == OR_STEEL02) {
        double Fy = d->p[0], E0 = d->p[1];
        double epsy = Fy / E0;
        s->s2_epsmax_c = epsy;
        s->s2_epsmax_t = epsy;
        s->s2_epsmin_c = -epsy;
        s->s2_epsmin_t = -epsy;
        double sigini = d->p[10];
        if (sigini != 0.0) {
            double epsini = sigini / E0;
            s->eps_c = epsini;
            s->eps_t = epsini;
            s->sig_c = sigini;
            s->sig_t = sigini;
        }
    }
}

/* steel02.mojo:3-114 */
static void steel02_set_trial(const or_mat_def *d, or_mat_state *s, double trial_eps) {
    double Fy = d->p[0], E0 = d->p[1], b = d->p[2], R0 = d->p[3], cR1 = d->p[4], cR2 = d->p[5];
    double a1 = d->p[6], a2 = d->p[7], a3 = d->p[8], a4 = d->p[9], sigini = d->p[10];
    double Esh = b * E0;
    double epsy = Fy / E0;
    double eps = trial_eps;
    if (sigini != 0.0) {
        double epsini = sigini / E0;
        eps = trial_eps + epsini;
    }
    s->eps_t = eps;
    double deps = eps - s->eps_c;
    double epsmax = s->s2_epsmax_c, epsmin = s->s2_epsmin_c, epspl = s->s2_epspl_c;
    double epss0 = s->s2_epss0_c, sigs0 = s->s2_sigs0_c, epsr = s->s2_epsr_c, sigr = s->s2_sigr_c;
    long kon = s->s2_kon_c;
    if (kon == 0 || kon == 3) {
        if (or_abs(deps) < 2.220446049250313e-15) {
            s->tangent_t = E0;
            s->sig_t = sigini;
            s->s2_kon_t = 3;
            s->s2_epsmin_t = epsmin;
            s->s2_epsmax_t = epsmax;
            s->s2_epspl_t = epspl;
            s->s2_epss0_t = epss0;
            s->s2_sigs0_t = sigs0;
            s->s2_epsr_t = epsr;
            s->s2_sigr_t = sigr;
            s->eps_p_t = s->eps_p_c;
            s->alpha_t = s->alpha_c;
            return;
        }
        epsmax = epsy;
        epsmin = -epsy;
        if (deps < 0.0) {
            kon = 2;
            epss0 = epsmin;
            sigs0 = -Fy;
            epspl = epsmin;
        } else {
            kon = 1;
            epss0 = epsmax;
            sigs0 = Fy;
            epspl = epsmax;
        }
    }
    if (kon == 2 && deps > 0.0) {
        kon = 1;
        epsr = s->eps_c;
        sigr = s->sig_c;
        if (s->eps_c < epsmin) epsmin = s->eps_c;
        double d1 = (epsmax - epsmin) / (2.0 * (a4 * epsy));
        double shft = 1.0 + a3 * pow(d1, 0.8);
        epss0 = (Fy * shft - Esh * epsy * shft - sigr + E0 * epsr) / (E0 - Esh);
        sigs0 = Fy * shft + Esh * (epss0 - epsy * shft);
        epspl = epsmax;
    } else if (kon == 1 && deps < 0.0) {
        kon = 2;
        epsr = s->eps_c;
        sigr = s->sig_c;
        if (s->eps_c > epsmax) epsmax = s->eps_c;
        double d1 = (epsmax - epsmin) / (2.0 * (a2 * epsy));
        double shft = 1.0 + a1 * pow(d1, 0.8);
        epss0 = (-Fy * shft + Esh * epsy * shft - sigr + E0 * epsr) / (E0 - Esh);
        sigs0 = -Fy * shft + Esh * (epss0 + epsy * shft);
        epspl = epsmin;
    }
    double xi = or_abs((epspl - epss0) / epsy);
    double R = R0 * (1.0 - (cR1 * xi) / (cR2 + xi));
    double epsrat = (eps - epsr) / (epss0 - epsr);
    double dum1 = 1.0 + or_pow_abs(epsrat, R);
    double dum2 = pow(dum1, 1.0 / R);
    double sig = b * epsrat + (1.0 - b) * epsrat / dum2;
    sig = sig * (sigs0 - sigr) + sigr;
    double e = b + (1.0 - b) / (dum1 * dum2);
    e = e * (sigs0 - sigr) / (epss0 - epsr);
    s->sig_t = sig;
    s->tangent_t = e;
    s->eps_p_t = s->eps_p_c;
    s->alpha_t = s->alpha_c;
    s->s2_epsmin_t = epsmin;
    s->s2_epsmax_t = epsmax;
    s->s2_epspl_t = epspl;
    s->s2_epss0_t = epss0;
    s->s2_sigs0_t = sigs0;
    s->s2_epsr_t = epsr;
    s->s2_sigr_t = sigr;
    s->s2_kon_t = kon;
}

/* concrete02.mojo:3-14 */
static void c02_tens_envlp(double fc, double epsc0, double ft, double Ets, double strain,
                           double *sig, double *tan) {
    double Ec0 = (2.0 * fc) / epsc0;
    double eps0 = ft / Ec0;
    double epsu = ft * (1.0 / Ets + 1.0 / Ec0);
    if (strain <= eps0) {
        *sig = strain * Ec0;
        *tan = Ec0;
    } else if (strain <= epsu) {
        *sig = ft - Ets * (strain - eps0);
        *tan = -Ets;
    } else {
        *sig = 0.0;
        *tan = 1.0e-10;
    }
}

/* concrete02.mojo:17-31 */
static void c02_compr_envlp(double fc, double epsc0, double fcu, double epscu, double strain,
                            double *sig, double *tan) {
    double Ec0 = (2.0 * fc) / epsc0;
    double rat = strain / epsc0;
    if (strain >= epsc0) {
        *sig = fc * rat * (2.0 - rat);
        *tan = Ec0 * (1.0 - rat);
    } else if (strain > epscu) {
        *sig = (fcu - fc) * (strain - epsc0) / (epscu - epsc0) + fc;
        *tan = (fcu - fc) / (epscu - epsc0);
    } else {
        *sig = fcu;
        *tan = 1.0e-10;
    }
}

/* concrete02.mojo:34-109 */
static void concrete02_set_trial(const or_mat_def *d, or_mat_state *s, double strain) {
    double fc = d->p[0], epsc0 = d->p[1], fcu = d->p[2], epscu = d->p[3];
    double rat = d->p[4], ft = d->p[5], Ets = d->p[6];
    double Ec0 = (2.0 * fc) / epsc0;
    s->c2_ecmin_t = s->c2_ecmin_c;
    s->c2_dept_t = s->c2_dept_c;
    s->eps_t = strain;
    double deps = strain - s->eps_c;
    if (or_abs(deps) < 2.220446049250313e-16) {
        s->sig_t = s->sig_c;
        s->tangent_t = s->tangent_c;
        s->eps_p_t = s->eps_p_c;
        s->alpha_t = s->alpha_c;
        return;
    }
    if (strain < s->c2_ecmin_t) {
        double sg, tg;
        c02_compr_envlp(fc, epsc0, fcu, epscu, strain, &sg, &tg);
        s->sig_t = sg;
        s->tangent_t = tg;
        s->c2_ecmin_t = strain;
        s->eps_p_t = s->eps_p_c;
        s->alpha_t = s->alpha_c;
        return;
    }
    double epsr = (fcu - rat * Ec0 * epscu) / (Ec0 * (1.0 - rat));
    double sigmr = Ec0 * epsr;
    double sigmm, tdummy;
    c02_compr_envlp(fc, epsc0, fcu, epscu, s->c2_ecmin_t, &sigmm, &tdummy);
    double er = (sigmm - sigmr) / (s->c2_ecmin_t - epsr);
    double ept = s->c2_ecmin_t - sigmm / er;
    if (strain <= ept) {
        double sigmin = sigmm + er * (strain - s->c2_ecmin_t);
        double sigmax = 0.5 * er * (strain - ept);
        s->sig_t = s->sig_c + Ec0 * deps;
        s->tangent_t = Ec0;
        if (s->sig_t <= sigmin) {
            s->sig_t = sigmin;
            s->tangent_t = er;
        }
        if (s->sig_t >= sigmax) {
            s->sig_t = sigmax;
            s->tangent_t = 0.5 * er;
        }
        s->eps_p_t = s->eps_p_c;
        s->alpha_t = s->alpha_c;
        return;
    }
    double epn = ept + s->c2_dept_t;
    if (strain <= epn) {
        double sicn, tdum2;
        c02_tens_envlp(fc, epsc0, ft, Ets, s->c2_dept_t, &sicn, &tdum2);
        double e = Ec0;
        if (s->c2_dept_t != 0.0) e = sicn / s->c2_dept_t;
        s->tangent_t = e;
        s->sig_t = e * (strain - ept);
    } else {
        double sg, tg;
        c02_tens_envlp(fc, epsc0, ft, Ets, strain - ept, &sg, &tg);
        s->sig_t = sg;
        s->tangent_t = tg;
        s->c2_dept_t = strain - ept;
    }
    s->eps_p_t = s->eps_p_c;
    s->alpha_t = s->alpha_c;
}

/* concrete01.mojo:3-18 */
static void c01_envelope(double fpc, double epsc0, double fpcu, double epscu, double strain,
                         double *sig, double *tan) {
    if (strain > epsc0) {
        double eta = strain / epsc0;
        *sig = fpc * (2.0 * eta - eta * eta);
        double Ec0 = (2.0 * fpc) / epsc0;
        *tan = Ec0 * (1.0 - eta);
    } else if (strain > epscu) {
        double t = (fpc - fpcu) / (epsc0 - epscu);
        *tan = t;
        *sig = fpc + t * (strain - epsc0);
    } else {
        *sig = fpcu;
        *tan = 0.0;
    }
}

/* concrete01.mojo:20-45 */
static void c01_unload(double fpc, double epsc0, double epscu, double min_strain, double stress,
                       double *end_strain, double *unload_slope) {
    double temp_strain = min_strain;
    if (temp_strain < epscu) temp_strain = epscu;
    double eta = temp_strain / epsc0;
    double ratio = 0.707 * (eta - 2.0) + 0.834;
    if (eta < 2.0) ratio = 0.145 * eta * eta + 0.13 * eta;
    double es = ratio * epsc0;
    double temp1 = min_strain - es;
    double Ec0 = (2.0 * fpc) / epsc0;
    double temp2 = stress / Ec0;
    if (temp1 > -1.0e-14) {
        *end_strain = es;
        *unload_slope = Ec0;
    } else if (temp1 <= temp2) {
        *end_strain = min_strain - temp1;
        *unload_slope = stress / temp1;
    } else {
        *end_strain = min_strain - temp2;
        *unload_slope = Ec0;
    }
}

/* concrete01.mojo:48-70 */
static void c01_reload(double fpc, double epsc0, double fpcu, double epscu, double strain,
                       double *min_strain, double *end_strain, double *unload_slope,
                       double *sig, double *tan) {
    if (strain <= *min_strain) {
        *min_strain = strain;
        c01_envelope(fpc, epsc0, fpcu, epscu, strain, sig, tan);
        c01_unload(fpc, epsc0, epscu, *min_strain, *sig, end_strain, unload_slope);
    } else if (strain <= *end_strain) {
        *tan = *unload_slope;
        *sig = *tan * (strain - *end_strain);
    } else {
        *sig = 0.0;
        *tan = 0.0;
    }
}

/* ops.mojo:10-21 */
static void elastic_trial(const or_mat_def *d, or_mat_state *s, double eps) {
    if (eps == s->eps_t) return;
    s->eps_t = eps;
    double E = d->p[0];
    s->sig_t = E * eps;
    s->tangent_t = E;
    s->eps_p_t = s->eps_p_c;
    s->alpha_t = s->alpha_c;
}

/* ops.mojo:24-53 without the eps == eps_t early return, which is also
 * fiber2d.mojo:1020-1045 (_fiber_section2d_runtime_apply_steel01) */
void or_steel01_core(const or_mat_def *d, or_mat_state *s, double eps) {
    s->eps_t = eps;
    double Fy = d->p[0], E0 = d->p[1], b = d->p[2];
    double H = (b * E0) / (1.0 - b);
    double eps_p = s->eps_p_c, alpha = s->alpha_c;
    double sigma_trial = E0 * (eps - eps_p);
    double xi = sigma_trial - alpha;
    double f = or_abs(xi) - Fy;
    if (f <= 0.0) {
        s->sig_t = sigma_trial;
        s->tangent_t = E0;
        s->eps_p_t = eps_p;
        s->alpha_t = alpha;
        return;
    }
    double dg = f / (E0 + H);
    double sgn = or_sign(xi);
    s->eps_p_t = eps_p + dg * sgn;
    s->alpha_t = alpha + H * dg * sgn;
    s->sig_t = sigma_trial - E0 * dg * sgn;
    s->tangent_t = (E0 * H) / (E0 + H);
}

/* ops.mojo:56-114 without the eps == eps_t early return, which is also
 * fiber2d.mojo:1048-1108 (_fiber_section2d_runtime_apply_concrete01) */
void or_concrete01_core(const or_mat_def *d, or_mat_state *s, double eps) {
    s->eps_t = eps;
    double fpc = d->p[0], epsc0 = d->p[1], fpcu = d->p[2], epscu = d->p[3];
    double d_strain = eps - s->eps_c;
    if (or_abs(d_strain) < 1.0e-14) {
        s->sig_t = s->sig_c;
        s->tangent_t = s->tangent_c;
        s->eps_p_t = s->eps_p_c;
        s->alpha_t = s->alpha_c;
        s->min_strain_t = s->min_strain_c;
        s->end_strain_t = s->end_strain_c;
        s->unload_slope_t = s->unload_slope_c;
        return;
    }
    s->min_strain_t = s->min_strain_c;
    s->end_strain_t = s->end_strain_c;
    s->unload_slope_t = s->unload_slope_c;
    if (eps > 0.0) {
        s->sig_t = 0.0;
        s->tangent_t = 0.0;
        s->eps_p_t = s->eps_p_c;
        s->alpha_t = s->alpha_c;
        return;
    }
    double temp_stress = s->sig_c + s->unload_slope_t * eps - s->unload_slope_t * s->eps_c;
    if (eps <= s->eps_c) {
        double sg, tg;
        c01_reload(fpc, epsc0, fpcu, epscu, eps, &s->min_strain_t, &s->end_strain_t,
                   &s->unload_slope_t, &sg, &tg);
        s->sig_t = sg;
        s->tangent_t = tg;
        if (temp_stress > s->sig_t) {
            s->sig_t = temp_stress;
            s->tangent_t = s->unload_slope_t;
        }
    } else if (temp_stress <= 0.0) {
        s->sig_t = temp_stress;
        s->tangent_t = s->unload_slope_t;
    } else {
        s->sig_t = 0.0;
        s->tangent_t = 0.0;
    }
    s->eps_p_t = s->eps_p_c;
    s->alpha_t = s->alpha_c;
}

static void steel01_trial(const or_mat_def *d, or_mat_state *s, double eps) {
    if (eps == s->eps_t) return; /* ops.mojo:27 */
    or_steel01_core(d, s, eps);
}

static void concrete01_trial(const or_mat_def *d, or_mat_state *s, double eps) {
    if (eps == s->eps_t) return; /* ops.mojo:59 */
    or_concrete01_core(d, s, eps);
}

/* fiber2d.mojo:1110-1194: the SoA runtime Steel02 law is steel02.mojo:3-114
 * verbatim; exported for the 2D section family loop. */
void or_steel02_core(const or_mat_def *d, or_mat_state *s, double eps) { steel02_set_trial(d, s, eps); }

/* fiber2d.mojo:1197-1264: the SoA runtime Concrete02 law. */
void or_concrete02_core(const or_mat_def *d, or_mat_state *s, double eps) { concrete02_set_trial(d, s, eps); }

/* ops.mojo:117-151 (dispatch with the AoS early return on eps == eps_t) */
void or_uniaxial_set_trial(const or_mat_def *d, or_mat_state *s, double eps) {
    switch (d->type) {
    case OR_ELASTIC: elastic_trial(d, s, eps); return;
    case OR_STEEL01: steel01_trial(d, s, eps); return;
    case OR_CONCRETE01: concrete01_trial(d, s, eps); return;
    case OR_STEEL02:
        if (eps == s->eps_t) return;
        steel02_set_trial(d, s, eps);
        return;
    case OR_CONCRETE02:
        if (eps == s->eps_t) return;
        concrete02_set_trial(d, s, eps);
        return;
    default: return;
    }
}

/* ops.mojo:154-174 */
void or_uniaxial_commit(or_mat_state *s) {
    s->eps_c = s->eps_t;
    s->sig_c = s->sig_t;
    s->tangent_c = s->tangent_t;
    s->eps_p_c = s->eps_p_t;
    s->alpha_c = s->alpha_t;
    s->min_strain_c = s->min_strain_t;
    s->end_strain_c = s->end_strain_t;
    s->unload_slope_c = s->unload_slope_t;
    s->s2_epsmin_c = s->s2_epsmin_t;
    s->s2_epsmax_c = s->s2_epsmax_t;
    s->s2_epspl_c = s->s2_epspl_t;
    s->s2_epss0_c = s->s2_epss0_t;
    s->s2_sigs0_c = s->s2_sigs0_t;
    s->s2_epsr_c = s->s2_epsr_t;
    s->s2_sigr_c = s->s2_sigr_t;
    s->s2_kon_c = s->s2_kon_t;
    s->c2_ecmin_c = s->c2_ecmin_t;
    s->c2_dept_c = s->c2_dept_t;
}

/* ops.mojo:177-197 */
void or_uniaxial_revert(or_mat_state *s) {
    s->eps_t = s->eps_c;
    s->sig_t = s->sig_c;
    s->tangent_t = s->tangent_c;
    s->eps_p_t = s->eps_p_c;
    s->alpha_t = s->alpha_c;
    s->min_strain_t = s->min_strain_c;
    s->end_strain_t = s->end_strain_c;
    s->unload_slope_t = s->unload_slope_c;
    s->s2_epsmin_t = s->s2_epsmin_c;
    s->s2_epsmax_t = s->s2_epsmax_c;
    s->s2_epspl_t = s->s2_epspl_c;
    s->s2_epss0_t = s->s2_epss0_c;
    s->s2_sigs0_t = s->s2_sigs0_c;
    s->s2_epsr_t = s->s2_epsr_c;
    s->s2_sigr_t = s->s2_sigr_c;
    s->s2_kon_t = s->s2_kon_c;
    s->c2_ecmin_t = s->c2_ecmin_c;
    s->c2_dept_t = s->c2_dept_c;
}

int or_mat_state_size(void) { return (int)sizeof(or_mat_state); }

/* material_path.mojo:228-259: drive one material through a strain path,
 * committing after every point when commit_each is set. */
void or_material_path(int type, const double *params, int n, const double *strains,
                      int commit_each, double *out_sig, double *out_tan) {
    or_mat_def d;
    d.type = type;
    for (int i = 0; i < 11; ++i) d.p[i] = params[i];
    or_mat_state s;
    or_mat_state_init(&d, &s);
    for (int i = 0; i < n; ++i) {
        or_uniaxial_set_trial(&d, &s, strains[i]);
        out_sig[i] = s.sig_t;
        out_tan[i] = s.tangent_t;
        if (commit_each) or_uniaxial_commit(&s);
    }
}
