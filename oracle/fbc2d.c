/*
 * forceBeamColumn2d of the oracle (test infrastructure only).
 * Restates src/mojo/elements/force_beam_column2d.mojo for the Linear
 * geometric transformation without element loads (the transient fiber
 * frames): try_increment (:804-1189), the 3x3 inverse (:223-272), the exact
 * elastic state (:608-732), the initial basic tangent (:735-801) and the
 * element entry (:1418-1919) with its subdivision / three flexibility
 * schemes / initial-tangent fallback, plus the 2D transforms of
 * elements/utils.mojo:142-238.
 *
 * Element state q = [q0 q1 q2 | eps0[nip] | kappa[nip] | v_basic[3]]
 * (loader.mojo:1763-1771).  Section instances of the element are
 * first_inst + ip on its section definition.
 */
#include <math.h>
#include <string.h>
#include "strut_oracle.h"
#include "oracle_internal.h"

/* :223-272 */
static int invert3(double a, double b, double c, double d, double e, double f, double g, double h,
                   double i, double out[9]) {
    double co00 = e * i - f * h;
    double co01 = -(d * i - f * g);
    double co02 = d * h - e * g;
    double co10 = -(b * i - c * h);
    double co11 = a * i - c * g;
    double co12 = -(a * h - b * g);
    double co20 = b * f - c * e;
    double co21 = -(a * f - c * d);
    double co22 = a * e - b * d;
    double det = a * co00 + b * co01 + c * co02;
    if (fabs(det) <= 1.0e-40) return 0;
    double inv_det = 1.0 / det;
    out[0] = co00 * inv_det;
    out[1] = co10 * inv_det;
    out[2] = co20 * inv_det;
    out[3] = co01 * inv_det;
    out[4] = co11 * inv_det;
    out[5] = co21 * inv_det;
    out[6] = co02 * inv_det;
    out[7] = co12 * inv_det;
    out[8] = co22 * inv_det;
    return 1;
}

static double max_abs3(double a, double b, double c) { /* :275-284 */
    double m = fabs(a);
    if (fabs(b) > m) m = fabs(b);
    if (fabs(c) > m) m = fabs(c);
    return m;
}

typedef struct {
    double vs_eps0[OR_MAX_NIP], vs_kappa[OR_MAX_NIP];
    double vs_sub_eps0[OR_MAX_NIP], vs_sub_kappa[OR_MAX_NIP];
    double ssr_ax[OR_MAX_NIP], ssr_mo[OR_MAX_NIP];
    double ssr_sub_ax[OR_MAX_NIP], ssr_sub_mo[OR_MAX_NIP];
    double fs00[OR_MAX_NIP], fs01[OR_MAX_NIP], fs11[OR_MAX_NIP];
    double fs_sub00[OR_MAX_NIP], fs_sub01[OR_MAX_NIP], fs_sub11[OR_MAX_NIP];
} fbc_scratch;

/* one batch of section points at (eps0[ip], kappa[ip]); returns 0 at the
 * first singular flexibility like fiber_section2d_batch_eval_same_def. */
static int eval_sections(const or_elem2d *el, const double *eps0, const double *kappa, double *out8) {
    int insts[OR_MAX_NIP];
    int valid[OR_MAX_NIP];
    for (int ip = 0; el->nip > ip; ++ip) insts[ip] = el->first_inst + ip;
    el->stats->section_evals += el->nip;
    int rc = or_sec2d_batch_eval(el->sec, el->nip, insts, eps0, kappa, out8, valid);
    return rc == 0;
}

/* :804-1189; returns 1 and best[10] = {q0,q1,q2,k00,k01,k02,k10,k11,k12,k22} */
static int try_increment(const or_elem2d *el, double L, double *q_state, fbc_scratch *sc, double base_v0,
                         double base_v1, double base_v2, double dv0, double dv1, double dv2,
                         int use_initial_section_flexibility, double best[10]) {
    const int nip = el->nip;
    const double *xis = el->xi, *weights = el->w;
    double q0 = q_state[0], q1 = q_state[1], q2 = q_state[2];
    const int eps0_off = 3, kappa_off = 3 + nip;
    const double elem_tol = 1.0e-12;
    const int max_elem_iters = 10;
    double target_v0 = base_v0 + dv0, target_v1 = base_v1 + dv1, target_v2 = base_v2 + dv2;
    double if3[3];
    if (!or_sec2d_initial_flex(el->sec, if3)) return 0;
    double pe[OR_MAX_NIP], pk[OR_MAX_NIP], out8[8 * OR_MAX_NIP];
    for (int ip = 0; ip < nip; ++ip) {
        pe[ip] = q_state[eps0_off + ip];
        pk[ip] = q_state[kappa_off + ip];
        sc->vs_eps0[ip] = pe[ip];
        sc->vs_kappa[ip] = pk[ip];
    }
    if (!eval_sections(el, pe, pk, out8)) return 0;
    for (int ip = 0; ip < nip; ++ip) {
        const double *r = out8 + 8 * ip;
        sc->ssr_ax[ip] = r[0];
        sc->ssr_mo[ip] = r[1];
        sc->fs00[ip] = r[5];
        sc->fs01[ip] = r[6];
        sc->fs11[ip] = r[7];
        sc->vs_sub_eps0[ip] = sc->vs_eps0[ip];
        sc->vs_sub_kappa[ip] = sc->vs_kappa[ip];
        sc->ssr_sub_ax[ip] = r[0];
        sc->ssr_sub_mo[ip] = r[1];
        sc->fs_sub00[ip] = r[5];
        sc->fs_sub01[ip] = r[6];
        sc->fs_sub11[ip] = r[7];
    }
    double pf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int ip = 0; ip < nip; ++ip) {
        double xi = xis[ip], wL = weights[ip] * L, b_mi = xi - 1.0, b_mj = xi;
        double f00 = sc->fs00[ip], f01 = sc->fs01[ip], f10 = f01, f11 = sc->fs11[ip];
        pf[0] += wL * f00;
        pf[1] += wL * f01 * b_mi;
        pf[2] += wL * f01 * b_mj;
        pf[3] += wL * b_mi * f10;
        pf[4] += wL * b_mi * f11 * b_mi;
        pf[5] += wL * b_mi * f11 * b_mj;
        pf[6] += wL * b_mj * f10;
        pf[7] += wL * b_mj * f11 * b_mi;
        pf[8] += wL * b_mj * f11 * b_mj;
    }
    double pk3[9];
    if (invert3(pf[0], pf[1], pf[2], pf[3], pf[4], pf[5], pf[6], pf[7], pf[8], pk3)) {
        q0 += pk3[0] * dv0 + pk3[1] * dv1 + pk3[2] * dv2;
        q1 += pk3[3] * dv0 + pk3[4] * dv1 + pk3[5] * dv2;
        q2 += pk3[6] * dv0 + pk3[7] * dv1 + pk3[8] * dv2;
    }
    int num_elem_iters = max_elem_iters;
    if (use_initial_section_flexibility == 1) num_elem_iters = 10 * max_elem_iters;
    for (int it = 0; it < num_elem_iters; ++it) {
        el->stats->local_iters++;
        double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double vr0 = 0.0, vr1 = 0.0, vr2 = 0.0;
        int use_init = use_initial_section_flexibility == 1 || (use_initial_section_flexibility == 2 && it == 0);
        for (int ip = 0; ip < nip; ++ip) {
            double xi = xis[ip], b_mi = xi - 1.0, b_mj = xi;
            double ss0 = q0 + 0.0;
            double ss1 = b_mi * q1 + b_mj * q2 + 0.0;
            double dss0 = ss0 - sc->ssr_sub_ax[ip];
            double dss1 = ss1 - sc->ssr_sub_mo[ip];
            double s00, s01, s11;
            if (use_init) {
                s00 = if3[0];
                s01 = if3[1];
                s11 = if3[2];
            } else {
                s00 = sc->fs_sub00[ip];
                s01 = sc->fs_sub01[ip];
                s11 = sc->fs_sub11[ip];
            }
            double dvs0 = s00 * dss0 + s01 * dss1;
            double dvs1 = s01 * dss0 + s11 * dss1;
            pe[ip] = sc->vs_sub_eps0[ip] + dvs0;
            pk[ip] = sc->vs_sub_kappa[ip] + dvs1;
        }
        if (!eval_sections(el, pe, pk, out8)) return 0;
        for (int ip = 0; ip < nip; ++ip) {
            double xi = xis[ip], wL = weights[ip] * L, b_mi = xi - 1.0, b_mj = xi;
            double ss0 = q0 + 0.0;
            double ss1 = b_mi * q1 + b_mj * q2 + 0.0;
            const double *r = out8 + 8 * ip;
            sc->vs_sub_eps0[ip] = pe[ip];
            sc->vs_sub_kappa[ip] = pk[ip];
            sc->ssr_sub_ax[ip] = r[0];
            sc->ssr_sub_mo[ip] = r[1];
            double f00 = r[5], f01 = r[6], f11 = r[7];
            sc->fs_sub00[ip] = f00;
            sc->fs_sub01[ip] = f01;
            sc->fs_sub11[ip] = f11;
            double f10 = f01;
            double dr0 = ss0 - sc->ssr_sub_ax[ip];
            double dr1 = ss1 - sc->ssr_sub_mo[ip];
            double dv_0 = f00 * dr0 + f01 * dr1;
            double dv_1 = f01 * dr0 + f11 * dr1;
            f[0] += wL * f00;
            f[1] += wL * f01 * b_mi;
            f[2] += wL * f01 * b_mj;
            f[3] += wL * b_mi * f10;
            f[4] += wL * b_mi * f11 * b_mi;
            f[5] += wL * b_mi * f11 * b_mj;
            f[6] += wL * b_mj * f10;
            f[7] += wL * b_mj * f11 * b_mi;
            f[8] += wL * b_mj * f11 * b_mj;
            vr0 += wL * (pe[ip] + dv_0);
            vr1 += wL * b_mi * (pk[ip] + dv_1);
            vr2 += wL * b_mj * (pk[ip] + dv_1);
        }
        double k[9];
        if (!invert3(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], k)) return 0;
        double r0 = target_v0 - vr0, r1 = target_v1 - vr1, r2 = target_v2 - vr2;
        double dq0 = k[0] * r0 + k[1] * r1 + k[2] * r2;
        double dq1 = k[3] * r0 + k[4] * r1 + k[5] * r2;
        double dq2 = k[6] * r0 + k[7] * r1 + k[8] * r2;
        double work = fabs(r0 * dq0 + r1 * dq1 + r2 * dq2);
        q0 += dq0;
        q1 += dq1;
        q2 += dq2;
        if (work < elem_tol) {
            q_state[0] = q0;
            q_state[1] = q1;
            q_state[2] = q2;
            for (int ip = 0; ip < nip; ++ip) {
                q_state[eps0_off + ip] = sc->vs_sub_eps0[ip];
                q_state[kappa_off + ip] = sc->vs_sub_kappa[ip];
                sc->vs_eps0[ip] = sc->vs_sub_eps0[ip];
                sc->vs_kappa[ip] = sc->vs_sub_kappa[ip];
            }
            best[0] = q0; best[1] = q1; best[2] = q2;
            best[3] = k[0]; best[4] = k[1]; best[5] = k[2];
            best[6] = k[3]; best[7] = k[4]; best[8] = k[5];
            best[9] = k[8];
            return 1;
        }
    }
    return 0;
}

/* :735-801 */
static int initial_basic_tangent(const or_elem2d *el, double L, double k_out[9]) {
    double if3[3];
    if (!or_sec2d_initial_flex(el->sec, if3)) return 0;
    double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int ip = 0; ip < el->nip; ++ip) {
        double xi = el->xi[ip], wL = el->w[ip] * L, b_mi = xi - 1.0, b_mj = xi;
        double f00 = if3[0], f01 = if3[1], f10 = f01, f11 = if3[2];
        f[0] += wL * f00;
        f[1] += wL * f01 * b_mi;
        f[2] += wL * f01 * b_mj;
        f[3] += wL * b_mi * f10;
        f[4] += wL * b_mi * f11 * b_mi;
        f[5] += wL * b_mi * f11 * b_mj;
        f[6] += wL * b_mj * f10;
        f[7] += wL * b_mj * f11 * b_mi;
        f[8] += wL * b_mj * f11 * b_mj;
    }
    return invert3(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], k_out);
}

/* :608-732 (all materials elastic) */
static int exact_elastic(const or_elem2d *el, double L, double *q_state, double tv0, double tv1, double tv2,
                         double best[10]) {
    double if3[3];
    if (!or_sec2d_initial_flex(el->sec, if3)) return 0;
    const int nip = el->nip;
    double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double vl0 = 0.0, vl1 = 0.0, vl2 = 0.0;
    double s00 = if3[0], s01 = if3[1], s10 = s01, s11 = if3[2];
    for (int ip = 0; ip < nip; ++ip) {
        double xi = el->xi[ip], wL = el->w[ip] * L, b_mi = xi - 1.0, b_mj = xi;
        f[0] += wL * s00;
        f[1] += wL * s01 * b_mi;
        f[2] += wL * s01 * b_mj;
        f[3] += wL * b_mi * s10;
        f[4] += wL * b_mi * s11 * b_mi;
        f[5] += wL * b_mi * s11 * b_mj;
        f[6] += wL * b_mj * s10;
        f[7] += wL * b_mj * s11 * b_mi;
        f[8] += wL * b_mj * s11 * b_mj;
        vl0 += wL * (s00 * 0.0 + s01 * 0.0);
        vl1 += wL * b_mi * (s10 * 0.0 + s11 * 0.0);
        vl2 += wL * b_mj * (s10 * 0.0 + s11 * 0.0);
    }
    double k[9];
    if (!invert3(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], k)) return 0;
    double r0 = tv0 - vl0, r1 = tv1 - vl1, r2 = tv2 - vl2;
    double q0 = k[0] * r0 + k[1] * r1 + k[2] * r2;
    double q1 = k[3] * r0 + k[4] * r1 + k[5] * r2;
    double q2 = k[6] * r0 + k[7] * r1 + k[8] * r2;
    double pe[OR_MAX_NIP], pk[OR_MAX_NIP];
    for (int ip = 0; ip < nip; ++ip) {
        double xi = el->xi[ip], b_mi = xi - 1.0, b_mj = xi;
        double axial = q0 + 0.0;
        double moment = b_mi * q1 + b_mj * q2 + 0.0;
        pe[ip] = s00 * axial + s01 * moment;
        pk[ip] = s10 * axial + s11 * moment;
        q_state[3 + ip] = pe[ip];
        q_state[3 + nip + ip] = pk[ip];
        double out5[5];
        or_sec2d_set_trial(el->sec, el->first_inst + ip, pe[ip], pk[ip], out5);
        el->stats->section_evals++;
    }
    q_state[0] = q0;
    q_state[1] = q1;
    q_state[2] = q2;
    best[0] = q0; best[1] = q1; best[2] = q2;
    best[3] = k[0]; best[4] = k[1]; best[5] = k[2];
    best[6] = k[3]; best[7] = k[4]; best[8] = k[5];
    best[9] = k[8];
    return 1;
}

static void restore_predictor(double *q_state, int nip, double q0, double q1, double q2, const double *ve,
                              const double *vk) { /* :334-352 */
    q_state[0] = q0;
    q_state[1] = q1;
    q_state[2] = q2;
    for (int ip = 0; ip < nip; ++ip) {
        q_state[3 + ip] = ve[ip];
        q_state[3 + nip + ip] = vk[ip];
    }
}

/* elements/utils.mojo:172-238 */
static void transform_k_to_global(double k[36], double c, double s) {
    for (int i = 0; i < 6; ++i) {
        double k0 = k[6 * i + 0], k1 = k[6 * i + 1], k3 = k[6 * i + 3], k4 = k[6 * i + 4];
        k[6 * i + 0] = c * k0 - s * k1;
        k[6 * i + 1] = s * k0 + c * k1;
        k[6 * i + 3] = c * k3 - s * k4;
        k[6 * i + 4] = s * k3 + c * k4;
    }
    for (int pr = 0; pr < 2; ++pr) {
        int a = pr == 0 ? 0 : 3, b = a + 1;
        double ra[6], rb[6];
        for (int j = 0; j < 6; ++j) {
            ra[j] = k[6 * a + j];
            rb[j] = k[6 * b + j];
        }
        for (int j = 0; j < 6; ++j) {
            k[6 * a + j] = c * ra[j] - s * rb[j];
            k[6 * b + j] = s * ra[j] + c * rb[j];
        }
    }
}

/* :1418-1919 with the Linear transformation and no element loads */
int or_fbc2d_state(const or_elem2d *el, double *q_state, const double u_global[6], double k_out[36],
                   double f_out[6]) {
    const int nip = el->nip;
    const double L = el->L, inv_L = el->inv_L, c = el->c, s = el->s;
    double ul[6];
    ul[0] = c * u_global[0] + s * u_global[1];
    ul[1] = -s * u_global[0] + c * u_global[1];
    ul[2] = u_global[2];
    ul[3] = c * u_global[3] + s * u_global[4];
    ul[4] = -s * u_global[3] + c * u_global[4];
    ul[5] = u_global[5];
    const int basic_off = 3 + 2 * nip;
    double chord = (ul[4] - ul[1]) / L;
    double v0 = ul[3] - ul[0], v1 = ul[2] - chord, v2 = ul[5] - chord;
    double acc_b0 = q_state[basic_off], acc_b1 = q_state[basic_off + 1], acc_b2 = q_state[basic_off + 2];
    double acc_q0 = q_state[0], acc_q1 = q_state[1], acc_q2 = q_state[2];
    fbc_scratch sc;
    memset(&sc, 0, sizeof(sc));
    for (int ip = 0; ip < nip; ++ip) {
        sc.vs_eps0[ip] = q_state[3 + ip];
        sc.vs_kappa[ip] = q_state[3 + nip + ip];
    }
    double rem0 = v0 - acc_b0, rem1 = v1 - acc_b1, rem2 = v2 - acc_b2;
    double att0 = rem0, att1 = rem1, att2 = rem2;
    int converged = 0, have_best = 0;
    double best[10];
    const double tolerance = 1.0e-12;
    const int max_subdivisions = 4;
    int num_subdivide = 1;
    const double cutback = 10.0;
    if (or_sec2d_all_elastic(el->sec)) { /* :1577-1620 */
        double solved[10];
        if (exact_elastic(el, L, q_state, v0, v1, v2, solved)) {
            acc_b0 = v0; acc_b1 = v1; acc_b2 = v2;
            q_state[basic_off] = acc_b0;
            q_state[basic_off + 1] = acc_b1;
            q_state[basic_off + 2] = acc_b2;
            acc_q0 = solved[0]; acc_q1 = solved[1]; acc_q2 = solved[2];
            for (int ip = 0; ip < nip; ++ip) {
                sc.vs_eps0[ip] = q_state[3 + ip];
                sc.vs_kappa[ip] = q_state[3 + nip + ip];
            }
            converged = 1;
            rem0 = rem1 = rem2 = 0.0;
            memcpy(best, solved, sizeof(best));
            have_best = 1;
        }
    }
    if (!converged && max_abs3(rem0, rem1, rem2) <= 1.0e-16) {
        for (int use_initial = 0; use_initial < 3; ++use_initial) {
            restore_predictor(q_state, nip, acc_q0, acc_q1, acc_q2, sc.vs_eps0, sc.vs_kappa);
            q_state[basic_off] = acc_b0;
            q_state[basic_off + 1] = acc_b1;
            q_state[basic_off + 2] = acc_b2;
            double solved[10];
            if (!try_increment(el, L, q_state, &sc, acc_b0, acc_b1, acc_b2, 0.0, 0.0, 0.0, use_initial, solved))
                continue;
            memcpy(best, solved, sizeof(best));
            have_best = 1;
            converged = 1;
            break;
        }
    } else if (!converged) {
        while (1) {
            if (num_subdivide > max_subdivisions) break;
            double tv0 = acc_b0 + att0, tv1 = acc_b1 + att1, tv2 = acc_b2 + att2;
            int scheme_success = 0;
            for (int use_initial = 0; use_initial < 3; ++use_initial) {
                restore_predictor(q_state, nip, acc_q0, acc_q1, acc_q2, sc.vs_eps0, sc.vs_kappa);
                q_state[basic_off] = acc_b0;
                q_state[basic_off + 1] = acc_b1;
                q_state[basic_off + 2] = acc_b2;
                double solved[10];
                if (!try_increment(el, L, q_state, &sc, acc_b0, acc_b1, acc_b2, att0, att1, att2, use_initial,
                                   solved))
                    continue;
                acc_b0 = tv0; acc_b1 = tv1; acc_b2 = tv2;
                q_state[basic_off] = acc_b0;
                q_state[basic_off + 1] = acc_b1;
                q_state[basic_off + 2] = acc_b2;
                acc_q0 = solved[0]; acc_q1 = solved[1]; acc_q2 = solved[2];
                for (int ip = 0; ip < nip; ++ip) {
                    sc.vs_eps0[ip] = q_state[3 + ip];
                    sc.vs_kappa[ip] = q_state[3 + nip + ip];
                }
                memcpy(best, solved, sizeof(best));
                have_best = 1;
                rem0 = v0 - acc_b0; rem1 = v1 - acc_b1; rem2 = v2 - acc_b2;
                if (max_abs3(rem0, rem1, rem2) <= tolerance) {
                    converged = 1;
                } else {
                    att0 = rem0; att1 = rem1; att2 = rem2;
                    num_subdivide = 1;
                }
                scheme_success = 1;
                break;
            }
            if (converged) break;
            if (scheme_success) continue;
            att0 /= cutback;
            att1 /= cutback;
            att2 /= cutback;
            num_subdivide += 1;
            el->stats->subdivisions++;
        }
    }
    if (!converged) {
        restore_predictor(q_state, nip, acc_q0, acc_q1, acc_q2, sc.vs_eps0, sc.vs_kappa);
        q_state[basic_off] = acc_b0;
        q_state[basic_off + 1] = acc_b1;
        q_state[basic_off + 2] = acc_b2;
        for (int ip = 0; ip < nip; ++ip) {
            double out5[5];
            or_sec2d_set_trial(el->sec, el->first_inst + ip, sc.vs_eps0[ip], sc.vs_kappa[ip], out5);
            el->stats->section_evals++;
        }
        el->stats->unconverged++;
    }
    if (!have_best) {
        double kt[9];
        if (!initial_basic_tangent(el, L, kt)) return -1; /* reference aborts */
        best[0] = q_state[0]; best[1] = q_state[1]; best[2] = q_state[2];
        best[3] = kt[0]; best[4] = kt[1]; best[5] = kt[2];
        best[6] = kt[3]; best[7] = kt[4]; best[8] = kt[5];
        best[9] = kt[8];
        el->stats->fallbacks++;
    }
    double q0 = best[0], q1 = best[1], q2 = best[2];
    double k00 = best[3], k01 = best[4], k02 = best[5], k10 = best[6], k11 = best[7], k12 = best[8];
    double k20 = k02, k21 = k12, k22 = best[9];
    /* basic map, :387-408 */
    double amap[6][3] = {{-1.0, 0.0, 0.0}, {0.0, inv_L, inv_L}, {0.0, 1.0, 0.0},
                         {1.0, 0.0, 0.0},  {0.0, -inv_L, -inv_L}, {0.0, 0.0, 1.0}};
    for (int a = 0; a < 6; ++a) {
        double a0 = amap[a][0], a1 = amap[a][1], a2 = amap[a][2];
        for (int b = 0; b < 6; ++b) {
            double b0 = amap[b][0], b1 = amap[b][1], b2 = amap[b][2];
            k_out[6 * a + b] = a0 * (k00 * b0 + k01 * b1 + k02 * b2) + a1 * (k10 * b0 + k11 * b1 + k12 * b2) +
                               a2 * (k20 * b0 + k21 * b1 + k22 * b2);
        }
    }
    f_out[0] = -q0;
    f_out[1] = (q1 + q2) * inv_L;
    f_out[2] = q1;
    f_out[3] = q0;
    f_out[4] = -(q1 + q2) * inv_L;
    f_out[5] = q2;
    f_out[0] += 0.0;
    f_out[1] += 0.0;
    f_out[4] += 0.0;
    transform_k_to_global(k_out, c, s);
    /* utils.mojo:157-169 */
    double f0 = f_out[0], f1 = f_out[1], f3 = f_out[3], f4 = f_out[4];
    f_out[0] = c * f0 - s * f1;
    f_out[1] = s * f0 + c * f1;
    f_out[3] = c * f3 - s * f4;
    f_out[4] = s * f3 + c * f4;
    return converged ? 0 : 1;
}

/* One forceBeamColumn2d element between (x1, y1) and (x2, y2) on the
 * section instances first_inst .. first_inst + nip - 1, with its basic-force
 * state q_state (3 + 2 nip + 3 values), for the checks in tests/. */
int or_fbc2d_eval(or_sec2d *sec, int first_inst, int nip, const double *xi, const double *w, double x1, double y1,
                  double x2, double y2, double *q_state, const double u_global[6], double k_out[36],
                  double f_out[6]) {
    or_stats stats = {0, 0, 0, 0, 0};
    or_elem2d el;
    el.sec = sec;
    el.first_inst = first_inst;
    el.nip = nip;
    for (int ip = 0; ip < nip; ++ip) {
        el.xi[ip] = xi[ip];
        el.w[ip] = w[ip];
    }
    const double dx = x2 - x1, dy = y2 - y1;
    el.L = hypot(dx, dy);
    el.inv_L = 1.0 / el.L;
    el.c = dx * el.inv_L;
    el.s = dy * el.inv_L;
    el.type = 0;
    el.dc = dx / el.L;
    el.ds = dy / el.L;
    el.stats = &stats;
    return or_fbc2d_state(&el, q_state, u_global, k_out, f_out);
}
