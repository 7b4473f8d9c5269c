/* 3D fiber beam-column elements of the oracle (test infrastructure only).
 *
 * Restates, for fiber sections without element loads and with the Linear or
 * PDelta geometric transformation:
 *   - beam_column3d_fiber_global_tangent_and_internal
 *     (elements/beam_column3d_nonlinear.mojo:357-575), the displacement-based
 *     3D element (dispBeamColumn3d);
 *   - force_beam_column3d_fiber_global_tangent_and_internal
 *     (elements/force_beam_column3d.mojo:2549-2990) with its try_increment
 *     (:1203-1607), the force-based 3D element (forceBeamColumn3d);
 *   - the local axes (elements/utils.mojo:70-139) and the rotation
 *     (force_beam_column3d.mojo:236-284, beam_column3d_nonlinear.mojo:34-78).
 * Every expression keeps the reference's operation order, including the
 * terms whose factors are zero (they can turn -0 into +0). */
#include <math.h>
#include <string.h>

#include "oracle_internal.h"

/* elements/utils.mojo:48-139 */
static void or_cross(double ax, double ay, double az, double bx, double by, double bz, double *o) {
    o[0] = ay * bz - az * by;
    o[1] = az * bx - ax * bz;
    o[2] = ax * by - ay * bx;
}

int or_beam3d_rotation(double x1, double y1, double z1, double x2, double y2, double z2, int has_vecxz,
                       const double *vecxz, double R[9]) {
    const double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
    const double L = sqrt(dx * dx + dy * dy + dz * dz);
    if (L == 0.0) return -1;
    const double n = sqrt(dx * dx + dy * dy + dz * dz); /* _normalize */
    const double lx = dx / n, ly = dy / n, lz = dz / n;
    double vx, vy, vz;
    if (has_vecxz) {
        vx = vecxz[0];
        vy = vecxz[1];
        vz = vecxz[2];
    } else { /* _beam3d_default_vecxz */
        vx = 1.0;
        vy = 0.0;
        vz = 0.0;
        if (fabs(lx * vx + ly * vy + lz * vz) >= 0.9) {
            vx = 0.0;
            vy = 1.0;
            vz = 0.0;
            if (fabs(lx * vx + ly * vy + lz * vz) >= 0.9) {
                vx = 0.0;
                vy = 0.0;
                vz = 1.0;
            }
        }
    }
    double yv[3], zv[3];
    or_cross(vx, vy, vz, lx, ly, lz, yv);
    const double ynorm = sqrt(yv[0] * yv[0] + yv[1] * yv[1] + yv[2] * yv[2]);
    if (ynorm == 0.0) return -2;
    yv[0] /= ynorm;
    yv[1] /= ynorm;
    yv[2] /= ynorm;
    or_cross(lx, ly, lz, yv[0], yv[1], yv[2], zv);
    R[0] = lx;
    R[1] = ly;
    R[2] = lz;
    R[3] = yv[0];
    R[4] = yv[1];
    R[5] = yv[2];
    R[6] = zv[0];
    R[7] = zv[1];
    R[8] = zv[2];
    return 0;
}

/* ---------------------------------------------------------------------
 * dispBeamColumn3d: beam_column3d_nonlinear.mojo:357-575 */

/* :81-90 (12 x 12 T from R, row major) */
static void transform_matrix(const double R[9], double T[144]) {
    for (int k = 0; k < 144; ++k) T[k] = 0.0;
    for (int b = 0; b < 4; ++b)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) T[(3 * b + i) * 12 + 3 * b + j] = R[3 * i + j];
}

/* linalg.mojo:20-36: out = a b, terms with a[i][k] == 0 skipped */
static void matmul12(const double *a, const double *b, double *out) {
    for (int k = 0; k < 144; ++k) out[k] = 0.0;
    for (int i = 0; i < 12; ++i)
        for (int k = 0; k < 12; ++k) {
            const double aik = a[i * 12 + k];
            if (aik == 0.0) continue;
            for (int j = 0; j < 12; ++j) out[i * 12 + j] += aik * b[k * 12 + j];
        }
}

int or_dbc3d_state(const or_elem3d *el, const double u_global[12], double k_out[144], double f_out[12]) {
    const double dx = el->x2 - el->x1, dy = el->y2 - el->y1, dz = el->z2 - el->z1;
    const double L = sqrt(dx * dx + dy * dy + dz * dz);
    if (L == 0.0) return -1;
    double R[9], T[144];
    if (or_beam3d_rotation(el->x1, el->y1, el->z1, el->x2, el->y2, el->z2, el->has_vecxz, el->vecxz, R)) return -1;
    transform_matrix(R, T);
    double ul[12]; /* :92-102 */
    for (int i = 0; i < 12; ++i) {
        double sum = 0.0;
        for (int j = 0; j < 12; ++j) sum += T[i * 12 + j] * u_global[j];
        ul[i] = sum;
    }
    double kl[144], fl[12], qb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < 144; ++k) kl[k] = 0.0;
    for (int k = 0; k < 12; ++k) fl[k] = 0.0;
    double axial_force_avg = 0.0;
    for (int ip = 0; ip < el->nip; ++ip) {
        const double xi = el->xi[ip], weight = el->w[ip], wL = weight * L;
        /* _section_kinematics_at_xi (:154-171) */
        const double inv_L = 1.0 / L;
        const double eps0 = (-inv_L) * ul[0] + inv_L * ul[6];
        const double kappa_y = ((-6.0 + 12.0 * xi) / (L * L)) * ul[2] + ((4.0 - 6.0 * xi) / L) * ul[4] +
                               ((6.0 - 12.0 * xi) / (L * L)) * ul[8] + ((2.0 - 6.0 * xi) / L) * ul[10];
        const double kappa_z = ((-6.0 + 12.0 * xi) / (L * L)) * ul[1] + ((-4.0 + 6.0 * xi) / L) * ul[5] +
                               ((6.0 - 12.0 * xi) / (L * L)) * ul[7] + ((-2.0 + 6.0 * xi) / L) * ul[11];
        double r[9];
        if (or_sec3d_set_trial(el->sec, el->first_inst + ip, eps0, kappa_y, kappa_z, r)) return -1;
        if (el->stats) el->stats->section_evals++;
        const double N = r[0], My = r[1], Mz = r[2];
        const double k11 = r[3], k12 = r[4], k13 = r[5], k22 = r[6], k23 = r[7], k33 = r[8];
        axial_force_avg += weight * N;
        double b_eps[12], b_ky[12], b_kz[12];
        for (int a = 0; a < 12; ++a) b_eps[a] = b_ky[a] = b_kz[a] = 0.0;
        b_eps[0] = -1.0 / L;
        b_eps[6] = 1.0 / L;
        b_ky[2] = (-6.0 + 12.0 * xi) / (L * L);
        b_ky[4] = (4.0 - 6.0 * xi) / L;
        b_ky[8] = (6.0 - 12.0 * xi) / (L * L);
        b_ky[10] = (2.0 - 6.0 * xi) / L;
        b_kz[1] = (-6.0 + 12.0 * xi) / (L * L);
        b_kz[5] = (-4.0 + 6.0 * xi) / L;
        b_kz[7] = (6.0 - 12.0 * xi) / (L * L);
        b_kz[11] = (-2.0 + 6.0 * xi) / L;
        for (int a = 0; a < 12; ++a) {
            const double ba_n = b_eps[a], ba_my = b_ky[a], ba_mz = b_kz[a];
            fl[a] += wL * (ba_n * N + ba_my * My + ba_mz * Mz);
            for (int b = 0; b < 12; ++b) {
                const double bb_n = b_eps[b], bb_my = b_ky[b], bb_mz = b_kz[b];
                const double dN = k11 * bb_n + k12 * bb_my + k13 * bb_mz;
                const double dMy = k12 * bb_n + k22 * bb_my + k23 * bb_mz;
                const double dMz = k13 * bb_n + k23 * bb_my + k33 * bb_mz;
                kl[a * 12 + b] += wL * (ba_n * dN + ba_my * dMy + ba_mz * dMz);
            }
        }
    }
    if (el->G > 0.0 && el->J > 0.0) { /* :489-496 */
        const double torsion_k = el->G * el->J / L;
        kl[3 * 12 + 3] += torsion_k;
        kl[3 * 12 + 9] -= torsion_k;
        kl[9 * 12 + 3] -= torsion_k;
        kl[9 * 12 + 9] += torsion_k;
        qb[5] = torsion_k * (ul[9] - ul[3]);
    }
    if (el->transf == OR_TRANSF_PDELTA) { /* :118-131 */
        const double N_over_L = axial_force_avg / L;
        kl[1 * 12 + 1] += N_over_L;
        kl[1 * 12 + 7] -= N_over_L;
        kl[7 * 12 + 1] -= N_over_L;
        kl[7 * 12 + 7] += N_over_L;
        kl[2 * 12 + 2] += N_over_L;
        kl[2 * 12 + 8] -= N_over_L;
        kl[8 * 12 + 2] -= N_over_L;
        kl[8 * 12 + 8] += N_over_L;
    }
    /* :503-526: f_local[a] += A[5][a] q5 (rows 0..4 only with element loads) */
    for (int a = 0; a < 12; ++a) {
        const double a5 = a == 3 ? -1.0 : (a == 9 ? 1.0 : 0.0);
        fl[a] += a5 * qb[5];
    }
    /* :528-529 */
    double Tt[144], kT[144];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) Tt[j * 12 + i] = T[i * 12 + j];
    matmul12(kl, T, kT);
    matmul12(Tt, kT, k_out);
    for (int i = 0; i < 12; ++i) { /* :105-115 */
        double sum = 0.0;
        for (int j = 0; j < 12; ++j) sum += T[j * 12 + i] * fl[j];
        f_out[i] = sum;
    }
    return 0;
}

/* ---------------------------------------------------------------------
 * forceBeamColumn3d: force_beam_column3d.mojo */

/* :679-727 via _fiber_section3d_response_flexibility (:1093-1117) */
static int sec_flex(const double r[9], double f[6]) { return or_sec3d_flexibility(r + 3, f); }

/* :788-833 (Gauss-Jordan with partial pivoting on a copy of a) */
int or_invert5(const double a_in[25], double inv[25]) {
    double a[25];
    for (int k = 0; k < 25; ++k) a[k] = a_in[k];
    for (int k = 0; k < 25; ++k) inv[k] = 0.0;
    for (int i = 0; i < 5; ++i) inv[i * 5 + i] = 1.0;
    for (int i = 0; i < 5; ++i) {
        int pivot = i;
        double max_val = fabs(a[i * 5 + i]);
        for (int r = i + 1; r < 5; ++r) {
            const double v = fabs(a[r * 5 + i]);
            if (v > max_val) {
                max_val = v;
                pivot = r;
            }
        }
        if (max_val <= 1.0e-40) return 0;
        if (pivot != i)
            for (int c = 0; c < 5; ++c) {
                double t = a[i * 5 + c];
                a[i * 5 + c] = a[pivot * 5 + c];
                a[pivot * 5 + c] = t;
                t = inv[i * 5 + c];
                inv[i * 5 + c] = inv[pivot * 5 + c];
                inv[pivot * 5 + c] = t;
            }
        const double piv = a[i * 5 + i];
        for (int c = 0; c < 5; ++c) {
            a[i * 5 + c] /= piv;
            inv[i * 5 + c] /= piv;
        }
        for (int r = 0; r < 5; ++r) {
            if (r == i) continue;
            const double factor = a[r * 5 + i];
            if (factor == 0.0) continue;
            for (int c = 0; c < 5; ++c) {
                a[r * 5 + c] -= factor * a[i * 5 + c];
                inv[r * 5 + c] -= factor * inv[i * 5 + c];
            }
        }
    }
    return 1;
}

/* :419-463 */
static double basic_coeff(int row, int col, double inv_L) {
    switch (row) {
    case 0: return col == 0 ? -1.0 : (col == 6 ? 1.0 : 0.0);
    case 1: return col == 1 ? inv_L : (col == 5 ? 1.0 : (col == 7 ? -inv_L : 0.0));
    case 2: return col == 1 ? inv_L : (col == 7 ? -inv_L : (col == 11 ? 1.0 : 0.0));
    case 3: return col == 2 ? -inv_L : (col == 4 ? 1.0 : (col == 8 ? inv_L : 0.0));
    case 4: return col == 2 ? -inv_L : (col == 8 ? inv_L : (col == 10 ? 1.0 : 0.0));
    default: return col == 3 ? -1.0 : (col == 9 ? 1.0 : 0.0);
    }
}

/* Per-element scratch of the local iteration (ForceBeamColumn3dScratch). */
typedef struct {
    double vs_e[OR_MAX_NIP], vs_ky[OR_MAX_NIP], vs_kz[OR_MAX_NIP];
    double sub_e[OR_MAX_NIP], sub_ky[OR_MAX_NIP], sub_kz[OR_MAX_NIP];
    double sub_n[OR_MAX_NIP], sub_my[OR_MAX_NIP], sub_mz[OR_MAX_NIP];
    double sub_f[OR_MAX_NIP][6];
    double k_basic[25];
} fbc3_scratch;

/* :1149-1173 */
static void restore_predictor(double *q, int nip, const double *qa, const fbc3_scratch *sc) {
    for (int k = 0; k < 5; ++k) q[k] = qa[k];
    for (int ip = 0; ip < nip; ++ip) {
        q[5 + ip] = sc->vs_e[ip];
        q[5 + nip + ip] = sc->vs_ky[ip];
        q[5 + 2 * nip + ip] = sc->vs_kz[ip];
    }
}

/* b(xi) rows of the five basic forces: (N, My, Mz) components */
static void basic_b(int a, double b_mi, double b_mj, double *bn, double *bmy, double *bmz) {
    *bn = 0.0;
    *bmy = 0.0;
    *bmz = 0.0;
    if (a == 0)
        *bn = 1.0;
    else if (a == 1)
        *bmz = b_mi;
    else if (a == 2)
        *bmz = b_mj;
    else if (a == 3)
        *bmy = b_mi;
    else
        *bmy = b_mj;
}

/* :1203-1607; returns 1 on convergence with q_out[5] */
static int try_increment(const or_elem3d *el, double L, double *q, fbc3_scratch *sc, const double *base_v,
                         const double *dv, int scheme, double *q_out) {
    const int nip = el->nip;
    double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
    double target[5];
    for (int k = 0; k < 5; ++k) target[k] = base_v[k] + dv[k];
    double init_f[OR_MAX_NIP][6];
    for (int ip = 0; ip < nip; ++ip) { /* :1263-1322 */
        const double e0 = q[5 + ip], ky = q[5 + nip + ip], kz = q[5 + 2 * nip + ip];
        sc->vs_e[ip] = e0;
        sc->vs_ky[ip] = ky;
        sc->vs_kz[ip] = kz;
        double r[9], f6[6];
        or_sec3d_set_trial(el->sec, el->first_inst + ip, e0, ky, kz, r);
        if (el->stats) el->stats->section_evals++;
        if (!sec_flex(r, f6)) return 0;
        sc->sub_e[ip] = e0;
        sc->sub_ky[ip] = ky;
        sc->sub_kz[ip] = kz;
        sc->sub_n[ip] = r[0];
        sc->sub_my[ip] = r[1];
        sc->sub_mz[ip] = r[2];
        for (int k = 0; k < 6; ++k) sc->sub_f[ip][k] = f6[k];
        or_sec3d_set_trial(el->sec, el->first_inst + ip, 0.0, 0.0, 0.0, r);
        if (el->stats) el->stats->section_evals++;
        if (!sec_flex(r, init_f[ip])) return 0;
    }
    double pf[25];
    for (int k = 0; k < 25; ++k) pf[k] = 0.0;
    for (int ip = 0; ip < nip; ++ip) { /* :1323-1356 */
        const double xi = el->xi[ip], wL = el->w[ip] * L, b_mi = xi - 1.0, b_mj = xi;
        const double *f = sc->sub_f[ip];
        const double f11 = f[0], f12 = f[1], f13 = f[2], f22 = f[3], f23 = f[4], f33 = f[5];
        pf[0] += wL * f11;
        pf[1] += wL * f13 * b_mi;
        pf[2] += wL * f13 * b_mj;
        pf[3] += wL * f12 * b_mi;
        pf[4] += wL * f12 * b_mj;
        pf[5] += wL * b_mi * f13;
        pf[6] += wL * b_mi * f33 * b_mi;
        pf[7] += wL * b_mi * f33 * b_mj;
        pf[8] += wL * b_mi * f23 * b_mi;
        pf[9] += wL * b_mi * f23 * b_mj;
        pf[10] += wL * b_mj * f13;
        pf[11] += wL * b_mj * f33 * b_mi;
        pf[12] += wL * b_mj * f33 * b_mj;
        pf[13] += wL * b_mj * f23 * b_mi;
        pf[14] += wL * b_mj * f23 * b_mj;
        pf[15] += wL * b_mi * f12;
        pf[16] += wL * b_mi * f23 * b_mi;
        pf[17] += wL * b_mi * f23 * b_mj;
        pf[18] += wL * b_mi * f22 * b_mi;
        pf[19] += wL * b_mi * f22 * b_mj;
        pf[20] += wL * b_mj * f12;
        pf[21] += wL * b_mj * f23 * b_mi;
        pf[22] += wL * b_mj * f23 * b_mj;
        pf[23] += wL * b_mj * f22 * b_mi;
        pf[24] += wL * b_mj * f22 * b_mj;
    }
    if (or_invert5(pf, sc->k_basic)) { /* :1357-1393 */
        const double *k = sc->k_basic;
        q0 += k[0] * dv[0] + k[1] * dv[1] + k[2] * dv[2] + k[3] * dv[3] + k[4] * dv[4];
        q1 += k[5] * dv[0] + k[6] * dv[1] + k[7] * dv[2] + k[8] * dv[3] + k[9] * dv[4];
        q2 += k[10] * dv[0] + k[11] * dv[1] + k[12] * dv[2] + k[13] * dv[3] + k[14] * dv[4];
        q3 += k[15] * dv[0] + k[16] * dv[1] + k[17] * dv[2] + k[18] * dv[3] + k[19] * dv[4];
        q4 += k[20] * dv[0] + k[21] * dv[1] + k[22] * dv[2] + k[23] * dv[3] + k[24] * dv[4];
    }
    const int n_iters = scheme == 1 ? 100 : 10;
    for (int it = 0; it < n_iters; ++it) { /* :1395-1605 */
        if (el->stats) el->stats->local_iters++;
        double fb[25], vf[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < 25; ++k) fb[k] = 0.0;
        const int use_init = scheme == 1 || (scheme == 2 && it == 0);
        for (int ip = 0; ip < nip; ++ip) {
            const double xi = el->xi[ip], wL = el->w[ip] * L, b_mi = xi - 1.0, b_mj = xi;
            const double axial_target = q0 + 0.0;
            const double mz_target = b_mi * q1 + b_mj * q2 + 0.0;
            const double my_target = b_mi * q3 + b_mj * q4 + 0.0;
            const double dss0 = axial_target - sc->sub_n[ip];
            const double dss1 = my_target - sc->sub_my[ip];
            const double dss2 = mz_target - sc->sub_mz[ip];
            const double *sf = use_init ? init_f[ip] : sc->sub_f[ip];
            const double s11 = sf[0], s12 = sf[1], s13 = sf[2], s22 = sf[3], s23 = sf[4], s33 = sf[5];
            const double dvs0 = s11 * dss0 + s12 * dss1 + s13 * dss2;
            const double dvs1 = s12 * dss0 + s22 * dss1 + s23 * dss2;
            const double dvs2 = s13 * dss0 + s23 * dss1 + s33 * dss2;
            const double v0 = sc->sub_e[ip] + dvs0, v1 = sc->sub_ky[ip] + dvs1, v2 = sc->sub_kz[ip] + dvs2;
            double r[9], f6[6];
            or_sec3d_set_trial(el->sec, el->first_inst + ip, v0, v1, v2, r);
            if (el->stats) el->stats->section_evals++;
            sc->sub_e[ip] = v0;
            sc->sub_ky[ip] = v1;
            sc->sub_kz[ip] = v2;
            sc->sub_n[ip] = r[0];
            sc->sub_my[ip] = r[1];
            sc->sub_mz[ip] = r[2];
            if (!sec_flex(r, f6)) return 0;
            for (int k = 0; k < 6; ++k) sc->sub_f[ip][k] = f6[k];
            const double dr0 = axial_target - sc->sub_n[ip];
            const double dr1 = my_target - sc->sub_my[ip];
            const double dr2 = mz_target - sc->sub_mz[ip];
            const double e0 = f6[0] * dr0 + f6[1] * dr1 + f6[2] * dr2;
            const double e1 = f6[1] * dr0 + f6[3] * dr1 + f6[4] * dr2;
            const double e2 = f6[2] * dr0 + f6[4] * dr1 + f6[5] * dr2;
            vf[0] += wL * (v0 + e0);
            vf[1] += wL * b_mi * (v2 + e2);
            vf[2] += wL * b_mj * (v2 + e2);
            vf[3] += wL * b_mi * (v1 + e1);
            vf[4] += wL * b_mj * (v1 + e1);
            for (int a = 0; a < 5; ++a) {
                double an, amy, amz;
                basic_b(a, b_mi, b_mj, &an, &amy, &amz);
                for (int b = 0; b < 5; ++b) {
                    double bn, bmy, bmz;
                    basic_b(b, b_mi, b_mj, &bn, &bmy, &bmz);
                    const double dEps = f6[0] * bn + f6[1] * bmy + f6[2] * bmz;
                    const double dKy = f6[1] * bn + f6[3] * bmy + f6[4] * bmz;
                    const double dKz = f6[2] * bn + f6[4] * bmy + f6[5] * bmz;
                    fb[a * 5 + b] += wL * (an * dEps + amy * dKy + amz * dKz);
                }
            }
        }
        if (!or_invert5(fb, sc->k_basic)) return 0;
        double res[5], dq[5];
        for (int k = 0; k < 5; ++k) res[k] = target[k] - vf[k];
        const double *k = sc->k_basic;
        for (int r = 0; r < 5; ++r)
            dq[r] = k[r * 5] * res[0] + k[r * 5 + 1] * res[1] + k[r * 5 + 2] * res[2] + k[r * 5 + 3] * res[3] +
                    k[r * 5 + 4] * res[4];
        const double work =
            fabs(res[0] * dq[0] + res[1] * dq[1] + res[2] * dq[2] + res[3] * dq[3] + res[4] * dq[4]);
        q0 += dq[0];
        q1 += dq[1];
        q2 += dq[2];
        q3 += dq[3];
        q4 += dq[4];
        if (work < 1.0e-12) {
            q[0] = q0;
            q[1] = q1;
            q[2] = q2;
            q[3] = q3;
            q[4] = q4;
            for (int ip = 0; ip < nip; ++ip) {
                q[5 + ip] = sc->sub_e[ip];
                q[5 + nip + ip] = sc->sub_ky[ip];
                q[5 + 2 * nip + ip] = sc->sub_kz[ip];
                sc->vs_e[ip] = sc->sub_e[ip];
                sc->vs_ky[ip] = sc->sub_ky[ip];
                sc->vs_kz[ip] = sc->sub_kz[ip];
            }
            q_out[0] = q0;
            q_out[1] = q1;
            q_out[2] = q2;
            q_out[3] = q3;
            q_out[4] = q4;
            return 1;
        }
    }
    return 0;
}

/* force_beam_column3d_fiber_global_tangent_and_internal (:2549-2990).
 * q_state = [q0..q4 | eps0[nip] | kappa_y[nip] | kappa_z[nip]] (trial,
 * updated in place).  Returns 0, or -2 when the element compatibility does
 * not converge (where the reference aborts, :2846). */
int or_fbc3d_state(const or_elem3d *el, double *q_state, const double u_global[12], double k_out[144],
                   double f_out[12]) {
    const int nip = el->nip;
    const double dx = el->x2 - el->x1, dy = el->y2 - el->y1, dz = el->z2 - el->z1;
    const double L = sqrt(dx * dx + dy * dy + dz * dz);
    if (L == 0.0) return -1;
    double R[9];
    if (or_beam3d_rotation(el->x1, el->y1, el->z1, el->x2, el->y2, el->z2, el->has_vecxz, el->vecxz, R)) return -1;
    double ul[12]; /* :340-355 */
    for (int i = 0; i < 12; ++i) ul[i] = 0.0;
    for (int blk = 0; blk < 4; ++blk) {
        const int o = 3 * blk;
        const double g0 = u_global[o], g1 = u_global[o + 1], g2 = u_global[o + 2];
        for (int i = 0; i < 3; ++i) ul[o + i] = R[3 * i] * g0 + R[3 * i + 1] * g1 + R[3 * i + 2] * g2;
    }
    /* :1610-1622 */
    const double chord_z = (ul[7] - ul[1]) / L, chord_y = (ul[8] - ul[2]) / L;
    const double vb[5] = {ul[6] - ul[0], ul[5] - chord_z, ul[11] - chord_z, ul[4] + chord_y, ul[10] + chord_y};
    const double torsion_basic = ul[9] - ul[3];
    fbc3_scratch sc;
    memset(&sc, 0, sizeof(sc));
    /* :1176-1200 */
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int ip = 0; ip < nip; ++ip) {
        const double xi = el->xi[ip], wL = el->w[ip] * L, b_mi = xi - 1.0, b_mj = xi;
        acc[0] += wL * q_state[5 + ip];
        acc[1] += wL * b_mi * q_state[5 + 2 * nip + ip];
        acc[2] += wL * b_mj * q_state[5 + 2 * nip + ip];
        acc[3] += wL * b_mi * q_state[5 + nip + ip];
        acc[4] += wL * b_mj * q_state[5 + nip + ip];
    }
    double qa[5];
    for (int k = 0; k < 5; ++k) qa[k] = q_state[k];
    for (int ip = 0; ip < nip; ++ip) {
        sc.vs_e[ip] = q_state[5 + ip];
        sc.vs_ky[ip] = q_state[5 + nip + ip];
        sc.vs_kz[ip] = q_state[5 + 2 * nip + ip];
    }
    /* :2733-2831 */
    int converged = 0;
    double rem[5], att[5];
    for (int k = 0; k < 5; ++k) rem[k] = att[k] = vb[k] - acc[k];
    int num_subdivide = 1;
    while (num_subdivide <= 4) {
        double target[5];
        for (int k = 0; k < 5; ++k) target[k] = acc[k] + att[k];
        int success = 0;
        for (int scheme = 0; scheme < 3; ++scheme) {
            restore_predictor(q_state, nip, qa, &sc);
            double qs[5];
            if (!try_increment(el, L, q_state, &sc, acc, att, scheme, qs)) continue;
            for (int k = 0; k < 5; ++k) {
                acc[k] = target[k];
                qa[k] = qs[k];
            }
            for (int ip = 0; ip < nip; ++ip) {
                sc.vs_e[ip] = q_state[5 + ip];
                sc.vs_ky[ip] = q_state[5 + nip + ip];
                sc.vs_kz[ip] = q_state[5 + 2 * nip + ip];
            }
            for (int k = 0; k < 5; ++k) rem[k] = vb[k] - acc[k];
            const double m01 = fmax(fabs(rem[0]), fabs(rem[1]));
            const double m23 = fmax(fabs(rem[2]), fabs(rem[3]));
            const double rn = fmax(m01, fmax(m23, fabs(rem[4])));
            if (rn <= 1.0e-12) {
                converged = 1;
            } else {
                for (int k = 0; k < 5; ++k) att[k] = rem[k];
                num_subdivide = 1;
            }
            success = 1;
            break;
        }
        if (converged) break;
        if (success) continue;
        for (int k = 0; k < 5; ++k) att[k] /= 10.0;
        num_subdivide += 1;
        if (el->stats) el->stats->subdivisions++;
    }
    if (!converged) {
        restore_predictor(q_state, nip, qa, &sc);
        if (el->stats) el->stats->unconverged++;
        return -2;
    }
    const double q0 = qa[0], q1 = qa[1], q2 = qa[2], q3 = qa[3], q4 = qa[4];
    const double inv_L = 1.0 / L;
    double kl[144];
    for (int k = 0; k < 144; ++k) kl[k] = 0.0;
    for (int a = 0; a < 12; ++a) /* :2867-2893 */
        for (int b = 0; b < 12; ++b)
            for (int i = 0; i < 5; ++i)
                for (int j = 0; j < 5; ++j)
                    kl[a * 12 + b] += basic_coeff(i, a, inv_L) * sc.k_basic[i * 5 + j] * basic_coeff(j, b, inv_L);
    if (el->G > 0.0 && el->J > 0.0) { /* :2894-2912 */
        const double torsion_k = el->G * el->J / L;
        for (int a = 0; a < 12; ++a)
            for (int b = 0; b < 12; ++b)
                kl[a * 12 + b] += basic_coeff(5, a, inv_L) * torsion_k * basic_coeff(5, b, inv_L);
    }
    if (el->transf == OR_TRANSF_PDELTA) { /* :2914-2922 */
        kl[1 * 12 + 1] += q0 / L;
        kl[1 * 12 + 7] -= q0 / L;
        kl[7 * 12 + 1] -= q0 / L;
        kl[7 * 12 + 7] += q0 / L;
        kl[2 * 12 + 2] += q0 / L;
        kl[2 * 12 + 8] -= q0 / L;
        kl[8 * 12 + 2] -= q0 / L;
        kl[8 * 12 + 8] += q0 / L;
    }
    double fl[12]; /* :2924-2953 */
    const double q5 = el->G > 0.0 && el->J > 0.0 ? (el->G * el->J / L) * torsion_basic : 0.0;
    for (int a = 0; a < 12; ++a) {
        fl[a] = 0.0;
        fl[a] += basic_coeff(0, a, inv_L) * q0;
        fl[a] += basic_coeff(1, a, inv_L) * q1;
        fl[a] += basic_coeff(2, a, inv_L) * q2;
        fl[a] += basic_coeff(3, a, inv_L) * q3;
        fl[a] += basic_coeff(4, a, inv_L) * q4;
        fl[a] += basic_coeff(5, a, inv_L) * q5;
    }
    fl[0] += 0.0; /* fixed-end reactions of element loads (none) */
    fl[1] += 0.0;
    fl[2] += 0.0;
    fl[7] += 0.0;
    fl[8] += 0.0;
    double work[144]; /* :836-867 */
    for (int rb = 0; rb < 4; ++rb)
        for (int cb = 0; cb < 4; ++cb)
            for (int i = 0; i < 3; ++i) {
                const int base = (3 * rb + i) * 12 + 3 * cb;
                const double s0 = kl[base], s1 = kl[base + 1], s2 = kl[base + 2];
                for (int j = 0; j < 3; ++j) work[base + j] = s0 * R[j] + s1 * R[3 + j] + s2 * R[6 + j];
            }
    for (int rb = 0; rb < 4; ++rb)
        for (int col = 0; col < 12; ++col)
            for (int i = 0; i < 3; ++i)
                k_out[(3 * rb + i) * 12 + col] = R[i] * work[(3 * rb) * 12 + col] +
                                                 R[3 + i] * work[(3 * rb + 1) * 12 + col] +
                                                 R[6 + i] * work[(3 * rb + 2) * 12 + col];
    for (int blk = 0; blk < 4; ++blk) { /* :371-385 */
        const int o = 3 * blk;
        const double l0 = fl[o], l1 = fl[o + 1], l2 = fl[o + 2];
        for (int i = 0; i < 3; ++i) f_out[o + i] = R[i] * l0 + R[3 + i] * l1 + R[6 + i] * l2;
    }
    return 0;
}
