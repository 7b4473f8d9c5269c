/* dispBeamColumn2d state determination (test infrastructure only): a
 * restatement of disp_beam_column2d_global_tangent_and_internal
 * (elements/disp_beam_column2d.mojo:114-254) for the Linear transformation
 * without element loads.  Every expression keeps the reference's operation
 * order. */
#include <math.h>

#include "oracle_internal.h"

/* disp_beam_column2d.mojo:25-31 */
static double axial_shape(int idx, double inv_L) {
    if (idx == 0) return -inv_L;
    if (idx == 3) return inv_L;
    return 0.0;
}

/* disp_beam_column2d.mojo:35-46 */
static double curvature_shape(int idx, double inv_L, double inv_L_sq, double xi) {
    if (idx == 1) return (-6.0 + 12.0 * xi) * inv_L_sq;
    if (idx == 2) return (-4.0 + 6.0 * xi) * inv_L;
    if (idx == 4) return (6.0 - 12.0 * xi) * inv_L_sq;
    if (idx == 5) return (-2.0 + 6.0 * xi) * inv_L;
    return 0.0;
}

int or_dbc2d_state(const or_elem2d *el, const double u_global[6], double k_out[36], double f_out[6]) {
    const double L = el->L, c = el->dc, s = el->ds;
    const double inv_L = 1.0 / L;
    const double inv_L_sq = inv_L * inv_L;
    double ul[6]; /* elements/utils.mojo:142-153 */
    ul[0] = c * u_global[0] + s * u_global[1];
    ul[1] = -s * u_global[0] + c * u_global[1];
    ul[2] = u_global[2];
    ul[3] = c * u_global[3] + s * u_global[4];
    ul[4] = -s * u_global[3] + c * u_global[4];
    ul[5] = u_global[5];
    for (int k = 0; k < 36; ++k) k_out[k] = 0.0;
    for (int k = 0; k < 6; ++k) f_out[k] = 0.0;
    for (int ip = 0; ip < el->nip; ++ip) { /* :167-205 */
        const double xi = el->xi[ip], weight = el->w[ip];
        const double b1 = curvature_shape(1, inv_L, inv_L_sq, xi);
        const double b2 = curvature_shape(2, inv_L, inv_L_sq, xi);
        const double b4 = curvature_shape(4, inv_L, inv_L_sq, xi);
        const double b5 = curvature_shape(5, inv_L, inv_L_sq, xi);
        const double eps0 = (-inv_L) * ul[0] + inv_L * ul[3];
        const double kappa = b1 * ul[1] + b2 * ul[2] + b4 * ul[4] + b5 * ul[5];
        double r[5];
        or_sec2d_set_trial(el->sec, el->first_inst + ip, eps0, kappa, r);
        el->stats->section_evals++;
        const double N = r[0], Mz = r[1], k11 = r[2], k12 = r[3], k22 = r[4];
        const double wL = weight * L;
        for (int a = 0; a < 6; ++a) {
            const double ba_n = axial_shape(a, inv_L), ba_m = curvature_shape(a, inv_L, inv_L_sq, xi);
            f_out[a] += wL * (ba_n * N + ba_m * Mz);
            for (int b = 0; b < 6; ++b) {
                const double bb_n = axial_shape(b, inv_L), bb_m = curvature_shape(b, inv_L, inv_L_sq, xi);
                const double dN = k11 * bb_n + k12 * bb_m;
                const double dM = k12 * bb_n + k22 * bb_m;
                k_out[6 * a + b] += wL * (ba_n * dN + ba_m * dM);
            }
        }
    }
    /* elements/utils.mojo:171-240: columns 0,1 and 3,4 of every row, then rows 0,1 and 3,4 */
    for (int i = 0; i < 6; ++i) {
        double *row = k_out + 6 * i;
        const double k0 = row[0], k1 = row[1], k3 = row[3], k4 = row[4];
        row[0] = c * k0 - s * k1;
        row[1] = s * k0 + c * k1;
        row[3] = c * k3 - s * k4;
        row[4] = s * k3 + c * k4;
    }
    for (int p = 0; p < 2; ++p) {
        double *r0 = k_out + 6 * (3 * p), *r1 = k_out + 6 * (3 * p + 1);
        double a0[6], a1[6];
        for (int j = 0; j < 6; ++j) {
            a0[j] = r0[j];
            a1[j] = r1[j];
        }
        for (int j = 0; j < 6; ++j) {
            r0[j] = c * a0[j] - s * a1[j];
            r1[j] = s * a0[j] + c * a1[j];
        }
    }
    const double f0 = f_out[0], f1 = f_out[1], f3 = f_out[3], f4 = f_out[4]; /* utils.mojo:156-168 */
    f_out[0] = c * f0 - s * f1;
    f_out[1] = s * f0 + c * f1;
    f_out[3] = c * f3 - s * f4;
    f_out[4] = s * f3 + c * f4;
    return 0;
}

/* One dispBeamColumn2d element between (x1, y1) and (x2, y2) on the section
 * instances first_inst .. first_inst + nip - 1, for the checks in tests/. */
int or_dbc2d_eval(or_sec2d *sec, int first_inst, int nip, const double *xi, const double *w, double x1, double y1,
                  double x2, double y2, const double u_global[6], double k_out[36], double f_out[6]) {
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
    el.type = 1;
    el.dc = dx / el.L;
    el.ds = dy / el.L;
    el.stats = &stats;
    return or_dbc2d_state(&el, u_global, k_out, f_out);
}
