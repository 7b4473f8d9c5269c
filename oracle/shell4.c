/* Elastic four-node shell of the oracle (test infrastructure only).
 *
 * Restates shell4_mindlin_stiffness (src/mojo/elements/shell4.mojo:10-330):
 * the MITC4 shell with an ElasticMembranePlateSection (E, nu, h), 2 x 2
 * Gauss points, membrane / bending / assumed transverse shear and the drilling
 * term G h, returned as the 24 x 24 stiffness in global coordinates (node
 * order as given, DOFs ux uy uz rx ry rz per node).  Every expression keeps
 * the reference's operation order; the helper vectors follow
 * elements/utils.mojo:48-74. */
#include <math.h>
#include <string.h>

#include "oracle_internal.h"

static int or_normalize(double x, double y, double z, double o[3]) {
    double n = sqrt(x * x + y * y + z * z);
    if (n == 0.0) return -1;
    o[0] = x / n;
    o[1] = y / n;
    o[2] = z / n;
    return 0;
}

int or_shell4_stiffness(double E, double nu, double h, const double x[4], const double y[4], const double z[4],
                        double K[576]) {
    /* :22-49 basis g1, g2 (Gram-Schmidt), g3 = g1 x g2 */
    double v1x = (x[2] + x[1] - x[3] - x[0]) * 0.5;
    double v1y = (y[2] + y[1] - y[3] - y[0]) * 0.5;
    double v1z = (z[2] + z[1] - z[3] - z[0]) * 0.5;
    double v2x = (x[3] + x[2] - x[1] - x[0]) * 0.5;
    double v2y = (y[3] + y[2] - y[1] - y[0]) * 0.5;
    double v2z = (z[3] + z[2] - z[1] - z[0]) * 0.5;
    double g1[3], g2[3], g3[3];
    if (or_normalize(v1x, v1y, v1z, g1)) return -1;
    double alpha = v2x * g1[0] + v2y * g1[1] + v2z * g1[2];
    v2x -= alpha * g1[0];
    v2y -= alpha * g1[1];
    v2z -= alpha * g1[2];
    if (or_normalize(v2x, v2y, v2z, g2)) return -1;
    g3[0] = g1[1] * g2[2] - g1[2] * g2[1];
    g3[1] = g1[2] * g2[0] - g1[0] * g2[2];
    g3[2] = g1[0] * g2[1] - g1[1] * g2[0];
    /* :53-61 local in-plane coordinates */
    double xl[2][4];
    for (int i = 0; i < 4; ++i) {
        xl[0][i] = x[i] * g1[0] + y[i] * g1[1] + z[i] * g1[2];
        xl[1][i] = x[i] * g2[0] + y[i] * g2[1] + z[i] * g2[2];
    }
    /* :63-84 section matrices */
    double factor = E / (1.0 - nu * nu);
    double Dm[3][3] = {{factor, factor * nu, 0.0}, {factor * nu, factor, 0.0}, {0.0, 0.0, factor * (1.0 - nu) / 2.0}};
    double bend_factor = E * h * h * h / (12.0 * (1.0 - nu * nu));
    double Db[3][3] = {{bend_factor, bend_factor * nu, 0.0},
                       {bend_factor * nu, bend_factor, 0.0},
                       {0.0, 0.0, bend_factor * (1.0 - nu) / 2.0}};
    double G = E / (2.0 * (1.0 + nu));
    double kappa = 5.0 / 6.0;
    double Ds = kappa * G * h;
    double Ktt = G * h;
    double dd[8][8];
    memset(dd, 0, sizeof(dd));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            dd[i][j] = Dm[i][j] * h;
            dd[i + 3][j + 3] = -Db[i][j];
        }
    dd[6][6] = Ds;
    dd[7][7] = Ds;
    /* :86-126 MITC4 shear interpolation */
    double dx34 = xl[0][2] - xl[0][3], dy34 = xl[1][2] - xl[1][3];
    double dx21 = xl[0][1] - xl[0][0], dy21 = xl[1][1] - xl[1][0];
    double dx32 = xl[0][2] - xl[0][1], dy32 = xl[1][2] - xl[1][1];
    double dx41 = xl[0][3] - xl[0][0], dy41 = xl[1][3] - xl[1][0];
    double Gm[4][12];
    memset(Gm, 0, sizeof(Gm));
    const double q4 = 0.25;
    Gm[0][0] = -0.5;
    Gm[0][1] = -dy41 * q4;
    Gm[0][2] = dx41 * q4;
    Gm[0][9] = 0.5;
    Gm[0][10] = -dy41 * q4;
    Gm[0][11] = dx41 * q4;
    Gm[1][0] = -0.5;
    Gm[1][1] = -dy21 * q4;
    Gm[1][2] = dx21 * q4;
    Gm[1][3] = 0.5;
    Gm[1][4] = -dy21 * q4;
    Gm[1][5] = dx21 * q4;
    Gm[2][3] = -0.5;
    Gm[2][4] = -dy32 * q4;
    Gm[2][5] = dx32 * q4;
    Gm[2][6] = 0.5;
    Gm[2][7] = -dy32 * q4;
    Gm[2][8] = dx32 * q4;
    Gm[3][6] = 0.5;
    Gm[3][7] = -dy34 * q4;
    Gm[3][8] = dx34 * q4;
    Gm[3][9] = -0.5;
    Gm[3][10] = -dy34 * q4;
    Gm[3][11] = dx34 * q4;
    double Ax = -xl[0][0] + xl[0][1] + xl[0][2] - xl[0][3];
    double Bx = xl[0][0] - xl[0][1] + xl[0][2] - xl[0][3];
    double Cx = -xl[0][0] - xl[0][1] + xl[0][2] + xl[0][3];
    double Ay = -xl[1][0] + xl[1][1] + xl[1][2] - xl[1][3];
    double By = xl[1][0] - xl[1][1] + xl[1][2] - xl[1][3];
    double Cy = -xl[1][0] - xl[1][1] + xl[1][2] + xl[1][3];
    double alph = atan2(Ay, Ax);
    double beta = 0.5 * 3.141592653589793 - atan2(Cx, Cy);
    double Rot[2][2] = {{sin(beta), -sin(alph)}, {-cos(beta), cos(alph)}};
    /* :142-146 Gauss points and node signs */
    const double r3 = 1.0 / sqrt(3.0);
    const double sg[4] = {-r3, r3, r3, -r3}, tg[4] = {-r3, -r3, r3, r3};
    const double wg[4] = {1.0, 1.0, 1.0, 1.0};
    const double s[4] = {-0.5, 0.5, 0.5, -0.5}, t[4] = {-0.5, -0.5, 0.5, 0.5};
    const double Gmem[2][3] = {{g1[0], g1[1], g1[2]}, {g2[0], g2[1], g2[2]}};
    double Gshear[3][6];
    memset(Gshear, 0, sizeof(Gshear));
    for (int q = 0; q < 3; ++q) {
        Gshear[0][q] = g3[q];
        Gshear[1][3 + q] = g1[q];
        Gshear[2][3 + q] = g2[q];
    }
    for (int k = 0; k < 576; ++k) K[k] = 0.0;
    for (int gp = 0; gp < 4; ++gp) {
        double ss = sg[gp], tt = tg[gp];
        double shp[3][4];
        for (int i = 0; i < 4; ++i) {
            shp[2][i] = (0.5 + s[i] * ss) * (0.5 + t[i] * tt);
            shp[0][i] = s[i] * (0.5 + t[i] * tt);
            shp[1][i] = t[i] * (0.5 + s[i] * ss);
        }
        double xs[2][2];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                xs[i][j] = 0.0;
                for (int k = 0; k < 4; ++k) xs[i][j] += xl[i][k] * shp[j][k];
            }
        double xsj = xs[0][0] * xs[1][1] - xs[0][1] * xs[1][0];
        if (xsj == 0.0) return -2;
        double jinv = 1.0 / xsj;
        double sx00 = xs[1][1] * jinv, sx11 = xs[0][0] * jinv, sx01 = -xs[0][1] * jinv, sx10 = -xs[1][0] * jinv;
        for (int i = 0; i < 4; ++i) {
            double temp = shp[0][i] * sx00 + shp[1][i] * sx10;
            shp[1][i] = shp[0][i] * sx01 + shp[1][i] * sx11;
            shp[0][i] = temp;
        }
        double dvol = wg[gp] * xsj;
        double Ms[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
        Ms[1][0] = 1.0 - ss;
        Ms[0][1] = 1.0 - tt;
        Ms[1][2] = 1.0 + ss;
        Ms[0][3] = 1.0 + tt;
        double Bsv[2][12], Bs[2][12];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 12; ++j) {
                double sum = 0.0;
                for (int k = 0; k < 4; ++k) sum += Ms[i][k] * Gm[k][j];
                Bsv[i][j] = sum;
            }
        double r1 = Cx + ss * Bx;
        double r3v = Cy + ss * By;
        r1 = sqrt(r1 * r1 + r3v * r3v);
        double r2 = Ax + tt * Bx;
        r3v = Ay + tt * By;
        r2 = sqrt(r2 * r2 + r3v * r3v);
        for (int j = 0; j < 12; ++j) {
            Bsv[0][j] = Bsv[0][j] * r1 / (8.0 * xsj);
            Bsv[1][j] = Bsv[1][j] * r2 / (8.0 * xsj);
        }
        for (int j = 0; j < 12; ++j) {
            Bs[0][j] = Rot[0][0] * Bsv[0][j] + Rot[0][1] * Bsv[1][j];
            Bs[1][j] = Rot[1][0] * Bsv[0][j] + Rot[1][1] * Bsv[1][j];
        }
        /* :253-303 per-node strain-displacement blocks */
        double saveB[4][8][6], Bdrill[4][6];
        for (int j = 0; j < 4; ++j) {
            double Bmem[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}}, Bb[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
            Bmem[0][0] = shp[0][j];
            Bmem[1][1] = shp[1][j];
            Bmem[2][0] = shp[1][j];
            Bmem[2][1] = shp[0][j];
            Bb[0][1] = -shp[0][j];
            Bb[1][0] = shp[1][j];
            Bb[2][0] = shp[0][j];
            Bb[2][1] = -shp[1][j];
            double Bshear[2][3];
            for (int p = 0; p < 3; ++p) {
                Bshear[0][p] = Bs[0][j * 3 + p];
                Bshear[1][p] = Bs[1][j * 3 + p];
            }
            double BmemShell[3][3], BbendShell[3][3], BshearShell[2][6];
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) {
                    double sum_m = 0.0, sum_b = 0.0;
                    for (int r = 0; r < 2; ++r) {
                        sum_m += Bmem[p][r] * Gmem[r][q];
                        sum_b += Bb[p][r] * Gmem[r][q];
                    }
                    BmemShell[p][q] = sum_m;
                    BbendShell[p][q] = sum_b;
                }
            for (int p = 0; p < 2; ++p)
                for (int q = 0; q < 6; ++q) {
                    double sum_s = 0.0;
                    for (int r = 0; r < 3; ++r) sum_s += Bshear[p][r] * Gshear[r][q];
                    BshearShell[p][q] = sum_s;
                }
            memset(saveB[j], 0, sizeof(saveB[j]));
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) {
                    saveB[j][p][q] = BmemShell[p][q];
                    saveB[j][p + 3][q + 3] = BbendShell[p][q];
                }
            for (int p = 0; p < 2; ++p)
                for (int q = 0; q < 6; ++q) saveB[j][p + 6][q] = BshearShell[p][q];
            double B1 = -0.5 * shp[1][j], B2 = 0.5 * shp[0][j], B6 = -shp[2][j];
            Bdrill[j][0] = B1 * g1[0] + B2 * g2[0];
            Bdrill[j][1] = B1 * g1[1] + B2 * g2[1];
            Bdrill[j][2] = B1 * g1[2] + B2 * g2[2];
            Bdrill[j][3] = B6 * g3[0];
            Bdrill[j][4] = B6 * g3[1];
            Bdrill[j][5] = B6 * g3[2];
        }
        /* :305-330 K += B_j^T D B_k dvol + drilling */
        for (int j = 0; j < 4; ++j) {
            double BJ[8][6], BJtranD[6][8], BdrillJ[6];
            memcpy(BJ, saveB[j], sizeof(BJ));
            for (int p = 3; p < 6; ++p)
                for (int q = 3; q < 6; ++q) BJ[p][q] *= -1.0;
            for (int p = 0; p < 6; ++p)
                for (int q = 0; q < 8; ++q) {
                    double sum = 0.0;
                    for (int r = 0; r < 8; ++r) sum += BJ[r][p] * dd[r][q];
                    BJtranD[p][q] = sum * dvol;
                }
            for (int p = 0; p < 6; ++p) BdrillJ[p] = Bdrill[j][p] * (Ktt * dvol);
            for (int k = 0; k < 4; ++k)
                for (int p = 0; p < 6; ++p)
                    for (int q = 0; q < 6; ++q) {
                        double sum = 0.0;
                        for (int r = 0; r < 8; ++r) sum += BJtranD[p][r] * saveB[k][r][q];
                        K[(j * 6 + p) * 24 + k * 6 + q] += sum + BdrillJ[p] * Bdrill[k][q];
                    }
        }
    }
    return 0;
}
