// Uniaxial material laws as gfx950 device functions.
//
// Each law maps (committed history, trial strain) -> (trial history, stress,
// tangent) exactly as Strut's SoA fiber runtime does; it never reads a
// previous trial, so repeated evaluations of one point agree bit for bit.
//   Steel02    materials/uniaxial/steel02.mojo:3-114, fiber2d.mojo:2301-2516
//   Concrete02 materials/uniaxial/concrete02.mojo:3-109, fiber2d.mojo:1785-1972
//   Steel01    materials/uniaxial/ops.mojo:24-53, fiber2d.mojo:1020-1045
//   Concrete01 materials/uniaxial/ops.mojo:56-114, concrete01.mojo:3-70
//   Elastic    fiber2d.mojo:3459-3501 (no runtime state)
// Every expression keeps the reference's operation order; the library is
// built with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "strut_gpu_types.h"

namespace sgpu {

// History carried per fiber slot.  Two doubles (h0, h1) live in the common
// slot arrays; the five aux doubles and kon live in the compact aux arrays
// that only Steel02 and Concrete01 slots own.
//   Concrete02: h0 = ecmin, h1 = dept
//   Steel02:    h0 = epsmin, h1 = epsmax, a = {epspl, epss0, sigs0, epsr, sigr}, kon
//   Steel01:    h0 = eps_p, h1 = alpha
//   Concrete01: h0 = min_strain, h1 = end_strain, a[0] = unload_slope
struct FiberHist {
    double eps, sig, tan, h0, h1;
    double a0, a1, a2, a3, a4;
    int kon;
};

__device__ __forceinline__ void concrete02_compr_envlp(const MatDev &m, double strain, double &sig,
                                                       double &tan) {
    const double fc = m.p[0], epsc0 = m.p[1], fcu = m.p[2], epscu = m.p[3];
    const double ec0 = m.pre[0];
    if (strain >= epsc0) {
        double rat = strain / epsc0;
        sig = fc * rat * (2.0 - rat);
        tan = ec0 * (1.0 - rat);
    } else if (strain > epscu) {
        sig = (fcu - fc) * (strain - epsc0) / (epscu - epsc0) + fc;
        tan = (fcu - fc) / (epscu - epsc0);
    } else {
        sig = fcu;
        tan = 1.0e-10;
    }
}

__device__ __forceinline__ void concrete02_tens_envlp(const MatDev &m, double strain, double &sig,
                                                      double &tan) {
    const double ft = m.p[5], ets = m.p[6];
    const double ec0 = m.pre[0], eps0 = m.pre[1], epsu = m.pre[2];
    if (strain <= eps0) {
        sig = strain * ec0;
        tan = ec0;
    } else if (strain <= epsu) {
        sig = ft - ets * (strain - eps0);
        tan = -ets;
    } else {
        sig = 0.0;
        tan = 1.0e-10;
    }
}

// concrete02.mojo:34-109
__device__ __forceinline__ void concrete02_trial(const MatDev &m, const FiberHist &c, double strain,
                                                 FiberHist &t) {
    const double ec0 = m.pre[0];
    t.h0 = c.h0;  // ecmin
    t.h1 = c.h1;  // dept
    t.eps = strain;
    const double deps = strain - c.eps;
    if (fabs(deps) < 2.220446049250313e-16) {
        t.sig = c.sig;
        t.tan = c.tan;
        return;
    }
    if (strain < c.h0) {
        concrete02_compr_envlp(m, strain, t.sig, t.tan);
        t.h0 = strain;
        return;
    }
    const double epsr = m.pre[3], sigmr = m.pre[4];
    double sigmm, dummy;
    concrete02_compr_envlp(m, c.h0, sigmm, dummy);
    const double er = (sigmm - sigmr) / (c.h0 - epsr);
    const double ept = c.h0 - sigmm / er;
    if (strain <= ept) {
        const double sigmin = sigmm + er * (strain - c.h0);
        const double sigmax = 0.5 * er * (strain - ept);
        double sig = c.sig + ec0 * deps;
        double tan = ec0;
        if (sig <= sigmin) {
            sig = sigmin;
            tan = er;
        }
        if (sig >= sigmax) {
            sig = sigmax;
            tan = 0.5 * er;
        }
        t.sig = sig;
        t.tan = tan;
        return;
    }
    const double epn = ept + c.h1;
    if (strain <= epn) {
        double sicn, dummy2;
        concrete02_tens_envlp(m, c.h1, sicn, dummy2);
        double e = ec0;
        if (c.h1 != 0.0) e = sicn / c.h1;
        t.tan = e;
        t.sig = e * (strain - ept);
    } else {
        concrete02_tens_envlp(m, strain - ept, t.sig, t.tan);
        t.h1 = strain - ept;
    }
}

// steel02.mojo:3-114
__device__ __forceinline__ void steel02_trial(const MatDev &m, const FiberHist &c, double trial_eps,
                                              FiberHist &t) {
    const double Fy = m.p[0], E0 = m.p[1], b = m.p[2], R0 = m.p[3], cR1 = m.p[4], cR2 = m.p[5];
    const double a1 = m.p[6], a2 = m.p[7], a3 = m.p[8], a4 = m.p[9], sigini = m.p[10];
    const double Esh = m.pre[0], epsy = m.pre[1];
    double eps = trial_eps;
    if (sigini != 0.0) eps = trial_eps + m.pre[2];
    t.eps = eps;
    const double deps = eps - c.eps;
    double epsmin = c.h0, epsmax = c.h1;
    double epspl = c.a0, epss0 = c.a1, sigs0 = c.a2, epsr = c.a3, sigr = c.a4;
    int kon = c.kon;
    if (kon == 0 || kon == 3) {
        if (fabs(deps) < 2.220446049250313e-15) {
            t.tan = E0;
            t.sig = sigini;
            t.kon = 3;
            t.h0 = epsmin;
            t.h1 = epsmax;
            t.a0 = epspl;
            t.a1 = epss0;
            t.a2 = sigs0;
            t.a3 = epsr;
            t.a4 = sigr;
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
        epsr = c.eps;
        sigr = c.sig;
        if (c.eps < epsmin) epsmin = c.eps;
        const double d1 = (epsmax - epsmin) / (2.0 * (a4 * epsy));
        const double shft = 1.0 + a3 * pow(d1, 0.8);
        epss0 = (Fy * shft - Esh * epsy * shft - sigr + E0 * epsr) / (E0 - Esh);
        sigs0 = Fy * shft + Esh * (epss0 - epsy * shft);
        epspl = epsmax;
    } else if (kon == 1 && deps < 0.0) {
        kon = 2;
        epsr = c.eps;
        sigr = c.sig;
        if (c.eps > epsmax) epsmax = c.eps;
        const double d1 = (epsmax - epsmin) / (2.0 * (a2 * epsy));
        const double shft = 1.0 + a1 * pow(d1, 0.8);
        epss0 = (-Fy * shft + Esh * epsy * shft - sigr + E0 * epsr) / (E0 - Esh);
        sigs0 = -Fy * shft + Esh * (epss0 + epsy * shft);
        epspl = epsmin;
    }
    const double xi = fabs((epspl - epss0) / epsy);
    const double R = R0 * (1.0 - (cR1 * xi) / (cR2 + xi));
    const double epsrat = (eps - epsr) / (epss0 - epsr);
    const double dum1 = 1.0 + pow(fabs(epsrat), R);
    const double dum2 = pow(dum1, 1.0 / R);
    double sig = b * epsrat + (1.0 - b) * epsrat / dum2;
    sig = sig * (sigs0 - sigr) + sigr;
    double e = b + (1.0 - b) / (dum1 * dum2);
    e = e * (sigs0 - sigr) / (epss0 - epsr);
    t.sig = sig;
    t.tan = e;
    t.h0 = epsmin;
    t.h1 = epsmax;
    t.a0 = epspl;
    t.a1 = epss0;
    t.a2 = sigs0;
    t.a3 = epsr;
    t.a4 = sigr;
    t.kon = kon;
}

// ops.mojo:24-53 (fiber2d.mojo:1020-1045)
__device__ __forceinline__ void steel01_trial(const MatDev &m, const FiberHist &c, double eps,
                                              FiberHist &t) {
    const double Fy = m.p[0], E0 = m.p[1], b = m.p[2];
    const double H = (b * E0) / (1.0 - b);
    const double eps_p = c.h0, alpha = c.h1;
    const double sigma_trial = E0 * (eps - eps_p);
    const double xi = sigma_trial - alpha;
    const double f = fabs(xi) - Fy;
    t.eps = eps;
    if (f <= 0.0) {
        t.sig = sigma_trial;
        t.tan = E0;
        t.h0 = eps_p;
        t.h1 = alpha;
        return;
    }
    const double dg = f / (E0 + H);
    const double sgn = xi < 0.0 ? -1.0 : 1.0;
    t.h0 = eps_p + dg * sgn;
    t.h1 = alpha + H * dg * sgn;
    t.sig = sigma_trial - E0 * dg * sgn;
    t.tan = (E0 * H) / (E0 + H);
}

__device__ __forceinline__ void concrete01_envelope(const MatDev &m, double strain, double &sig,
                                                    double &tan) {
    const double fpc = m.p[0], epsc0 = m.p[1], fpcu = m.p[2], epscu = m.p[3];
    if (strain > epsc0) {
        const double eta = strain / epsc0;
        sig = fpc * (2.0 * eta - eta * eta);
        const double Ec0 = (2.0 * fpc) / epsc0;
        tan = Ec0 * (1.0 - eta);
    } else if (strain > epscu) {
        const double tg = (fpc - fpcu) / (epsc0 - epscu);
        tan = tg;
        sig = fpc + tg * (strain - epsc0);
    } else {
        sig = fpcu;
        tan = 0.0;
    }
}

// ops.mojo:56-114 with concrete01.mojo:20-70
__device__ __forceinline__ void concrete01_trial(const MatDev &m, const FiberHist &c, double eps,
                                                 FiberHist &t) {
    const double fpc = m.p[0], epsc0 = m.p[1], epscu = m.p[3];
    t.eps = eps;
    const double d_strain = eps - c.eps;
    if (fabs(d_strain) < 1.0e-14) {
        t.sig = c.sig;
        t.tan = c.tan;
        t.h0 = c.h0;
        t.h1 = c.h1;
        t.a0 = c.a0;
        return;
    }
    double min_strain = c.h0, end_strain = c.h1, unload_slope = c.a0;
    if (eps > 0.0) {
        t.sig = 0.0;
        t.tan = 0.0;
        t.h0 = min_strain;
        t.h1 = end_strain;
        t.a0 = unload_slope;
        return;
    }
    const double temp_stress = c.sig + unload_slope * eps - unload_slope * c.eps;
    double sig, tan;
    if (eps <= c.eps) {
        if (eps <= min_strain) {
            min_strain = eps;
            concrete01_envelope(m, eps, sig, tan);
            // concrete01.mojo:20-45 (_concrete01_unload)
            double temp_strain = min_strain;
            if (temp_strain < epscu) temp_strain = epscu;
            const double eta = temp_strain / epsc0;
            double ratio = 0.707 * (eta - 2.0) + 0.834;
            if (eta < 2.0) ratio = 0.145 * eta * eta + 0.13 * eta;
            const double es = ratio * epsc0;
            const double temp1 = min_strain - es;
            const double Ec0 = (2.0 * fpc) / epsc0;
            const double temp2 = sig / Ec0;
            if (temp1 > -1.0e-14) {
                end_strain = es;
                unload_slope = Ec0;
            } else if (temp1 <= temp2) {
                end_strain = min_strain - temp1;
                unload_slope = sig / temp1;
            } else {
                end_strain = min_strain - temp2;
                unload_slope = Ec0;
            }
        } else if (eps <= end_strain) {
            tan = unload_slope;
            sig = tan * (eps - end_strain);
        } else {
            sig = 0.0;
            tan = 0.0;
        }
        if (temp_stress > sig) {
            sig = temp_stress;
            tan = unload_slope;
        }
    } else if (temp_stress <= 0.0) {
        sig = temp_stress;
        tan = unload_slope;
    } else {
        sig = 0.0;
        tan = 0.0;
    }
    t.sig = sig;
    t.tan = tan;
    t.h0 = min_strain;
    t.h1 = end_strain;
    t.a0 = unload_slope;
}

}  // namespace sgpu
