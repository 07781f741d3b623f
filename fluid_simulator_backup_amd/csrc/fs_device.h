// fs_device.h -- device-side math shared by the step kernels.
//
// Every expression keeps the operation order of the reference text it cites
// (the library is built with -ffp-contract=off), so results match the FP64
// reference bit for bit wherever no libm call or reduction intervenes.
#pragma once
#include <hip/hip_runtime.h>

#include "fs_layout.h"

#define FS_PI 3.14159265358979323846  // MACROS.h:13

// WENO5-JS one-sided derivatives at c[0] from c[-3..3]
// (ADVECTION_METHOD_3D.h:55-183).
__device__ __forceinline__ void fs_weno5(const double* c, double inv, double eps, double& dmo, double& dpo) {
    double m1 = inv * (c[-2] - c[-3]), m2 = inv * (c[-1] - c[-2]), m3 = inv * (c[0] - c[-1]);
    double m4 = inv * (c[1] - c[0]), m5 = inv * (c[2] - c[1]);
    double p1 = inv * (c[3] - c[2]);
    // p2..p5 equal m5..m2 term by term (same operands, same operation)
    double p2 = m5, p3 = m4, p4 = m3, p5 = m2;
    const double k13 = (double)13 / 12, q = (double)1 / 4;
    double sm1 = k13 * (m1 - 2 * m2 + m3) * (m1 - 2 * m2 + m3) + q * (m1 - 4 * m2 + 3 * m3) * (m1 - 4 * m2 + 3 * m3);
    double sm2 = k13 * (m2 - 2 * m3 + m4) * (m2 - 2 * m3 + m4) + q * (m2 - m4) * (m2 - m4);
    double sm3 = k13 * (m3 - 2 * m4 + m5) * (m3 - 2 * m4 + m5) + q * (3 * m3 - 4 * m4 + m5) * (3 * m3 - 4 * m4 + m5);
    double sp1 = k13 * (p1 - 2 * p2 + p3) * (p1 - 2 * p2 + p3) + q * (p1 - 4 * p2 + 3 * p3) * (p1 - 4 * p2 + 3 * p3);
    double sp2 = k13 * (p2 - 2 * p3 + p4) * (p2 - 2 * p3 + p4) + q * (p2 - p4) * (p2 - p4);
    double sp3 = k13 * (p3 - 2 * p4 + p5) * (p3 - 2 * p4 + p5) + q * (3 * p3 - 4 * p4 + p5) * (3 * p3 - 4 * p4 + p5);
    double am1 = (double)1 / 10 * (double)1 / ((eps + sm1) * (eps + sm1));
    double am2 = (double)6 / 10 * (double)1 / ((eps + sm2) * (eps + sm2));
    double am3 = (double)3 / 10 * (double)1 / ((eps + sm3) * (eps + sm3));
    double summ = am1 + am2 + am3;
    double wm1 = am1 / summ, wm2 = am2 / summ, wm3 = am3 / summ;
    double ap1 = (double)1 / 10 * (double)1 / ((eps + sp1) * (eps + sp1));
    double ap2 = (double)6 / 10 * (double)1 / ((eps + sp2) * (eps + sp2));
    double ap3 = (double)3 / 10 * (double)1 / ((eps + sp3) * (eps + sp3));
    double sump = ap1 + ap2 + ap3;
    double wp1 = ap1 / sump, wp2 = ap2 / sump, wp3 = ap3 / sump;
    dmo = wm1 * (m1 * (double)1 / 3 - m2 * (double)7 / 6 + m3 * (double)11 / 6) +
          wm2 * (m2 * (-(double)1 / 6) + m3 * (double)5 / 6 + m4 * (double)1 / 3) +
          wm3 * (m3 * (double)1 / 3 + m4 * (double)5 / 6 - m5 * (double)1 / 6);
    dpo = wp1 * (p1 * (double)1 / 3 - p2 * (double)7 / 6 + p3 * (double)11 / 6) +
          wp2 * (p2 * (-(double)1 / 6) + p3 * (double)5 / 6 + p4 * (double)1 / 3) +
          wp3 * (p3 * (double)1 / 3 + p4 * (double)5 / 6 - p5 * (double)1 / 6);
}

// ENO3 one-sided derivatives (ADVECTION_METHOD_3D.h:1653-1718).
__device__ __forceinline__ void fs_eno3(const double* c, double dx, double inv, double inv2, double inv3, double& dmo,
                                        double& dpo) {
    double d1n3 = inv * (c[-2] - c[-3]), d1n2 = inv * (c[-1] - c[-2]), d1n1 = inv * (c[0] - c[-1]);
    double d10 = inv * (c[1] - c[0]), d1p1 = inv * (c[2] - c[1]), d1p2 = inv * (c[3] - c[2]);
    double d2n2 = inv2 * (d1n2 - d1n3), d2n1 = inv2 * (d1n1 - d1n2), d20 = inv2 * (d10 - d1n1);
    double d2p1 = inv2 * (d1p1 - d10), d2p2 = inv2 * (d1p2 - d1p1);
    double d3n2 = inv3 * (d2n1 - d2n2), d3n1 = inv3 * (d20 - d2n1), d30 = inv3 * (d2p1 - d20);
    double d3p1 = inv3 * (d2p2 - d2p1);
    double m, p;
    if (fabs(d2n1) <= fabs(d20)) {
        if (fabs(d3n2) <= fabs(d3n1)) m = d1n1 + d2n1 * dx + 2 * d3n2 * dx * dx;
        else m = d1n1 + d2n1 * dx + 2 * d3n1 * dx * dx;
    } else {
        if (fabs(d3n1) <= fabs(d30)) m = d1n1 + d20 * dx - d3n1 * dx * dx;
        else m = d1n1 + d20 * dx - d30 * dx * dx;
    }
    if (fabs(d20) <= fabs(d2p1)) {
        if (fabs(d3n1) <= fabs(d30)) p = d10 - d20 * dx - d3n1 * dx * dx;
        else p = d10 - d20 * dx - d30 * dx * dx;
    } else {
        if (fabs(d30) <= fabs(d3p1)) p = d10 - d2p1 * dx + 2 * d30 * dx * dx;
        else p = d10 - d2p1 * dx + 2 * d3p1 * dx * dx;
    }
    dmo = m;
    dpo = p;
}

// VISCOSITY_3D::HeavisideFunction, VISCOSITY_3D.h:1234-1256.
__device__ __forceinline__ double fs_heaviside_visc(double phi, double eps) {
    double one_over_epsilon = (double)1 / eps, one_over_pi = (double)1 / FS_PI;
    if (phi < -eps) return 0;
    if (phi > eps) return 1;
    return (double)0.5 * ((double)1 + phi * one_over_epsilon + one_over_pi * sin(FS_PI * phi * one_over_epsilon));
}

// PROJECTION_3D::HeavisideFunction, PROJECTION_3D.h:1452-1472.
__device__ __forceinline__ double fs_heaviside_proj(double phi, double eps) {
    double one_over_epsilon = (double)1 / eps, one_over_pi = (double)1 / FS_PI;
    if (phi < -eps) return 0;
    if (phi > eps) return 1;
    return (double)0.5 + phi * one_over_epsilon * (double)0.5 + (double)0.5 * one_over_pi * sin(FS_PI * phi * one_over_epsilon);
}

// Two-phase material constants, in the order the mixing formulas use them
// (VISCOSITY_3D.h:1264-1347, PROJECTION_3D.h:1474-1519).
struct FsMaterial {
    int air_bubble_rising;
    double rho_w, rho_a, mu_w, mu_a;
    __device__ __forceinline__ double density(double H) const {
        return air_bubble_rising ? rho_a + (rho_w - rho_a) * H : rho_w + (rho_a - rho_w) * H;
    }
    __device__ __forceinline__ double viscosity(double H) const {
        return air_bubble_rising ? mu_a + (mu_w - mu_a) * H : mu_w + (mu_a - mu_w) * H;
    }
};

// Godunov case of one axis (ADVECTION_METHOD_3D.h:908-1622).
enum { FS_AX_NONE = 0, FS_AX_M = 1, FS_AX_P = 2, FS_AX_DROP = 3, FS_AX_SM = 4, FS_AX_SP = 5 };
__device__ __forceinline__ int fs_godunov_axis(double sg, double m, double p, double& val) {
    double sm = sg * m, sp = sg * p;
    if (sp >= 0 && sm >= 0) { val = m; return FS_AX_M; }
    if (sp <= 0 && sm <= 0) { val = p; return FS_AX_P; }
    if (sp > 0 && sm < 0) { val = 0; return FS_AX_DROP; }
    if (sp < 0 && sm > 0) {
        double ss = sg * (fabs(p) - fabs(m)) / (p - m);
        if (ss > 0) { val = m; return FS_AX_SM; }
        val = p;
        return FS_AX_SP;
    }
    return FS_AX_NONE;
}
