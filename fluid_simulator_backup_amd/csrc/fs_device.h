// fs_device.h -- device-side math shared by the step kernels.
//
// Every expression keeps the operation order of the reference text it cites
// (the library is built with -ffp-contract=off), so results match the FP64
// reference bit for bit wherever no libm call or reduction intervenes.
#pragma once
#include <hip/hip_runtime.h>

#include "fs_layout.h"

#define FS_PI 3.14159265358979323846  // MACROS.h:13

// Correctly rounded quotients without a division instruction sequence per
// quotient.  With y = RN(1/d), q0 = RN(x y), the remainder x - d q0 taken by
// one fma and q = RN(q0 + r y): Markstein's theorem (IBM J. Res. Dev. 34(1),
// 1990; Muller et al., Handbook of Floating-Point Arithmetic, ch. on
// division) gives q = RN(x/d) whenever q0 is within one ulp of x/d and
// nothing underflows or overflows.
//
// fs_div3/fs_div6: d = 3 or 6, whose RN(1/d) has relative error 2^-54, so
// q0 is within one ulp.  copysign restores the sign of a zero quotient
// (the correction of -0 gives +0).  For finite x of magnitude above 2^-1000.
__device__ __forceinline__ double fs_div_const(double x, double d, double y) {
    double q0 = x * y;
    double r = fma(-d, q0, x);
    return copysign(fma(r, y, q0), x);
}
__device__ __forceinline__ double fs_div3(double x) { return fs_div_const(x, (double)3, (double)1 / 3); }
__device__ __forceinline__ double fs_div6(double x) { return fs_div_const(x, (double)6, (double)1 / 6); }
// a / b for a >= 0, b > 0 with y = RN(1/b) (an IEEE division): RN(a y) may
// be off by up to 1.5 ulps, so one correction brings it within half an ulp
// plus a term of order u^2 and a second one, now under Markstein's
// hypothesis, rounds it correctly.
__device__ __forceinline__ double fs_div_recip(double a, double b, double y) {
    double q0 = a * y;
    double q1 = fma(fma(-b, q0, a), y, q0);
    return fma(fma(-b, q1, a), y, q1);
}

// WENO5-JS one-sided derivatives at c[0] from c[-3..3]
// (ADVECTION_METHOD_3D.h:55-183).  Bit for bit the reference's arithmetic:
// its `m * (T)7 / 6` is (m * 7) / 6 and its `m * (T)1 / 3` is m / 3, both
// correctly rounded divisions, taken here by fs_div3/fs_div6; the weights
// a_k / sum share one reciprocal of the sum (fs_div_recip).
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
    double summ = am1 + am2 + am3, rsumm = (double)1 / summ;
    double wm1 = fs_div_recip(am1, summ, rsumm), wm2 = fs_div_recip(am2, summ, rsumm);
    double wm3 = fs_div_recip(am3, summ, rsumm);
    double ap1 = (double)1 / 10 * (double)1 / ((eps + sp1) * (eps + sp1));
    double ap2 = (double)6 / 10 * (double)1 / ((eps + sp2) * (eps + sp2));
    double ap3 = (double)3 / 10 * (double)1 / ((eps + sp3) * (eps + sp3));
    double sump = ap1 + ap2 + ap3, rsump = (double)1 / sump;
    double wp1 = fs_div_recip(ap1, sump, rsump), wp2 = fs_div_recip(ap2, sump, rsump);
    double wp3 = fs_div_recip(ap3, sump, rsump);
    dmo = wm1 * (fs_div3(m1) - fs_div6(m2 * 7) + fs_div6(m3 * 11)) +
          wm2 * (m2 * (-(double)1 / 6) + fs_div6(m3 * 5) + fs_div3(m4)) +
          wm3 * (fs_div3(m3) + fs_div6(m4 * 5) - fs_div6(m5));
    dpo = wp1 * (fs_div3(p1) - fs_div6(p2 * 7) + fs_div6(p3 * 11)) +
          wp2 * (p2 * (-(double)1 / 6) + fs_div6(p3 * 5) + fs_div3(p4)) +
          wp3 * (fs_div3(p3) + fs_div6(p4 * 5) - fs_div6(p5));
}

// The same WENO5-JS one-sided derivatives in fewer operations, for the
// fast_weno option: not the reference's operation order, so the results
// differ from fs_weno5 in the last bits.  Every rounding is written out
// (explicit fma, the build contracts nothing), so the per-cell form below
// and the z-marching reinitialization sweep, which computes the same
// quantities once per triple and slides them along z, give the same bits.
//
// With first differences d_j = (c[j+1] - c[j]) / dx, j = -3..2, the six
// candidate stencils of the two sides are the triples T_j = (d_j, d_j+1,
// d_j+2), j = -3..0, read forwards (side -) or backwards (side +).  Each
// triple has one smoothness part A = 13/12 (a - 2b + c)^2 and three ways of
// closing it (Jiang & Shu 1996): left (a - 4b + 3c), middle (a - c) and
// right (3a - 4b + c), the right one of a triple being the left one of the
// triple read backwards.  X = (eps + A + f^2/4)^2.  Side - takes
// XL(T-3), XM(T-2), XR(T-1); side + takes XR(T0), XM(T-1), XL(T-2).  The
// weights w_k = (d_k / X_k) / sum_j (d_j / X_j) take the algebraically equal
// form n_k / sum_j n_j with n_1 = d_1 X_2 X_3, n_2 = d_2 X_1 X_3, n_3 = d_3
// X_1 X_2 (all terms positive, so each weight is within a few ulps), and
// both sides share one division: D- = N-/S-, D+ = N+/S+ = (N- S+, N+ S-) /
// (S- S+).  X_k lies in [eps^2, ~1e22] for a level set on any grid, so the
// products stay far inside the FP64 range.
struct FsW5Tri {
    double xl, xm, xr;
};
__device__ __forceinline__ FsW5Tri fs_w5_tri(double a, double b, double c, double eps) {
    const double e = fma(-2.0, b, a) + c;
    const double A = ((double)13 / 12) * (e * e);
    const double fl = fma(3.0, c, fma(-4.0, b, a)), fm = a - c, fr = fma(-4.0, b, fma(3.0, a, c));
    FsW5Tri t;
    t.xl = eps + fma(0.25 * fl, fl, A);
    t.xm = eps + fma(0.25 * fm, fm, A);
    t.xr = eps + fma(0.25 * fr, fr, A);
    t.xl *= t.xl;
    t.xm *= t.xm;
    t.xr *= t.xr;
    return t;
}
// d[0..5] = d_-3 .. d_2; t[0..3] = the triples T-3 .. T0
__device__ __forceinline__ void fs_w5_combine(const double* d, const FsW5Tri* t, double& dmo, double& dpo) {
    const double r3 = (double)1 / 3, r6 = (double)1 / 6, r56 = (double)5 / 6, r76 = (double)7 / 6,
                 r116 = (double)11 / 6;
    const double X1 = t[0].xl, X2 = t[1].xm, X3 = t[2].xr;
    const double nm1 = 0.1 * (X2 * X3), nm2 = 0.6 * (X1 * X3), nm3 = 0.3 * (X1 * X2);
    const double qm1 = fma(r116, d[2], fma(-r76, d[1], r3 * d[0]));
    const double qm2 = fma(r3, d[3], fma(r56, d[2], -r6 * d[1]));
    const double qm3 = fma(-r6, d[4], fma(r56, d[3], r3 * d[2]));
    const double Nm = fma(nm3, qm3, fma(nm2, qm2, nm1 * qm1)), Sm = (nm1 + nm2) + nm3;
    const double Y1 = t[3].xr, Y2 = t[2].xm, Y3 = t[1].xl;
    const double np1 = 0.1 * (Y2 * Y3), np2 = 0.6 * (Y1 * Y3), np3 = 0.3 * (Y1 * Y2);
    const double qp1 = fma(r116, d[3], fma(-r76, d[4], r3 * d[5]));
    const double qp2 = fma(r3, d[2], fma(r56, d[3], -r6 * d[4]));
    const double qp3 = fma(-r6, d[1], fma(r56, d[2], r3 * d[3]));
    const double Np = fma(np3, qp3, fma(np2, qp2, np1 * qp1)), Sp = (np1 + np2) + np3;
    const double r = (double)1 / (Sm * Sp);
    dmo = (Nm * Sp) * r;
    dpo = (Np * Sm) * r;
}
__device__ __forceinline__ void fs_weno5_fast(const double* c, double inv, double eps, double& dmo, double& dpo) {
    double d[6];
#pragma unroll
    for (int q = 0; q < 6; q++) d[q] = inv * (c[q - 2] - c[q - 3]);
    FsW5Tri t[4];
#pragma unroll
    for (int q = 0; q < 4; q++) t[q] = fs_w5_tri(d[q], d[q + 1], d[q + 2], eps);
    fs_w5_combine(d, t, dmo, dpo);
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

// Velocity at the sample point from the MAC ghost copies,
// ADVECTION_METHOD_3D.h:201-224 (same for ENO3rd, :1862-1879).
__device__ __forceinline__ void vel_at(const double* __restrict__ gu, const double* __restrict__ gv,
                                       const double* __restrict__ gw, const FsLayout& L, int comp, int i, int j,
                                       int k, double& uu, double& vv, double& ww) {
    long c = L.idx(i, j, k);
    const long sx = L.sx, sy = L.sy;
    if (comp == 0) {
        uu = (gu[c + 1] + gu[c]) * (double)0.5;
        vv = (gv[c + sx] + gv[c]) * (double)0.5;
        ww = (gw[c + sy] + gw[c]) * (double)0.5;
    } else if (comp == 1) {
        uu = gu[c];
        vv = (gv[c] + gv[c + sx] + gv[c - 1] + gv[c - 1 + sx]) * (double)0.25;
        ww = (gw[c] + gw[c + sy] + gw[c - 1] + gw[c - 1 + sy]) * (double)0.25;
    } else if (comp == 2) {
        uu = (gu[c] + gu[c + 1] + gu[c - sx] + gu[c + 1 - sx]) * (double)0.25;
        vv = gv[c];
        ww = (gw[c] + gw[c + sy] + gw[c - sx] + gw[c - sx + sy]) * (double)0.25;
    } else {
        uu = (gu[c] + gu[c + 1] + gu[c - sy] + gu[c + 1 - sy]) * (double)0.25;
        vv = (gv[c] + gv[c + sx] + gv[c - sy] + gv[c + sx - sy]) * (double)0.25;
        ww = gw[c];
    }
}

__device__ __forceinline__ void gather7(const double* __restrict__ a, long c, long stride, double* f) {
#pragma unroll
    for (int o = -3; o <= 3; o++) f[o + 3] = a[c + o * stride];
}

// Gradient of the one-sided WENO5 value D(v1..v5) (fs_weno5's dmo with
// v = m, dpo with v = p) with respect to v.
__device__ __forceinline__ void weno5_grad(const double v[5], double eps, double g[5]) {
    const double k13 = (double)13 / 12, q4 = (double)1 / 4;
    double e1 = v[0] - 2 * v[1] + v[2], f1 = v[0] - 4 * v[1] + 3 * v[2];
    double e2 = v[1] - 2 * v[2] + v[3], f2 = v[1] - v[3];
    double e3 = v[2] - 2 * v[3] + v[4], f3 = 3 * v[2] - 4 * v[3] + v[4];
    double s[3] = {k13 * e1 * e1 + q4 * f1 * f1, k13 * e2 * e2 + q4 * f2 * f2, k13 * e3 * e3 + q4 * f3 * f3};
    const double d[3] = {(double)1 / 10, (double)6 / 10, (double)3 / 10};
    double a[3], S = 0;
    for (int k = 0; k < 3; k++) {
        a[k] = d[k] / ((eps + s[k]) * (eps + s[k]));
        S += a[k];
    }
    double q[3] = {v[0] / 3 - v[1] * 7 / 6 + v[2] * 11 / 6, -v[1] / 6 + v[2] * 5 / 6 + v[3] / 3,
                   v[2] / 3 + v[3] * 5 / 6 - v[4] / 6};
    double w[3] = {a[0] / S, a[1] / S, a[2] / S};
    double D = w[0] * q[0] + w[1] * q[1] + w[2] * q[2];
    // dD/ds_k = -2 w_k (q_k - D) / (eps + s_k)
    double ds[3];
    for (int k = 0; k < 3; k++) ds[k] = -2 * w[k] * (q[k] - D) / (eps + s[k]);
    // w_k dq_k/dv
    g[0] = w[0] / 3;
    g[1] = -w[0] * 7 / 6 - w[1] / 6;
    g[2] = w[0] * 11 / 6 + w[1] * 5 / 6 + w[2] / 3;
    g[3] = w[1] / 3 + w[2] * 5 / 6;
    g[4] = -w[2] / 6;
    // ds_k/dv
    double a1 = ds[0] * 2 * k13 * e1, b1 = ds[0] * 2 * q4 * f1;
    g[0] += a1 + b1;
    g[1] += -2 * a1 - 4 * b1;
    g[2] += a1 + 3 * b1;
    double a2 = ds[1] * 2 * k13 * e2, b2 = ds[1] * 2 * q4 * f2;
    g[1] += a2 + b2;
    g[2] += -2 * a2;
    g[3] += a2 - b2;
    double a3 = ds[2] * 2 * k13 * e3, b3 = ds[2] * 2 * q4 * f3;
    g[2] += a3 + 3 * b3;
    g[3] += -2 * a3 - 4 * b3;
    g[4] += a3 + b3;
}

// GFM face coefficient of BuildLinearSystemNodeJumpConditionVaribleCoefficient,
// POISSON_SOLVER_3D.h:1080-1085, written from the cell on the + side (beta_l
// form).  The - side cell's beta_r of the same face evaluates the same
// products and sums, so both rows see one value.
__device__ __forceinline__ double gfm_beta(double vc, double vn, double fn, double fc) {
    return vc * vn * (fabs(fn) + fabs(fc)) / (vc * fabs(fn) + vn * fabs(fc));
}
