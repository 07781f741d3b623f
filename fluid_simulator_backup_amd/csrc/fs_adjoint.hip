// fs_adjoint.hip -- reverse-mode kernels of the fixed-schedule step.
//
// Each kernel is the transpose of one forward kernel's derivative at the
// forward's inputs (recomputed from the recorded fields).  Where several
// outputs read one input the contributions are scattered with FP64 atomics
// into a padded accumulator, ghost positions included; a fold then maps
// each ghost contribution onto the interior values that ghost was made from.
// The reference has no adjoint; the forward kernels cite the reference.
#include <hip/hip_runtime.h>

#include "fs_adjoint.h"

#define BX 64
#define BY 4

static inline dim3 grid_for(FsBox b) { return dim3((b.ni + BX - 1) / BX, (b.nj + BY - 1) / BY, b.nk); }

// d/dx of the smoothed Heaviside of fs_heaviside_visc / fs_heaviside_proj:
// (1 + cos(pi x/eps)) / (2 eps) inside [-eps, eps], 0 outside.
__device__ __forceinline__ double heaviside_slope(double x, double eps) {
    if (x < -eps || x > eps) return 0;
    return (double)0.5 / eps * ((double)1 + cos(FS_PI * x / eps));
}

// ---- ghost folds --------------------------------------------------------------

__global__ void k_fold_linear(double* a, FsLayout L, FsBox b, int d, int axis) {
    int t = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    long st;
    int n;
    long base;
    if (axis == 0) {
        if (t >= b.nj) return;
        st = 1, n = b.ni, base = L.idx(0, t, q);
    } else if (axis == 1) {
        if (t >= b.ni) return;
        st = L.sx, n = b.nj, base = L.idx(t, 0, q);
    } else {
        if (t >= b.ni) return;
        st = L.sy, n = b.nk, base = L.idx(t, q, 0);
    }
    double lo0 = 0, lo1 = 0, hi0 = 0, hi1 = 0;
    for (int g = 1; g <= d; g++) {
        long gl = base - g * st, gh = base + (n - 1 + g) * st;
        double x = a[gl], y = a[gh];
        a[gl] = 0;
        a[gh] = 0;
        lo0 += (g + 1) * x;
        lo1 -= g * x;
        hi0 += (g + 1) * y;
        hi1 -= g * y;
    }
    a[base] += lo0;
    a[base + st] += lo1;
    a[base + (n - 1) * st] += hi0;
    a[base + (n - 2) * st] += hi1;
}

void fs_adj_fold_linear(hipStream_t st, double* a, const FsLayout& L, int d) {
    FsBox b = fs_box(L, 0);
    hipLaunchKernelGGL(k_fold_linear, dim3((b.nj + 63) / 64, b.nk), dim3(64), 0, st, a, L, b, d, 0);
    hipLaunchKernelGGL(k_fold_linear, dim3((b.ni + 63) / 64, b.nk), dim3(64), 0, st, a, L, b, d, 1);
    hipLaunchKernelGGL(k_fold_linear, dim3((b.ni + 63) / 64, b.nj), dim3(64), 0, st, a, L, b, d, 2);
}

// Gather form: each boundary cell adds the ghosts that clamp onto it.
__global__ void __launch_bounds__(BX* BY) k_fold_const(double* a, FsLayout L, FsBox b, int d) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    bool edge = i == 0 || j == 0 || k == 0 || i == b.ni - 1 || j == b.nj - 1 || k == b.nk - 1;
    if (!edge) return;
    int i0 = i == 0 ? -d : i, i1 = i == b.ni - 1 ? b.ni - 1 + d : i;
    int j0 = j == 0 ? -d : j, j1 = j == b.nj - 1 ? b.nj - 1 + d : j;
    int k0 = k == 0 ? -d : k, k1 = k == b.nk - 1 ? b.nk - 1 + d : k;
    double s = 0;
    for (int kk = k0; kk <= k1; kk++)
        for (int jj = j0; jj <= j1; jj++)
            for (int ii = i0; ii <= i1; ii++) {
                bool inside = ii >= 0 && ii < b.ni && jj >= 0 && jj < b.nj && kk >= 0 && kk < b.nk;
                if (!inside) s += a[L.idx(ii, jj, kk)];
            }
    a[L.idx(i, j, k)] += s;
}

__global__ void k_zero_outside(double* a, FsLayout L, FsBox b) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < L.total; t += (long)gridDim.x * blockDim.x) {
        long q = t - L.off;
        bool inside = false;
        if (q >= 0) {
            long k = q / L.sy - FS_G, r = q % L.sy;
            long j = r / L.sx - FS_G, i = r % L.sx - FS_G;
            inside = i >= 0 && i < b.ni && j >= 0 && j < b.nj && k >= 0 && k < b.nk;
        }
        if (!inside) a[t] = 0;
    }
}

void fs_zero_outside(hipStream_t st, double* a, const FsLayout& L, FsBox b) {
    hipLaunchKernelGGL(k_zero_outside, dim3(1024), dim3(256), 0, st, a, L, b);
}

void fs_adj_fold_const(hipStream_t st, double* a, const FsLayout& L, int d) {
    FsBox b = fs_box(L, 0);
    hipLaunchKernelGGL(k_fold_const, grid_for(b), dim3(BX, BY), 0, st, a, L, b, d);
    fs_zero_outside(st, a, L, b);
}

// ---- WENO5 upwind stage ----------------------------------------------------------

// Scatter bar * dD/d(samples) of the selected one-sided WENO5 derivative
// along one axis: f[0..6] are the samples at c - 3 st .. c + 3 st; D uses
// the differences m_q = inv (f[q] - f[q-1]) (minus side: f[0..5]; plus side:
// p1 = inv (f6 - f5), p2..p5 = m5..m2).
__device__ __forceinline__ void scatter_weno5(double* agh, long c, long st, const double* f, double inv, double eps,
                                              bool minus, double bar) {
    double d[5], g[5];
    if (minus) {
        for (int q = 0; q < 5; q++) d[q] = inv * (f[q + 1] - f[q]);
    } else {
        d[0] = inv * (f[6] - f[5]);
        for (int q = 1; q < 5; q++) d[q] = inv * (f[6 - q] - f[5 - q]);
    }
    weno5_grad(d, eps, g);
    double sc = bar * inv;
    if (minus) {
        for (int q = 0; q < 5; q++) {  // d_q = inv (f[q+1] - f[q])
            atomicAdd(&agh[c + (q - 2) * st], sc * g[q]);
            atomicAdd(&agh[c + (q - 3) * st], -sc * g[q]);
        }
    } else {
        atomicAdd(&agh[c + 3 * st], sc * g[0]);
        atomicAdd(&agh[c + 2 * st], -sc * g[0]);
        for (int q = 1; q < 5; q++) {  // d_q = inv (f[6-q] - f[5-q])
            atomicAdd(&agh[c + (3 - q) * st], sc * g[q]);
            atomicAdd(&agh[c + (2 - q) * st], -sc * g[q]);
        }
    }
}

// Transpose of vel_at: bar of the sample velocity (uu, vv, ww) onto the MAC
// values it averaged.
__device__ __forceinline__ void scatter_vel(double* au, double* av, double* aw, const FsLayout& L, int comp, long c,
                                            double bu, double bv, double bw) {
    const long sx = L.sx, sy = L.sy;
    if (comp == 0) {
        atomicAdd(&au[c + 1], (double)0.5 * bu);
        atomicAdd(&au[c], (double)0.5 * bu);
        atomicAdd(&av[c + sx], (double)0.5 * bv);
        atomicAdd(&av[c], (double)0.5 * bv);
        atomicAdd(&aw[c + sy], (double)0.5 * bw);
        atomicAdd(&aw[c], (double)0.5 * bw);
    } else if (comp == 1) {
        atomicAdd(&au[c], bu);
        const long pv[4] = {c, c + sx, c - 1, c - 1 + sx}, pw[4] = {c, c + sy, c - 1, c - 1 + sy};
        for (int q = 0; q < 4; q++) {
            atomicAdd(&av[pv[q]], (double)0.25 * bv);
            atomicAdd(&aw[pw[q]], (double)0.25 * bw);
        }
    } else if (comp == 2) {
        atomicAdd(&av[c], bv);
        const long pu[4] = {c, c + 1, c - sx, c + 1 - sx}, pw[4] = {c, c + sy, c - sx, c - sx + sy};
        for (int q = 0; q < 4; q++) {
            atomicAdd(&au[pu[q]], (double)0.25 * bu);
            atomicAdd(&aw[pw[q]], (double)0.25 * bw);
        }
    } else {
        atomicAdd(&aw[c], bw);
        const long pu[4] = {c, c + 1, c - sy, c + 1 - sy}, pv[4] = {c, c + sx, c - sy, c + sx - sy};
        for (int q = 0; q < 4; q++) {
            atomicAdd(&au[pu[q]], (double)0.25 * bu);
            atomicAdd(&av[pv[q]], (double)0.25 * bv);
        }
    }
}

__global__ void __launch_bounds__(BX* BY) k_adj_weno5_stage(const double* __restrict__ lam, const double* __restrict__ gh,
                                                            const double* __restrict__ gu, const double* __restrict__ gv,
                                                            const double* __restrict__ gw, double* agh, double* au,
                                                            double* av, double* aw, FsLayout L, FsBox b, int comp,
                                                            double ix, double iy, double iz, double eps, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    long c = L.idx(i, j, k);
    double l = lam[c];
    if (l == 0) return;
    double uu, vv, ww;
    vel_at(gu, gv, gw, L, comp, i, j, k, uu, vv, ww);
    const double vel[3] = {uu, vv, ww}, inv[3] = {ix, iy, iz};
    const long st[3] = {1, L.sx, L.sy};
    double bvel[3];
    for (int axis = 0; axis < 3; axis++) {
        double f[7], dm, dp;
        gather7(gh, c, st[axis], f);
        fs_weno5(f + 3, inv[axis], eps, dm, dp);
        bool minus = vel[axis] > 0;  // upwind's choice
        scatter_weno5(agh, c, st[axis], f, inv[axis], eps, minus, -dt * vel[axis] * l);
        bvel[axis] = -dt * l * (minus ? dm : dp);
    }
    scatter_vel(au, av, aw, L, comp, c, bvel[0], bvel[1], bvel[2]);
}

void fs_adj_weno5_stage(hipStream_t st, const double* lam, const double* gh, const double* gu, const double* gv,
                        const double* gw, double* agh, double* au, double* av, double* aw, const FsLayout& L,
                        int comp, const FsGridConst& gc, double eps, double dt) {
    FsBox b = fs_box(L, comp);
    hipLaunchKernelGGL(k_adj_weno5_stage, grid_for(b), dim3(BX, BY), 0, st, lam, gh, gu, gv, gw, agh, au, av, aw, L,
                       b, comp, gc.x.inv, gc.y.inv, gc.z.inv, eps, dt);
}

// ---- explicit viscosity, level-set part ----------------------------------------------
// k_viscosity's update is coef * sum_q mu_q A_q with coef = dt / rho_h:
// each mu_q = mu(H(arg_q)) and rho_h = rho(H(lh)), arg_q a cell value or a
// 4-cell mean of phi.  d upd/d mu_q = coef A_q, d upd/d rho_h = -upd/rho_h.

__global__ void __launch_bounds__(BX* BY) k_adj_viscosity_phi(int comp, const double* __restrict__ lam,
                                                              const double* __restrict__ U,
                                                              const double* __restrict__ V,
                                                              const double* __restrict__ W,
                                                              const double* __restrict__ f, double* aphi, FsLayout L,
                                                              FsBox b, FsMaterial m, double eps, double odx,
                                                              double ody, double odz, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    const long c = L.idx(i, j, k), sx = L.sx, sy = L.sy;
    double l = lam[c];
    if (l == 0) return;
    long cells[6][4];
    int ncell[6];
    double A[6];
    long h0, h1;  // the two cells of the face density
    auto four = [&](int q, long a0, long a1, long a2, long a3) {
        cells[q][0] = a0, cells[q][1] = a1, cells[q][2] = a2, cells[q][3] = a3, ncell[q] = 4;
    };
    auto one = [&](int q, long a0) { cells[q][0] = a0, ncell[q] = 1; };
    if (comp == 1) {
        h0 = c - 1, h1 = c;
        one(0, c);
        A[0] = (double)2 * odx * odx * (U[c + 1] - U[c]);
        one(1, c - 1);
        A[1] = -(double)2 * odx * ody * (U[c] - U[c - 1]);
        four(2, c - 1, c, c - 1 + sx, c + sx);
        A[2] = ody * (ody * (U[c + sx] - U[c]) + odx * (V[c + sx] - V[c - 1 + sx]));
        four(3, c - 1, c, c - 1 - sx, c - sx);
        A[3] = -ody * (ody * (U[c] - U[c - sx]) + odx * (V[c] - V[c - 1]));
        four(4, c - 1, c, c - 1 + sy, c + sy);
        A[4] = odz * (odz * (U[c + sy] - U[c]) + odx * (W[c + sy] - W[c - 1 + sy]));
        four(5, c - 1, c, c - 1 - sy, c - sy);
        A[5] = -odz * (odz * (U[c] - U[c - sy]) + odx * (W[c] - W[c - 1]));
    } else if (comp == 2) {
        h0 = c - sx, h1 = c;
        four(0, c - sx, c + 1 - sx, c, c + 1);
        A[0] = odx * (ody * (U[c + 1] - U[c + 1 - sx]) + odx * (V[c + 1] - V[c]));
        four(1, c - 1 - sx, c - sx, c - 1, c);
        A[1] = -odx * (ody * (U[c] - U[c - sx]) + odx * (V[c] - V[c - 1]));
        one(2, c);
        A[2] = (double)2 * ody * ody * (V[c + sx] - V[c]);
        one(3, c - sx);
        A[3] = -(double)2 * ody * ody * (V[c] - V[c - sx]);
        four(4, c - sx + sy, c - sx, c + sy, c);
        A[4] = odz * (odz * (V[c + sy] - V[c]) + ody * (W[c] - W[c - sx]));
        four(5, c - sx - sy, c - sx, c - sy, c);
        A[5] = -odz * (odz * (V[c] - V[c - sy]) + ody * (W[c - sy] - W[c - sx - sy]));
    } else {
        h0 = c - sy, h1 = c;
        four(0, c - sy, c + 1 - sy, c, c + 1);
        A[0] = odx * (odz * (U[c] - U[c - sy]) + odx * (W[c + 1] - W[c]));
        four(1, c - sy, c - 1 - sy, c, c - 1);
        A[1] = -odx * (odz * (U[c - 1] - U[c - 1 - sy]) + odx * (W[c] - W[c - 1]));
        four(2, c - sy, c + sx - sy, c, c + sx);
        A[2] = ody * (odz * (V[c] - V[c - sy]) + ody * (W[c + sx] - W[c]));
        four(3, c - sy, c - sx - sy, c, c - sx);
        A[3] = -ody * (odz * (V[c - sx] - V[c - sx - sy]) + ody * (W[c] - W[c - sx]));
        one(4, c);
        A[4] = (double)2 * odz * odz * (W[c + sy] - W[c]);
        one(5, c - sy);
        A[5] = -(double)2 * odz * odz * (W[c] - W[c - sy]);
    }
    const double dmu = m.air_bubble_rising ? m.mu_w - m.mu_a : m.mu_a - m.mu_w;
    const double drho = m.air_bubble_rising ? m.rho_w - m.rho_a : m.rho_a - m.rho_w;
    double lh = (double)0.5 * (f[h0] + f[h1]);
    double rho_h = m.density(fs_heaviside_visc(lh, eps));
    double coef = dt / rho_h;
    double upd = 0;
    for (int q = 0; q < 6; q++) {
        double arg = 0;
        for (int e = 0; e < ncell[q]; e++) arg += f[cells[q][e]];
        if (ncell[q] == 4) arg *= (double)0.25;
        double mu = m.viscosity(fs_heaviside_visc(arg, eps));
        upd += coef * mu * A[q];
        double g = l * coef * A[q] * dmu * heaviside_slope(arg, eps);
        if (g == 0) continue;
        double w = ncell[q] == 4 ? (double)0.25 : (double)1;
        for (int e = 0; e < ncell[q]; e++) atomicAdd(&aphi[cells[q][e]], w * g);
    }
    double gr = l * (-upd / rho_h) * drho * heaviside_slope(lh, eps) * (double)0.5;
    if (gr != 0) {
        atomicAdd(&aphi[h0], gr);
        atomicAdd(&aphi[h1], gr);
    }
}

void fs_adj_viscosity_phi(hipStream_t st, int comp, const double* lam, const double* U, const double* V,
                          const double* W, const double* phi, double* aphi, const FsLayout& L, const FsMaterial& m,
                          const FsGridConst& cell, double dt) {
    FsBox b = fs_box(L, comp);
    double eps = 1.5 * cell.x.d;  // as fs_viscosity
    hipLaunchKernelGGL(k_adj_viscosity_phi, grid_for(b), dim3(BX, BY), 0, st, comp, lam, U, V, W, phi, aphi, L, b, m,
                       eps, cell.x.inv, cell.y.inv, cell.z.inv, dt);
}

// ---- projection, level-set part ----------------------------------------------------

// d beta / d f_self of gfm_beta(vc, vn, fn, fc) for the face's cell `self`
// (fc when upper, fn when lower) with the other cell `oth`:
// beta = vc vn (|fn| + |fc|) / (vc |fn| + vn |fc|).
__device__ __forceinline__ double dbeta_dfc(double vc, double vn, double fn, double fc) {
    double D = vc * fabs(fn) + vn * fabs(fc);
    double s = (double)((fc > 0) - (fc < 0));
    return vc * vn * s * fabs(fn) * (vc - vn) / (D * D);
}
__device__ __forceinline__ double dbeta_dfn(double vc, double vn, double fn, double fc) {
    double D = vc * fabs(fn) + vn * fabs(fc);
    double s = (double)((fn > 0) - (fn < 0));
    return vc * vn * s * fabs(fc) * (vn - vc) / (D * D);
}

__global__ void __launch_bounds__(BX* BY) k_adj_projection_phi(const double* __restrict__ f,
                                                               const double* __restrict__ v,
                                                               const double* __restrict__ p,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ lu,
                                                               const double* __restrict__ lv,
                                                               const double* __restrict__ lw, double* aphi, FsLayout L,
                                                               double dt, double od0, double od1, double od2,
                                                               double diag2, double off0, double off1, double off2) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const long c = L.idx(i, j, k);
    const long st[3] = {1, L.sx, L.sy};
    const int idx[3] = {i, j, k}, n[3] = {L.nx, L.ny, L.nz};
    const double od[3] = {od0, od1, od2}, off[3] = {off0, off1, off2};
    const double* lam[3] = {lu, lv, lw};
    double acc = 0;
    for (int a = 0; a < 3; a++) {
        // face below c: upper cell c, lower cell c - st
        if (idx[a] > 0) {
            long lo = c - st[a];
            double vc = v[c], vn = v[lo];
            if (vc != vn) {
                double dp = p[c] - p[lo];
                double yAp = diag2 * (y[c] * p[c] + y[lo] * p[lo]) - off[a] * (y[c] * p[lo] + y[lo] * p[c]);
                double term = -dt * od[a] * lam[a][c] * dp - yAp;
                acc += term * dbeta_dfc(vc, vn, f[lo], f[c]);
            }
        }
        // face above c: upper cell c + st, lower cell c
        if (idx[a] + 1 < n[a]) {
            long up = c + st[a];
            double vc = v[up], vn = v[c];
            if (vc != vn) {
                double dp = p[up] - p[c];
                double yAp = diag2 * (y[up] * p[up] + y[c] * p[c]) - off[a] * (y[up] * p[c] + y[c] * p[up]);
                double term = -dt * od[a] * lam[a][up] * dp - yAp;
                acc += term * dbeta_dfn(vc, vn, f[c], f[up]);
            }
        }
    }
    aphi[c] += acc;
}

void fs_adj_projection_phi(hipStream_t st, const double* phi, const double* irho, const double* p, const double* y,
                           const double* lu, const double* lv, const double* lw, double* aphi, const FsLayout& L,
                           const FsGridConst& cell, double dt) {
    // A's diagonal uses 1/dx^2 for every face (k_diag_rhs), its
    // off-diagonals the face's own axis
    double diag2 = (double)1 / (cell.x.d * cell.x.d);
    hipLaunchKernelGGL(k_adj_projection_phi, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, phi, irho, p, y, lu, lv, lw,
                       aphi, L, dt, cell.x.inv, cell.y.inv, cell.z.inv, diag2, (double)1 / (cell.x.d * cell.x.d),
                       (double)1 / (cell.y.d * cell.y.d), (double)1 / (cell.z.d * cell.z.d));
}

// ---- reinitialization ----------------------------------------------------------------

__global__ void __launch_bounds__(BX* BY) k_adj_reinit(const double* __restrict__ lam, const double* __restrict__ phi,
                                                       const double* __restrict__ gh, const double* __restrict__ sgn,
                                                       double* out, double* lsg, FsLayout L, double ix, double iy,
                                                       double iz, double eps, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    double l = lam[c];
    if (l == 0) return;
    const double inv[3] = {ix, iy, iz};
    const long st[3] = {1, L.sx, L.sy};
    double f[3][7], m[3], p[3], sel[3];
    int cs[3];
    for (int a = 0; a < 3; a++) {
        gather7(gh, c, st[a], f[a]);
        fs_weno5(f[a] + 3, inv[a], eps, m[a], p[a]);
    }
    double sg = sgn[c];
    cs[0] = fs_godunov_axis(sg, m[0], p[0], sel[0]);
    cs[1] = cs[0] ? fs_godunov_axis(sg, m[1], p[1], sel[1]) : FS_AX_NONE;
    cs[2] = cs[1] ? fs_godunov_axis(sg, m[2], p[2], sel[2]) : FS_AX_NONE;
    bool update = cs[0] && cs[1] && cs[2];
    double G = 0, g = 0, rho = phi[c];
    if (update) {
        G = sqrt(sel[0] * sel[0] + sel[1] * sel[1] + sel[2] * sel[2]);
        g = G - 1;
        rho = rho - dt * sg * g;
    }
    if (sg * rho < (double)0) return;  // clamped to +-eps: no dependence
    atomicAdd(&out[c], l);
    if (!update) return;
    lsg[c] += -dt * g * l;
    if (!(G > 0)) return;
    double barG = -dt * sg * l;
    for (int a = 0; a < 3; a++) {
        if (cs[a] == FS_AX_DROP) continue;
        bool minus = cs[a] == FS_AX_M || cs[a] == FS_AX_SM;
        scatter_weno5(out, c, st[a], f[a], inv[a], eps, minus, barG * sel[a] / G);
    }
}

void fs_adj_reinit_sweep(hipStream_t st, const double* lam, const double* phi, const double* gh, const double* sgn,
                         double* out, double* lsg, const FsLayout& L, const FsGridConst& cell, double eps, double dt) {
    hipLaunchKernelGGL(k_adj_reinit, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, lam, phi, gh, sgn, out, lsg, L,
                       cell.x.inv, cell.y.inv, cell.z.inv, eps, dt);
}

__global__ void __launch_bounds__(BX* BY) k_adj_sign(const double* __restrict__ phi, const double* __restrict__ lsg,
                                                     double* __restrict__ lphi, FsLayout L, double dx2) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    double f = phi[c], q = f * f + dx2;
    lphi[c] += lsg[c] * dx2 / (q * sqrt(q));
}

void fs_adj_sign(hipStream_t st, const double* phi, const double* lsg, double* lphi, const FsLayout& L, double dx2) {
    hipLaunchKernelGGL(k_adj_sign, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, phi, lsg, lphi, L, dx2);
}

// ---- level-set loss ---------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_levelset_loss(const double* __restrict__ phi,
                                                       const double* __restrict__ target, double* lphi,
                                                       double* __restrict__ part, FsLayout L, double eps) {
    __shared__ double red[256];
    const long n = (long)L.nx * L.ny * L.nz;
    double s = 0;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        int i = (int)(t % L.nx);
        long r = t / L.nx;
        int j = (int)(r % L.ny), k = (int)(r / L.ny);
        long c = L.idx(i, j, k);
        double d = fs_heaviside_proj(phi[c], eps) - fs_heaviside_proj(target[c], eps);
        s += d * d;
        if (lphi) lphi[c] = (double)2 * d * heaviside_slope(phi[c], eps);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_loss_total(double* part) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0;
    for (int b = 0; b < FS_LOSS_BLOCKS; b++) s += part[b];
    part[FS_LOSS_BLOCKS] = s;
}

void fs_levelset_loss(hipStream_t st, const double* phi, const double* target, double* lphi, double* part,
                      const FsLayout& L, double eps) {
    hipLaunchKernelGGL(k_levelset_loss, dim3(FS_LOSS_BLOCKS), dim3(256), 0, st, phi, target, lphi, part, L, eps);
    hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(64), 0, st, part);
}
