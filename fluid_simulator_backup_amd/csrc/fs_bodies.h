// fs_bodies.h -- the per-cell (or per-row) work of the reinitialization
// stage's kernels, shared by the one-scene kernels (fs_kernels.hip) and the
// scene-table kernels of batches (fs_batch.hip), so both compute every
// value the same way.
#pragma once
#include <hip/hip_runtime.h>

#include "fs_device.h"
#include "fs_layout.h"

struct FsBox {
    int ni, nj, nk;
};
__host__ __device__ inline FsBox fs_box(const FsLayout& L, int comp) {
    return FsBox{L.nx + (comp == 1), L.ny + (comp == 2), L.nz + (comp == 3)};
}

__device__ __forceinline__ int fs_clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// FillGhostCellsFrom, FIELD_STRUCTURE_3D.h:663-1086: every ghost cell of the
// depth-d shell takes its clamped interior value (reads interior only); one
// call per (jj, kk) row of the grown box, jj, kk in [-d, n + d).
__device__ __forceinline__ void fs_fill_const_row(double* dst, const double* src, const FsLayout& L, FsBox b, int d,
                                                  int jj, int kk) {
    bool row_inside = jj >= 0 && jj < b.nj && kk >= 0 && kk < b.nk;
    int jc = fs_clampi(jj, 0, b.nj - 1), kc = fs_clampi(kk, 0, b.nk - 1);
    if (row_inside) {
        for (int g = 1; g <= d; g++) {
            dst[L.idx(-g, jj, kk)] = src[L.idx(0, jc, kc)];
            dst[L.idx(b.ni - 1 + g, jj, kk)] = src[L.idx(b.ni - 1, jc, kc)];
        }
    } else {
        for (int ii = -d; ii < b.ni + d; ii++) dst[L.idx(ii, jj, kk)] = src[L.idx(fs_clampi(ii, 0, b.ni - 1), jc, kc)];
    }
}

// FillGhostCellsContinuousDerivativesFrom face blocks,
// FIELD_STRUCTURE_3D.h:1806-1894: 2a-b outwards, deeper layers reading the
// layers just written when src == dst (one call walks its own row).  S(x):
// src inside the box along the axis, gsrc in its ghost layers.  Row (a, c)
// of face axis `axis`: a runs over the first in-face index, c the second.
__device__ __forceinline__ void fs_linear_faces_row(double* dst, const double* src, const double* gsrc,
                                                    const FsLayout& L, FsBox b, int d, int axis, int a, int c) {
    if (c >= (axis == 2 ? b.nj : b.nk)) return;
    if (axis == 0) {
        if (a >= b.nj) return;
        auto S = [&](int i) { return (i >= 0 && i < b.ni ? src : gsrc)[L.idx(i, a, c)]; };
        for (int i = -1; i >= -d; i--) dst[L.idx(i, a, c)] = (double)2 * S(i + 1) - S(i + 2);
        for (int i = b.ni; i <= b.ni - 1 + d; i++) dst[L.idx(i, a, c)] = (double)2 * S(i - 1) - S(i - 2);
    } else if (axis == 1) {
        if (a >= b.ni) return;
        auto S = [&](int j) { return (j >= 0 && j < b.nj ? src : gsrc)[L.idx(a, j, c)]; };
        for (int j = -1; j >= -d; j--) dst[L.idx(a, j, c)] = (double)2 * S(j + 1) - S(j + 2);
        for (int j = b.nj; j <= b.nj - 1 + d; j++) dst[L.idx(a, j, c)] = (double)2 * S(j - 1) - S(j - 2);
    } else {
        if (a >= b.ni) return;
        auto S = [&](int k) { return (k >= 0 && k < b.nk ? src : gsrc)[L.idx(a, c, k)]; };
        for (int k = -1; k >= -d; k--) dst[L.idx(a, c, k)] = (double)2 * S(k + 1) - S(k + 2);
        for (int k = b.nk; k <= b.nk - 1 + d; k++) dst[L.idx(a, c, k)] = (double)2 * S(k - 1) - S(k - 2);
    }
}

// Sign function S = phi/sqrt(phi^2 + dx^2), EULERIAN_FLUID_SOLVER_3D.h:2211-2215.
__device__ __forceinline__ void fs_sign_cell(const double* __restrict__ phi, double* __restrict__ sgn, long c,
                                             double dx2) {
    double f = phi[c];
    sgn[c] = f / sqrt(f * f + dx2);
}

// WENO5 one-sided derivatives in the reference's operation order (FAST =
// false, bit for bit) or in fs_weno5_fast's fewer operations
template <bool FAST>
__device__ __forceinline__ void fs_weno5_of(const double* c, double inv, double eps, double& dmo, double& dpo) {
    if constexpr (FAST)
        fs_weno5_fast(c, inv, eps, dmo, dpo);
    else
        fs_weno5(c, inv, eps, dmo, dpo);
}

// The Godunov update and clamp of one Sussman sweep at cell c from its six
// one-sided derivatives (WENO5thReinitialization, ADVECTION_METHOD_3D.h:
// 745-1637): Godunov case per axis, the misplaced parenthesis of :1463/:1605
// (x shock, y shock towards +, z expansion; the Euclidean norm instead when
// `exact` is set), and the sign-preserving clamp (:1624-1634).
__device__ __forceinline__ void fs_reinit_finish(double* __restrict__ out, const double* __restrict__ gh,
                                                 const double* __restrict__ sgn, double mx, double px, double my,
                                                 double py, double mz, double pz, double eps, double dt, int exact,
                                                 long c) {
    double X, Y, Z;
    double sg = sgn[c];
    double rho = gh[c];  // the ghost copy's box is phi
    int cx = fs_godunov_axis(sg, mx, px, X);
    int cy = cx ? fs_godunov_axis(sg, my, py, Y) : FS_AX_NONE;
    int cz = cy ? fs_godunov_axis(sg, mz, pz, Z) : FS_AX_NONE;
    if (cx && cy && cz) {
        double g;
        if (!exact && (cx == FS_AX_SM || cx == FS_AX_SP) && cy == FS_AX_SP && cz == FS_AX_DROP)
            g = sqrt(X * X) + Y * Y - 1;
        else
            g = sqrt(X * X + Y * Y + Z * Z) - 1;
        rho = rho - dt * sg * g;
    }
    if (sg * rho < (double)0) {
        if (sg < 0) rho = -eps;
        if (sg > 0) rho = eps;
    }
    out[c] = rho;
}

// One Sussman sweep at cell c: WENO5 one-sided derivatives of the ghost copy
// gh along each axis, then fs_reinit_finish.
template <bool FAST>
__device__ __forceinline__ void fs_reinit_cell(double* __restrict__ out, const double* __restrict__ gh,
                                               const double* __restrict__ sgn, const FsLayout& L, double ix, double iy,
                                               double iz, double eps, double dt, int exact, long c) {
    double f[7], mx, px, my, py, mz, pz;
    gather7(gh, c, 1, f);
    fs_weno5_of<FAST>(f + 3, ix, eps, mx, px);
    gather7(gh, c, L.sx, f);
    fs_weno5_of<FAST>(f + 3, iy, eps, my, py);
    gather7(gh, c, L.sy, f);
    fs_weno5_of<FAST>(f + 3, iz, eps, mz, pz);
    fs_reinit_finish(out, gh, sgn, mx, px, my, py, mz, pz, eps, dt, exact, c);
}
