/*
 * fs_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU checker for the HIP path.
 *
 * A plain-C, FP64, sequential-order restatement of the reference's per-step
 * hot path in air/water mode (EULERIAN_FLUID_SOLVER_3D::AdvanceOneTimeStepThread,
 * EULERIAN_FLUID_SOLVER_3D.h:803-1111, order_for_time_advancing 1, 2, 3).  Each
 * function cites the reference lines it restates.  Operation order inside every
 * expression follows the reference text (no FMA contraction: build with
 * -ffp-contract=off) so that the HIP kernels, which keep the same order, can be
 * compared bit for bit where no libm call or reduction intervenes.
 *
 * State model.  The reference keeps persistent ghosted arrays whose ghost
 * cells are refreshed only at specific points; several stencils read ghost
 * copies made earlier in the step (stale).  This file keeps the same arrays
 * with the same refresh points:
 *   phi     level set, ghost width 2 (LEVELSET_3D, EULERIAN_FLUID_SOLVER_3D.h:457-459)
 *   sghost  shared scalar ghost copy, width 3 (scalar_field_ghost, :435), written
 *           only by reinitialization sweeps (ADVECTION_3D.h:375)
 *   gu,gv,gw  MAC ghost copies, width 3 (:436-438), written only by the
 *           projection's divergence (PROJECTION_3D.h:686-688)
 *   u,v,w   MAC velocity, width 2; nothing writes their ghosts (stay 0)
 *   p       pressure, width 1 (PROJECTION_3D.h:362), ghosts stay 0
 * Every array shares one padded layout of ghost width 3, x fastest.
 *
 * Reductions (dot products, the Neumann mean) are summed per z-plane in
 * row order and the plane partials in plane order: the reference's
 * MULTITHREADING::SyncSum (MULTITHREADING.h:229-250) with one slab per plane.
 *
 * Parity unpinned here: the libm boundary (sin in the smoothed Heaviside,
 * sqrt in reinitialization / sign function) is whatever the host libm gives.
 */
#include "fs_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* A statistic read out of a field value: the value itself here; the
 * reverse-mode build (fs_oracle_ad.cpp) defines it as the value part of its
 * active type. */
#ifndef FSO_VALUE
#define FSO_VALUE(x) (x)
#endif

#define G 3
#define FS_BC_FULL 0
#define FS_BC_NEUM (-4) /* MACROS.h:10 */
#define FS_PI 3.14159265358979323846 /* MACROS.h:13 */

typedef struct {
    double d, d2, inv, inv2d, inv_d2;
} ax_t; /* GRID_STRUCTURE_3D.h:114-133 */
typedef struct {
    ax_t x, y, z;
} gc_t;

struct fso_state {
    fs_params P;
    fs_grid gr;
    int nx, ny, nz;
    long sx, sy, sz, ntot;
    gc_t cell, mx, my, mz;
    double *phi, *sghost, *u, *v, *w, *gu, *gv, *gw, *p, *div, *rho, *irho, *sgn, *tmp, *pg, *vtmp;
    double *urk, *vrk, *wrk; /* water_velocity_field_*_rk (orders 2 and 3) */
    double *fbx, *fby, *fbz; /* face betas of the last assembly (inspection only) */
    int* bc;
    /* CSR of the last assembly (POISSON_SOLVER_3D.h:1036-1038) */
    int N, nnz;
    int *rp, *ci;
    double *val, *b, *x, *r, *zv, *pp, *Ap;
    double* plane_part; /* per-plane partial sums of plane_sum_dot */
    int visc_inplace;
    double max_vel[3];
    int last_iters;
    int visc_iters; /* semi-implicit viscosity PCG iterations since the step began */
    double last_res_abs, last_res_rel;
    /* surface tension: LEVELSET_3D::curvature, PROJECTION_3D::jc_on_solution
     * (zero ghosts), abs_max_curvature, and whether the last curvature pass
     * met a cell whose gradient vanishes in both difference forms */
    double *kappa, *jc;
    double kmax;
    int curv_fail;
    /* the scene's spherical level-set sources and the frame counter */
    fs_levelset_source src[64];
    int nsrc, frame;
    /* test hook (fso_solve_poisson): a jump field that replaces sigma*kappa */
    const double* jc_override;
    /* PROJECTION_3D::CSF_model, set by the first Sourcing with the CSF model */
    int csf_on;
    /* LEVELSET_3D::normal (unit normals, zero until the first ComputeNormals) */
    double* nrm[3];
};

/* A CSR system with rows numbered plane by plane (per rows in each of
 * nplanes planes) and its PCG work vectors. */
typedef struct {
    int N, per, nplanes;
    const int *rp, *ci;
    const double *val, *b;
    double *x, *r, *zv, *pp, *Ap;
} sys_t;

static int g_threads = 1;
void fso_set_threads(int n) { g_threads = n > 0 ? n : 1; }

#define IX(s, i, j, k) ((((long)(k) + G) * (s)->sy + ((long)(j) + G)) * (s)->sx + ((long)(i) + G))
#define A3(a, i, j, k) ((a)[IX(s, i, j, k)])

static ax_t mk_ax(double mn, double mx, int res) {
    ax_t a;
    a.d = (mx - mn) / (res - 1);
    a.d2 = a.d * a.d;
    a.inv = (double)1 / a.d;
    a.inv2d = (double)0.5 / a.d;
    a.inv_d2 = (double)1 / a.d2;
    return a;
}

static double* zalloc(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

int fso_create(const fs_params* P, const fs_grid* g, fso_state** out) {
    if (!P || !g || !out || g->nx < 3 || g->ny < 3 || g->nz < 3) return FS_ERR_BAD_ARG;
    if (P->order_for_time_advancing < 1 || P->order_for_time_advancing > 3) return FS_ERR_BAD_ARG;
    fso_state* s = (fso_state*)calloc(1, sizeof(fso_state));
    s->P = *P;
    s->gr = *g;
    s->nx = g->nx;
    s->ny = g->ny;
    s->nz = g->nz;
    s->sx = g->nx + 1 + 2 * G;
    s->sy = g->ny + 1 + 2 * G;
    s->sz = g->nz + 1 + 2 * G;
    s->ntot = s->sx * s->sy * s->sz;
    s->cell.x = mk_ax(g->x_min, g->x_max, g->nx);
    s->cell.y = mk_ax(g->y_min, g->y_max, g->ny);
    s->cell.z = mk_ax(g->z_min, g->z_max, g->nz);
    /* MAC grids, EULERIAN_FLUID_SOLVER_3D.h:407,415,423 */
    double dx = s->cell.x.d, dy = s->cell.y.d, dz = s->cell.z.d;
    s->mx = s->cell;
    s->mx.x = mk_ax(g->x_min - (double)0.5 * dx, g->x_max + (double)0.5 * dx, g->nx + 1);
    s->my = s->cell;
    s->my.y = mk_ax(g->y_min - (double)0.5 * dy, g->y_max + (double)0.5 * dy, g->ny + 1);
    s->mz = s->cell;
    s->mz.z = mk_ax(g->z_min - (double)0.5 * dz, g->z_max + (double)0.5 * dz, g->nz + 1);
    long n = s->ntot;
    s->phi = zalloc(n); s->sghost = zalloc(n); s->u = zalloc(n); s->v = zalloc(n); s->w = zalloc(n);
    s->gu = zalloc(n); s->gv = zalloc(n); s->gw = zalloc(n); s->p = zalloc(n); s->div = zalloc(n);
    s->rho = zalloc(n); s->irho = zalloc(n); s->sgn = zalloc(n); s->tmp = zalloc(n); s->pg = zalloc(n);
    s->vtmp = zalloc(n);
    s->urk = zalloc(n); s->vrk = zalloc(n); s->wrk = zalloc(n);
    s->fbx = zalloc(n); s->fby = zalloc(n); s->fbz = zalloc(n);
    if (P->surface_tension != 0.0) {
        s->kappa = zalloc(n);
        s->jc = zalloc(n);
        s->nrm[0] = zalloc(n); s->nrm[1] = zalloc(n); s->nrm[2] = zalloc(n);
    }
    s->bc = (int*)calloc((size_t)n, sizeof(int));
    s->plane_part = (double*)calloc((size_t)s->nz + 1, sizeof(double));
    s->visc_inplace = 1;
    *out = s;
    return FS_OK;
}

static void free_csr(fso_state* s) {
    free(s->rp); free(s->ci); free(s->val); free(s->b); free(s->x); free(s->r); free(s->zv); free(s->pp); free(s->Ap);
    s->rp = s->ci = NULL;
    s->val = s->b = s->x = s->r = s->zv = s->pp = s->Ap = NULL;
}

void fso_destroy(fso_state* s) {
    if (!s) return;
    free(s->phi); free(s->sghost); free(s->u); free(s->v); free(s->w); free(s->gu); free(s->gv); free(s->gw);
    free(s->p); free(s->div); free(s->rho); free(s->irho); free(s->sgn); free(s->tmp); free(s->pg); free(s->vtmp); free(s->fbx); free(s->fby); free(s->fbz);
    free(s->kappa); free(s->jc);
    free(s->nrm[0]); free(s->nrm[1]); free(s->nrm[2]);
    free(s->bc);
    free(s->urk); free(s->vrk); free(s->wrk);
    free(s->plane_part);
    free_csr(s);
    free(s);
}

void fso_set_viscosity_inplace(fso_state* s, int inplace) { s->visc_inplace = inplace; }

/* ------------------------------------------------------------------------- */
/* Ghost fills (FIELD_STRUCTURE_3D.h)                                        */
/* ------------------------------------------------------------------------- */

static void copy_interior(fso_state* s, double* dst, const double* src, int ni, int nj, int nk) {
    if (dst == src) return;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) A3(dst, i, j, k) = A3(src, i, j, k);
}

/* FillGhostCellsZero, FIELD_STRUCTURE_3D.h:538-661: copy the interior and
 * zero the six face blocks; edge and vertex blocks are left as they are. */
static void fill_zero(fso_state* s, double* dst, const double* src, int ni, int nj, int nk, int d) {
    copy_interior(s, dst, src, ni, nj, nk);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int g = 1; g <= d; g++) {
                A3(dst, -g, j, k) = 0;
                A3(dst, ni - 1 + g, j, k) = 0;
            }
    for (int k = 0; k < nk; k++)
        for (int g = 1; g <= d; g++)
            for (int i = 0; i < ni; i++) {
                A3(dst, i, -g, k) = 0;
                A3(dst, i, nj - 1 + g, k) = 0;
            }
    for (int g = 1; g <= d; g++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                A3(dst, i, j, -g) = 0;
                A3(dst, i, j, nk - 1 + g) = 0;
            }
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* FillGhostCellsFrom, FIELD_STRUCTURE_3D.h:663-1086: constant extrapolation
 * of faces, edges and vertices (every ghost takes its clamped interior value). */
static void fill_const(fso_state* s, double* dst, const double* src, int ni, int nj, int nk, int d, int copy) {
    if (copy) copy_interior(s, dst, src, ni, nj, nk);
    for (int k = -d; k < nk + d; k++)
        for (int j = -d; j < nj + d; j++)
            for (int i = -d; i < ni + d; i++) {
                if (i >= 0 && i < ni && j >= 0 && j < nj && k >= 0 && k < nk) continue;
                A3(dst, i, j, k) = A3(src, clampi(i, 0, ni - 1), clampi(j, 0, nj - 1), clampi(k, 0, nk - 1));
            }
}

/* FillGhostCellsContinuousDerivativesFrom, FIELD_STRUCTURE_3D.h:1776-2199:
 * linear extrapolation 2a-b on the six faces in the reference's block and
 * loop order (src may equal dst: later ghost layers then read earlier ones);
 * with edges != 0 also the three filled edge blocks normal to z (1898-1946).
 * The other edges and all vertices are commented out in the reference. */
static void fill_linear(fso_state* s, double* dst, const double* src, int ni, int nj, int nk, int d, int copy,
                        int edges) {
    if (copy) copy_interior(s, dst, src, ni, nj, nk);
    int i, j, k;
    for (k = 0; k < nk; k++)
        for (j = 0; j < nj; j++)
            for (i = -1; i >= -d; i--) A3(dst, i, j, k) = (double)2 * A3(src, i + 1, j, k) - A3(src, i + 2, j, k);
    for (k = 0; k < nk; k++)
        for (j = 0; j < nj; j++)
            for (i = ni; i <= ni - 1 + d; i++) A3(dst, i, j, k) = (double)2 * A3(src, i - 1, j, k) - A3(src, i - 2, j, k);
    for (k = 0; k < nk; k++)
        for (j = -1; j >= -d; j--)
            for (i = 0; i < ni; i++) A3(dst, i, j, k) = (double)2 * A3(src, i, j + 1, k) - A3(src, i, j + 2, k);
    for (k = 0; k < nk; k++)
        for (j = nj; j <= nj - 1 + d; j++)
            for (i = 0; i < ni; i++) A3(dst, i, j, k) = (double)2 * A3(src, i, j - 1, k) - A3(src, i, j - 2, k);
    for (k = -1; k >= -d; k--)
        for (j = 0; j < nj; j++)
            for (i = 0; i < ni; i++) A3(dst, i, j, k) = (double)2 * A3(src, i, j, k + 1) - A3(src, i, j, k + 2);
    for (k = nk; k <= nk - 1 + d; k++)
        for (j = 0; j < nj; j++)
            for (i = 0; i < ni; i++) A3(dst, i, j, k) = (double)2 * A3(src, i, j, k - 1) - A3(src, i, j, k - 2);
    if (!edges) return;
    for (k = 0; k < nk; k++)
        for (j = -1; j >= -d; j--)
            for (i = -1; i >= -d; i--)
                A3(dst, i, j, k) = (double)0.5 * (((double)2 * A3(src, i + 1, j, k) - A3(src, i + 2, j, k)) +
                                                  ((double)2 * A3(src, i, j + 1, k) - A3(src, i, j + 2, k)));
    for (k = 0; k < nk; k++)
        for (j = nj; j <= nj - 1 + d; j++)
            for (i = ni; i <= ni - 1 + d; i++)
                A3(dst, i, j, k) = (double)0.5 * (((double)2 * A3(src, i - 1, j, k) - A3(src, i - 2, j, k)) +
                                                  ((double)2 * A3(src, i, j - 1, k) - A3(src, i, j - 2, k)));
    for (k = 0; k < nk; k++)
        for (j = nj; j <= nj - 1 + d; j++)
            for (i = -1; i >= -d; i--)
                A3(dst, i, j, k) = (double)0.5 * (((double)2 * A3(src, i + 1, j, k) - A3(src, i + 2, j, k)) +
                                                  ((double)2 * A3(src, i, j - 1, k) - A3(src, i, j - 2, k)));
}

/* ------------------------------------------------------------------------- */
/* 1-D derivative kernels                                                     */
/* ------------------------------------------------------------------------- */

/* WENO5-JS (Jiang-Shu 1996) one-sided first derivatives at v[3], restating
 * ADVECTION_METHOD_3D.h:55-183 term by term. */
void fso_weno5_1d(const double* f, double inv, double eps, double* dmo, double* dpo) {
    const double* c = f + 3;
    double m1 = inv * (c[-2] - c[-3]), m2 = inv * (c[-1] - c[-2]), m3 = inv * (c[0] - c[-1]);
    double m4 = inv * (c[1] - c[0]), m5 = inv * (c[2] - c[1]);
    double p1 = inv * (c[3] - c[2]), p2 = inv * (c[2] - c[1]), p3 = inv * (c[1] - c[0]);
    double p4 = inv * (c[0] - c[-1]), p5 = inv * (c[-1] - c[-2]);
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
    *dmo = wm1 * (m1 * (double)1 / 3 - m2 * (double)7 / 6 + m3 * (double)11 / 6) +
           wm2 * (m2 * (-(double)1 / 6) + m3 * (double)5 / 6 + m4 * (double)1 / 3) +
           wm3 * (m3 * (double)1 / 3 + m4 * (double)5 / 6 - m5 * (double)1 / 6);
    *dpo = wp1 * (p1 * (double)1 / 3 - p2 * (double)7 / 6 + p3 * (double)11 / 6) +
           wp2 * (p2 * (-(double)1 / 6) + p3 * (double)5 / 6 + p4 * (double)1 / 3) +
           wp3 * (p3 * (double)1 / 3 + p4 * (double)5 / 6 - p5 * (double)1 / 6);
}

/* ENO3 one-sided first derivatives at v[3] with the reference's divided
 * differences and stencil coefficients, ADVECTION_METHOD_3D.h:1653-1718. */
void fso_eno3_1d(const double* f, double dx, double inv, double inv2, double inv3, double* dmo, double* dpo) {
    const double* c = f + 3;
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
    *dmo = m;
    *dpo = p;
}

/* gather 7 samples along an axis */
static void gather7(fso_state* s, const double* a, int i, int j, int k, int axis, double* f) {
    for (int o = -3; o <= 3; o++)
        f[o + 3] = axis == 0 ? A3(a, i + o, j, k) : axis == 1 ? A3(a, i, j + o, k) : A3(a, i, j, k + o);
}

/* component: 0 scalar (cell), 1 x-face, 2 y-face, 3 z-face; velocity at the
 * sample point from the MAC ghost arrays (ADVECTION_METHOD_3D.h:201-224). */
static void vel_at(fso_state* s, int comp, int i, int j, int k, double* uu, double* vv, double* ww) {
    const double *gu = s->gu, *gv = s->gv, *gw = s->gw;
    if (comp == 0) {
        *uu = (A3(gu, i + 1, j, k) + A3(gu, i, j, k)) * (double)0.5;
        *vv = (A3(gv, i, j + 1, k) + A3(gv, i, j, k)) * (double)0.5;
        *ww = (A3(gw, i, j, k + 1) + A3(gw, i, j, k)) * (double)0.5;
    } else if (comp == 1) {
        *uu = A3(gu, i, j, k);
        *vv = (A3(gv, i, j, k) + A3(gv, i, j + 1, k) + A3(gv, i - 1, j, k) + A3(gv, i - 1, j + 1, k)) * (double)0.25;
        *ww = (A3(gw, i, j, k) + A3(gw, i, j, k + 1) + A3(gw, i - 1, j, k) + A3(gw, i - 1, j, k + 1)) * (double)0.25;
    } else if (comp == 2) {
        *uu = (A3(gu, i, j, k) + A3(gu, i + 1, j, k) + A3(gu, i, j - 1, k) + A3(gu, i + 1, j - 1, k)) * (double)0.25;
        *vv = A3(gv, i, j, k);
        *ww = (A3(gw, i, j, k) + A3(gw, i, j, k + 1) + A3(gw, i, j - 1, k) + A3(gw, i, j - 1, k + 1)) * (double)0.25;
    } else {
        *uu = (A3(gu, i, j, k) + A3(gu, i + 1, j, k) + A3(gu, i, j, k - 1) + A3(gu, i + 1, j, k - 1)) * (double)0.25;
        *vv = (A3(gv, i, j, k) + A3(gv, i, j + 1, k) + A3(gv, i, j, k - 1) + A3(gv, i, j + 1, k - 1)) * (double)0.25;
        *ww = A3(gw, i, j, k);
    }
}

/* Upwind Euler update, ADVECTION_METHOD_3D.h:226-275. */
static double upwind_update(double rho, double dt, double uu, double vv, double ww, double mx, double px, double my,
                            double py, double mz, double pz) {
    double X = uu > 0 ? mx : px;
    double Y = vv > 0 ? my : py;
    double Z = ww > 0 ? mz : pz;
    return rho - dt * (uu * X + vv * Y + ww * Z);
}

static void extents(fso_state* s, int comp, int* ni, int* nj, int* nk) {
    *ni = s->nx + (comp == 1);
    *nj = s->ny + (comp == 2);
    *nk = s->nz + (comp == 3);
}

static const gc_t* grid_of(fso_state* s, int comp) {
    return comp == 0 ? &s->cell : comp == 1 ? &s->mx : comp == 2 ? &s->my : &s->mz;
}

/* ADVECTION_METHOD_3D::WENO5th, ADVECTION_METHOD_3D.h:42-279: one Euler
 * stage on field `rho` (component comp) reading the ghost copy `ghost`. */
static void weno5_stage(fso_state* s, double* rho, const double* ghost, int comp, double dt, double eps) {
    int ni, nj, nk;
    extents(s, comp, &ni, &nj, &nk);
    const gc_t* gc = grid_of(s, comp);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                double f[7], mx, px, my, py, mz, pz, uu, vv, ww;
                gather7(s, ghost, i, j, k, 0, f);
                fso_weno5_1d(f, gc->x.inv, eps, &mx, &px);
                gather7(s, ghost, i, j, k, 1, f);
                fso_weno5_1d(f, gc->y.inv, eps, &my, &py);
                gather7(s, ghost, i, j, k, 2, f);
                fso_weno5_1d(f, gc->z.inv, eps, &mz, &pz);
                vel_at(s, comp, i, j, k, &uu, &vv, &ww);
                A3(rho, i, j, k) = upwind_update(A3(rho, i, j, k), dt, uu, vv, ww, mx, px, my, py, mz, pz);
            }
}

/* ADVECTION_METHOD_3D::ENO3rd, ADVECTION_METHOD_3D.h:1639-1934 (note
 * one_over_3dz is built from one_over_dx, :1649). */
static void eno3_stage(fso_state* s, double* rho, const double* ghost, int comp, double dt) {
    int ni, nj, nk;
    extents(s, comp, &ni, &nj, &nk);
    const gc_t* gc = grid_of(s, comp);
    double inv3x = gc->x.inv * (double)1 / 3, inv3y = gc->y.inv * (double)1 / 3, inv3z = gc->x.inv * (double)1 / 3;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                double f[7], mx, px, my, py, mz, pz, uu, vv, ww;
                gather7(s, ghost, i, j, k, 0, f);
                fso_eno3_1d(f, gc->x.d, gc->x.inv, gc->x.inv2d, inv3x, &mx, &px);
                gather7(s, ghost, i, j, k, 1, f);
                fso_eno3_1d(f, gc->y.d, gc->y.inv, gc->y.inv2d, inv3y, &my, &py);
                gather7(s, ghost, i, j, k, 2, f);
                fso_eno3_1d(f, gc->z.d, gc->z.inv, gc->z.inv2d, inv3z, &mz, &pz);
                vel_at(s, comp, i, j, k, &uu, &vv, &ww);
                A3(rho, i, j, k) = upwind_update(A3(rho, i, j, k), dt, uu, vv, ww, mx, px, my, py, mz, pz);
            }
}

/* ------------------------------------------------------------------------- */
/* Phases                                                                     */
/* ------------------------------------------------------------------------- */

static int is_fixed_schedule(const fso_state* s) { return s->P.schedule == FS_SCHEDULE_FIXED; }

static int curv_normals(const fso_state* s) {
    return s->P.curvature_method == FS_CURV_NORMALS || s->P.curvature_method == FS_CURV_BOTH;
}
static int curv_levelset(const fso_state* s) {
    return s->P.curvature_method == FS_CURV_LEVELSET || s->P.curvature_method == FS_CURV_BOTH;
}

/* The unit normals of ComputeNormals, LEVELSET_3D.h:933-955: central
 * differences of phi, forward ones where those vanish, each vector scaled by
 * the reciprocal of its magnitude unless that is zero (VECTOR_3D::Normalize,
 * VECTOR_3D.h:153-165).  phi's ghosts are read as its constant fill (:931). */
static void normals_from_phi(fso_state* s) {
    const gc_t* g = &s->cell;
    const double* phi = s->phi;
#define PC(a, b, c) A3(phi, clampi(a, 0, s->nx - 1), clampi(b, 0, s->ny - 1), clampi(c, 0, s->nz - 1))
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double x = (PC(i + 1, j, k) - PC(i - 1, j, k)) * g->x.inv2d;
                double y = (PC(i, j + 1, k) - PC(i, j - 1, k)) * g->y.inv2d;
                double z = (PC(i, j, k + 1) - PC(i, j, k - 1)) * g->z.inv2d;
                double mag = sqrt(x * x + y * y + z * z);
                if (!(mag != 0)) {
                    x = (PC(i + 1, j, k) - PC(i, j, k)) * g->x.inv;
                    y = (PC(i, j + 1, k) - PC(i, j, k)) * g->y.inv;
                    z = (PC(i, j, k + 1) - PC(i, j, k)) * g->z.inv;
                }
                double m = sqrt(x * x + y * y + z * z);
                if (m != 0) {
                    double r = (double)1 / m;
                    x *= r;
                    y *= r;
                    z *= r;
                }
                A3(s->nrm[0], i, j, k) = x;
                A3(s->nrm[1], i, j, k) = y;
                A3(s->nrm[2], i, j, k) = z;
            }
#undef PC
}

/* LEVELSET_3D::ComputeCurvatures (LEVELSET_3D.h:963-1082): by the normals
 * (:974-996) when curvature_method asks for them, then by the level set
 * (:1001-1080), on phi with the ghosts the caller left.  The derivative fields
 * phi_x .. phi_yz of :1005-1013 are folded into the cell loop (each is read
 * only at its own cell).  The compatible schedule keeps the reference's
 * cross term "py pz pyz" (:1033, the mean-curvature formula has 2 py pz pyz)
 * and its one-sided z difference scaled by dz (:1037, 1/dz intended); the
 * fixed schedule uses 2 and 1/dz.  Where the one-sided gradient vanishes as
 * well the reference stops the program (:1040-1044): here the curvature is
 * 0 and curv_fail is set.  jc (may be NULL) = sigma * curvature, at interior
 * cells only (PROJECTION_3D.h:752-767). */
static void curvature(fso_state* s, double* jc) {
    const gc_t* g = &s->cell;
    const double* phi = s->phi;
    double tolerance = (double)1 / fmin(fmin(g->x.d, g->y.d), g->z.d); /* MIN3 of finite spacings */
    int exact = is_fixed_schedule(s);
    double abs_max_curvature = 0;
    s->curv_fail = 0;
    if (curv_normals(s)) {
        /* LEVELSET_3D.h:974-996: the divergence of the normals (their ghost
         * layer the constant fill, :976), |curv| above the tolerance taken as
         * -tolerance, the maximum of |curv| over every cell */
        if (exact) normals_from_phi(s);
        double* const* n = s->nrm;
#define NC(f, a, b, c) A3(f, clampi(a, 0, s->nx - 1), clampi(b, 0, s->ny - 1), clampi(c, 0, s->nz - 1))
        for (int k = 0; k < s->nz; k++)
            for (int j = 0; j < s->ny; j++)
                for (int i = 0; i < s->nx; i++) {
                    double curv = (NC(n[0], i + 1, j, k) - NC(n[0], i - 1, j, k)) * g->x.inv2d;
                    curv += (NC(n[1], i, j + 1, k) - NC(n[1], i, j - 1, k)) * g->y.inv2d;
                    curv += (NC(n[2], i, j, k + 1) - NC(n[2], i, j, k - 1)) * g->z.inv2d;
                    double kv = fabs(curv) <= tolerance ? curv : -tolerance;
                    A3(s->kappa, i, j, k) = kv;
                    if (jc) A3(jc, i, j, k) = s->P.surface_tension * kv;
                    abs_max_curvature = fabs(curv) > abs_max_curvature ? fabs(curv) : abs_max_curvature;
                }
#undef NC
    }
    if (!curv_levelset(s)) {
        s->kmax = abs_max_curvature;
        return;
    }
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double px = (A3(phi, i + 1, j, k) - A3(phi, i - 1, j, k)) * g->x.inv2d;
                double py = (A3(phi, i, j + 1, k) - A3(phi, i, j - 1, k)) * g->y.inv2d;
                double pz = (A3(phi, i, j, k + 1) - A3(phi, i, j, k - 1)) * g->z.inv2d;
                double pxx = (A3(phi, i + 1, j, k) - 2 * A3(phi, i, j, k) + A3(phi, i - 1, j, k)) * g->x.inv_d2;
                double pyy = (A3(phi, i, j + 1, k) - 2 * A3(phi, i, j, k) + A3(phi, i, j - 1, k)) * g->y.inv_d2;
                double pzz = (A3(phi, i, j, k + 1) - 2 * A3(phi, i, j, k) + A3(phi, i, j, k - 1)) * g->z.inv_d2;
                double pxy = (A3(phi, i + 1, j + 1, k) - A3(phi, i + 1, j - 1, k) - A3(phi, i - 1, j + 1, k) +
                              A3(phi, i - 1, j - 1, k)) * g->x.inv2d * g->y.inv2d;
                double pxz = (A3(phi, i + 1, j, k + 1) - A3(phi, i + 1, j, k - 1) - A3(phi, i - 1, j, k + 1) +
                              A3(phi, i - 1, j, k - 1)) * g->x.inv2d * g->z.inv2d;
                double pyz = (A3(phi, i, j + 1, k + 1) - A3(phi, i, j + 1, k - 1) - A3(phi, i, j - 1, k + 1) +
                              A3(phi, i, j - 1, k - 1)) * g->y.inv2d * g->z.inv2d;
                double magnitude = sqrt(px * px + py * py + pz * pz);
                double deno = magnitude * magnitude * magnitude;
                double curv = 0;
                if (deno == 0) {
                    px = (A3(phi, i + 1, j, k) - A3(phi, i, j, k)) * g->x.inv;
                    py = (A3(phi, i, j + 1, k) - A3(phi, i, j, k)) * g->y.inv;
                    pz = (A3(phi, i, j, k + 1) - A3(phi, i, j, k)) * (exact ? g->z.inv : g->z.d);
                    magnitude = sqrt(px * px + py * py + pz * pz);
                    deno = magnitude * magnitude * magnitude;
                    if (deno == 0) s->curv_fail = 1;
                }
                if (deno != 0) {
                    double one_over_deno = (double)1 / deno;
                    double cross_yz = exact ? (double)2 * py * pz * pyz : py * pz * pyz;
                    curv = -(px * px * pyy - (double)2 * px * py * pxy + py * py * pxx + px * px * pzz -
                             (double)2 * px * pz * pxz + pz * pz * pxx + py * py * pzz - cross_yz + pz * pz * pyy) *
                           one_over_deno;
                }
                double kv = curv < -tolerance ? -tolerance : curv; /* :1061-1068 */
                A3(s->kappa, i, j, k) = kv;
                if (jc) A3(jc, i, j, k) = s->P.surface_tension * kv;
                if (fabs(A3(phi, i, j, k)) < 1e-1) /* :1070-1077, MAX skips NaN */
                    abs_max_curvature = fabs(kv) > abs_max_curvature ? fabs(kv) : abs_max_curvature;
            }
    s->kmax = abs_max_curvature;
}

/* LEVELSET_3D::ComputeCurvatures, LEVELSET_3D.h:963-1082: linear
 * re-extrapolation of phi's own ghost cells (:998), then the curvature.  With
 * zero surface tension nothing reads the curvature; the ghost refresh is the
 * call's lasting effect. */
int fso_compute_curvatures(fso_state* s) {
    fill_linear(s, s->phi, s->phi, s->nx, s->ny, s->nz, 2, 0, 1);
    if (s->P.surface_tension != 0.0) curvature(s, NULL);
    return FS_OK;
}

static void get(fso_state* s, const double* src, double* dst, int ni, int nj, int nk);

int fso_get_curvature(fso_state* s, double* kappa, double* abs_max_curvature) {
    if (s->P.surface_tension == 0.0) return FS_ERR_UNSUPPORTED;
    fso_compute_curvatures(s);
    if (s->curv_fail) return FS_ERR_NONFINITE;
    if (kappa) get(s, s->kappa, kappa, s->nx, s->ny, s->nz);
    if (abs_max_curvature) *abs_max_curvature = s->kmax;
    return FS_OK;
}

/* jc_on_solution as the pressure solve reads it, or NULL where it stays at
 * its initial zeros (no surface tension, or use_delta_function_formulation
 * off: PROJECTION_3D.h:750). */
static const double* jump_field(const fso_state* s) {
    if (s->jc_override) return s->jc_override;
    return s->P.surface_tension != 0.0 && s->P.use_delta_function_formulation && !s->csf_on &&
                   s->P.curvature_method != FS_CURV_NONE
               ? s->jc
               : NULL;
}

/* LEVELSET_3D::ComputeNormals, LEVELSET_3D.h:929-956: lasting effect is the
 * constant re-extrapolation of phi's ghost cells (:931); normals only feed
 * jump terms that vanish with zero surface tension. */
/* LEVELSET_3D::ComputeNormals, LEVELSET_3D.h:929-956: the lasting effect on
 * phi is its constant ghost fill; the normals themselves are kept only where
 * the curvature is taken from them (compatible schedule: read by the next
 * ComputeCurvatures). */
static void compute_normals(fso_state* s) {
    fill_const(s, s->phi, s->phi, s->nx, s->ny, s->nz, 2, 0);
    if (s->P.surface_tension != 0.0 && curv_normals(s) && s->P.schedule != FS_SCHEDULE_FIXED) normals_from_phi(s);
}

static int is_fixed(const fso_state* s) { return s->P.schedule == FS_SCHEDULE_FIXED; }

/* Fixed schedule: ghost copies made from the current fields.  The level
 * set's ghost layers extrapolate its own interior linearly (faces only, each
 * layer from the layers written before it); velocities get zero faces. */
static void refresh_levelset_ghost(fso_state* s) {
    copy_interior(s, s->sghost, s->phi, s->nx, s->ny, s->nz);
    fill_linear(s, s->sghost, s->sghost, s->nx, s->ny, s->nz, 3, 0, 0);
}
static void refresh_velocity_ghosts(fso_state* s) {
    fill_zero(s, s->gu, s->u, s->nx + 1, s->ny, s->nz, 3);
    fill_zero(s, s->gv, s->v, s->nx, s->ny + 1, s->nz, 3);
    fill_zero(s, s->gw, s->w, s->nx, s->ny, s->nz + 1, 3);
}
/* Fixed schedule: the two normal wall planes of a MAC component are zero. */
static void zero_walls(fso_state* s, int comp) {
    int ni, nj, nk;
    extents(s, comp, &ni, &nj, &nk);
    double* vel = comp == 1 ? s->u : comp == 2 ? s->v : s->w;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                int wall = comp == 1 ? (i == 0 || i == ni - 1) : comp == 2 ? (j == 0 || j == nj - 1)
                                                                           : (k == 0 || k == nk - 1);
                if (wall) A3(vel, i, j, k) = 0;
            }
}

/* ADVECTION_3D::Solve_Levelset, ADVECTION_3D.h:204-242 (air/water, WENO5):
 * the compat schedule reads the shared ghost copy and the MAC ghost copies
 * as they are; the fixed schedule refreshes them first. */
int fso_levelset_advection(fso_state* s, double dt) {
    if (is_fixed(s)) {
        refresh_levelset_ghost(s);
        refresh_velocity_ghosts(s);
    }
    weno5_stage(s, s->phi, s->sghost, 0, dt, s->P.levelset_epsilon);
    return FS_OK;
}

/* ADVECTION_3D::Solve_Velocity, ADVECTION_3D.h:244-358 (air/water branch).
 * ENO3 reads each component's own ghost copy; the WENO5 branch passes the
 * x ghost copy for all three components (:314-316), the fixed schedule each
 * component's own, refreshed first. */
int fso_velocity_advection(fso_state* s, double dt) {
    if (is_fixed(s)) refresh_velocity_ghosts(s);
    if (s->P.use_5th_weno_v) {
        int own = is_fixed(s);
        weno5_stage(s, s->u, s->gu, 1, dt, s->P.velocity_epsilon);
        weno5_stage(s, s->v, own ? s->gv : s->gu, 2, dt, s->P.velocity_epsilon);
        weno5_stage(s, s->w, own ? s->gw : s->gu, 3, dt, s->P.velocity_epsilon);
    } else {
        eno3_stage(s, s->u, s->gu, 1, dt);
        eno3_stage(s, s->v, s->gv, 2, dt);
        eno3_stage(s, s->w, s->gw, 3, dt);
    }
    return FS_OK;
}

/* EULERIAN_FLUID_SOLVER_3D::ApplyGravity, :2237-2270: every face. */
int fso_apply_gravity(fso_state* s, double dt) {
    double gx = dt * s->P.gravity[0], gy = dt * s->P.gravity[1], gz = dt * s->P.gravity[2];
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i <= s->nx; i++) A3(s->u, i, j, k) += gx;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j <= s->ny; j++)
            for (int i = 0; i < s->nx; i++) A3(s->v, i, j, k) += gy;
    for (int k = 0; k <= s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) A3(s->w, i, j, k) += gz;
    if (is_fixed(s))
        for (int c = 1; c <= 3; c++) zero_walls(s, c);
    return FS_OK;
}

/* VISCOSITY_3D::HeavisideFunction, VISCOSITY_3D.h:1234-1256. */
static double heaviside_visc(double phi, double eps) {
    double one_over_epsilon = (double)1 / eps, one_over_pi = (double)1 / FS_PI;
    if (phi < -eps) return 0;
    if (phi > eps) return 1;
    return (double)0.5 * ((double)1 + phi * one_over_epsilon + one_over_pi * sin(FS_PI * phi * one_over_epsilon));
}

/* PROJECTION_3D::HeavisideFunction, PROJECTION_3D.h:1452-1472. */
static double heaviside_proj(double phi, double eps) {
    double one_over_epsilon = (double)1 / eps, one_over_pi = (double)1 / FS_PI;
    if (phi < -eps) return 0;
    if (phi > eps) return 1;
    return (double)0.5 + phi * one_over_epsilon * (double)0.5 + (double)0.5 * one_over_pi * sin(FS_PI * phi * one_over_epsilon);
}

static double mix_density(const fs_params* P, double H) {
    return P->air_bubble_rising ? P->air_density + (P->water_density - P->air_density) * H
                                : P->water_density + (P->air_density - P->water_density) * H;
}
static double mix_viscosity(const fs_params* P, double H) {
    return P->air_bubble_rising ? P->air_viscosity + (P->water_viscosity - P->air_viscosity) * H
                                : P->water_viscosity + (P->air_viscosity - P->water_viscosity) * H;
}

/* One face of the explicit viscous update; reads velocities from (ux,vy,wz). */
static double visc_face(fso_state* s, int comp, int i, int j, int k, double dt, const double* ux, const double* vy,
                        const double* wz) {
    const fs_params* P = &s->P;
    const double* f = s->phi;
    const double eps = (double)1.5 * s->cell.x.d; /* epsilon_for_mollification, EULERIAN_FLUID_SOLVER_3D.h:393 */
    const double odx = s->cell.x.inv, ody = s->cell.y.inv, odz = s->cell.z.inv;
#define PH(a, b, c) A3(f, a, b, c)
#define U(a, b, c) A3(ux, a, b, c)
#define V(a, b, c) A3(vy, a, b, c)
#define W(a, b, c) A3(wz, a, b, c)
#define MU(x) mix_viscosity(P, heaviside_visc((x), eps))
    if (comp == 1) { /* VISCOSITY_3D.h:462-491 */
        double lh = (double)0.5 * (PH(i - 1, j, k) + PH(i, j, k));
        double lcu = (double)0.25 * (PH(i - 1, j, k) + PH(i, j, k) + PH(i - 1, j + 1, k) + PH(i, j + 1, k));
        double lcd = (double)0.25 * (PH(i - 1, j, k) + PH(i, j, k) + PH(i - 1, j - 1, k) + PH(i, j - 1, k));
        double lct = (double)0.25 * (PH(i - 1, j, k) + PH(i, j, k) + PH(i - 1, j, k + 1) + PH(i, j, k + 1));
        double lcb = (double)0.25 * (PH(i - 1, j, k) + PH(i, j, k) + PH(i - 1, j, k - 1) + PH(i, j, k - 1));
        double rho_h = mix_density(P, heaviside_visc(lh, eps));
        double mcu = MU(lcu), mcd = MU(lcd), mr = MU(PH(i, j, k)), mc = MU(PH(i - 1, j, k)), mct = MU(lct), mcb = MU(lcb);
        double coef = dt * ((double)1 / rho_h);
        double first = (double)2 * coef * odx * (mr * odx * (U(i + 1, j, k) - U(i, j, k)) - mc * ody * (U(i, j, k) - U(i - 1, j, k)));
        double second = coef * ody * (mcu * (ody * (U(i, j + 1, k) - U(i, j, k)) + odx * (V(i, j + 1, k) - V(i - 1, j + 1, k))) -
                                      mcd * (ody * (U(i, j, k) - U(i, j - 1, k)) + odx * (V(i, j, k) - V(i - 1, j, k))));
        double third = coef * odz * (mct * (odz * (U(i, j, k + 1) - U(i, j, k)) + odx * (W(i, j, k + 1) - W(i - 1, j, k + 1))) -
                                     mcb * (odz * (U(i, j, k) - U(i, j, k - 1)) + odx * (W(i, j, k) - W(i - 1, j, k))));
        return first + second + third;
    } else if (comp == 2) { /* VISCOSITY_3D.h:496-524 */
        double lh = (double)0.5 * (PH(i, j - 1, k) + PH(i, j, k));
        double lcu = (double)0.25 * (PH(i, j - 1, k) + PH(i + 1, j - 1, k) + PH(i, j, k) + PH(i + 1, j, k));
        double lcl = (double)0.25 * (PH(i - 1, j - 1, k) + PH(i, j - 1, k) + PH(i - 1, j, k) + PH(i, j, k));
        double lct = (double)0.25 * (PH(i, j - 1, k + 1) + PH(i, j - 1, k) + PH(i, j, k + 1) + PH(i, j, k));
        double lcb = (double)0.25 * (PH(i, j - 1, k - 1) + PH(i, j - 1, k) + PH(i, j, k - 1) + PH(i, j, k));
        double rho_h = mix_density(P, heaviside_visc(lh, eps));
        double mcu = MU(lcu), mcl = MU(lcl), mu_ = MU(PH(i, j, k)), mc = MU(PH(i, j - 1, k)), mct = MU(lct), mcb = MU(lcb);
        double coef = dt * ((double)1 / rho_h);
        double first = coef * odx * (mcu * (ody * (U(i + 1, j, k) - U(i + 1, j - 1, k)) + odx * (V(i + 1, j, k) - V(i, j, k))) -
                                     mcl * (ody * (U(i, j, k) - U(i, j - 1, k)) + odx * (V(i, j, k) - V(i - 1, j, k))));
        double second = (double)2 * coef * ody * (mu_ * ody * (V(i, j + 1, k) - V(i, j, k)) - mc * ody * (V(i, j, k) - V(i, j - 1, k)));
        double third = coef * odz * (mct * (odz * (V(i, j, k + 1) - V(i, j, k)) + ody * (W(i, j, k) - W(i, j - 1, k))) -
                                     mcb * (odz * (V(i, j, k) - V(i, j, k - 1)) + ody * (W(i, j, k - 1) - W(i, j - 1, k - 1))));
        return first + second + third;
    } else { /* VISCOSITY_3D.h:530-559 */
        double lh = (double)0.5 * (PH(i, j, k - 1) + PH(i, j, k));
        double lct = (double)0.25 * (PH(i, j, k - 1) + PH(i + 1, j, k - 1) + PH(i, j, k) + PH(i + 1, j, k));
        double lcb = (double)0.25 * (PH(i, j, k - 1) + PH(i - 1, j, k - 1) + PH(i, j, k) + PH(i - 1, j, k));
        double lcu = (double)0.25 * (PH(i, j, k - 1) + PH(i, j + 1, k - 1) + PH(i, j, k) + PH(i, j + 1, k));
        double lcd = (double)0.25 * (PH(i, j, k - 1) + PH(i, j - 1, k - 1) + PH(i, j, k) + PH(i, j - 1, k));
        double rho_h = mix_density(P, heaviside_visc(lh, eps));
        double mct = MU(lct), mcb = MU(lcb), mcu = MU(lcu), mcd = MU(lcd), mt = MU(PH(i, j, k)), mc = MU(PH(i, j, k - 1));
        double coef = dt * ((double)1 / rho_h);
        double first = coef * odx * (mct * (odz * (U(i, j, k) - U(i, j, k - 1)) + odx * (W(i + 1, j, k) - W(i, j, k))) -
                                     mcb * (odz * (U(i - 1, j, k) - U(i - 1, j, k - 1)) + odx * (W(i, j, k) - W(i - 1, j, k))));
        double second = coef * ody * (mcu * (odz * (V(i, j, k) - V(i, j, k - 1)) + ody * (W(i, j + 1, k) - W(i, j, k))) -
                                      mcd * (odz * (V(i, j - 1, k) - V(i, j - 1, k - 1)) + ody * (W(i, j, k) - W(i, j - 1, k))));
        double third = (double)2 * coef * odz * (mt * odz * (W(i, j, k + 1) - W(i, j, k)) - mc * odz * (W(i, j, k) - W(i, j, k - 1)));
        return first + second + third;
    }
#undef PH
#undef U
#undef V
#undef W
#undef MU
}

typedef struct {
    int N, per, nplanes;
    int *rp, *ci;
    double *val, *b, *x, *r, *zv, *pp, *Ap;
} vsys_t;

static sys_t as_sys(vsys_t* V) {
    sys_t S = {V->N, V->per, V->nplanes, V->rp, V->ci, V->val, V->b, V->x, V->r, V->zv, V->pp, V->Ap};
    return S;
}
static void pcg_sys(fso_state* s, const sys_t* S);

/* Rows of the semi-implicit viscosity system of component comp on its face
 * box (ni, nj, nk), numbered i fastest:
 *   coefficients and explicit term, VISCOSITY_3D.h:853-989 (air/water
 *   density and viscosity mixes, :1258-1347), built from the velocities at
 *   entry (ou, ov, ow: the reference reads ghost copies made before the first
 *   solve); live = the component itself (x0 and the Dirichlet values);
 *   rows, BuildLinearSystemNodeForSemiImplicitViscosity,
 *   POISSON_SOLVER_3D.h:3359-3496, with every box face an unknown and the
 *   ring beyond it Dirichlet (regular_boundary, VISCOSITY_3D.h:1403-1417).
 * The cross terms of x and z are those of :887 and :975; the y term's z part
 * is the explicit update's (:522), where :930 differences w across x. */
static void visc_rows(fso_state* s, int comp, double dt, const double* ou, const double* ov, const double* ow,
                      const double* live, vsys_t* V) {
    const fs_params* P = &s->P;
    const double* f = s->phi;
    const double eps = (double)1.5 * s->cell.x.d;
    const double odx = s->cell.x.inv, ody = s->cell.y.inv, odz = s->cell.z.inv;
    const double idxx = (double)1 / (s->cell.x.d * s->cell.x.d), idyy = (double)1 / (s->cell.y.d * s->cell.y.d),
                 idzz = (double)1 / (s->cell.z.d * s->cell.z.d);
    int ni, nj, nk;
    extents(s, comp, &ni, &nj, &nk);
#define PH(a, b, c) A3(f, a, b, c)
#define U(a, b, c) A3(ou, a, b, c)
#define V_(a, b, c) A3(ov, a, b, c)
#define W(a, b, c) A3(ow, a, b, c)
#define MU(x) mix_viscosity(P, heaviside_visc((x), eps))
    int nnz = 0;
    V->rp[0] = 0;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                int row = i + ni * (j + nj * k);
                double c1, c2, c3, c4, c5, c6, ex;
                if (comp == 1) {
                    double lh = (double)0.5 * (PH(i - 1, j, k) + PH(i, j, k));
                    double lcu = (double)0.25 * (PH(i - 1, j, k) + PH(i, j, k) + PH(i - 1, j + 1, k) + PH(i, j + 1, k));
                    double lcd = (double)0.25 * (PH(i - 1, j, k) + PH(i, j, k) + PH(i - 1, j - 1, k) + PH(i, j - 1, k));
                    double lct = (double)0.25 * (PH(i - 1, j, k) + PH(i, j, k) + PH(i - 1, j, k + 1) + PH(i, j, k + 1));
                    double lcb = (double)0.25 * (PH(i - 1, j, k) + PH(i, j, k) + PH(i - 1, j, k - 1) + PH(i, j, k - 1));
                    double rh = mix_density(P, heaviside_visc(lh, eps));
                    double mcu = MU(lcu), mcd = MU(lcd), mr = MU(PH(i, j, k)), mc = MU(PH(i - 1, j, k)), mct = MU(lct),
                           mcb = MU(lcb);
                    c1 = (double)2 * dt / rh * mr;
                    c2 = (double)2 * dt / rh * mc;
                    c3 = dt / rh * mcu;
                    c4 = dt / rh * mcd;
                    c5 = dt / rh * mct;
                    c6 = dt / rh * mcb;
                    ex = U(i, j, k) +
                         dt / rh * (ody * (mcu * odx * (V_(i, j + 1, k) - V_(i - 1, j + 1, k)) - mcd * odx * (V_(i, j, k) - V_(i - 1, j, k)))) +
                         dt / rh * (odz * (mct * odx * (W(i, j, k + 1) - W(i - 1, j, k + 1)) - mcb * odx * (W(i, j, k) - W(i - 1, j, k))));
                } else if (comp == 2) {
                    double lh = (double)0.5 * (PH(i, j - 1, k) + PH(i, j, k));
                    double lcu = (double)0.25 * (PH(i, j - 1, k) + PH(i + 1, j - 1, k) + PH(i, j, k) + PH(i + 1, j, k));
                    double lcl = (double)0.25 * (PH(i - 1, j - 1, k) + PH(i, j - 1, k) + PH(i - 1, j, k) + PH(i, j, k));
                    double lct = (double)0.25 * (PH(i, j - 1, k + 1) + PH(i, j - 1, k) + PH(i, j, k + 1) + PH(i, j, k));
                    double lcb = (double)0.25 * (PH(i, j - 1, k - 1) + PH(i, j - 1, k) + PH(i, j, k - 1) + PH(i, j, k));
                    double rh = mix_density(P, heaviside_visc(lh, eps));
                    double mcu = MU(lcu), mcl = MU(lcl), mu_ = MU(PH(i, j, k)), mc = MU(PH(i, j - 1, k)), mct = MU(lct),
                           mcb = MU(lcb);
                    c1 = dt / rh * mcu;
                    c2 = dt / rh * mcl;
                    c3 = (double)2 * dt / rh * mu_;
                    c4 = (double)2 * dt / rh * mc;
                    c5 = dt / rh * mct;
                    c6 = dt / rh * mcb;
                    ex = V_(i, j, k) +
                         dt / rh * (odx * (mcu * ody * (U(i + 1, j, k) - U(i + 1, j - 1, k)) - mcl * ody * (U(i, j, k) - U(i, j - 1, k)))) +
                         dt / rh * (odz * (mct * ody * (W(i, j, k) - W(i, j - 1, k)) - mcb * ody * (W(i, j, k - 1) - W(i, j - 1, k - 1))));
                } else {
                    double lh = (double)0.5 * (PH(i, j, k - 1) + PH(i, j, k));
                    double lct = (double)0.25 * (PH(i, j, k - 1) + PH(i + 1, j, k - 1) + PH(i, j, k) + PH(i + 1, j, k));
                    double lcb = (double)0.25 * (PH(i, j, k - 1) + PH(i - 1, j, k - 1) + PH(i, j, k) + PH(i - 1, j, k));
                    double lcu = (double)0.25 * (PH(i, j, k - 1) + PH(i, j + 1, k - 1) + PH(i, j, k) + PH(i, j + 1, k));
                    double lcd = (double)0.25 * (PH(i, j, k - 1) + PH(i, j - 1, k - 1) + PH(i, j, k) + PH(i, j - 1, k));
                    double rh = mix_density(P, heaviside_visc(lh, eps));
                    double mct = MU(lct), mcb = MU(lcb), mcu = MU(lcu), mcd = MU(lcd), mt = MU(PH(i, j, k)),
                           mc = MU(PH(i, j, k - 1));
                    c1 = dt / rh * mct;
                    c2 = dt / rh * mcb;
                    c3 = dt / rh * mcu;
                    c4 = dt / rh * mcd;
                    c5 = (double)2 * dt / rh * mt;
                    c6 = (double)2 * dt / rh * mc;
                    ex = A3(live, i, j, k) +
                         dt / rh * (odx * odz * (mct * (U(i, j, k) - U(i, j, k - 1)) - mcb * (U(i - 1, j, k) - U(i - 1, j, k - 1)))) +
                         dt / rh * (ody * odz * (mcu * (V_(i, j, k) - V_(i, j, k - 1)) - mcd * (V_(i, j - 1, k) - V_(i, j - 1, k - 1))));
                }
                /* neighbours in the order the reference assigns them */
                const double nc[6] = {c2, c1, c4, c3, c6, c5};
                const double nv[6] = {idxx, idxx, idyy, idyy, idzz, idzz};
                const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
                double coef = 0;
                for (int e = 0; e < 6; e++) {
                    int a = i + di[e], bb = j + dj[e], cc = k + dk[e];
                    if (a >= 0 && a < ni && bb >= 0 && bb < nj && cc >= 0 && cc < nk) {
                        coef += nc[e];
                        V->ci[nnz] = a + ni * (bb + nj * cc);
                        V->val[nnz++] = -nc[e] * nv[e];
                    }
                }
                for (int e = 0; e < 6; e++) {
                    int a = i + di[e], bb = j + dj[e], cc = k + dk[e];
                    if (!(a >= 0 && a < ni && bb >= 0 && bb < nj && cc >= 0 && cc < nk)) coef += nc[e];
                }
                if (coef == 0) coef = 1;
                V->ci[nnz] = row;
                V->val[nnz++] = (double)1 + idxx * coef;
                V->rp[row + 1] = nnz;
                double bv = ex;
                for (int e = 0; e < 6; e++) {
                    int a = i + di[e], bb = j + dj[e], cc = k + dk[e];
                    if (!(a >= 0 && a < ni && bb >= 0 && bb < nj && cc >= 0 && cc < nk))
                        bv += nc[e] * nv[e] * A3(live, a, bb, cc);
                }
                V->b[row] = bv;
                V->x[row] = A3(live, i, j, k);
            }
#undef PH
#undef U
#undef V_
#undef W
#undef MU
}

/* VISCOSITY_3D::ApplyViscosity, semi-implicit branch (VISCOSITY_3D.h:853-989)
 * in the air/water box: components x, y, z solved in turn
 * (POISSON_SOLVER_3D::SolveForViscosity, :254-266) with the Jacobi-PCG and
 * stop rule of PROJECTION_WATER; the systems read the velocities at entry. */
static int semi_implicit_viscosity(fso_state* s, double dt, int* iterations) {
    int rc = FS_OK;
    long n = s->ntot;
    double* old[3];
    for (int c = 0; c < 3; c++) {
        old[c] = zalloc(n);
        memcpy(old[c], c == 0 ? s->u : (c == 1 ? s->v : s->w), sizeof(double) * (size_t)n);
    }
    double* comps[3] = {s->u, s->v, s->w};
    for (int comp = 1; comp <= 3; comp++) {
        int ni, nj, nk;
        extents(s, comp, &ni, &nj, &nk);
        vsys_t V;
        V.N = ni * nj * nk;
        V.per = ni * nj;
        V.nplanes = nk;
        V.rp = (int*)calloc((size_t)V.N + 1, sizeof(int));
        V.ci = (int*)calloc((size_t)V.N * 7, sizeof(int));
        V.val = (double*)calloc((size_t)V.N * 7, sizeof(double));
        double** vecs[6] = {&V.b, &V.x, &V.r, &V.zv, &V.pp, &V.Ap};
        for (int q = 0; q < 6; q++) *vecs[q] = (double*)calloc((size_t)V.N, sizeof(double));
        visc_rows(s, comp, dt, old[0], old[1], old[2], comps[comp - 1], &V);
        sys_t S = as_sys(&V);
        pcg_sys(s, &S);
        if (iterations) *iterations += s->last_iters;
        if (s->last_iters >= s->P.max_iteration) rc = FS_ERR_NOT_CONVERGED;
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) A3(comps[comp - 1], i, j, k) = V.x[i + ni * (j + nj * k)];
        free(V.rp); free(V.ci); free(V.val);
        for (int q = 0; q < 6; q++) free(*vecs[q]);
    }
    for (int c = 0; c < 3; c++) free(old[c]);
    return rc;
}

/* EULERIAN_FLUID_SOLVER_3D::Sourcing (:2076-2100) -> VISCOSITY_3D::ApplyViscosity
 * air/water explicit branch, VISCOSITY_3D.h:448-563: components x, y, z in
 * turn, each updated in place (in-place mode) or from a copy of itself;
 * the semi-implicit branch when semi_implicit_viscosity is set.  The fixed
 * schedule reads phi's ghosts as the constant extrapolation of the current
 * level set and zeroes each component's wall faces after its update. */
/* EULERIAN_FLUID_SOLVER_3D::SourcingFromObjectList, :2592-2671, for
 * spherical objects (SIMULATION_WORLD.cpp:355-364).  Each object's cells are
 * those of its AABB grid (SIMULATION_OBJECT::BuildAABBGrid,
 * SIMULATION_OBJECT.h:149-160: ClampedCell of the box corners,
 * GRID_STRUCTURE_3D.h:310-319); its signed distance is taken at CellCenter,
 * which adds half a cell (GRID_STRUCTURE_3D.h:269-272). */
static void sourcing_from_object_list(fso_state* s) {
    const ax_t* ax[3] = {&s->cell.x, &s->cell.y, &s->cell.z};
    const double mn[3] = {s->gr.x_min, s->gr.y_min, s->gr.z_min};
    const int res[3] = {s->nx, s->ny, s->nz};
    for (int o = 0; o < s->nsrc; o++) {
        const fs_levelset_source* b = &s->src[o];
        if (b->mode == FS_SOURCE_NONE || !(b->start_frame <= s->frame && s->frame <= b->stop_frame)) continue;
        int lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            double aabb_min = b->center[a] - b->half_width, aabb_max = b->center[a] + b->half_width;
            double fl = floor((aabb_min - mn[a]) * ax[a]->inv), fh = floor((aabb_max - mn[a]) * ax[a]->inv);
            lo[a] = fl < 0 ? 0 : fl > res[a] - 1 ? res[a] - 1 : (int)fl;
            hi[a] = fh < 0 ? 0 : fh > res[a] - 1 ? res[a] - 1 : (int)fh;
        }
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    double cx = mn[0] + ((double)0.5 + (double)i) * s->cell.x.d;
                    double cy = mn[1] + ((double)0.5 + (double)j) * s->cell.y.d;
                    double cz = mn[2] + ((double)0.5 + (double)k) * s->cell.z.d;
                    double dx = cx - b->center[0], dy = cy - b->center[1], dz = cz - b->center[2];
                    double phi_obj = sqrt(dx * dx + dy * dy + dz * dz) - b->radius; /* SPHERE.h:29-32 */
                    if (phi_obj <= s->cell.x.d * (double)3) {
                        double* f = &A3(s->phi, i, j, k);
                        if (b->mode == FS_SOURCE_ADD) *f = phi_obj > *f ? *f : phi_obj;  /* MIN */
                        else *f = -phi_obj > *f ? -phi_obj : *f;                         /* MAX */
                    }
                }
    }
}

int fso_set_levelset_sources(fso_state* s, int n, const fs_levelset_source* src) {
    if (n < 0 || n > 64 || (n > 0 && !src)) return FS_ERR_BAD_ARG;
    for (int i = 0; i < n; i++) s->src[i] = src[i];
    s->nsrc = n;
    return FS_OK;
}
int fso_set_frame(fso_state* s, int frame) {
    s->frame = frame;
    return FS_OK;
}
int fso_apply_levelset_sources(fso_state* s) {
    sourcing_from_object_list(s);
    return FS_OK;
}

static int apply_viscosity(fso_state* s, double dt);

/* EULERIAN_FLUID_SOLVER_3D::Sourcing, :2076-2100: viscosity, then the objects. */
int fso_sourcing(fso_state* s, double dt) {
    int rc = apply_viscosity(s, dt);
    if (s->nsrc > 0) sourcing_from_object_list(s);
    /* CSF_model, EULERIAN_FLUID_SOLVER_3D.h:2089-2093: ApplySurfaceTension
     * computes the curvature (LEVELSET_3D::ComputeCurvatures, with its ghost
     * refresh of phi, :2104); its force is oil/water-only (:2106-2195).  It
     * sets the projection's CSF flag, which zeroes jc_on_solution from the
     * next projection on (PROJECTION_3D.h:754-757). */
    if (s->P.csf_model) {
        fso_compute_curvatures(s);
        s->csf_on = 1;
    }
    return rc;
}

/* VISCOSITY_3D::ApplyViscosity, VISCOSITY_3D.h:448-563 (and the semi-implicit
 * systems, :853-989). */
static int apply_viscosity(fso_state* s, double dt) {
    if (!s->P.use_delta_function_formulation) return FS_OK;
    if (is_fixed(s)) compute_normals(s);
    if (s->P.semi_implicit_viscosity) {
        int rc = semi_implicit_viscosity(s, dt, &s->visc_iters);
        if (is_fixed(s))
            for (int c = 1; c <= 3; c++) zero_walls(s, c);
        return rc;
    }
    double* comps[3] = {s->u, s->v, s->w};
    for (int c = 1; c <= 3; c++) {
        int ni, nj, nk;
        extents(s, c, &ni, &nj, &nk);
        double* fld = comps[c - 1];
        if (s->visc_inplace) {
            for (int k = 0; k < nk; k++)
                for (int j = 0; j < nj; j++)
                    for (int i = 0; i < ni; i++) A3(fld, i, j, k) += visc_face(s, c, i, j, k, dt, s->u, s->v, s->w);
        } else {
            memcpy(s->vtmp, fld, sizeof(double) * (size_t)s->ntot);
            const double* ux = c == 1 ? s->vtmp : s->u;
            const double* vy = c == 2 ? s->vtmp : s->v;
            const double* wz = c == 3 ? s->vtmp : s->w;
#pragma omp parallel for num_threads(g_threads) schedule(static)
            for (int k = 0; k < nk; k++)
                for (int j = 0; j < nj; j++)
                    for (int i = 0; i < ni; i++) A3(fld, i, j, k) = A3(s->vtmp, i, j, k) + visc_face(s, c, i, j, k, dt, ux, vy, wz);
        }
        if (is_fixed(s)) zero_walls(s, c);
    }
    return FS_OK;
}

/* ------------------------------------------------------------------------- */
/* Projection                                                                 */
/* ------------------------------------------------------------------------- */

/* rows are numbered plane by plane (per rows each); per-plane partials
 * (each summed in row order, planes in parallel), then added in plane order */
static double planes_dot(fso_state* s, int per, int nplanes, const double* a, const double* b) {
    double* part = s->plane_part;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int k = 0; k < nplanes; k++) {
        double acc = 0;
        const double* aa = a + (long)k * per;
        const double* bb = b + (long)k * per;
        for (int r = 0; r < per; r++) acc += aa[r] * bb[r];
        part[k] = acc;
    }
    double total = 0;
    for (int k = 0; k < nplanes; k++) total += part[k];
    return total;
}


/* PROJECTION_3D::DetermineProjectionDensity, PROJECTION_3D.h:596-679. */
static void projection_density(fso_state* s) {
    const fs_params* P = &s->P;
    double light = P->air_bubble_rising ? P->air_density : P->water_density;
    double heavy = P->air_bubble_rising ? P->water_density : P->air_density;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double f = A3(s->phi, i, j, k);
                if (f <= 0) A3(s->rho, i, j, k) = light;
                else if (f > 0) A3(s->rho, i, j, k) = heavy;
            }
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) A3(s->irho, i, j, k) = 1 / A3(s->rho, i, j, k);
}

/* PROJECTION_3D::DetermineDivergence, PROJECTION_3D.h:681-744 (air/water). */
static void projection_divergence(fso_state* s, double dt) {
    fill_zero(s, s->gu, s->u, s->nx + 1, s->ny, s->nz, 3);
    fill_zero(s, s->gv, s->v, s->nx, s->ny + 1, s->nz, 3);
    fill_zero(s, s->gw, s->w, s->nx, s->ny, s->nz + 1, 3);
    double odx = s->cell.x.inv, ody = s->cell.y.inv, odz = s->cell.z.inv;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++)
                A3(s->div, i, j, k) = odx * (A3(s->gu, i + 1, j, k) - A3(s->gu, i, j, k)) +
                                      ody * (A3(s->gv, i, j + 1, k) - A3(s->gv, i, j, k)) +
                                      odz * (A3(s->gw, i, j, k + 1) - A3(s->gw, i, j, k));
    const double one_over_dt = 1 / dt;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) A3(s->div, i, j, k) *= one_over_dt;
}

/* PROJECTION_3D::SetupBoundaryConditions (Neumann, air/water, :1000-1016) and
 * POISSON_SOLVER_3D::AssignSequentialindicesToFullCells (:527-552). */
static void setup_bc(fso_state* s) {
    int row = 0;
    for (int k = -1; k <= s->nz; k++)
        for (int j = -1; j <= s->ny; j++)
            for (int i = -1; i <= s->nx; i++) {
                int inside = i >= 0 && i < s->nx && j >= 0 && j < s->ny && k >= 0 && k < s->nz;
                s->bc[IX(s, i, j, k)] = inside ? FS_BC_FULL : FS_BC_NEUM;
            }
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++)
                if (s->bc[IX(s, i, j, k)] > -1) s->bc[IX(s, i, j, k)] = row++;
    s->N = row;
}

#define GFM_BETA(vc, vn, fn, fc) ((vc) * (vn) * (fabs(fn) + fabs(fc)) / ((vc) * fabs(fn) + (vn) * fabs(fc)))
#define GFM_BETA_R(vc, vn, fn, fc) ((vc) * (vn) * (fabs(fn) + fabs(fc)) / ((vn) * fabs(fc) + (vc) * fabs(fn)))

/* One arm of the jump stencil, POISSON_SOLVER_3D.h:1217-1324: F_L, F_R, ...
 * for the face between cell c and neighbour n.  jc_on_derivative is zero in
 * this solver, so the b_* terms vanish and only +-one_over_d2 a beta stays;
 * + when phi_c <= 0 < phi_n, - when phi_n <= 0 < phi_c, 0 within a phase. */
static double jump_arm(double fc, double fn, double jcc, double jcn, double one_over_d2, double beta) {
    double a = (jcc * fabs(fn) + jcn * fabs(fc)) / (fabs(fc) + fabs(fn));
    if ((fc <= 0 && fn <= 0) || (fc > 0 && fn > 0)) return 0;
    if (fc <= 0 && fn > 0) return one_over_d2 * a * beta;
    return -one_over_d2 * a * beta;
}

/* PROJECTION_3D::DeterminePressure (:794-834) up to the solve:
 * POISSON_SOLVER_3D::BuildLinearSystemNodeJumpConditionVaribleCoefficient
 * (:1030-1358, air/water) and the Neumann mean removal (:225-246). */
static void assemble(fso_state* s) {
    const fs_params* P = &s->P;
    if (!P->use_previous_pressure)
        for (long n = 0; n < s->ntot; n++) s->p[n] = 0;
    setup_bc(s);
    compute_normals(s);
    fill_const(s, s->irho, s->irho, s->nx, s->ny, s->nz, 2, 0); /* :813 */
    free_csr(s);
    int N = s->N;
    s->rp = (int*)calloc((size_t)N + 1, sizeof(int));
    s->ci = (int*)calloc((size_t)N * 7, sizeof(int));
    s->val = (double*)calloc((size_t)N * 7, sizeof(double));
    s->b = (double*)calloc((size_t)N, sizeof(double));
    s->x = (double*)calloc((size_t)N, sizeof(double));
    s->r = (double*)calloc((size_t)N, sizeof(double));
    s->zv = (double*)calloc((size_t)N, sizeof(double));
    s->pp = (double*)calloc((size_t)N, sizeof(double));
    s->Ap = (double*)calloc((size_t)N, sizeof(double));
    const double inv_dxdx = (double)1 / (s->cell.x.d * s->cell.x.d);
    const double inv_dydy = (double)1 / (s->cell.y.d * s->cell.y.d);
    const double inv_dzdz = (double)1 / (s->cell.z.d * s->cell.z.d);
    const double* vv = s->irho;
    const double* f = s->phi;
    const double* jc = jump_field(s);
    /* the fixed schedule on unequal spacings corrects the diagonal */
    const int axis = is_fixed_schedule(s) && !(inv_dxdx == inv_dydy && inv_dxdx == inv_dzdz);
    int nz_ix = 0;
    int i0 = 0, j0 = 0, k0 = 0;
    memset(s->fbx, 0, sizeof(double) * (size_t)s->ntot);
    memset(s->fby, 0, sizeof(double) * (size_t)s->ntot);
    memset(s->fbz, 0, sizeof(double) * (size_t)s->ntot);
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                int row = s->bc[IX(s, i, j, k)];
                if (row < 0) continue;
                double vc = A3(vv, i, j, k), fc = A3(f, i, j, k);
                double bl = GFM_BETA(vc, A3(vv, i - 1, j, k), A3(f, i - 1, j, k), fc);
                double br = GFM_BETA_R(vc, A3(vv, i + 1, j, k), A3(f, i + 1, j, k), fc);
                double bb = GFM_BETA(vc, A3(vv, i, j - 1, k), A3(f, i, j - 1, k), fc);
                double bt = GFM_BETA_R(vc, A3(vv, i, j + 1, k), A3(f, i, j + 1, k), fc);
                double bd = GFM_BETA(vc, A3(vv, i, j, k - 1), A3(f, i, j, k - 1), fc);
                double bu = GFM_BETA_R(vc, A3(vv, i, j, k + 1), A3(f, i, j, k + 1), fc);
                int cl = s->bc[IX(s, i - 1, j, k)], cr = s->bc[IX(s, i + 1, j, k)];
                int cb = s->bc[IX(s, i, j - 1, k)], ct = s->bc[IX(s, i, j + 1, k)];
                int cd = s->bc[IX(s, i, j, k - 1)], cu = s->bc[IX(s, i, j, k + 1)];
                double coef = 0;
                s->rp[row] = nz_ix;
                if (cl > -1) A3(s->fbx, i, j, k) = bl;
                if (cb > -1) A3(s->fby, i, j, k) = bb;
                if (cd > -1) A3(s->fbz, i, j, k) = bd;
                if (cl > -1) { coef += bl; s->ci[nz_ix] = cl; s->val[nz_ix++] = -bl * inv_dxdx; }
                if (cr > -1) { coef += br; s->ci[nz_ix] = cr; s->val[nz_ix++] = -br * inv_dxdx; }
                if (cb > -1) { coef += bb; s->ci[nz_ix] = cb; s->val[nz_ix++] = -bb * inv_dydy; }
                if (ct > -1) { coef += bt; s->ci[nz_ix] = ct; s->val[nz_ix++] = -bt * inv_dydy; }
                if (cd > -1) { coef += bd; s->ci[nz_ix] = cd; s->val[nz_ix++] = -bd * inv_dzdz; }
                if (cu > -1) { coef += bu; s->ci[nz_ix] = cu; s->val[nz_ix++] = -bu * inv_dzdz; }
                if (row == 0) { i0 = i; j0 = j; k0 = k; }
                if (i == i0 && j == j0 && k == k0) { /* Neumann pin, :1145-1179 */
                    if (cb == FS_BC_NEUM) coef += bb;
                    if (cd == FS_BC_NEUM) coef += bd;
                    if (cu == FS_BC_NEUM) coef += bu;
                }
                double dval;
                if (axis) {
                    /* fixed schedule, unequal spacings: each face's beta over
                     * its own d^2, pin extras on their axes (y for beta_b,
                     * z for beta_d and beta_u) */
                    double cx = 0, cy = 0, cz = 0;
                    if (cl > -1) cx += bl;
                    if (cr > -1) cx += br;
                    if (cb > -1) cy += bb;
                    if (ct > -1) cy += bt;
                    if (cd > -1) cz += bd;
                    if (cu > -1) cz += bu;
                    if (i == i0 && j == j0 && k == k0) {
                        if (cb == FS_BC_NEUM) cy += bb;
                        if (cd == FS_BC_NEUM) cz += bd;
                        if (cu == FS_BC_NEUM) cz += bu;
                    }
                    dval = cx + cy + cz == 0 ? s->cell.x.inv_d2
                                             : s->cell.x.inv_d2 * cx + s->cell.y.inv_d2 * cy + s->cell.z.inv_d2 * cz;
                } else {
                    if (coef == 0) coef = 1;
                    dval = inv_dxdx * coef; /* :1213, 1/dx^2 for every face */
                }
                s->ci[nz_ix] = row;
                s->val[nz_ix++] = dval;
                if (jc) { /* :1215-1328 */
                    double jcc = A3(jc, i, j, k);
                    double F_L = jump_arm(fc, A3(f, i - 1, j, k), jcc, A3(jc, i - 1, j, k), s->cell.x.inv_d2, bl);
                    double F_R = jump_arm(fc, A3(f, i + 1, j, k), jcc, A3(jc, i + 1, j, k), s->cell.x.inv_d2, br);
                    double F_B = jump_arm(fc, A3(f, i, j - 1, k), jcc, A3(jc, i, j - 1, k), s->cell.y.inv_d2, bb);
                    double F_T = jump_arm(fc, A3(f, i, j + 1, k), jcc, A3(jc, i, j + 1, k), s->cell.y.inv_d2, bt);
                    double F_D = jump_arm(fc, A3(f, i, j, k - 1), jcc, A3(jc, i, j, k - 1), s->cell.z.inv_d2, bd);
                    double F_U = jump_arm(fc, A3(f, i, j, k + 1), jcc, A3(jc, i, j, k + 1), s->cell.z.inv_d2, bu);
                    double F_X = F_L + F_R, F_Y = F_B + F_T, F_Z = F_D + F_U;
                    s->b[row] = -A3(s->div, i, j, k) - F_X - F_Y - F_Z;
                } else {
                    s->b[row] = -A3(s->div, i, j, k);
                }
                s->x[row] = A3(s->p, i, j, k);
            }
    s->rp[N] = nz_ix;
    s->nnz = nz_ix;
    /* Neumann compatibility: b -= mean(b), POISSON_SOLVER_3D.h:225-246.  The
     * reference adds per-thread slabs (its order depends on the thread
     * count); here each row in x order, the rows of a plane in order, the
     * planes in order. */
    double sum = 0;
    for (int k = 0; k < s->nz; k++) {
        double part = 0;
        for (int j = 0; j < s->ny; j++) {
            double rowsum = 0;
            for (int r = (k * s->ny + j) * s->nx; r < (k * s->ny + j + 1) * s->nx; r++) rowsum += s->b[r];
            part += rowsum;
        }
        sum += part;
    }
    double one_over_b_dimension = 1 / (double)N;
    for (int r = 0; r < N; r++) s->b[r] -= one_over_b_dimension * sum;
}

static void get(fso_state* s, const double* src, double* dst, int ni, int nj, int nk);

static double csr_diag(const fso_state* s, int row) {
    for (int e = s->rp[row]; e < s->rp[row + 1]; e++)
        if (s->ci[e] == row) return s->val[e];
    return 0;
}

/* CSR_MATRIX::Multiply, CSR_MATRIX.h:236-260 */
static void sys_mul(const sys_t* S, const double* x, double* y) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int row = 0; row < S->N; row++) {
        double v = 0;
        for (int e = S->rp[row]; e < S->rp[row + 1]; e++) v += S->val[e] * x[S->ci[e]];
        y[row] = v;
    }
}
static double sys_diag(const sys_t* S, int row) {
    for (int e = S->rp[row]; e < S->rp[row + 1]; e++)
        if (S->ci[e] == row) return S->val[e];
    return 0;
}
static sys_t pressure_sys(fso_state* s) {
    sys_t S = {s->N, s->nx * s->ny, s->nz, s->rp, s->ci, s->val, s->b, s->x, s->r, s->zv, s->pp, s->Ap};
    return S;
}
static void csr_mul(fso_state* s, const double* x, double* y) {
    sys_t S = pressure_sys(s);
    sys_mul(&S, x, y);
}

/* PCG_METHOD::PCGMethod, PCG_METHOD.cpp:4-126, with the Jacobi
 * preconditioner of MultiplicationByMinverseAsDiagonal (:128-160), on any
 * system; the stop rule, tolerances and max_iteration of PROJECTION_WATER. */
static void pcg_sys(fso_state* s, const sys_t* S) {
    const fs_params* P = &s->P;
    int N = S->N;
    double *x = S->x, *r = S->r, *sv = S->zv, *p = S->pp, *Ap = S->Ap;
    const double* b = S->b;
    double* dg = (double*)malloc(sizeof(double) * (size_t)N);
    for (int i = 0; i < N; i++) dg[i] = sys_diag(S, i);
    sys_mul(S, x, Ap);
    for (int i = 0; i < N; i++) r[i] = b[i] - Ap[i];
    for (int i = 0; i < N; i++) p[i] = r[i] / dg[i];
    double res_new = planes_dot(s, S->per, S->nplanes, r, p);
    double bnorm = sqrt(planes_dot(s, S->per, S->nplanes, b, b));
    double sqr_tol = P->tolerance * P->tolerance;
    int it = 0;
    double rr = planes_dot(s, S->per, S->nplanes, r, r);
    int rel = P->stop_rule == FS_STOP_REL_RESIDUAL;
    if (rel && (bnorm == 0 || sqrt(rr) <= P->rel_tolerance * bnorm)) goto done;
    while (it < P->max_iteration) {
        sys_mul(S, p, Ap);
        double dot_result = planes_dot(s, S->per, S->nplanes, p, Ap);
        double alpha = res_new / dot_result;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int i = 0; i < N; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
            sv[i] = r[i] / dg[i];
        }
        double res_old = res_new;
        res_new = planes_dot(s, S->per, S->nplanes, r, sv);
        if (rel) {
            rr = planes_dot(s, S->per, S->nplanes, r, r);
            if (sqrt(rr) < P->rel_tolerance * bnorm) break;
        } else if (res_new < sqr_tol)
            break;
        double beta = res_new / res_old;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int i = 0; i < N; i++) {
            p[i] *= beta;
            p[i] += sv[i];
        }
        it++;
    }
done:
    s->last_iters = it;
    s->last_res_abs = sqrt(res_new);
    /* FP64 true residual for the report */
    sys_mul(S, x, Ap);
    double tr = 0;
    for (int i = 0; i < N; i++) {
        double d = b[i] - Ap[i];
        tr += d * d;
    }
    s->last_res_rel = bnorm > 0 ? sqrt(tr) / bnorm : 0;
    free(dg);
}
static void pcg(fso_state* s) {
    sys_t S = pressure_sys(s);
    pcg_sys(s, &S);
}

/* PROJECTION_3D::UpdateVelocityByPressureGradientVariableDensity,
 * PROJECTION_3D.h:1099-1449 (air/water).  comp 1,2,3 = x,y,z faces. */
static double gradient_update(fso_state* s, int comp, double dt, int apply) {
    const fs_params* P = &s->P;
    int ni, nj, nk;
    extents(s, comp, &ni, &nj, &nk);
    int di = comp == 1, dj = comp == 2, dk = comp == 3;
    double* vel = comp == 1 ? s->u : comp == 2 ? s->v : s->w;
    double od = comp == 1 ? s->cell.x.inv : comp == 2 ? s->cell.y.inv : s->cell.z.inv;
    double dgrid = comp == 1 ? s->cell.x.d : comp == 2 ? s->cell.y.d : s->cell.z.d;
    double eps = (double)1.5 * dgrid; /* :1126, :1137, :1148 */
    const double* jc = jump_field(s);
    double vmax = 0;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                int bcc = s->bc[IX(s, i, j, k)];
                if (bcc < 0) continue;
                double lface = (double)0.5 * (A3(s->pg, i, j, k) + A3(s->pg, i - di, j - dj, k - dk));
                double density_half = 0;
                /* PROJECTION_3D::DetermineDensityField, :1474-1519 */
                double H = heaviside_proj(lface, eps);
                density_half = P->air_bubble_rising ? P->air_density + (P->water_density - P->air_density) * H
                                                    : P->water_density + (P->air_density - P->water_density) * H;
                double density_ijk = A3(s->rho, i, j, k);
                double one_over_density_ijk = (double)1 / density_ijk;
                double lc = A3(s->phi, i, j, k), ln = A3(s->phi, i - di, j - dj, k - dk);
                /* :1178-1181; jc_on_solution's ghosts stay 0 */
                double jc_c = jc ? A3(jc, i, j, k) : (double)0, jc_n = jc ? A3(jc, i - di, j - dj, k - dk) : (double)0;
                double jump_gamma = (jc_c * fabs(ln) + jc_n * fabs(lc)) / (fabs(ln) + fabs(lc));
                double one_over_density_half = 1 / density_half;
                double bcd = (double)bcc, bcn = (double)s->bc[IX(s, i - di, j - dj, k - dk)];
                double* vr = &A3(vel, i, j, k);
                double pc = A3(s->p, i, j, k), pn = A3(s->p, i - di, j - dj, k - dk);
                if (apply && bcd * bcn >= 0) {
                    if (lc * ln > 0) {
                        *vr -= dt * (pc - pn) * one_over_density_ijk * od;
                    } else {
                        if (lc <= 0) *vr -= dt * (pc - pn + jump_gamma) * one_over_density_half * od;
                        else if (lc > 0) *vr -= dt * (pc - pn - jump_gamma) * one_over_density_half * od;
                    }
                }
                double a = fabs(*vr);
                vmax = vmax > a ? vmax : a;
            }
    return vmax;
}

/* Fixed schedule: every face corrected with the pressure matrix's own face
 * coefficient, u -= dt beta (p_c - p_n + s_c a) / dx (beta is 0 on wall
 * faces), where a face between the phases carries the right-hand side's jump
 * a = (jc_c |phi_n| + jc_n |phi_c|) / (|phi_c| + |phi_n|) with s_c = +1 when
 * phi_c <= 0, else -1: the corrected velocity is then discretely
 * divergence-free with the jump included. */
static double gradient_update_beta(fso_state* s, int comp, double dt) {
    int ni, nj, nk;
    extents(s, comp, &ni, &nj, &nk);
    int di = comp == 1, dj = comp == 2, dk = comp == 3;
    double* vel = comp == 1 ? s->u : comp == 2 ? s->v : s->w;
    const double* beta = comp == 1 ? s->fbx : comp == 2 ? s->fby : s->fbz;
    const double* jc = jump_field(s);
    double od = comp == 1 ? s->cell.x.inv : comp == 2 ? s->cell.y.inv : s->cell.z.inv;
    double vmax = 0;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                double* vr = &A3(vel, i, j, k);
                double bf = A3(beta, i, j, k);
                if (bf != 0) {
                    double dp = A3(s->p, i, j, k) - A3(s->p, i - di, j - dj, k - dk);
                    if (jc) {
                        double fc = A3(s->phi, i, j, k), fn = A3(s->phi, i - di, j - dj, k - dk);
                        if (!((fc <= 0 && fn <= 0) || (fc > 0 && fn > 0))) {
                            double a = (A3(jc, i, j, k) * fabs(fn) + A3(jc, i - di, j - dj, k - dk) * fabs(fc)) /
                                       (fabs(fc) + fabs(fn));
                            dp = fc <= 0 ? dp + a : dp - a;
                        }
                    }
                    *vr -= dt * bf * dp * od;
                }
                double a = fabs(*vr);
                vmax = vmax > a ? vmax : a;
            }
    return vmax;
}

/* PROJECTION_3D::DetermineJumpConditionField, :746-792: ComputeCurvatures,
 * then jc_on_solution = surface_tension * curvature when
 * use_delta_function_formulation (its ghosts stay 0). */
static void determine_jump_condition(fso_state* s) {
    fill_linear(s, s->phi, s->phi, s->nx, s->ny, s->nz, 2, 0, 1);
    if (s->P.surface_tension != 0.0) curvature(s, s->P.use_delta_function_formulation && !s->csf_on ? s->jc : NULL);
}

static void projection_solve(fso_state* s, double dt) {
    projection_density(s);
    projection_divergence(s, dt);
    determine_jump_condition(s);
    assemble(s);
    pcg(s);
    /* POISSON_SOLVER_3D::VectorToGrid, :4060-4078 */
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) A3(s->p, i, j, k) = s->x[s->bc[IX(s, i, j, k)]];
    /* UpdateVelocity, :836-841; the projection's own ghost copy of phi is
     * refreshed by linear extrapolation before each component (:1115) */
    fill_linear(s, s->pg, s->phi, s->nx, s->ny, s->nz, 2, 1, 0);
    /* air/water corrects z-faces only: the x and y updates sit inside the
     * pipe-mode branches (:1188-1232, :1297-1341) while the max |u| is still
     * taken over every visited face */
    if (is_fixed(s)) {
        for (int c = 1; c <= 3; c++) s->max_vel[c - 1] = gradient_update_beta(s, c, dt);
        return;
    }
    s->max_vel[0] = gradient_update(s, 1, dt, 0);
    s->max_vel[1] = gradient_update(s, 2, dt, 0);
    s->max_vel[2] = gradient_update(s, 3, dt, 1);
}

int fso_projection(fso_state* s, double dt, fs_step_stats* st) {
    if (!(dt > 0)) return FS_ERR_BAD_ARG;
    projection_solve(s, dt);
    if (s->P.surface_tension != 0.0 && s->curv_fail) return FS_ERR_NONFINITE;
    if (st) {
        st->pcg_iterations = s->last_iters;
        st->refinement_sweeps = 0;
        st->residual_precond = FSO_VALUE(s->last_res_abs);
        st->residual_rel = FSO_VALUE(s->last_res_rel);
        st->max_velocity[0] = FSO_VALUE(s->max_vel[0]);
        st->max_velocity[1] = FSO_VALUE(s->max_vel[1]);
        st->max_velocity[2] = FSO_VALUE(s->max_vel[2]);
    }
    return s->last_iters >= s->P.max_iteration ? FS_ERR_NOT_CONVERGED : FS_OK;
}

int fso_assemble_pressure_system(fso_state* s, double dt, double* diag, double* bx, double* by, double* bz, double* rhs) {
    if (!(dt > 0)) return FS_ERR_BAD_ARG;
    projection_density(s);
    projection_divergence(s, dt);
    determine_jump_condition(s);
    assemble(s);
    int nx = s->nx, ny = s->ny, nz = s->nz;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int row = s->bc[IX(s, i, j, k)];
                long c = i + (long)nx * (j + (long)ny * k);
                if (diag) diag[c] = csr_diag(s, row);
                if (rhs) rhs[c] = s->b[row];
            }
    /* face betas between two full cells, as each row computed them (beta_l,
     * beta_b, beta_d of the cell on the + side of the face) */
    if (bx) get(s, s->fbx, bx, nx + 1, ny, nz);
    if (by) get(s, s->fby, by, nx, ny + 1, nz);
    if (bz) get(s, s->fbz, bz, nx, ny, nz + 1);
    return FS_OK;
}

int fso_apply_pressure_operator(fso_state* s, const double* x, double* q) {
    if (!s->rp) return FS_ERR_BAD_ARG;
    csr_mul(s, x, q);
    return FS_OK;
}

int fso_csr_info(fso_state* s, int* n, int* nnz) {
    if (!s->rp) return FS_ERR_BAD_ARG;
    *n = s->N;
    *nnz = s->nnz;
    return FS_OK;
}

int fso_csr_copy(fso_state* s, int* row_ptr, int* col, double* val, double* b) {
    if (!s->rp) return FS_ERR_BAD_ARG;
    if (row_ptr) memcpy(row_ptr, s->rp, sizeof(int) * (size_t)(s->N + 1));
    if (col) memcpy(col, s->ci, sizeof(int) * (size_t)s->nnz);
    if (val) memcpy(val, s->val, sizeof(double) * (size_t)s->nnz);
    if (b) memcpy(b, s->b, sizeof(double) * (size_t)s->N);
    return FS_OK;
}

/* ------------------------------------------------------------------------- */
/* Reinitialization                                                           */
/* ------------------------------------------------------------------------- */

enum { AX_M = 1, AX_P = 2, AX_DROP = 3, AX_SM = 4, AX_SP = 5, AX_NONE = 0 };

/* One axis of the Godunov case tree, ADVECTION_METHOD_3D.h:908-1622. */
static int godunov_axis(double sg, double m, double p, double* val) {
    double sm = sg * m, sp = sg * p;
    if (sp >= 0 && sm >= 0) { *val = m; return AX_M; }
    if (sp <= 0 && sm <= 0) { *val = p; return AX_P; }
    if (sp > 0 && sm < 0) { *val = 0; return AX_DROP; }
    if (sp < 0 && sm > 0) {
        double ss = sg * (fabs(p) - fabs(m)) / (p - m);
        if (ss > 0) { *val = m; return AX_SM; }
        *val = p;
        return AX_SP;
    }
    return AX_NONE;
}

/* |grad phi| - 1 of the Godunov case tree (ADVECTION_METHOD_3D.h:908-1622)
 * from the one-sided derivatives of the three axes; 0 when a case is not
 * met (only NaN operands).  The compatible form keeps the two leaves whose
 * norm the reference writes as sqrt(px^2) + py^2 (:1463, :1605). */
static int godunov_norm(double sg, double mx, double px, double my, double py, double mz, double pz, int exact,
                        double* g) {
    double X, Y, Z;
    int cx = godunov_axis(sg, mx, px, &X);
    int cy = cx ? godunov_axis(sg, my, py, &Y) : AX_NONE;
    int cz = cy ? godunov_axis(sg, mz, pz, &Z) : AX_NONE;
    if (!(cx && cy && cz)) return 0;
    if (!exact && (cx == AX_SM || cx == AX_SP) && cy == AX_SP && cz == AX_DROP)
        *g = sqrt(X * X) + Y * Y - 1; /* :1463, :1605 */
    else
        *g = sqrt(X * X + Y * Y + Z * Z) - 1;
    return 1;
}

int fso_godunov_axis(double sg, double m, double p, double* val) { return godunov_axis(sg, m, p, val); }
int fso_godunov_norm(double sg, double mx, double px, double my, double py, double mz, double pz, int exact,
                     double* g) {
    return godunov_norm(sg, mx, px, my, py, mz, pz, exact, g);
}

/* ADVECTION_METHOD_3D::WENO5thReinitialization, ADVECTION_METHOD_3D.h:745-1637
 * (the fixed schedule takes the Euclidean norm in every Godunov case). */
static void reinit_sweep(fso_state* s, double dt) {
    const int exact = is_fixed(s);
    const double eps = s->P.levelset_epsilon;
    const gc_t* gc = &s->cell;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double f[7], mx, px, my, py, mz, pz;
                gather7(s, s->sghost, i, j, k, 0, f);
                fso_weno5_1d(f, gc->x.inv, eps, &mx, &px);
                gather7(s, s->sghost, i, j, k, 1, f);
                fso_weno5_1d(f, gc->y.inv, eps, &my, &py);
                gather7(s, s->sghost, i, j, k, 2, f);
                fso_weno5_1d(f, gc->z.inv, eps, &mz, &pz);
                double sg = A3(s->sgn, i, j, k);
                double* rho = &A3(s->phi, i, j, k);
                double g;
                if (godunov_norm(sg, mx, px, my, py, mz, pz, exact, &g)) *rho = *rho - dt * sg * g;
                if (sg * *rho < (double)0) { /* :1624-1634 */
                    if (sg < 0) *rho = -eps;
                    if (sg > 0) *rho = eps;
                }
            }
}

/* EULERIAN_FLUID_SOLVER_3D::Reinitialization, :2198-2225, Sussman branch;
 * each sweep: ADVECTION_3D::ReinitializationBySussman, ADVECTION_3D.h:360-378. */
int fso_reinitialization(fso_state* s, double dt) {
    (void)dt;
    if (!s->P.sussmanstyle_reinitialization) return FS_OK;
    double fictious_dt = s->cell.x.d / s->P.scaling_number_for_reinitialization;
    double dx2 = s->cell.x.d * s->cell.x.d;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double f = A3(s->phi, i, j, k);
                A3(s->sgn, i, j, k) = f / sqrt(f * f + dx2);
            }
    for (int it = 0; it < s->P.iteration_number_for_reinitialization; it++) {
        if (is_fixed(s))
            refresh_levelset_ghost(s);
        else
            fill_linear(s, s->sghost, s->phi, s->nx, s->ny, s->nz, 3, 1, 0);
        reinit_sweep(s, fictious_dt);
    }
    return FS_OK;
}

/* ------------------------------------------------------------------------- */
/* Step                                                                       */
/* ------------------------------------------------------------------------- */

static void save_phi(fso_state* s) {
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) A3(s->tmp, i, j, k) = A3(s->phi, i, j, k);
}
static void average_phi(fso_state* s) {
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++)
                A3(s->phi, i, j, k) = (double)0.5 * (A3(s->tmp, i, j, k) + A3(s->phi, i, j, k));
}

/* dst = ca*a + cb*dst over an ni x nj x nk interior, each product rounded
 * before the sum: the stage blends of EULERIAN_FLUID_SOLVER_3D.h:939-951,
 * :1041-1053, :1061-1073, :1099, :1107. */
static void blend(fso_state* s, double* dst, const double* a, double ca, double cb, int ni, int nj, int nk) {
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                double x = ca * A3(a, i, j, k);
                double y = cb * A3(dst, i, j, k);
                A3(dst, i, j, k) = x + y;
            }
}
static void blend_velocity(fso_state* s, double ca, double cb) {
    blend(s, s->u, s->urk, ca, cb, s->nx + 1, s->ny, s->nz);
    blend(s, s->v, s->vrk, ca, cb, s->nx, s->ny + 1, s->nz);
    blend(s, s->w, s->wrk, ca, cb, s->nx, s->ny, s->nz + 1);
}
/* the water_velocity_field_*_rk copies, :913-932 / :1015-1034 */
static void save_velocity(fso_state* s) {
    copy_interior(s, s->urk, s->u, s->nx + 1, s->ny, s->nz);
    copy_interior(s, s->vrk, s->v, s->nx, s->ny + 1, s->nz);
    copy_interior(s, s->wrk, s->w, s->nx, s->ny, s->nz + 1);
}

/* EULERIAN_FLUID_SOLVER_3D::SolveForRK, :1453-1476.  Statistics: PCG
 * iterations add up over the step's projections, the rest is the last. */
static void solve_for_rk(fso_state* s, double dt, fs_step_stats* st, int* rc) {
    const fs_params* P = &s->P;
    if (P->is_velocity_advection_active) fso_velocity_advection(s, dt);
    if (P->is_external_force_active) fso_apply_gravity(s, dt);
    if (P->is_sourcing_active) {
        int r = fso_sourcing(s, dt);
        if (r != FS_OK) *rc = r;
    }
    if (P->is_projection_active) {
        int its = st ? st->pcg_iterations : 0;
        int r = fso_projection(s, dt, st);
        if (st) st->pcg_iterations += its;
        if (r != FS_OK) *rc = r;
    }
}

/* EULERIAN_FLUID_SOLVER_3D::AdvanceOneTimeStepThread, :803-1111: order 1
 * (:806-858), order 2 (:886-980), order 3 (:1008-1109). */
int fso_step(fso_state* s, double dt, fs_step_stats* st) {
    if (!(dt > 0)) return FS_ERR_BAD_ARG;
    const fs_params* P = &s->P;
    const int order = P->order_for_time_advancing;
    int rc = FS_OK;
    if (st) memset(st, 0, sizeof(*st));
    s->visc_iters = 0;
    if (P->is_levelset_advection_active) {
        if (order == 3) {
            fso_levelset_advection(s, dt);
        } else {
            save_phi(s);
            fso_levelset_advection(s, dt);
            fso_levelset_advection(s, dt);
            average_phi(s);
        }
    }
    if (order == 1) {
        solve_for_rk(s, dt, st, &rc);
    } else {
        save_velocity(s);
        solve_for_rk(s, dt, st, &rc);
        solve_for_rk(s, dt, st, &rc);
        if (order == 2) {
            blend_velocity(s, (double)0.5, (double)0.5);
        } else {
            blend_velocity(s, (double)0.75, (double)0.25);
            solve_for_rk(s, dt, st, &rc);
            blend_velocity(s, (double)1 / (double)3, (double)2 / (double)3);
        }
    }
    save_phi(s);
    fso_reinitialization(s, dt);
    fso_reinitialization(s, dt);
    if (order == 3) {
        blend(s, s->phi, s->tmp, (double)0.75, (double)0.25, s->nx, s->ny, s->nz);
        fso_reinitialization(s, dt);
        blend(s, s->phi, s->tmp, (double)1 / (double)3, (double)2 / (double)3, s->nx, s->ny, s->nz);
    } else {
        average_phi(s);
    }
    if (st) st->viscosity_iterations = s->visc_iters;
    return rc;
}

/* SIMULATION_WORLD::AdvanceOneFrame, SIMULATION_WORLD.h:320-330. */
int fso_advance_frame(fso_state* s, double dt, fs_step_stats* st) {
    fso_compute_curvatures(s);
    return fso_step(s, dt, st);
}

/* EULERIAN_FLUID_SOLVER_3D::CFLOneTimeStep, :2673-2785 (air/water). */
int fso_cfl_dt(fso_state* s, double cfl, int gravity_norm, double* dt) {
    const fs_params* P = &s->P;
    double um = 0, vm = 0, wm = 0;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i <= s->nx; i++)
                if (um <= fabs(A3(s->u, i, j, k))) um = fabs(A3(s->u, i, j, k));
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j <= s->ny; j++)
            for (int i = 0; i < s->nx; i++)
                if (vm <= fabs(A3(s->v, i, j, k))) vm = fabs(A3(s->v, i, j, k));
    for (int k = 0; k <= s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++)
                if (wm <= fabs(A3(s->w, i, j, k))) wm = fabs(A3(s->w, i, j, k));
    /* the reference's max skips NaN (a comparison with NaN is false); the
     * build reports non-finite velocities instead */
    for (long n = 0; n < s->ntot; n++)
        if (!isfinite(s->u[n]) || !isfinite(s->v[n]) || !isfinite(s->w[n])) return FS_ERR_NONFINITE;
    double c_f = um * s->mx.x.inv + vm * s->my.y.inv + wm * s->mz.z.inv;
    double g_f = 0;
    if (P->is_external_force_active) {
        double gmag = gravity_norm ? sqrt(P->gravity[0] * P->gravity[0] + P->gravity[1] * P->gravity[1] +
                                          P->gravity[2] * P->gravity[2])
                                   : fabs(P->gravity[1]);
        g_f = sqrt(gmag * s->my.y.inv);
    }
    /* :2740-2744 (air/water), abs_max_curvature of the level set as
     * SIMULATION_WORLD::AdvanceOneFrame leaves it right before the rule
     * (SIMULATION_WORLD.h:324, :289) */
    double s_f = 0;
    if (P->surface_tension != 0.0) {
        fso_compute_curvatures(s);
        if (s->curv_fail) return FS_ERR_NONFINITE;
        double dmin = fmin(fmin(s->mx.x.d, s->my.y.d), s->mz.z.d);
        double rmin = P->water_density > P->air_density ? P->air_density : P->water_density;
        s_f = sqrt((P->surface_tension * s->kmax) / (rmin * (dmin * dmin)));
    }
    double a = P->water_viscosity / P->water_density, b = P->air_viscosity / P->air_density;
    double v_f = (a > b ? a : b) * ((double)2 * s->cell.x.inv_d2 + (double)2 * s->cell.y.inv_d2 + (double)2 * s->cell.z.inv_d2);
    if (P->semi_implicit_viscosity) v_f = 0; /* :2770-2773 */
    double c = (double)0.5 * (c_f + v_f + sqrt((c_f + v_f) * (c_f + v_f) + (double)4 * (g_f * g_f) + (double)4 * (s_f * s_f)));
    *dt = cfl * ((double)1 / c);
    return FS_OK;
}

/* ------------------------------------------------------------------------- */
/* Test hooks                                                                 */
/* ------------------------------------------------------------------------- */

static void put(fso_state* s, double* dst, const double* src, int ni, int nj, int nk);

/* The projection's linear system for a given source and jump field at the
 * current level set (DeterminePressure, PROJECTION_3D.h:794-834, with the
 * divergence field replaced by `source` -- div u / dt, so that the solve is
 * div(beta grad p) = source -- and jc_on_solution by `jump`, the jump
 * p(phi > 0 side) - p(phi <= 0 side); NULL = no jump), solved from p = 0;
 * p receives the solution (interior, x fastest).  This is the reference's
 * own Poisson test hook (PROJECTION_3D.h:564-593, 727-740, 771-774) with the
 * flux jump [beta dp/dn] held at zero, as this solver's assembly has it. */
int fso_solve_poisson(fso_state* s, const double* source, const double* jump, double* p, fs_step_stats* st) {
    if (!source || !p) return FS_ERR_BAD_ARG;
    double* jz = NULL;
    if (jump) {
        jz = zalloc(s->ntot); /* zero ghosts, as jc_on_solution's */
        put(s, jz, jump, s->nx, s->ny, s->nz);
    }
    projection_density(s);
    for (long n = 0; n < s->ntot; n++) s->div[n] = 0;
    put(s, s->div, source, s->nx, s->ny, s->nz);
    for (long n = 0; n < s->ntot; n++) s->p[n] = 0;
    s->jc_override = jz;
    assemble(s);
    pcg(s);
    s->jc_override = NULL;
    free(jz);
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) A3(s->p, i, j, k) = s->x[s->bc[IX(s, i, j, k)]];
    get(s, s->p, p, s->nx, s->ny, s->nz);
    if (st) {
        memset(st, 0, sizeof(*st));
        st->pcg_iterations = s->last_iters;
        st->residual_precond = FSO_VALUE(s->last_res_abs);
        st->residual_rel = FSO_VALUE(s->last_res_rel);
    }
    return s->last_iters >= s->P.max_iteration ? FS_ERR_NOT_CONVERGED : FS_OK;
}

/* The persistent ghost copies, each over its box grown by the ghost width
 * (x fastest): the shared scalar copy (sghost, (nx+6)(ny+6)(nz+6)) and the
 * MAC copies gu, gv, gw ((nx+7)(ny+6)(nz+6), ...).  NULL skips one. */
static void get_ghosted(fso_state* s, const double* src, double* dst, int ni, int nj, int nk) {
    const int gi = ni + 2 * G, gj = nj + 2 * G, gk = nk + 2 * G;
    for (int k = 0; k < gk; k++)
        for (int j = 0; j < gj; j++)
            for (int i = 0; i < gi; i++) dst[i + (long)gi * (j + (long)gj * k)] = A3(src, i - G, j - G, k - G);
}
int fso_get_ghost_copies(fso_state* s, double* sghost, double* gu, double* gv, double* gw) {
    if (sghost) get_ghosted(s, s->sghost, sghost, s->nx, s->ny, s->nz);
    if (gu) get_ghosted(s, s->gu, gu, s->nx + 1, s->ny, s->nz);
    if (gv) get_ghosted(s, s->gv, gv, s->nx, s->ny + 1, s->nz);
    if (gw) get_ghosted(s, s->gw, gw, s->nx, s->ny, s->nz + 1);
    return FS_OK;
}

/* ------------------------------------------------------------------------- */
/* Field exchange                                                             */
/* ------------------------------------------------------------------------- */

static void put(fso_state* s, double* dst, const double* src, int ni, int nj, int nk) {
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) A3(dst, i, j, k) = src[i + (long)ni * (j + (long)nj * k)];
}
static void get(fso_state* s, const double* src, double* dst, int ni, int nj, int nk) {
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) dst[i + (long)ni * (j + (long)nj * k)] = A3(src, i, j, k);
}

int fso_set_fields(fso_state* s, const double* phi, const double* u, const double* v, const double* w, const double* p) {
    if (phi) {
        put(s, s->phi, phi, s->nx, s->ny, s->nz);
        fill_const(s, s->phi, s->phi, s->nx, s->ny, s->nz, 2, 0); /* EULERIAN_FLUID_SOLVER_3D.h:757 */
    }
    if (u) put(s, s->u, u, s->nx + 1, s->ny, s->nz);
    if (v) put(s, s->v, v, s->nx, s->ny + 1, s->nz);
    if (w) put(s, s->w, w, s->nx, s->ny, s->nz + 1);
    if (p) put(s, s->p, p, s->nx, s->ny, s->nz);
    return FS_OK;
}

int fso_get_fields(fso_state* s, double* phi, double* u, double* v, double* w, double* p) {
    if (phi) get(s, s->phi, phi, s->nx, s->ny, s->nz);
    if (u) get(s, s->u, u, s->nx + 1, s->ny, s->nz);
    if (v) get(s, s->v, v, s->nx, s->ny + 1, s->nz);
    if (w) get(s, s->w, w, s->nx, s->ny, s->nz + 1);
    if (p) get(s, s->p, p, s->nx, s->ny, s->nz);
    return FS_OK;
}

/* ------------------------------------------------------------------------- */
/* On-disk records: LEVELSET_3D::Write / Read, LEVELSET_3D.h:1132-1200        */
/* ------------------------------------------------------------------------- */

/* grid of component comp (0 cells, 1..3 MAC faces, EULERIAN_FLUID_SOLVER_3D.h
 * :405-431) as the reference records it */
static void rec_grid(const fso_state* s, int comp, int res[3], double mn[3], double mx[3]) {
    const int n[3] = {s->nx, s->ny, s->nz};
    const double lo[3] = {s->gr.x_min, s->gr.y_min, s->gr.z_min}, hi[3] = {s->gr.x_max, s->gr.y_max, s->gr.z_max};
    const double d[3] = {s->cell.x.d, s->cell.y.d, s->cell.z.d};
    for (int a = 0; a < 3; a++) {
        int face = comp == a + 1;
        res[a] = n[a] + face;
        mn[a] = face ? lo[a] - (double)0.5 * d[a] : lo[a];
        mx[a] = face ? hi[a] + (double)0.5 * d[a] : hi[a];
    }
}

static int write_rec(FILE* f, const fso_state* s, int comp, int ghost, const double* v) {
    int res[3], start = 0, ok = 1;
    double mn[3], mx[3];
    rec_grid(s, comp, res, mn, mx);
    for (int a = 0; a < 3; a++) ok &= fwrite(&start, sizeof(int), 1, f) == 1;
    for (int a = 0; a < 3; a++) ok &= fwrite(&res[a], sizeof(int), 1, f) == 1;
    for (int a = 0; a < 3; a++) ok &= fwrite(&mn[a], sizeof(double), 1, f) == 1;
    for (int a = 0; a < 3; a++) ok &= fwrite(&mx[a], sizeof(double), 1, f) == 1;
    ok &= fwrite(&ghost, sizeof(int), 1, f) == 1;
    size_t n = (size_t)res[0] * res[1] * res[2];
    ok &= fwrite(v, sizeof(double), n, f) == n;
    return ok ? FS_OK : FS_ERR_BAD_ARG;
}

static int read_rec(FILE* f, const fso_state* s, int comp, double* v) {
    int res[3], start[3], ghost;
    double mn[3], mx[3], wmn[3], wmx[3];
    int want[3];
    rec_grid(s, comp, want, wmn, wmx);
    if (fread(start, sizeof(int), 3, f) != 3 || fread(res, sizeof(int), 3, f) != 3 ||
        fread(mn, sizeof(double), 3, f) != 3 || fread(mx, sizeof(double), 3, f) != 3 ||
        fread(&ghost, sizeof(int), 1, f) != 1)
        return FS_ERR_BAD_ARG;
    for (int a = 0; a < 3; a++) {
        double tol = 1e-12 * (fabs(wmx[a] - wmn[a]) + 1.0);
        if (res[a] != want[a] || fabs(mn[a] - wmn[a]) > tol || fabs(mx[a] - wmx[a]) > tol) return FS_ERR_BAD_ARG;
    }
    size_t n = (size_t)res[0] * res[1] * res[2];
    return fread(v, sizeof(double), n, f) == n ? FS_OK : FS_ERR_BAD_ARG;
}

static size_t comp_size(const fso_state* s, int comp) {
    return (size_t)(s->nx + (comp == 1)) * (s->ny + (comp == 2)) * (s->nz + (comp == 3));
}

int fso_write_levelset(fso_state* s, const char* path) {
    FILE* f = path ? fopen(path, "wb") : NULL;
    if (!f) return FS_ERR_BAD_ARG;
    double* v = (double*)malloc(sizeof(double) * comp_size(s, 0));
    fso_get_fields(s, v, NULL, NULL, NULL, NULL);
    int rc = write_rec(f, s, 0, 2, v);
    free(v);
    fclose(f);
    return rc;
}

int fso_read_levelset(fso_state* s, const char* path) {
    FILE* f = path ? fopen(path, "rb") : NULL;
    if (!f) return FS_ERR_BAD_ARG;
    double* v = (double*)malloc(sizeof(double) * comp_size(s, 0));
    int rc = read_rec(f, s, 0, v);
    if (rc == FS_OK) rc = fso_set_fields(s, v, NULL, NULL, NULL, NULL);
    free(v);
    fclose(f);
    return rc;
}

/* phi, u, v, w, p: five records back to back (ghost widths 2, 2, 2, 2, 1) */
int fso_write_checkpoint(fso_state* s, const char* path) {
    FILE* f = path ? fopen(path, "wb") : NULL;
    if (!f) return FS_ERR_BAD_ARG;
    const int comp[5] = {0, 1, 2, 3, 0}, ghost[5] = {2, 2, 2, 2, 1};
    double* v[5];
    for (int q = 0; q < 5; q++) v[q] = (double*)malloc(sizeof(double) * comp_size(s, comp[q]));
    fso_get_fields(s, v[0], v[1], v[2], v[3], v[4]);
    int rc = FS_OK;
    for (int q = 0; q < 5 && rc == FS_OK; q++) rc = write_rec(f, s, comp[q], ghost[q], v[q]);
    for (int q = 0; q < 5; q++) free(v[q]);
    fclose(f);
    return rc;
}

int fso_read_checkpoint(fso_state* s, const char* path) {
    FILE* f = path ? fopen(path, "rb") : NULL;
    if (!f) return FS_ERR_BAD_ARG;
    const int comp[5] = {0, 1, 2, 3, 0};
    double* v[5];
    for (int q = 0; q < 5; q++) v[q] = (double*)malloc(sizeof(double) * comp_size(s, comp[q]));
    int rc = FS_OK;
    for (int q = 0; q < 5 && rc == FS_OK; q++) rc = read_rec(f, s, comp[q], v[q]);
    if (rc == FS_OK) rc = fso_set_fields(s, v[0], v[1], v[2], v[3], v[4]);
    for (int q = 0; q < 5; q++) free(v[q]);
    fclose(f);
    return rc;
}
