/*
 * fs_oracle_pipe.c -- TEST INFRASTRUCTURE ONLY: the CPU checker of the HIP
 * library's oil/water pipe mode (fs_create_pipe).
 *
 * A plain-C, FP64, single-thread restatement of
 * EULERIAN_FLUID_SOLVER_3D::AdvanceOneTimeStepThread (EULERIAN_FLUID_SOLVER_3D.h:
 * 803-1111) in oil_water_simulation mode with the pipe along x (is_parallel)
 * or y (is_vertical), cylindrical_boundary, Neumann pressure,
 * dimensionless_form, semi_implicit_approach and boundary_order_velocity 1 --
 * the configuration of the reference's shipped scripts (3D_script.sim:32-99,
 * 195-290 parallel; 3D_script_1/2/3.sim vertical).  Each function
 * cites the lines it restates; expressions keep the reference's operation
 * order (build with -ffp-contract=off).
 *
 * State model.  The reference's modules keep persistent ghosted arrays, some
 * shared (the solver's scalar_field_ghost and vector_field_mac_ghost_x/y/z are
 * written by ADVECTION_3D, VISCOSITY_3D and PROJECTION_3D alike), some private
 * (each module's velocity_field_mac_ghost_<c>_<axis>, PROJECTION_3D's own
 * scalar_field_ghost).  Their ghost cells and the cells outside the pipe are
 * refreshed only where the reference refreshes them, so later stencils read
 * what earlier phases left there; this file keeps every such array with the
 * same refresh points.  All arrays share one padded layout of ghost width 3,
 * zero-initialised as ARRAY_3D::Initialize(..., true) leaves them
 * (FIELD_STRUCTURE_3D.h:115-121).  The boundary_phi_<c> face arrays (ghost
 * width 2) are computed on their interior only, so their ghosts read 0; the
 * periodic fill's reads one layer beyond that width land on other ghost cells
 * of the same array (ARRAY_3D.h:90-99), which also hold 0 -- the zero ghosts
 * here give the same values.  Every module computes boundary_phi_<c> from the
 * same face coordinates, so one set serves all.
 *
 * Sequential semantics: loops run k, j, i (GRID_ITERATION_3D) on one thread.
 * The reference's in-place loops are restated in that order (what one thread
 * computes); the HIP kernels are built to give the same values.
 *
 * Parity unpinned here: the libm boundary (sin, log, sqrt, pow) is whatever
 * the host libm gives.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fs_oracle.h"

#define PG 3
#define BC_FULL 0    /* MACROS.h:6 */
#define BC_DIR (-1)  /* :7 */
#define BC_NEUM (-4) /* :10 */
#define BC_PER (-5)  /* :11 */
#define BC_OTHER (-6)
#define PI_ 3.14159265358979323846 /* MACROS.h:13 */

typedef struct {
    double d, d2, inv, inv2d, inv_d2;
} pax_t;
typedef struct {
    pax_t x, y, z;
} pgc_t;

/* one CSR system (rows in k, j, i order over the FULL cells) */
typedef struct {
    int N, nnz, cap;
    int *rp, *ci;
    double *val, *b, *x;
} psys_t;

struct fsop_state {
    fs_params P;
    fs_pipe_params Q;
    fs_grid gr;
    int nx, ny, nz;
    long sx, sy, ntot;
    pgc_t cell, mac[4]; /* mac[1..3]: the x-, y-, z-face grids */
    double x0f[4], y0f[4], z0f[4]; /* x_min, y_min, z_min of each component's grid */
    int vert;                      /* 1: the vertical pipe (along y), 0: the parallel pipe (along x) */
    /* derived constants, EULERIAN_FLUID_SOLVER_3D.h:496-520 */
    double R1, f, A, V0, J, Re, We, one_over_We, A0;
    /* state */
    double *phi, *bls, *u, *v, *w, *p;
    double *sfg, *gx, *gy, *gz;                    /* solver's shared ghost copies */
    double *ag[3][2];                              /* ADVECTION_3D: <c>_y, <c>_z */
    double *vg[3][2];                              /* VISCOSITY_3D: <c>_y, <c>_z */
    double *vgyx, *vgzx;                           /* VISCOSITY_3D: y_x, z_x (never written here) */
    double *bp[4];                                 /* boundary_phi_x/y/z (index 1..3) */
    double *psfg;                                  /* PROJECTION_3D::scalar_field_ghost */
    double *rho, *irho, *div, *jc, *kappa, *sgn, *tmp, *rk[4], *scr, *scr2;
    double *den_half, *lev, *visc[6], *coef[6], *expl;
    int *pbc, *vbc;
    double kmax;       /* LEVELSET_3D::abs_max_curvature */
    double max_vel[3];
    int last_iters, visc_iters;
    double last_res_abs, last_res_rel;
    psys_t sys;
};

#define PIX(s, i, j, k) ((((long)(k) + PG) * (s)->sy + ((long)(j) + PG)) * (s)->sx + ((long)(i) + PG))
#define P3(a, i, j, k) ((a)[PIX(s, i, j, k)])

static pax_t mkax(double mn, double mx, int res) {
    pax_t a;
    a.d = (mx - mn) / (res - 1);
    a.d2 = a.d * a.d;
    a.inv = (double)1 / a.d;
    a.inv2d = (double)0.5 / a.d;
    a.inv_d2 = (double)1 / a.d2;
    return a;
}
static double* pz(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

static void box(const fsop_state* s, int c, int* ni, int* nj, int* nk) {
    *ni = s->nx + (c == 1);
    *nj = s->ny + (c == 2);
    *nk = s->nz + (c == 3);
}

/* the centreline faces ADVECTION_3D::Solve_Velocity fixes on its first call
 * (ADVECTION_3D.h:248-307: along x for the parallel pipe, along y for the
 * vertical one); set from creation: every phase that reads them runs after
 * that call in a step */
static int fixed_face(const fsop_state* s, int c, int i, int j, int k) {
    int ie, je, ke;
    if (s->vert) {
        if (c == 1) {
            ie = s->nx, ke = s->nz - 1;
            return (i == ie / 2 || i == ie / 2 + 1) && k == ke / 2;
        }
        if (c == 2) {
            ie = s->nx - 1, ke = s->nz - 1;
            return i == ie / 2 && k == ke / 2;
        }
        if (c == 3) {
            ie = s->nx - 1, ke = s->nz;
            return i == ie / 2 && (k == ke / 2 || k == ke / 2 + 1);
        }
        return 0;
    }
    if (c == 1) {
        je = s->ny - 1, ke = s->nz - 1;
        return j == je / 2 && k == ke / 2;
    }
    if (c == 2) {
        je = s->ny, ke = s->nz - 1;
        return (j == je / 2 || j == je / 2 + 1) && k == ke / 2;
    }
    if (c == 3) {
        je = s->ny - 1, ke = s->nz;
        return j == je / 2 && (k == ke / 2 || k == ke / 2 + 1);
    }
    return 0;
}

static double* comp_field(fsop_state* s, int c) { return c == 1 ? s->u : c == 2 ? s->v : s->w; }
static double* ghost_comp(fsop_state* s, int c) { return c == 1 ? s->gx : c == 2 ? s->gy : s->gz; }

/* ---- x-periodic ghost fills, FIELD_STRUCTURE_3D.h:1088-1327 ---------------
 * comp 0 cell, 1..3 faces; gw the ghost width of the destination array.
 * Unmasked (:1088-1188): interior copy when `copy`, then the x ghosts (and
 * the first cell, or the two end faces of an x-face field) from the
 * node-periodic image.  Masked (:1190-1326, mask = boundary_phi of the
 * component): the interior copy only where the mask is <= 0, the ghost
 * region only where the mask at the written cell is < 1e-8. */
static void fill_periodic_x(fsop_state* s, double* dst, const double* src, int c, int gw, int copy, const double* mask) {
    int ni, nj, nk;
    box(s, c, &ni, &nj, &nk);
    const int ie = ni - 1;
    if (copy)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++)
                    if (!mask || P3(mask, i, j, k) <= 0) P3(dst, i, j, k) = P3(src, i, j, k);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++) {
            if (c == 1) {
                for (int i = 0; i >= -gw; i--)
                    if (!mask || P3(mask, i, j, k) < (double)1e-8) P3(dst, i, j, k) = P3(src, ie + (i - 1), j, k);
            } else {
                for (int i = 0; i >= -gw; i--)
                    if (!mask || P3(mask, i, j, k) < (double)1e-8) P3(dst, i, j, k) = P3(src, ie + i, j, k);
            }
        }
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++) {
            if (c == 1) {
                for (int i = ie; i <= ie + gw; i++)
                    if (!mask || P3(mask, i, j, k) < (double)1e-8) P3(dst, i, j, k) = P3(src, 0 + (i - (ie - 1)), j, k);
            } else {
                for (int i = ie + 1; i <= ie + gw; i++)
                    if (!mask || P3(mask, i, j, k) < (double)1e-8) P3(dst, i, j, k) = P3(src, 0 + (i - ie), j, k);
            }
        }
}

/* ---- y-periodic ghost fills, FIELD_STRUCTURE_3D.h:1328-1710 ---------------
 * Unmasked (:1328-1430): interior copy when `copy`, then the y ghosts from
 * the node-periodic image: a y-face field's two end faces and ghosts as the
 * x fill's x faces, other fields' ghosts j < 0 and j > je only (unlike the x
 * fill, the first row is not rewritten).  Masked (:1432-1710): the interior
 * copy where the mask is < 1e-8 and ZERO elsewhere, the ghost rows where
 * the mask at the written cell is < 1e-8, then four blocks of linear
 * extrapolation of the source beyond the wall along x and z, in the order
 * the reference's loop runs them on one thread: from each cell outside the
 * wall whose x neighbour below (block 2) or above (block 3) or z neighbour
 * below or above (the two blocks 4) is inside, every cell from it to the
 * box edge (ghosts included) takes 2 src(t -/+ 1) - src(t -/+ 2). */
static void fill_periodic_y(fsop_state* s, double* dst, const double* src, int c, int gw, int copy, const double* mask) {
    int ni, nj, nk;
    box(s, c, &ni, &nj, &nk);
    const int ie = ni - 1, je = nj - 1, ke = nk - 1;
    if (copy)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    if (!mask || P3(mask, i, j, k) < (double)1e-8)
                        P3(dst, i, j, k) = P3(src, i, j, k);
                    else
                        P3(dst, i, j, k) = 0;
                }
    for (int k = 0; k < nk; k++)
        for (int i = 0; i < ni; i++) {
            if (c == 2) {
                for (int j = 0; j >= -gw; j--)
                    if (!mask || P3(mask, i, j, k) < (double)1e-8) P3(dst, i, j, k) = P3(src, i, je + (j - (0 + 1)), k);
            } else {
                for (int j = 0 - 1; j >= -gw; j--)
                    if (!mask || P3(mask, i, j, k) < (double)1e-8) P3(dst, i, j, k) = P3(src, i, je + (j - 0), k);
            }
        }
    for (int k = 0; k < nk; k++)
        for (int i = 0; i < ni; i++) {
            if (c == 2) {
                for (int j = je; j <= je + gw; j++)
                    if (!mask || P3(mask, i, j, k) < (double)1e-8) P3(dst, i, j, k) = P3(src, i, 0 + (j - (je - 1)), k);
            } else {
                for (int j = je + 1; j <= je + gw; j++)
                    if (!mask || P3(mask, i, j, k) < (double)1e-8) P3(dst, i, j, k) = P3(src, i, 0 + (j - je), k);
            }
        }
    if (!mask) return;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                if (P3(mask, i - 1, j, k) < 0 && P3(mask, i, j, k) > 0)
                    for (int t = i; t <= ie + gw; t++)
                        P3(dst, t, j, k) = (double)2 * P3(src, t - 1, j, k) - P3(src, t - 2, j, k);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                if (P3(mask, i + 1, j, k) < 0 && P3(mask, i, j, k) > 0)
                    for (int t = i; t >= 0 - gw; t--)
                        P3(dst, t, j, k) = (double)2 * P3(src, t + 1, j, k) - P3(src, t + 2, j, k);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                if (P3(mask, i, j, k - 1) < 0 && P3(mask, i, j, k) > 0)
                    for (int t = k; t <= ke + gw; t++)
                        P3(dst, i, j, t) = (double)2 * P3(src, i, j, t - 1) - P3(src, i, j, t - 2);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                if (P3(mask, i, j, k + 1) < 0 && P3(mask, i, j, k) > 0)
                    for (int t = k; t >= 0 - gw; t--)
                        P3(dst, i, j, t) = (double)2 * P3(src, i, j, t + 1) - P3(src, i, j, t + 2);
}

/* the fill along the pipe's axis: x for the parallel pipe, y for the
 * vertical one (every periodic fill of the step has both branches) */
static void fill_periodic(fsop_state* s, double* dst, const double* src, int c, int gw, int copy, const double* mask) {
    if (s->vert)
        fill_periodic_y(s, dst, src, c, gw, copy, mask);
    else
        fill_periodic_x(s, dst, src, c, gw, copy, mask);
}

/* FillGhostCellsFrom (constant extrapolation, FIELD_STRUCTURE_3D.h:663-1086)
 * and FillGhostCellsContinuousDerivativesFrom (linear, :1776-2199) of a cell
 * field onto itself, depth d: faces, then (linear) the edges the reference
 * fills; the same loops as the air/water checker. */
static int clampi_(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static void fill_const_self(fsop_state* s, double* a, int d) {
    int ni = s->nx, nj = s->ny, nk = s->nz;
    for (int k = -d; k < nk + d; k++)
        for (int j = -d; j < nj + d; j++)
            for (int i = -d; i < ni + d; i++) {
                if (i >= 0 && i < ni && j >= 0 && j < nj && k >= 0 && k < nk) continue;
                P3(a, i, j, k) = P3(a, clampi_(i, 0, ni - 1), clampi_(j, 0, nj - 1), clampi_(k, 0, nk - 1));
            }
}
static void fill_linear_self(fsop_state* s, double* a, int d) {
    int ni = s->nx, nj = s->ny, nk = s->nz, i, j, k;
    for (k = 0; k < nk; k++)
        for (j = 0; j < nj; j++)
            for (i = -1; i >= -d; i--) P3(a, i, j, k) = (double)2 * P3(a, i + 1, j, k) - P3(a, i + 2, j, k);
    for (k = 0; k < nk; k++)
        for (j = 0; j < nj; j++)
            for (i = ni; i <= ni - 1 + d; i++) P3(a, i, j, k) = (double)2 * P3(a, i - 1, j, k) - P3(a, i - 2, j, k);
    for (k = 0; k < nk; k++)
        for (j = -1; j >= -d; j--)
            for (i = 0; i < ni; i++) P3(a, i, j, k) = (double)2 * P3(a, i, j + 1, k) - P3(a, i, j + 2, k);
    for (k = 0; k < nk; k++)
        for (j = nj; j <= nj - 1 + d; j++)
            for (i = 0; i < ni; i++) P3(a, i, j, k) = (double)2 * P3(a, i, j - 1, k) - P3(a, i, j - 2, k);
    for (k = -1; k >= -d; k--)
        for (j = 0; j < nj; j++)
            for (i = 0; i < ni; i++) P3(a, i, j, k) = (double)2 * P3(a, i, j, k + 1) - P3(a, i, j, k + 2);
    for (k = nk; k <= nk - 1 + d; k++)
        for (j = 0; j < nj; j++)
            for (i = 0; i < ni; i++) P3(a, i, j, k) = (double)2 * P3(a, i, j, k - 1) - P3(a, i, j, k - 2);
    for (k = 0; k < nk; k++)
        for (j = -1; j >= -d; j--)
            for (i = -1; i >= -d; i--)
                P3(a, i, j, k) = (double)0.5 * (((double)2 * P3(a, i + 1, j, k) - P3(a, i + 2, j, k)) +
                                                ((double)2 * P3(a, i, j + 1, k) - P3(a, i, j + 2, k)));
    for (k = 0; k < nk; k++)
        for (j = nj; j <= nj - 1 + d; j++)
            for (i = ni; i <= ni - 1 + d; i++)
                P3(a, i, j, k) = (double)0.5 * (((double)2 * P3(a, i - 1, j, k) - P3(a, i - 2, j, k)) +
                                                ((double)2 * P3(a, i, j - 1, k) - P3(a, i, j - 2, k)));
    for (k = 0; k < nk; k++)
        for (j = nj; j <= nj - 1 + d; j++)
            for (i = -1; i >= -d; i--)
                P3(a, i, j, k) = (double)0.5 * (((double)2 * P3(a, i + 1, j, k) - P3(a, i + 2, j, k)) +
                                                ((double)2 * P3(a, i, j - 1, k) - P3(a, i, j - 2, k)));
}

/* ---- set-up ----------------------------------------------------------------- */

int fsop_create(const fs_params* P, const fs_pipe_params* Q, const fs_grid* g, fsop_state** out) {
    if (!P || !Q || !g || !out) return FS_ERR_BAD_ARG;
    if ((Q->is_parallel != 0 && Q->is_parallel != 1) || !Q->dimensionless_form || !Q->cylindrical_boundary ||
        Q->boundary_order_velocity != 1)
        return FS_ERR_UNSUPPORTED;
    if (g->nx < 8 || g->ny < 8 || g->nz < 8) return FS_ERR_BAD_ARG;
    if (P->order_for_time_advancing < 1 || P->order_for_time_advancing > 3) return FS_ERR_BAD_ARG;
    fsop_state* s = (fsop_state*)calloc(1, sizeof(fsop_state));
    s->P = *P;
    s->Q = *Q;
    s->gr = *g;
    s->vert = !Q->is_parallel;
    s->nx = g->nx, s->ny = g->ny, s->nz = g->nz;
    s->sx = g->nx + 1 + 2 * PG;
    s->sy = g->ny + 1 + 2 * PG;
    s->ntot = s->sx * s->sy * (g->nz + 1 + 2 * PG);
    s->cell.x = mkax(g->x_min, g->x_max, g->nx);
    s->cell.y = mkax(g->y_min, g->y_max, g->ny);
    s->cell.z = mkax(g->z_min, g->z_max, g->nz);
    double dx = s->cell.x.d, dy = s->cell.y.d, dz = s->cell.z.d;
    for (int c = 0; c < 4; c++) {
        s->mac[c] = s->cell;
        s->x0f[c] = g->x_min;
        s->y0f[c] = g->y_min;
        s->z0f[c] = g->z_min;
    }
    /* MAC grids, EULERIAN_FLUID_SOLVER_3D.h:407,415,423 */
    s->mac[1].x = mkax(g->x_min - (double)0.5 * dx, g->x_max + (double)0.5 * dx, g->nx + 1);
    s->mac[2].y = mkax(g->y_min - (double)0.5 * dy, g->y_max + (double)0.5 * dy, g->ny + 1);
    s->mac[3].z = mkax(g->z_min - (double)0.5 * dz, g->z_max + (double)0.5 * dz, g->nz + 1);
    s->x0f[1] = g->x_min - (double)0.5 * dx;
    s->y0f[2] = g->y_min - (double)0.5 * dy;
    s->z0f[3] = g->z_min - (double)0.5 * dz;
    /* EULERIAN_FLUID_SOLVER_3D.h:496-520 */
    const double a = Q->a, m = Q->m, K = Q->K, gg = Q->g, rw = P->water_density, ro = Q->oil_density;
    s->R1 = Q->R2 / a;
    s->f = (double)1 / (K - (double)1) * gg * (ro - K * rw);
    s->A = (a * a - (double)1) + (m * K + (double)2 * (K - 1) * log(a));
    s->V0 = (s->f + rw * gg) * (s->R1 * s->R1) / ((double)4 * P->water_viscosity) * s->A;
    s->J = P->surface_tension * s->R1 * ro / (Q->oil_viscosity * Q->oil_viscosity);
    s->Re = (ro * s->R1 * fabs(s->V0)) / Q->oil_viscosity;
    s->We = ((double)1 / s->J) * (s->Re * s->Re);
    s->one_over_We = (double)1 / s->We;
    s->A0 = Q->A_0 / (Q->R2 / a); /* SIMULATION_WORLD.h:126-127 */
    long n = s->ntot;
    s->phi = pz(n); s->bls = pz(n); s->u = pz(n); s->v = pz(n); s->w = pz(n); s->p = pz(n);
    s->sfg = pz(n); s->gx = pz(n); s->gy = pz(n); s->gz = pz(n);
    for (int c = 0; c < 3; c++)
        for (int e = 0; e < 2; e++) {
            s->ag[c][e] = pz(n);
            s->vg[c][e] = pz(n);
        }
    s->vgyx = pz(n); s->vgzx = pz(n);
    for (int c = 1; c < 4; c++) s->bp[c] = pz(n);
    s->psfg = pz(n); s->rho = pz(n); s->irho = pz(n); s->div = pz(n); s->jc = pz(n); s->kappa = pz(n);
    s->sgn = pz(n); s->tmp = pz(n); s->scr = pz(n); s->scr2 = pz(n); s->den_half = pz(n); s->lev = pz(n);
    s->expl = pz(n);
    for (int c = 1; c < 4; c++) s->rk[c] = pz(n);
    for (int e = 0; e < 6; e++) {
        s->visc[e] = pz(n);
        s->coef[e] = pz(n);
    }
    s->pbc = (int*)calloc((size_t)n, sizeof(int));
    s->vbc = (int*)calloc((size_t)n, sizeof(int));
    /* boundary_phi_<c> on the interior of each face grid: sqrt(y^2 + z^2) - a
     * for the parallel pipe, sqrt(x^2 + z^2) - a for the vertical one
     * (ADVECTION_3D.h:380-432 and its copies in VISCOSITY_3D, PROJECTION_3D and
     * EULERIAN_FLUID_SOLVER_3D) */
    for (int c = 1; c < 4; c++) {
        int ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    double w1 = s->vert ? s->x0f[c] + i * s->mac[c].x.d : s->y0f[c] + j * s->mac[c].y.d;
                    double z = s->z0f[c] + k * s->mac[c].z.d;
                    P3(s->bp[c], i, j, k) = sqrt(w1 * w1 + z * z) - a;
                }
    }
    /* the pipe inside the box: the first and last two cells and faces of
     * every line across it (y and z for the parallel pipe, x and z for the
     * vertical one) are outside the wall */
    for (int c = 0; c < 4; c++) {
        int ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        int bad = 0;
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    int q = s->vert ? i : j, nq = s->vert ? ni : nj;
                    if (!(q < 2 || q >= nq - 2 || k < 2 || k >= nk - 2)) continue;
                    double w1 = s->vert ? s->x0f[c] + i * s->mac[c].x.d : s->y0f[c] + j * s->mac[c].y.d;
                    double z = s->z0f[c] + k * s->mac[c].z.d;
                    if (!(sqrt(w1 * w1 + z * z) - a > 0)) bad = 1;
                }
        if (bad) {
            fsop_destroy(s);
            return FS_ERR_BAD_ARG;
        }
    }
    *out = s;
    return FS_OK;
}

void fsop_destroy(fsop_state* s) {
    if (!s) return;
    double* all[] = {s->phi, s->bls, s->u, s->v, s->w, s->p, s->sfg, s->gx, s->gy, s->gz, s->vgyx, s->vgzx,
                     s->psfg, s->rho, s->irho, s->div, s->jc, s->kappa, s->sgn, s->tmp, s->scr, s->scr2,
                     s->den_half, s->lev, s->expl};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); q++) free(all[q]);
    for (int c = 0; c < 3; c++)
        for (int e = 0; e < 2; e++) {
            free(s->ag[c][e]);
            free(s->vg[c][e]);
        }
    for (int c = 1; c < 4; c++) {
        free(s->bp[c]);
        free(s->rk[c]);
    }
    for (int e = 0; e < 6; e++) {
        free(s->visc[e]);
        free(s->coef[e]);
    }
    free(s->pbc);
    free(s->vbc);
    free(s->sys.rp); free(s->sys.ci); free(s->sys.val); free(s->sys.b); free(s->sys.x);
    free(s);
}

int fsop_constants(fsop_state* s, double* out) {
    out[0] = s->R1, out[1] = s->f, out[2] = s->A, out[3] = s->V0, out[4] = s->J, out[5] = s->Re, out[6] = s->We;
    return FS_OK;
}

/* The scene's initial state: EULERIAN_FLUID_SOLVER_3D.h:536-615, the
 * dimensionless core-annular profile -- on the x faces with their ghost
 * layers of width 2 for the parallel pipe; on the y faces' interior, with
 * the sign of V0, for the vertical one --, :453-460 and :755-757 (the level
 * sets' ghosts from a constant 3 dx), then SIMULATION_WORLD.cpp:288-319:
 * phi = r - (A_0 cos(alpha s) + 1) on the interior (r the distance from the
 * pipe's axis, s the coordinate along it), its ghosts the linear
 * extrapolation of the pipe level set's values at that moment (:303, all
 * 3 dx, so 3 dx), and the pipe level set r - a with its ghosts linearly
 * extrapolated (:319). */
int fsop_initial_fields(fsop_state* s) {
    const double a = s->Q.a, A = s->A, K = s->Q.K, m = s->Q.m;
    memset(s->u, 0, sizeof(double) * s->ntot);
    memset(s->v, 0, sizeof(double) * s->ntot);
    memset(s->w, 0, sizeof(double) * s->ntot);
    memset(s->p, 0, sizeof(double) * s->ntot);
    if (s->vert) {
        for (int k = 0; k <= s->nz - 1; k++)
            for (int j = 0; j <= s->ny; j++)
                for (int i = 0; i <= s->nx - 1; i++) {
                    double x = s->gr.x_min + i * s->mac[2].x.d, z = s->gr.z_min + k * s->mac[2].z.d;
                    double magnitude = sqrt(x * x + z * z);
                    double val = 0;
                    if ((magnitude >= 1) && (magnitude <= a)) {
                        if (s->V0 > 0)
                            val = (double)1 / A * (a * a - (x * x + z * z) - (double)2 * (K - 1) * log(sqrt(x * x + z * z) / a));
                        else
                            val = -(double)1 / A * (a * a - (x * x + z * z) - (double)2 * (K - 1) * log(sqrt(x * x + z * z) / a));
                    } else if (magnitude < 1) {
                        if (s->V0 > 0)
                            val = (double)1 - (double)1 / A * (m * (x * x + z * z) * K);
                        else
                            val = -((double)1 - (double)1 / A * (m * (x * x + z * z) * K));
                    }
                    P3(s->v, i, j, k) = val;
                }
    } else {
        for (int k = -2; k <= s->nz - 1 + 2; k++)
            for (int j = -2; j <= s->ny - 1 + 2; j++)
                for (int i = -2; i <= s->nx + 2; i++) {
                    double y = s->gr.y_min + j * s->mac[1].y.d, z = s->gr.z_min + k * s->mac[1].z.d;
                    double magnitude = sqrt(y * y + z * z);
                    double val;
                    if ((magnitude >= 1) && (magnitude <= a))
                        val = (double)1 / A * (a * a - (y * y + z * z) - (double)2 * (K - 1) * log(sqrt(y * y + z * z) / a));
                    else if (magnitude < 1)
                        val = (double)1 - (double)1 / A * (m * (y * y + z * z) * K);
                    else
                        val = 0;
                    P3(s->u, i, j, k) = val;
                }
    }
    const double three_dx = s->cell.x.d * (double)3;
    for (long q = 0; q < s->ntot; q++) s->phi[q] = three_dx;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double x = s->gr.x_min + i * s->cell.x.d, y = s->gr.y_min + j * s->cell.y.d,
                       z = s->gr.z_min + k * s->cell.z.d;
                if (s->vert) {
                    P3(s->phi, i, j, k) = sqrt(x * x + z * z) - (s->A0 * cos(s->Q.alpha * y) + (double)1);
                    P3(s->bls, i, j, k) = sqrt(x * x + z * z) - a;
                } else {
                    P3(s->phi, i, j, k) = sqrt(y * y + z * z) - (s->A0 * cos(s->Q.alpha * x) + (double)1);
                    P3(s->bls, i, j, k) = sqrt(y * y + z * z) - a;
                }
            }
    fill_linear_self(s, s->bls, 2);
    return FS_OK;
}

int fsop_set_fields(fsop_state* s, const double* phi, const double* u, const double* v, const double* w,
                    const double* p) {
    const double* src[5] = {phi, u, v, w, p};
    double* dst[5] = {s->phi, s->u, s->v, s->w, s->p};
    const int comp[5] = {0, 1, 2, 3, 0};
    for (int f = 0; f < 5; f++) {
        if (!src[f]) continue;
        int ni, nj, nk;
        box(s, comp[f], &ni, &nj, &nk);
        long q = 0;
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) P3(dst[f], i, j, k) = src[f][q++];
    }
    return FS_OK;
}

int fsop_get_fields(fsop_state* s, double* phi, double* u, double* v, double* w, double* p) {
    double* dst[5] = {phi, u, v, w, p};
    const double* src[5] = {s->phi, s->u, s->v, s->w, s->p};
    const int comp[5] = {0, 1, 2, 3, 0};
    for (int f = 0; f < 5; f++) {
        if (!dst[f]) continue;
        int ni, nj, nk;
        box(s, comp[f], &ni, &nj, &nk);
        long q = 0;
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) dst[f][q++] = P3(src[f], i, j, k);
    }
    return FS_OK;
}

/* ---- boundary conditions of the velocity ghost copies ------------------------ */

static double theta_of(double near, double far) { return fabs(near) / (fabs(near) + fabs(far)); }

/* ADVECTION_3D::SetupBoundaryConditionForVelocity, level-set branch
 * (ADVECTION_3D.h:436-550); is_parallel (:494-550): masked periodic copies of u, v, w
 * into the shared ghost copies, then on y and z faces outside the pipe the
 * velocity is zeroed and the ghost copy extrapolated from inside.  On y
 * faces the second test repeats `phi(j-1) < 0` (:528) and takes the j+1
 * side, overwriting the first: the value read at j+1 is the one before this
 * loop reaches j+1. */
static void bc_levelset_advection(fsop_state* s) {
    fill_periodic(s, s->gx, s->u, 1, 3, 1, s->bp[1]);
    fill_periodic(s, s->gy, s->v, 2, 3, 1, s->bp[2]);
    fill_periodic(s, s->gz, s->w, 3, 3, 1, s->bp[3]);
    if (s->vert) {
        /* is_vertical (:438-491): x faces with both x tests as written, then
         * z faces, the all-outside test over x and z neighbours */
        const double* b = s->bp[1];
        for (int k = 0; k < s->nz; k++)
            for (int j = 0; j < s->ny; j++)
                for (int i = 0; i < s->nx + 1; i++) {
                    if (!(P3(b, i, j, k) > 0)) continue;
                    P3(s->u, i, j, k) = 0;
                    if (P3(b, i - 1, j, k) > 0 && P3(b, i + 1, j, k) > 0 && P3(b, i, j, k - 1) > 0 && P3(b, i, j, k + 1) > 0)
                        P3(s->gx, i, j, k) = 0;
                    if (P3(b, i - 1, j, k) < 0) {
                        double th = theta_of(P3(b, i - 1, j, k), P3(b, i, j, k));
                        P3(s->gx, i, j, k) = (th - 1) / th * P3(s->gx, i - 1, j, k);
                    }
                    if (P3(b, i + 1, j, k) < 0) {
                        double th = theta_of(P3(b, i + 1, j, k), P3(b, i, j, k));
                        P3(s->gx, i, j, k) = (th - 1) / th * P3(s->gx, i + 1, j, k);
                    }
                }
        b = s->bp[3];
        for (int k = 0; k < s->nz + 1; k++)
            for (int j = 0; j < s->ny; j++)
                for (int i = 0; i < s->nx; i++) {
                    if (!(P3(b, i, j, k) > 0)) continue;
                    P3(s->w, i, j, k) = 0;
                    if (P3(b, i - 1, j, k) > 0 && P3(b, i + 1, j, k) > 0 && P3(b, i, j, k - 1) > 0 && P3(b, i, j, k + 1) > 0)
                        P3(s->gz, i, j, k) = 0;
                    if (P3(b, i, j, k - 1) < 0) {
                        double th = theta_of(P3(b, i, j, k - 1), P3(b, i, j, k));
                        P3(s->gz, i, j, k) = (th - 1) / th * P3(s->gz, i, j, k - 1);
                    }
                    if (P3(b, i, j, k + 1) < 0) {
                        double th = theta_of(P3(b, i, j, k + 1), P3(b, i, j, k));
                        P3(s->gz, i, j, k) = (th - 1) / th * P3(s->gz, i, j, k + 1);
                    }
                }
        return;
    }
    const double* b = s->bp[2];
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny + 1; j++)
            for (int i = 0; i < s->nx; i++) {
                if (!(P3(b, i, j, k) > 0)) continue;
                P3(s->v, i, j, k) = 0;
                if (P3(b, i, j - 1, k) > 0 && P3(b, i, j + 1, k) > 0 && P3(b, i, j, k - 1) > 0 && P3(b, i, j, k + 1) > 0)
                    P3(s->gy, i, j, k) = 0;
                if (P3(b, i, j - 1, k) < 0) {
                    double th = theta_of(P3(b, i, j - 1, k), P3(b, i, j, k));
                    P3(s->gy, i, j, k) = (th - 1) / th * P3(s->gy, i, j - 1, k);
                }
                if (P3(b, i, j - 1, k) < 0) {
                    double th = theta_of(P3(b, i, j + 1, k), P3(b, i, j, k));
                    P3(s->gy, i, j, k) = (th - 1) / th * P3(s->gy, i, j + 1, k);
                }
            }
    b = s->bp[3];
    for (int k = 0; k < s->nz + 1; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                if (!(P3(b, i, j, k) > 0)) continue;
                P3(s->w, i, j, k) = 0;
                if (P3(b, i, j - 1, k) > 0 && P3(b, i, j + 1, k) > 0 && P3(b, i, j, k - 1) > 0 && P3(b, i, j, k + 1) > 0)
                    P3(s->gz, i, j, k) = 0;
                if (P3(b, i, j, k - 1) < 0) {
                    double th = theta_of(P3(b, i, j, k - 1), P3(b, i, j, k));
                    P3(s->gz, i, j, k) = (th - 1) / th * P3(s->gz, i, j, k - 1);
                }
                if (P3(b, i, j, k + 1) < 0) {
                    double th = theta_of(P3(b, i, j, k + 1), P3(b, i, j, k));
                    P3(s->gz, i, j, k) = (th - 1) / th * P3(s->gz, i, j, k + 1);
                }
            }
}

/* ADVECTION_3D::SetupBoundaryConditionForVelocity, velocity branch,
 * is_parallel, boundary_order_velocity 1 (ADVECTION_3D.h:702-851): masked
 * periodic copies into the shared and the per-axis ghost copies, then at
 * every wall crossing inside -> outside along y or z the first outside
 * value of the axis' copy is (theta - 1)/theta times the inside value and
 * the next three are pow(10 t, t), t = 2, 3, 4 (so ENO3 never picks a
 * stencil through the wall).  x faces, then z, then y (:716-851). */
static void bc_velocity_advection(fsop_state* s) {
    fill_periodic(s, s->gx, s->u, 1, 3, 1, s->bp[1]);
    fill_periodic(s, s->gy, s->v, 2, 3, 1, s->bp[2]);
    fill_periodic(s, s->gz, s->w, 3, 3, 1, s->bp[3]);
    const int order[3] = {1, 3, 2};
    for (int c = 1; c < 4; c++) {
        fill_periodic(s, s->ag[c - 1][0], comp_field(s, c), c, 3, 1, s->bp[c]);
        fill_periodic(s, s->ag[c - 1][1], comp_field(s, c), c, 3, 1, s->bp[c]);
    }
    /* the wall's first axis: y for the parallel pipe, x for the vertical one
     * (:553-701 writes the x copies <c>_x in place of the y copies) */
    const int ui = s->vert, uj = !s->vert;
    for (int oc = 0; oc < 3; oc++) {
        int c = order[oc], ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        const double* b = s->bp[c];
        const double* g = ghost_comp(s, c);
        double *ay = s->ag[c - 1][0], *az = s->ag[c - 1][1];
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    double bc0 = P3(b, i, j, k);
                    if (bc0 < 0 && P3(b, i + ui, j + uj, k) > 0) {
                        double th = theta_of(bc0, P3(b, i + ui, j + uj, k));
                        P3(ay, i + ui, j + uj, k) = (th - 1) / th * P3(g, i, j, k);
                        for (int t = 2; t <= 4; t++) P3(ay, i + t * ui, j + t * uj, k) = pow(10 * t, t);
                    }
                    if (bc0 < 0 && P3(b, i - ui, j - uj, k) > 0) {
                        double th = theta_of(bc0, P3(b, i - ui, j - uj, k));
                        P3(ay, i - ui, j - uj, k) = (th - 1) / th * P3(g, i, j, k);
                        for (int t = 2; t <= 4; t++) P3(ay, i - t * ui, j - t * uj, k) = pow(10 * t, t);
                    }
                    if (bc0 < 0 && P3(b, i, j, k + 1) > 0) {
                        double th = theta_of(bc0, P3(b, i, j, k + 1));
                        P3(az, i, j, k + 1) = (th - 1) / th * P3(g, i, j, k);
                        for (int t = 2; t <= 4; t++) P3(az, i, j, k + t) = pow(10 * t, t);
                    }
                    if (bc0 < 0 && P3(b, i, j, k - 1) > 0) {
                        double th = theta_of(bc0, P3(b, i, j, k - 1));
                        P3(az, i, j, k - 1) = (th - 1) / th * P3(g, i, j, k);
                        for (int t = 2; t <= 4; t++) P3(az, i, j, k - t) = pow(10 * t, t);
                    }
                }
    }
}

/* VISCOSITY_3D::SetupBoundaryConditionsForVelocity(thread_id), is_parallel,
 * boundary_order_velocity 1 (VISCOSITY_3D.h:2048-2106, 2235-2356): the
 * shared scalar ghost copy from phi, masked periodic copies, then each face
 * outside the pipe takes, in the reference's else-if order, zero (all four
 * y/z neighbours outside), the extrapolation from j-1, j+1, k-1 or k+1 (the
 * first inside one) into its y or z copy, or zero in both. */
static void bc_viscosity(fsop_state* s) {
    fill_periodic(s, s->sfg, s->phi, 0, 3, 1, NULL);
    fill_periodic(s, s->gx, s->u, 1, 3, 1, s->bp[1]);
    fill_periodic(s, s->gy, s->v, 2, 3, 1, s->bp[2]);
    fill_periodic(s, s->gz, s->w, 3, 3, 1, s->bp[3]);
    for (int c = 1; c < 4; c++) {
        fill_periodic(s, s->vg[c - 1][0], comp_field(s, c), c, 3, 1, s->bp[c]);
        fill_periodic(s, s->vg[c - 1][1], comp_field(s, c), c, 3, 1, s->bp[c]);
    }
    const int order[3] = {1, 3, 2};
    /* the wall's first axis: y (parallel), x (vertical, :2106-2233, the x
     * copies <c>_x; its y-face branch extrapolates from v itself, which
     * equals its shared copy on the inside faces it reads) */
    const int ui = s->vert, uj = !s->vert;
    for (int oc = 0; oc < 3; oc++) {
        int c = order[oc], ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        const double* b = s->bp[c];
        const double* g = s->vert && c == 2 ? s->v : ghost_comp(s, c);
        double *ay = s->vg[c - 1][0], *az = s->vg[c - 1][1];
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    double b0 = P3(b, i, j, k);
                    if (!(b0 > 0)) continue;
                    if (P3(b, i - ui, j - uj, k) > 0 && P3(b, i + ui, j + uj, k) > 0 && P3(b, i, j, k - 1) > 0 &&
                        P3(b, i, j, k + 1) > 0) {
                        P3(ay, i, j, k) = 0;
                        P3(az, i, j, k) = 0;
                    } else if (P3(b, i - ui, j - uj, k) < -0) {
                        double th = theta_of(P3(b, i - ui, j - uj, k), b0);
                        P3(ay, i, j, k) = (th - 1) / th * P3(g, i - ui, j - uj, k);
                    } else if (P3(b, i + ui, j + uj, k) < -0) {
                        double th = theta_of(P3(b, i + ui, j + uj, k), b0);
                        P3(ay, i, j, k) = (th - 1) / th * P3(g, i + ui, j + uj, k);
                    } else if (P3(b, i, j, k - 1) < -0) {
                        double th = theta_of(P3(b, i, j, k - 1), b0);
                        P3(az, i, j, k) = (th - 1) / th * P3(g, i, j, k - 1);
                    } else if (P3(b, i, j, k + 1) < -0) {
                        double th = theta_of(P3(b, i, j, k + 1), b0);
                        P3(az, i, j, k) = (th - 1) / th * P3(g, i, j, k + 1);
                    } else {
                        P3(ay, i, j, k) = 0;
                        P3(az, i, j, k) = 0;
                    }
                }
    }
}

/* PROJECTION_3D::SetupBoundaryConditionsForVelocityDivergence, is_parallel
 * (PROJECTION_3D.h:1548-1688): masked periodic copies, then z and y faces
 * outside the pipe: zero (all four outside), the extrapolation from the
 * first inside neighbour along the face's own axis, or zero. */
static void bc_divergence(fsop_state* s) {
    fill_periodic(s, s->gx, s->u, 1, 3, 1, s->bp[1]);
    fill_periodic(s, s->gy, s->v, 2, 3, 1, s->bp[2]);
    fill_periodic(s, s->gz, s->w, 3, 3, 1, s->bp[3]);
    /* parallel: z faces, then y; vertical (:1566-1627): x faces, then z */
    const int ui = s->vert, uj = !s->vert;
    const int comps[2] = {s->vert ? 1 : 3, s->vert ? 3 : 2};
    for (int q = 0; q < 2; q++) {
        int c = comps[q], ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        const double* b = s->bp[c];
        double* g = ghost_comp(s, c);
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    double b0 = P3(b, i, j, k);
                    if (!(b0 > 0)) continue;
                    int di = c == 1, dj = c == 2, dk = c == 3;
                    if (P3(b, i - ui, j - uj, k) > 0 && P3(b, i + ui, j + uj, k) > 0 && P3(b, i, j, k - 1) > 0 &&
                        P3(b, i, j, k + 1) > 0)
                        P3(g, i, j, k) = 0;
                    else if (P3(b, i - di, j - dj, k - dk) < 0) {
                        double th = theta_of(P3(b, i - di, j - dj, k - dk), b0);
                        P3(g, i, j, k) = (th - (double)1) / th * P3(g, i - di, j - dj, k - dk);
                    } else if (P3(b, i + di, j + dj, k + dk) < 0) {
                        double th = theta_of(P3(b, i + di, j + dj, k + dk), b0);
                        P3(g, i, j, k) = (th - (double)1) / th * P3(g, i + di, j + dj, k + dk);
                    } else
                        P3(g, i, j, k) = 0;
                }
    }
}

/* ---- advection ------------------------------------------------------------ */

static void gather7_(fsop_state* s, const double* a, int i, int j, int k, int axis, double* f) {
    for (int o = -3; o <= 3; o++)
        f[o + 3] = axis == 0 ? P3(a, i + o, j, k) : axis == 1 ? P3(a, i, j + o, k) : P3(a, i, j, k + o);
}

/* velocity at a sample point from the shared ghost copies
 * (ADVECTION_METHOD_3D.h:441-465, 2182-2205) */
static void vel_sample(fsop_state* s, int c, int i, int j, int k, double* uu, double* vv, double* ww) {
    const double *gu = s->gx, *gv = s->gy, *gw = s->gz;
    if (c == 0) {
        *uu = (P3(gu, i + 1, j, k) + P3(gu, i, j, k)) * (double)0.5;
        *vv = (P3(gv, i, j + 1, k) + P3(gv, i, j, k)) * (double)0.5;
        *ww = (P3(gw, i, j, k + 1) + P3(gw, i, j, k)) * (double)0.5;
    } else if (c == 1) {
        *uu = P3(gu, i, j, k);
        *vv = (P3(gv, i, j, k) + P3(gv, i, j + 1, k) + P3(gv, i - 1, j, k) + P3(gv, i - 1, j + 1, k)) * (double)0.25;
        *ww = (P3(gw, i, j, k) + P3(gw, i, j, k + 1) + P3(gw, i - 1, j, k) + P3(gw, i - 1, j, k + 1)) * (double)0.25;
    } else if (c == 2) {
        *uu = (P3(gu, i, j, k) + P3(gu, i + 1, j, k) + P3(gu, i, j - 1, k) + P3(gu, i + 1, j - 1, k)) * (double)0.25;
        *vv = P3(gv, i, j, k);
        *ww = (P3(gw, i, j, k) + P3(gw, i, j, k + 1) + P3(gw, i, j - 1, k) + P3(gw, i, j - 1, k + 1)) * (double)0.25;
    } else {
        *uu = (P3(gu, i, j, k) + P3(gu, i + 1, j, k) + P3(gu, i, j, k - 1) + P3(gu, i + 1, j, k - 1)) * (double)0.25;
        *vv = (P3(gv, i, j, k) + P3(gv, i, j + 1, k) + P3(gv, i, j, k - 1) + P3(gv, i, j + 1, k - 1)) * (double)0.25;
        *ww = P3(gw, i, j, k);
    }
}

/* the upwind update of ADVECTION_METHOD_3D.h:467-521 */
static double upwind_(double rho, double dt, double uu, double vv, double ww, double mx, double px, double my,
                      double py, double mz, double pz) {
    double X = uu > 0 ? mx : px, Y = vv > 0 ? my : py, Z = ww > 0 ? mz : pz;
    return rho - dt * (uu * X + vv * Y + ww * Z);
}

/* ADVECTION_3D::Solve_Levelset, oil/water (ADVECTION_3D.h:204-242): the
 * boundary conditions above, the shared scalar ghost copy from phi
 * (periodic along x; its y and z ghosts are never written, so they read 0),
 * and ADVECTION_METHOD_3D::WENO5th with the pipe's level set
 * (ADVECTION_METHOD_3D.h:281-523): cells outside the pipe are skipped. */
int fsop_levelset_advection(fsop_state* s, double dt) {
    bc_levelset_advection(s);
    fill_periodic(s, s->sfg, s->phi, 0, 3, 1, NULL);
    const double eps = s->P.levelset_epsilon;
    const pgc_t* gc = &s->cell;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                if (P3(s->bls, i, j, k) > 0) continue;
                double f[7], mx, px, my, py, mz, pz_, uu, vv, ww;
                gather7_(s, s->sfg, i, j, k, 0, f);
                fso_weno5_1d(f, gc->x.inv, eps, &mx, &px);
                gather7_(s, s->sfg, i, j, k, 1, f);
                fso_weno5_1d(f, gc->y.inv, eps, &my, &py);
                gather7_(s, s->sfg, i, j, k, 2, f);
                fso_weno5_1d(f, gc->z.inv, eps, &mz, &pz_);
                vel_sample(s, 0, i, j, k, &uu, &vv, &ww);
                P3(s->phi, i, j, k) = upwind_(P3(s->phi, i, j, k), dt, uu, vv, ww, mx, px, my, py, mz, pz_);
            }
    return FS_OK;
}

/* ADVECTION_3D::Solve_Velocity, oil/water, use_3rd_eno_v, is_parallel
 * (ADVECTION_3D.h:244-357): the boundary conditions above, then the ENO3
 * with the pipe (ADVECTION_METHOD_3D.h:1936-2243) on each component: fixed
 * faces and faces outside the pipe (boundary_phi >= 0) keep their value;
 * the x derivatives come from the shared copy, the y and z ones from the
 * component's per-axis copies; velocities from the shared copies. */
int fsop_velocity_advection(fsop_state* s, double dt) {
    bc_velocity_advection(s);
    for (int c = 1; c < 4; c++) {
        int ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        const pgc_t* g = &s->mac[c];
        const double inv3x = g->x.inv * (double)1 / 3, inv3y = g->y.inv * (double)1 / 3,
                     inv3z = g->z.inv * (double)1 / 3;
        /* the derivative along the pipe from the shared copy, across it from
         * the per-axis copies (:316-330) */
        const double *sx_ = s->vert ? s->ag[c - 1][0] : ghost_comp(s, c),
                     *sy_ = s->vert ? ghost_comp(s, c) : s->ag[c - 1][0], *sz_ = s->ag[c - 1][1];
        double* rho = comp_field(s, c);
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    if (fixed_face(s, c, i, j, k)) continue;
                    if (!(P3(s->bp[c], i, j, k) < (double)0)) continue;
                    double f[7], mx, px, my, py, mz, pz_, uu, vv, ww;
                    gather7_(s, sx_, i, j, k, 0, f);
                    fso_eno3_1d(f, g->x.d, g->x.inv, g->x.inv2d, inv3x, &mx, &px);
                    gather7_(s, sy_, i, j, k, 1, f);
                    fso_eno3_1d(f, g->y.d, g->y.inv, g->y.inv2d, inv3y, &my, &py);
                    gather7_(s, sz_, i, j, k, 2, f);
                    fso_eno3_1d(f, g->z.d, g->z.inv, g->z.inv2d, inv3z, &mz, &pz_);
                    vel_sample(s, c, i, j, k, &uu, &vv, &ww);
                    P3(rho, i, j, k) = upwind_(P3(rho, i, j, k), dt, uu, vv, ww, mx, px, my, py, mz, pz_);
                }
    }
    return FS_OK;
}

/* ---- forces ---------------------------------------------------------------- */

/* sin(x) for |x| <= pi (a hair beyond after rounding): the smoothed
 * Heaviside's one transcendental.  Host and device libm differ in the last
 * bit on about 3% of its arguments, so the pipe mode takes this polynomial
 * on both sides instead: reduction to [0, pi/2] by sin(x) = sin(pi - x) with
 * a two-part pi (the subtraction from the leading part exact by Sterbenz'
 * lemma), then the Taylor series of sin to x^23 (truncation below 1e-20 on
 * [0, pi/2]) by Horner's rule, coefficients (-1)^k/(2k+1)! rounded to
 * double.  Within 2 ulps of sin; no FMA, so every build gives the same bits. */
static double pipe_sin(double x) {
    static const double c[11] = {-0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13, 0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33, -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49, -0x1.2f49b46814157p-57, 0x1.71b8ef6dcf572p-66, -0x1.761b41316381ap-75};
    const double pi_hi = 0x1.921fb54442d18p+1, pi_lo = 0x1.1a62633145c07p-53;
    const int neg = x < 0;
    double a = neg ? -x : x;
    if (a > pi_hi * 0.5) a = (pi_hi - a) + pi_lo;
    double a2 = a * a, t = c[10];
    for (int k = 9; k >= 0; k--) t = t * a2 + c[k];
    double r = a + a * (a2 * t);
    return neg ? -r : r;
}

/* the polynomial alone, for the tests' accuracy check */
double fsop_sin(double x) { return pipe_sin(x); }

/* EULERIAN_FLUID_SOLVER_3D::HeavisideFunction (:2459-2479), the form
 * PROJECTION_3D::HeavisideFunction also uses (PROJECTION_3D.h:1452-1472) */
static double heav_solver(double phi, double eps) {
    double one_over_epsilon = (double)1 / eps, one_over_pi = (double)1 / PI_;
    if (phi < -eps) return 0;
    if (phi > eps) return 1;
    return (double)0.5 + phi * one_over_epsilon * (double)0.5 + (double)0.5 * one_over_pi * pipe_sin((double)PI_ * phi * one_over_epsilon);
}
/* VISCOSITY_3D::HeavisideFunction (VISCOSITY_3D.h:1234-1256) */
static double heav_visc(double phi, double eps) {
    double one_over_epsilon = (double)1 / eps, one_over_pi = (double)1 / PI_;
    if (phi < -eps) return (double)0;
    if (phi > eps) return (double)1;
    return (double)0.5 * ((double)1 + phi * one_over_epsilon + one_over_pi * pipe_sin(PI_ * phi * one_over_epsilon));
}

/* EULERIAN_FLUID_SOLVER_3D::ApplyGravity, oil/water, dimensionless,
 * is_parallel (:2271-2393): the shared scalar ghost copy from phi, the
 * smoothed density 1 + (eta - 1) H at the x faces (eps 1.5 dx, :2522-2561),
 * then inside the pipe, fixed faces excepted, u += R1/V0^2 (dt g_x +
 * 1/(rho_half rho_oil) dt f), v += R1/V0^2 dt g_y, w += R1/V0^2 dt g_z. */
int fsop_apply_gravity(fsop_state* s, double dt) {
    const double gdx = dt * s->P.gravity[0], gdy = dt * s->P.gravity[1], gdz = dt * s->P.gravity[2];
    const double drive = dt * s->f;
    const double eta = s->Q.eta, ro = s->Q.oil_density, R1 = s->R1, V0 = s->V0;
    const double eps = (double)1.5 * s->cell.x.d;
    fill_periodic(s, s->sfg, s->phi, 0, 3, 1, NULL);
    if (s->vert) {
        /* is_vertical (:2327-2360): the driving pressure along y, on the y
         * faces only, density at the y faces (eps 1.5 dy) */
        const double epsy = (double)1.5 * s->cell.y.d;
        for (int k = 0; k < s->nz; k++)
            for (int j = 0; j < s->ny + 1; j++)
                for (int i = 0; i < s->nx; i++) {
                    double lc = (double)0.5 * (P3(s->sfg, i, j, k) + P3(s->sfg, i, j - 1, k));
                    double dh = (double)1 + (eta - (double)1) * heav_solver(lc, epsy);
                    if (fixed_face(s, 2, i, j, k)) continue;
                    if (P3(s->bp[2], i, j, k) < 0)
                        P3(s->v, i, j, k) += R1 / (V0 * V0) * (gdy + (double)1 / (ro * dh) * drive);
                }
        (void)gdx, (void)gdz, (void)eps;
        return FS_OK;
    }
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx + 1; i++) {
                double lc = (double)0.5 * (P3(s->sfg, i, j, k) + P3(s->sfg, i - 1, j, k));
                double dh = (double)1 + (eta - (double)1) * heav_solver(lc, eps);
                if (fixed_face(s, 1, i, j, k)) continue;
                if (P3(s->bp[1], i, j, k) < 0)
                    P3(s->u, i, j, k) += R1 / (V0 * V0) * (gdx + ((double)1 / (dh * ro)) * drive);
            }
    for (int c = 2; c < 4; c++) {
        int ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        double gd = c == 2 ? gdy : gdz;
        double* vel = comp_field(s, c);
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    if (fixed_face(s, c, i, j, k)) continue;
                    if (P3(s->bp[c], i, j, k) < 0) P3(vel, i, j, k) += R1 / (V0 * V0) * gd;
                }
    }
    return FS_OK;
}

/* ---- linear systems --------------------------------------------------------- */

static void sys_reset(psys_t* S, int cap) {
    if (S->cap < cap) {
        free(S->rp); free(S->ci); free(S->val); free(S->b); free(S->x);
        S->rp = (int*)calloc((size_t)cap + 1, sizeof(int));
        S->ci = (int*)calloc((size_t)cap * 9, sizeof(int));
        S->val = (double*)calloc((size_t)cap * 9, sizeof(double));
        S->b = (double*)calloc((size_t)cap, sizeof(double));
        S->x = (double*)calloc((size_t)cap, sizeof(double));
        S->cap = cap;
    }
    S->N = 0;
    S->nnz = 0;
    S->rp[0] = 0;
}
static void sys_push(psys_t* S, int col, double v) {
    S->ci[S->nnz] = col;
    S->val[S->nnz] = v;
    S->nnz++;
}

/* CSR_MATRIX::Multiply (CSR_MATRIX.h:236-260): each row in its entry order */
static void sys_mul(const psys_t* S, const double* x, double* y) {
    for (int r = 0; r < S->N; r++) {
        double v = 0;
        for (int e = S->rp[r]; e < S->rp[r + 1]; e++) v += S->val[e] * x[S->ci[e]];
        y[r] = v;
    }
}
static double sys_diag(const psys_t* S, int r) {
    for (int e = S->rp[r]; e < S->rp[r + 1]; e++)
        if (S->ci[e] == r) return S->val[e];
    return 0;
}
/* The pipe mode's reductions in one fixed order, shared with the HIP
 * kernels (the reference's order depends on its thread count,
 * MULTITHREADING.h:229-250): rows in blocks of 256, each block summed by a
 * binary tree (element t += element t + h for h = 128, 64, ..., 1; rows past
 * the end count as +0) into a block sum; then 256 lanes, lane t adding the
 * block sums t, t + 256, t + 512, ... in order from +0, and the same binary
 * tree over the 256 lanes.  b == NULL sums a. */
#define RBLK 256
static void tree256(double* t) {
    for (int h = RBLK / 2; h >= 1; h /= 2)
        for (int q = 0; q < h; q++) t[q] += t[q + h];
}
static double dot_(const double* a, const double* b, int n) {
    const int nb = (n + RBLK - 1) / RBLK;
    double* part = (double*)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
    double t[RBLK];
    for (int blk = 0; blk < nb; blk++) {
        for (int q = 0; q < RBLK; q++) {
            int r = blk * RBLK + q;
            t[q] = r < n ? (b ? a[r] * b[r] : a[r]) : (double)0;
        }
        tree256(t);
        part[blk] = t[0];
    }
    for (int q = 0; q < RBLK; q++) {
        double v = 0;
        for (int p = q; p < nb; p += RBLK) v += part[p];
        t[q] = v;
    }
    tree256(t);
    free(part);
    return t[0];
}

/* PCG_METHOD::PCGMethod with the Jacobi preconditioner (PCG_METHOD.cpp:4-160)
 * and the PROJECTION_WATER stop rule; x holds x0 on entry. */
static void pcg_(fsop_state* s, psys_t* S, int* iters) {
    const fs_params* P = &s->P;
    int N = S->N;
    double *r = pz(N), *z = pz(N), *p = pz(N), *Ap = pz(N), *dg = pz(N);
    double *x = S->x, *b = S->b;
    for (int q = 0; q < N; q++) dg[q] = sys_diag(S, q);
    sys_mul(S, x, Ap);
    for (int q = 0; q < N; q++) r[q] = b[q] - Ap[q];
    for (int q = 0; q < N; q++) p[q] = r[q] / dg[q];
    double res_new = dot_(r, p, N);
    double bnorm = sqrt(dot_(b, b, N));
    double sqr_tol = P->tolerance * P->tolerance;
    int it = 0, rel = P->stop_rule == FS_STOP_REL_RESIDUAL;
    double rr = dot_(r, r, N);
    if (!(rel && (bnorm == 0 || sqrt(rr) <= P->rel_tolerance * bnorm))) {
        while (it < P->max_iteration) {
            sys_mul(S, p, Ap);
            double alpha = res_new / dot_(p, Ap, N);
            for (int q = 0; q < N; q++) {
                x[q] += alpha * p[q];
                r[q] -= alpha * Ap[q];
                z[q] = r[q] / dg[q];
            }
            double res_old = res_new;
            res_new = dot_(r, z, N);
            if (rel) {
                rr = dot_(r, r, N);
                if (sqrt(rr) < P->rel_tolerance * bnorm) break;
            } else if (res_new < sqr_tol)
                break;
            double beta = res_new / res_old;
            for (int q = 0; q < N; q++) {
                p[q] *= beta;
                p[q] += z[q];
            }
            it++;
        }
    }
    *iters = it;
    s->last_res_abs = sqrt(res_new);
    sys_mul(S, x, Ap);
    for (int q = 0; q < N; q++) r[q] = b[q] - Ap[q];
    double tr = dot_(r, r, N);
    s->last_res_rel = bnorm > 0 ? sqrt(tr) / bnorm : 0;
    free(r); free(z); free(p); free(Ap); free(dg);
}

/* AssignSequentialindicesToFullCells (POISSON_SOLVER_3D.h:527-545): FULL
 * cells of the component's interior renumbered 0..N-1 in k, j, i order */
static int number_rows(fsop_state* s, int* bc, int c) {
    int ni, nj, nk, n = 0;
    box(s, c, &ni, &nj, &nk);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++)
                if (P3(bc, i, j, k) > -1) P3(bc, i, j, k) = n++;
    return n;
}

/* ---- semi-implicit viscosity ---------------------------------------------- */

/* VISCOSITY_3D::SetupBoundaryConditionsForVelocity(velocity, bc),
 * cylindrical (VISCOSITY_3D.h:1350-1546): over the face box grown by one,
 * DIR outside the pipe, PER beyond the ends along the pipe (from the last
 * face on for faces normal to it), FULL elsewhere. */
static void viscosity_bc(fsop_state* s, int c) {
    int ni, nj, nk;
    box(s, c, &ni, &nj, &nk);
    const double* b = s->bp[c];
    for (int k = -1; k <= nk; k++)
        for (int j = -1; j <= nj; j++)
            for (int i = -1; i <= ni; i++) {
                int v;
                if (P3(b, i, j, k) > 0) v = BC_DIR;
                else if (s->vert ? j < 0 : i < 0) v = BC_PER;
                else if (s->vert ? (c == 2 ? j >= nj - 1 : j > nj - 1) : (c == 1 ? i >= ni - 1 : i > ni - 1)) v = BC_PER;
                else v = BC_FULL;
                P3(s->vbc, i, j, k) = v;
            }
}

/* POISSON_SOLVER_3D::BuildLinearSystemNodeForSemiImplicitViscosity with
 * per-axis ghost arrays (:3641-3828, periodic entries along x, then y),
 * then Solve and VectorToGrid (fixed faces keep their value, :4060-4078) and
 * the periodic copy of the last face normal to the pipe (:354-376). */
static void viscosity_solve(fsop_state* s, int c, const double* vx, const double* vy, const double* vz) {
    int ni, nj, nk;
    box(s, c, &ni, &nj, &nk);
    int* bc = s->vbc;
    int N = number_rows(s, bc, c);
    psys_t* S = &s->sys;
    sys_reset(S, N > 0 ? N : 1);
    const pgc_t* g = &s->mac[c];
    const double inv_dxdx = (double)1 / (g->x.d * g->x.d), inv_dydy = (double)1 / (g->y.d * g->y.d),
                 inv_dzdz = (double)1 / (g->z.d * g->z.d);
    const int ie = ni - 1, je = nj - 1;
    double* vel = comp_field(s, c);
    double* const* cf = s->coef;
#define BC(a, b, cc) P3(bc, a, b, cc)
#define CF(e) P3(cf[e], i, j, k)
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                if (BC(i, j, k) < 0) continue;
                int row = BC(i, j, k);
                double coef_ijk = 0;
                if (BC(i - 1, j, k) > -1) { coef_ijk += CF(1); sys_push(S, BC(i - 1, j, k), -CF(1) * inv_dxdx); }
                if (BC(i + 1, j, k) > -1) { coef_ijk += CF(0); sys_push(S, BC(i + 1, j, k), -CF(0) * inv_dxdx); }
                if (BC(i, j - 1, k) > -1) { coef_ijk += CF(3); sys_push(S, BC(i, j - 1, k), -CF(3) * inv_dydy); }
                if (BC(i, j + 1, k) > -1) { coef_ijk += CF(2); sys_push(S, BC(i, j + 1, k), -CF(2) * inv_dydy); }
                if (BC(i, j, k - 1) > -1) { coef_ijk += CF(5); sys_push(S, BC(i, j, k - 1), -CF(5) * inv_dzdz); }
                if (BC(i, j, k + 1) > -1) { coef_ijk += CF(4); sys_push(S, BC(i, j, k + 1), -CF(4) * inv_dzdz); }
                if (BC(i - 1, j, k) == BC_PER) {
                    coef_ijk += CF(1);
                    int col = c == 1 ? BC(ie + (i - 1 - (0 + 1)), j, k) : BC(ie + (i - 1 - 0), j, k);
                    sys_push(S, col, -CF(1) * inv_dxdx);
                }
                if (BC(i + 1, j, k) == BC_PER) {
                    coef_ijk += CF(0);
                    int col = c == 1 ? BC(0 + (i + 1 - (ie - 1)), j, k) : BC(0 + (i + 1 - ie), j, k);
                    sys_push(S, col, -CF(0) * inv_dxdx);
                }
                if (BC(i, j - 1, k) == BC_PER) {
                    coef_ijk += CF(3);
                    int col = c == 2 ? BC(i, je + (j - 1 - (0 + 1)), k) : BC(i, je + (j - 1 - 0), k);
                    sys_push(S, col, -CF(3) * inv_dydy);
                }
                if (BC(i, j + 1, k) == BC_PER) {
                    coef_ijk += CF(2);
                    int col = c == 2 ? BC(i, 0 + (j + 1 - (je - 1)), k) : BC(i, 0 + (j + 1 - je), k);
                    sys_push(S, col, -CF(2) * inv_dydy);
                }
                if (BC(i - 1, j, k) == BC_DIR) coef_ijk += CF(1);
                if (BC(i + 1, j, k) == BC_DIR) coef_ijk += CF(0);
                if (BC(i, j - 1, k) == BC_DIR) coef_ijk += CF(3);
                if (BC(i, j + 1, k) == BC_DIR) coef_ijk += CF(2);
                if (BC(i, j, k - 1) == BC_DIR) coef_ijk += CF(5);
                if (BC(i, j, k + 1) == BC_DIR) coef_ijk += CF(4);
                if (coef_ijk == 0) coef_ijk = 1;
                sys_push(S, row, (double)1 + inv_dxdx * (double)coef_ijk);
                double bv = P3(s->expl, i, j, k);
                if (BC(i - 1, j, k) == BC_DIR) bv += CF(1) * inv_dxdx * P3(vx, i - 1, j, k);
                if (BC(i + 1, j, k) == BC_DIR) bv += CF(0) * inv_dxdx * P3(vx, i + 1, j, k);
                if (BC(i, j - 1, k) == BC_DIR) bv += CF(3) * inv_dydy * P3(vy, i, j - 1, k);
                if (BC(i, j + 1, k) == BC_DIR) bv += CF(2) * inv_dydy * P3(vy, i, j + 1, k);
                if (BC(i, j, k - 1) == BC_DIR) bv += CF(5) * inv_dzdz * P3(vz, i, j, k - 1);
                if (BC(i, j, k + 1) == BC_DIR) bv += CF(4) * inv_dzdz * P3(vz, i, j, k + 1);
                S->b[row] = bv;
                S->x[row] = P3(vel, i, j, k);
                S->N = row + 1;
                S->rp[row + 1] = S->nnz;
            }
#undef CF
    int it = 0;
    pcg_(s, S, &it);
    s->visc_iters += it;
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                if (fixed_face(s, c, i, j, k)) continue;
                if (BC(i, j, k) < 0) continue;
                P3(vel, i, j, k) = S->x[BC(i, j, k)];
            }
#undef BC
    if (s->vert) {
        if (c == 2)
            for (int k = 0; k < nk; k++)
                for (int j = 0; j < nj; j++)
                    for (int i = 0; i < ni; i++) P3(vel, i, je, k) = P3(vel, i, 0 + 1, k);
    } else if (c == 1)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) P3(vel, ie, j, k) = P3(vel, 0 + 1, j, k);
}

/* VISCOSITY_3D::ApplyViscosity, oil/water, semi_implicit_approach,
 * dimensionless, is_parallel (VISCOSITY_3D.h:565-852): the boundary
 * conditions (once), then for x, y, z in turn: the smoothed level set at the
 * face and at the four edge centres and two cells around it from the shared
 * scalar ghost copy, density 1 + (eta - 1) H and viscosities 1 + (m - 1) H
 * (eps 1.5 dx, :1258-1347), the coefficients 2 dt/(Re rho) mu (normal) and
 * dt/(Re rho) mu (tangential), the explicit cross terms from the per-axis
 * ghost copies made before any solve, and the solve. */
static int viscosity(fsop_state* s, double dt) {
    bc_viscosity(s);
    const double eps = (double)1.5 * s->cell.x.d, eta = s->Q.eta, m = s->Q.m, Re = s->Re;
    const double* W = s->sfg;
    for (int c = 1; c < 4; c++) {
        int ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        const pgc_t* g = &s->mac[c];
        const double ox = g->x.inv, oy = g->y.inv, oz = g->z.inv;
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    double half, l[6]; /* l[0..5]: the levels of coef_1..coef_6's viscosities */
                    if (c == 1) {
                        half = (double)0.5 * (P3(W, i - 1, j, k) + P3(W, i, j, k));
                        double cu = (double)0.25 * (P3(W, i - 1, j, k) + P3(W, i, j, k) + P3(W, i - 1, j + 1, k) + P3(W, i, j + 1, k));
                        double cd = (double)0.25 * (P3(W, i - 1, j, k) + P3(W, i, j, k) + P3(W, i - 1, j - 1, k) + P3(W, i, j - 1, k));
                        double ct = (double)0.25 * (P3(W, i - 1, j, k) + P3(W, i, j, k) + P3(W, i - 1, j, k + 1) + P3(W, i, j, k + 1));
                        double cb = (double)0.25 * (P3(W, i - 1, j, k) + P3(W, i, j, k) + P3(W, i - 1, j, k - 1) + P3(W, i, j, k - 1));
                        l[0] = P3(W, i, j, k), l[1] = P3(W, i - 1, j, k), l[2] = cu, l[3] = cd, l[4] = ct, l[5] = cb;
                    } else if (c == 2) {
                        half = (double)0.5 * (P3(W, i, j - 1, k) + P3(W, i, j, k));
                        double cu = (double)0.25 * (P3(W, i, j - 1, k) + P3(W, i + 1, j - 1, k) + P3(W, i, j, k) + P3(W, i + 1, j, k));
                        double cl = (double)0.25 * (P3(W, i - 1, j - 1, k) + P3(W, i, j - 1, k) + P3(W, i - 1, j, k) + P3(W, i, j, k));
                        double ct = (double)0.25 * (P3(W, i, j - 1, k + 1) + P3(W, i, j - 1, k) + P3(W, i, j, k + 1) + P3(W, i, j, k));
                        double cb = (double)0.25 * (P3(W, i, j - 1, k - 1) + P3(W, i, j - 1, k) + P3(W, i, j, k - 1) + P3(W, i, j, k));
                        l[0] = cu, l[1] = cl, l[2] = P3(W, i, j, k), l[3] = P3(W, i, j - 1, k), l[4] = ct, l[5] = cb;
                    } else {
                        half = (double)0.5 * (P3(W, i, j, k - 1) + P3(W, i, j, k));
                        double ct = (double)0.25 * (P3(W, i, j, k - 1) + P3(W, i + 1, j, k - 1) + P3(W, i, j, k) + P3(W, i + 1, j, k));
                        double cb = (double)0.25 * (P3(W, i, j, k - 1) + P3(W, i - 1, j, k - 1) + P3(W, i, j, k) + P3(W, i - 1, j, k));
                        double cu = (double)0.25 * (P3(W, i, j, k - 1) + P3(W, i, j + 1, k - 1) + P3(W, i, j, k) + P3(W, i, j + 1, k));
                        double cd = (double)0.25 * (P3(W, i, j, k - 1) + P3(W, i, j - 1, k - 1) + P3(W, i, j, k) + P3(W, i, j - 1, k));
                        l[0] = ct, l[1] = cb, l[2] = cu, l[3] = cd, l[4] = P3(W, i, j, k), l[5] = P3(W, i, j, k - 1);
                    }
                    double rho_h = (double)1 + (eta - (double)1) * heav_visc(half, eps);
                    double mu[6];
                    for (int e = 0; e < 6; e++) mu[e] = (double)1 + (m - (double)1) * heav_visc(l[e], eps);
                    int normal0 = c == 1 ? 0 : c == 2 ? 2 : 4; /* the two normal coefficients take 2 dt */
                    for (int e = 0; e < 6; e++) {
                        int nrm = e == normal0 || e == normal0 + 1;
                        P3(s->coef[e], i, j, k) = nrm ? (double)2 * dt / (Re * rho_h) * mu[e] : dt / (Re * rho_h) * mu[e];
                    }
                    double ex;
                    /* the per-axis copies the cross terms read: the parallel pipe's
                     * x term reads y_x and z_x, which it never writes (zero); the
                     * vertical pipe's are its x copies (:615-621, 754-760, 821-827) */
                    if (c == 1) {
                        const double *Y = s->vert ? s->vg[1][0] : s->vgyx, *Z = s->vert ? s->vg[2][0] : s->vgzx;
                        ex = P3(s->gx, i, j, k) +
                             dt / (Re * rho_h) * (oy * (mu[2] * ox * (P3(Y, i, j + 1, k) - P3(Y, i - 1, j + 1, k)) - mu[3] * ox * (P3(Y, i, j, k) - P3(Y, i - 1, j, k)))) +
                             dt / (Re * rho_h) * (oz * (mu[4] * ox * (P3(Z, i, j, k + 1) - P3(Z, i - 1, j, k + 1)) - mu[5] * ox * (P3(Z, i, j, k) - P3(Z, i - 1, j, k))));
                    } else if (c == 2) {
                        const double *X = s->vg[0][0], *Z = s->vert ? s->vg[2][1] : s->vg[2][0];
                        ex = P3(s->gy, i, j, k) +
                             dt / (Re * rho_h) * (ox * (mu[0] * oy * (P3(X, i + 1, j, k) - P3(X, i + 1, j - 1, k)) - mu[1] * oy * (P3(X, i, j, k) - P3(X, i, j - 1, k)))) +
                             dt / (Re * rho_h) * (oz * (mu[4] * oy * (P3(Z, i, j, k + 1) - P3(Z, i, j - 1, k + 1)) - mu[5] * ox * (P3(Z, i, j, k) - P3(Z, i, j - 1, k))));
                    } else {
                        const double *X = s->vg[0][1], *Y = s->vg[1][1];
                        ex = P3(s->gz, i, j, k) +
                             dt / (Re * rho_h) * (ox * (oz * (mu[0] * (P3(X, i + 1, j, k) - P3(X, i + 1, j, k - 1))) - oz * (mu[1] * (P3(X, i, j, k) - P3(X, i, j, k - 1))))) +
                             dt / (Re * rho_h) * (oy * (oz * (mu[2] * (P3(Y, i, j + 1, k) - P3(Y, i, j + 1, k - 1))) - (s->vert ? oz : ox) * (mu[3] * (P3(Y, i, j, k) - P3(Y, i, j, k - 1)))));
                    }
                    P3(s->expl, i, j, k) = ex;
                }
        viscosity_bc(s, c);
        /* the Dirichlet values along x, y, z: the shared copy along the pipe,
         * the per-axis copies across it (:712-716, 778-782, 844-848) */
        if (s->vert)
            viscosity_solve(s, c, s->vg[c - 1][0], ghost_comp(s, c), s->vg[c - 1][1]);
        else
            viscosity_solve(s, c, ghost_comp(s, c), s->vg[c - 1][0], s->vg[c - 1][1]);
    }
    return FS_OK;
}

/* EULERIAN_FLUID_SOLVER_3D::Sourcing (:2076-2100) in this mode: the phase
 * density (read only by the CSF model, refused here), the viscosity; no
 * object sources the water level set. */
int fsop_sourcing(fsop_state* s, double dt, fs_step_stats* st) {
    s->visc_iters = 0;
    viscosity(s, dt);
    if (st) st->viscosity_iterations = s->visc_iters;
    return FS_OK;
}

/* ---- projection ----------------------------------------------------------- */

/* LEVELSET_3D::ComputeCurvatures by level set (LEVELSET_3D.h:963-1082):
 * phi's ghosts by linear extrapolation (:998), the curvature with the
 * one-sided fallback (its z difference times dz, :1037) and the clamp at
 * -1/min(d), the maximum |kappa| over |phi| < 0.1. */
static void curvatures(fsop_state* s) {
    fill_linear_self(s, s->phi, 2);
    const pgc_t* g = &s->cell;
    const double* phi = s->phi;
    double tolerance = (double)1 / fmin(fmin(g->x.d, g->y.d), g->z.d);
    double amax = 0;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double px = (P3(phi, i + 1, j, k) - P3(phi, i - 1, j, k)) * g->x.inv2d;
                double py = (P3(phi, i, j + 1, k) - P3(phi, i, j - 1, k)) * g->y.inv2d;
                double pz_ = (P3(phi, i, j, k + 1) - P3(phi, i, j, k - 1)) * g->z.inv2d;
                double pxx = (P3(phi, i + 1, j, k) - 2 * P3(phi, i, j, k) + P3(phi, i - 1, j, k)) * g->x.inv_d2;
                double pyy = (P3(phi, i, j + 1, k) - 2 * P3(phi, i, j, k) + P3(phi, i, j - 1, k)) * g->y.inv_d2;
                double pzz = (P3(phi, i, j, k + 1) - 2 * P3(phi, i, j, k) + P3(phi, i, j, k - 1)) * g->z.inv_d2;
                double pxy = (P3(phi, i + 1, j + 1, k) - P3(phi, i + 1, j - 1, k) - P3(phi, i - 1, j + 1, k) +
                              P3(phi, i - 1, j - 1, k)) * g->x.inv2d * g->y.inv2d;
                double pxz = (P3(phi, i + 1, j, k + 1) - P3(phi, i + 1, j, k - 1) - P3(phi, i - 1, j, k + 1) +
                              P3(phi, i - 1, j, k - 1)) * g->x.inv2d * g->z.inv2d;
                double pyz = (P3(phi, i, j + 1, k + 1) - P3(phi, i, j + 1, k - 1) - P3(phi, i, j - 1, k + 1) +
                              P3(phi, i, j - 1, k - 1)) * g->y.inv2d * g->z.inv2d;
                double magnitude = sqrt(px * px + py * py + pz_ * pz_);
                double deno = magnitude * magnitude * magnitude, curv = 0;
                if (deno == 0) {
                    px = (P3(phi, i + 1, j, k) - P3(phi, i, j, k)) * g->x.inv;
                    py = (P3(phi, i, j + 1, k) - P3(phi, i, j, k)) * g->y.inv;
                    pz_ = (P3(phi, i, j, k + 1) - P3(phi, i, j, k)) * g->z.d;
                    magnitude = sqrt(px * px + py * py + pz_ * pz_);
                    deno = magnitude * magnitude * magnitude;
                }
                if (deno != 0) {
                    double one_over_deno = (double)1 / deno;
                    curv = -(px * px * pyy - (double)2 * px * py * pxy + py * py * pxx + px * px * pzz -
                             (double)2 * px * pz_ * pxz + pz_ * pz_ * pxx + py * py * pzz - py * pz_ * pyz + pz_ * pz_ * pyy) *
                           one_over_deno;
                }
                double kv = curv < -tolerance ? -tolerance : curv;
                P3(s->kappa, i, j, k) = kv;
                if (fabs(P3(phi, i, j, k)) < 1e-1) amax = fabs(kv) > amax ? fabs(kv) : amax;
            }
    s->kmax = amax;
}

int fsop_compute_curvatures(fsop_state* s) {
    curvatures(s);
    return FS_OK;
}

/* the pressure system's bc field, PROJECTION_3D::SetupBoundaryConditions,
 * Neumann, cylindrical, is_parallel (:1017-1065): over the cell box grown by
 * one, cells outside the pipe are NEUM next to it (a y or z neighbour
 * inside) and OTHER beyond, the x ends NEUM, FULL elsewhere */
static void pressure_bc(fsop_state* s) {
    const double* b = s->bls;
    for (int k = -1; k <= s->nz; k++)
        for (int j = -1; j <= s->ny; j++)
            for (int i = -1; i <= s->nx; i++) {
                int v;
                if (s->vert) {
                    /* is_vertical (:1025-1043): every cell outside the pipe and
                     * every y ghost NEUM */
                    if (P3(b, i, j, k) > 0) v = BC_NEUM;
                    else if (j < 0) v = BC_NEUM;
                    else if (j > s->ny - 1) v = BC_NEUM;
                    else v = BC_FULL;
                    P3(s->pbc, i, j, k) = v;
                    continue;
                }
                if (P3(b, i, j, k) > 0) {
                    if (P3(b, i, j - 1, k) < 0 || P3(b, i, j + 1, k) < 0 || P3(b, i, j, k - 1) < 0 || P3(b, i, j, k + 1) < 0)
                        v = BC_NEUM;
                    else
                        v = BC_OTHER;
                } else if (i == -1)
                    v = BC_NEUM;
                else if (i == s->nx)
                    v = BC_NEUM;
                else
                    v = BC_FULL;
                P3(s->pbc, i, j, k) = v;
            }
}

/* POISSON_SOLVER_3D::BuildLinearSystemNodeJumpConditionVaribleCoefficient,
 * oil/water (:1360-1815) with jc_on_derivative = 0, then the Neumann mean
 * removal of Solve (:221-246, the sum in the fixed order of dot_), the PCG and
 * VectorToGrid. */
static void pressure_solve(fsop_state* s) {
    int* bc = s->pbc;
    int N = number_rows(s, bc, 0);
    psys_t* S = &s->sys;
    sys_reset(S, N > 0 ? N : 1);
    const pgc_t* g = &s->cell;
    const double inv_dxdx = (double)1 / (g->x.d * g->x.d), inv_dydy = (double)1 / (g->y.d * g->y.d),
                 inv_dzdz = (double)1 / (g->z.d * g->z.d);
    const double odx = g->x.inv, ody = g->y.inv, odz = g->z.inv, odx2 = g->x.inv_d2, ody2 = g->y.inv_d2,
                 odz2 = g->z.inv_d2;
    const double *L = s->phi, *V = s->irho, *JC = s->jc;
    int first = 1;
#define BC(a, b, cc) P3(bc, a, b, cc)
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                if (BC(i, j, k) < 0) continue;
                int row = BC(i, j, k);
                double coef_ijk = 0;
                double lc = P3(L, i, j, k), ll = P3(L, i - 1, j, k), lr = P3(L, i + 1, j, k), lb = P3(L, i, j - 1, k),
                       lt = P3(L, i, j + 1, k), ld = P3(L, i, j, k - 1), lu = P3(L, i, j, k + 1);
                double vc = P3(V, i, j, k);
                double beta_l = vc * P3(V, i - 1, j, k) * (fabs(ll) + fabs(lc)) / (vc * fabs(ll) + P3(V, i - 1, j, k) * fabs(lc));
                double beta_r = vc * P3(V, i + 1, j, k) * (fabs(lr) + fabs(lc)) / (P3(V, i + 1, j, k) * fabs(lc) + vc * fabs(lr));
                double beta_b = vc * P3(V, i, j - 1, k) * (fabs(lb) + fabs(lc)) / (vc * fabs(lb) + P3(V, i, j - 1, k) * fabs(lc));
                double beta_t = vc * P3(V, i, j + 1, k) * (fabs(lt) + fabs(lc)) / (P3(V, i, j + 1, k) * fabs(lc) + vc * fabs(lt));
                double beta_d = vc * P3(V, i, j, k - 1) * (fabs(ld) + fabs(lc)) / (vc * fabs(ld) + P3(V, i, j, k - 1) * fabs(lc));
                double beta_u = vc * P3(V, i, j, k + 1) * (fabs(lu) + fabs(lc)) / (P3(V, i, j, k + 1) * fabs(lc) + vc * fabs(lu));
                if (BC(i - 1, j, k) > -1) { coef_ijk += beta_l; sys_push(S, BC(i - 1, j, k), -beta_l * inv_dxdx); }
                if (BC(i + 1, j, k) > -1) { coef_ijk += beta_r; sys_push(S, BC(i + 1, j, k), -beta_r * inv_dxdx); }
                if (BC(i, j - 1, k) > -1) { coef_ijk += beta_b; sys_push(S, BC(i, j - 1, k), -beta_b * inv_dydy); }
                if (BC(i, j + 1, k) > -1) { coef_ijk += beta_t; sys_push(S, BC(i, j + 1, k), -beta_t * inv_dydy); }
                if (BC(i, j, k - 1) > -1) { coef_ijk += beta_d; sys_push(S, BC(i, j, k - 1), -beta_d * inv_dzdz); }
                if (BC(i, j, k + 1) > -1) { coef_ijk += beta_u; sys_push(S, BC(i, j, k + 1), -beta_u * inv_dzdz); }
                /* the pin of :1445-1478: the first full cell adds beta_d when its lower
                 * neighbour is NEUM (its other NEUM neighbours add 0) */
                if (first && row == 0) {
                    if (BC(i, j, k - 1) == BC_NEUM) coef_ijk += beta_d;
                    first = 0;
                }
                if (coef_ijk == 0) coef_ijk = 1;
                sys_push(S, row, inv_dxdx * (double)coef_ijk);
                /* jump terms, :1682-1792, with b_* = 0 */
                double F[6];
                const double ln[6] = {ll, lr, lb, lt, ld, lu};
                const double jn[6] = {P3(JC, i - 1, j, k), P3(JC, i + 1, j, k), P3(JC, i, j - 1, k), P3(JC, i, j + 1, k),
                                      P3(JC, i, j, k - 1), P3(JC, i, j, k + 1)};
                const double bt[6] = {beta_l, beta_r, beta_b, beta_t, beta_d, beta_u};
                const double o2[6] = {odx2, odx2, ody2, ody2, odz2, odz2}, o1[6] = {odx, odx, ody, ody, odz, odz};
                const double vn[6] = {P3(V, i - 1, j, k), P3(V, i + 1, j, k), P3(V, i, j - 1, k), P3(V, i, j + 1, k),
                                      P3(V, i, j, k - 1), P3(V, i, j, k + 1)};
                const double jcc = P3(JC, i, j, k);
                for (int e = 0; e < 6; e++) {
                    double subcell = fabs(ln[e]) / (fabs(lc) + fabs(ln[e]));
                    double aa = (jcc * fabs(ln[e]) + jn[e] * fabs(lc)) / (fabs(lc) + fabs(ln[e]));
                    double bb = (0 * fabs(ln[e]) + 0 * fabs(lc)) / (fabs(lc) + fabs(ln[e]));
                    int minus = e % 2 == 0; /* l, b, d: the neighbour below */
                    if ((lc <= 0 && ln[e] <= 0) || (lc > 0 && ln[e] > 0))
                        F[e] = 0;
                    else if (lc <= 0 && ln[e] > 0)
                        F[e] = minus ? o2[e] * aa * bt[e] - o1[e] * bt[e] * bb * subcell / vn[e]
                                     : o2[e] * aa * bt[e] + o1[e] * bt[e] * bb * subcell / vn[e];
                    else
                        F[e] = minus ? -o2[e] * aa * bt[e] + o1[e] * bt[e] * bb * subcell / vn[e]
                                     : -o2[e] * aa * bt[e] - o1[e] * bt[e] * bb * subcell / vn[e];
                }
                double FX = F[0] + F[1], FY = F[2] + F[3], FZ = F[4] + F[5];
                S->b[row] = -P3(s->div, i, j, k) - FX - FY - FZ;
                S->x[row] = P3(s->p, i, j, k);
                S->N = row + 1;
                S->rp[row + 1] = S->nnz;
            }
    /* Neumann compatibility, POISSON_SOLVER_3D.h:221-246 */
    double sum_for_b = dot_(S->b, NULL, S->N);
    double one_over_b_dimension = 1 / (double)S->N;
    for (int r = 0; r < S->N; r++) S->b[r] -= one_over_b_dimension * sum_for_b;
    int it = 0;
    pcg_(s, S, &it);
    s->last_iters = it;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++)
                if (BC(i, j, k) >= 0) P3(s->p, i, j, k) = S->x[BC(i, j, k)];
#undef BC
}

/* PROJECTION_3D::UpdateVelocityByPressureGradientVariableDensity, oil/water,
 * is_parallel (:1101-1449): the projection's own scalar ghost copy (periodic
 * in x), the smoothed face density 1 + (rho_w/rho_o - 1) H (eps 1.5 d,
 * :1474-1519), then each face not fixed whose cell is FULL: the same-phase
 * or interface update, the max |u|, and the x-periodic copies of the end
 * faces. */
static void velocity_update(fsop_state* s, double dt) {
    fill_periodic(s, s->psfg, s->phi, 0, 2, 1, NULL);
    const double rr = s->P.water_density / s->Q.oil_density;
    for (int c = 1; c < 4; c++) {
        int ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        const pgc_t* g = &s->mac[c];
        const double od = c == 1 ? g->x.inv : c == 2 ? g->y.inv : g->z.inv;
        const double eps = (double)1.5 * (c == 1 ? s->cell.x.d : c == 2 ? s->cell.y.d : s->cell.z.d);
        const int di = c == 1, dj = c == 2, dk = c == 3;
        double* vel = comp_field(s, c);
        double vmax = 0;
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    double lcf = (double)0.5 * (P3(s->psfg, i, j, k) + P3(s->psfg, i - di, j - dj, k - dk));
                    double dh = 1 + (rr - (double)1) * heav_solver(lcf, eps);
                    if (fixed_face(s, c, i, j, k)) continue;
                    if (P3(s->pbc, i, j, k) < 0) continue;
                    double* vp = &P3(vel, i, j, k);
                    const double density_ijk = P3(s->rho, i, j, k);
                    const double one_over_density_ijk = (double)1 / density_ijk;
                    const double lv = P3(s->phi, i, j, k), ln = P3(s->phi, i - di, j - dj, k - dk);
                    const double blv = P3(s->bls, i, j, k), bln = P3(s->bls, i - di, j - dj, k - dk);
                    const double jcv = P3(s->jc, i, j, k), jcn = P3(s->jc, i - di, j - dj, k - dk);
                    const double gamma = (jcv * fabs(ln) + jcn * fabs(lv)) / (fabs(ln) + fabs(lv));
                    const double one_over_half = 1 / dh;
                    const double bc = (double)P3(s->pbc, i, j, k), bcn = (double)P3(s->pbc, i - di, j - dj, k - dk);
                    const double dp = P3(s->p, i, j, k) - P3(s->p, i - di, j - dj, k - dk);
                    if (s->vert && c != 3) {
                        /* is_vertical (:1188-1205, 1297-1318); the y faces skip the
                         * first and the last face */
                        if (bc * bcn >= 0 && (c == 1 || (j != 0 && j <= nj - 1 - 1))) {
                            if (lv * ln > 0)
                                *vp -= dt * dp * one_over_density_ijk * od;
                            else if (lv <= 0)
                                *vp -= dt * (dp + gamma) * one_over_half * od;
                            else if (lv > 0)
                                *vp -= dt * (dp - gamma) * one_over_half * od;
                        }
                    } else if (c == 1) {
                        if (blv * bln >= 0) {
                            if (bc * bcn > 0)
                                *vp -= dt * dp * one_over_density_ijk * od;
                            else if (lv <= 0)
                                *vp -= dt * (dp + gamma) * one_over_half * od;
                            else if (lv > 0)
                                *vp -= dt * (dp - gamma) * one_over_half * od;
                        }
                    } else if (c == 2) {
                        if (bc * bcn >= 0) {
                            if (lv * ln > 0)
                                *vp -= dt * dp * one_over_density_ijk * od;
                            else if (lv < 0)
                                *vp -= dt * (dp + gamma) * one_over_half * od;
                            else if (lv > 0)
                                *vp -= dt * (dp - gamma) * one_over_half * od;
                        }
                    } else {
                        if (bc * bcn >= 0) {
                            if (lv * ln > 0)
                                *vp -= dt * dp * one_over_density_ijk * od;
                            else if (lv <= 0)
                                *vp -= dt * (dp + gamma) * one_over_half * od;
                            else if (lv > 0)
                                *vp -= dt * (dp - gamma) * one_over_half * od;
                        }
                    }
                    vmax = vmax > fabs(*vp) ? vmax : fabs(*vp);
                }
        s->max_vel[c - 1] = vmax;
        /* :1238-1256, :1345-1363, :1429-1444 */
        if (s->vert) {
            for (int k = 0; k < nk; k++)
                for (int i = 0; i < ni; i++) {
                    if (c == 2) {
                        P3(vel, i, nj - 1, k) = P3(vel, i, 0 + 1, k);
                        P3(vel, i, 0, k) = P3(vel, i, nj - 1 - 1, k);
                    } else
                        P3(vel, i, nj - 1, k) = P3(vel, i, 0, k);
                }
            continue;
        }
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++) {
                if (c == 1) {
                    P3(vel, ni - 1, j, k) = P3(vel, 0 + 1, j, k);
                    P3(vel, 0, j, k) = P3(vel, ni - 1 - 1, j, k);
                } else
                    P3(vel, ni - 1, j, k) = P3(vel, 0, j, k);
            }
    }
}

/* PROJECTION_3D::Solve, oil/water, is_parallel (PROJECTION_3D.h:414-432): */
int fsop_projection(fsop_state* s, double dt, fs_step_stats* st) {
    /* phi's own x ghosts and first cell from the periodic image (:424-427) */
    fill_periodic(s, s->phi, s->phi, 0, 2, 0, NULL);
    /* DetermineProjectionDensity, dimensionless (:642-668) */
    const double rr = s->P.water_density / s->Q.oil_density;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double d = P3(s->phi, i, j, k) <= 0 ? 1 : rr;
                P3(s->rho, i, j, k) = d;
                P3(s->irho, i, j, k) = 1 / d;
            }
    /* DetermineDivergence (:697-725) */
    bc_divergence(s);
    const double one_over_dt = 1 / dt;
    const pgc_t* g = &s->cell;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double dv;
                if (P3(s->bls, i, j, k) > 0)
                    dv = (double)0;
                else {
                    dv = g->x.inv * (P3(s->gx, i + 1, j, k) - P3(s->gx, i, j, k));
                    dv += g->y.inv * (P3(s->gy, i, j + 1, k) - P3(s->gy, i, j, k));
                    dv += g->z.inv * (P3(s->gz, i, j, k + 1) - P3(s->gz, i, j, k));
                }
                P3(s->div, i, j, k) = dv * one_over_dt;
            }
    /* DetermineJumpConditionField (:746-792): jc = kappa/We, x-periodic */
    curvatures(s);
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) P3(s->jc, i, j, k) = (double)1 / s->We * P3(s->kappa, i, j, k);
    fill_periodic(s, s->jc, s->jc, 0, 1, 1, NULL);
    /* DeterminePressure (:794-834) */
    pressure_bc(s);
    fill_const_self(s, s->phi, 2); /* water_levelset.ComputeNormals (LEVELSET_3D.h:931) */
    fill_periodic(s, s->irho, s->irho, 0, 2, 0, NULL);
    fill_const_self(s, s->bls, 2); /* boundary_levelset.ComputeNormals */
    pressure_solve(s);
    velocity_update(s, dt);
    if (st) {
        st->pcg_iterations += s->last_iters;
        st->residual_precond = s->last_res_abs;
        st->residual_rel = s->last_res_rel;
        for (int c = 0; c < 3; c++) st->max_velocity[c] = s->max_vel[c];
    }
    return FS_OK;
}

/* ---- reinitialization ------------------------------------------------------ */

/* EULERIAN_FLUID_SOLVER_3D::Reinitialization (:2198-2225) with
 * ADVECTION_3D::ReinitializationBySussman, oil/water (ADVECTION_3D.h:
 * 360-378): each sweep copies phi into the shared scalar ghost copy with its
 * x ghosts periodic (its y and z ghosts stay 0) and takes one
 * WENO5thReinitialization step (ADVECTION_METHOD_3D.h:745-1637) with the
 * reference's Godunov leaves. */
int fsop_reinitialization(fsop_state* s, double dt) {
    (void)dt;
    if (!s->P.sussmanstyle_reinitialization) return FS_OK;
    const pgc_t* gc = &s->cell;
    double fictious_dt = gc->x.d / s->P.scaling_number_for_reinitialization;
    double dx2 = gc->x.d * gc->x.d;
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) {
                double f = P3(s->phi, i, j, k);
                P3(s->sgn, i, j, k) = f / sqrt(f * f + dx2);
            }
    const double eps = s->P.levelset_epsilon;
    for (int it = 0; it < s->P.iteration_number_for_reinitialization; it++) {
        fill_periodic(s, s->sfg, s->phi, 0, 3, 1, NULL);
        for (int k = 0; k < s->nz; k++)
            for (int j = 0; j < s->ny; j++)
                for (int i = 0; i < s->nx; i++) {
                    double f[7], mx, px, my, py, mz, pz_, gn;
                    gather7_(s, s->sfg, i, j, k, 0, f);
                    fso_weno5_1d(f, gc->x.inv, eps, &mx, &px);
                    gather7_(s, s->sfg, i, j, k, 1, f);
                    fso_weno5_1d(f, gc->y.inv, eps, &my, &py);
                    gather7_(s, s->sfg, i, j, k, 2, f);
                    fso_weno5_1d(f, gc->z.inv, eps, &mz, &pz_);
                    double sg = P3(s->sgn, i, j, k);
                    double* rho = &P3(s->phi, i, j, k);
                    if (fso_godunov_norm(sg, mx, px, my, py, mz, pz_, 0, &gn)) *rho = *rho - fictious_dt * sg * gn;
                    if (sg * *rho < (double)0) {
                        if (sg < 0) *rho = -eps;
                        if (sg > 0) *rho = eps;
                    }
                }
    }
    return FS_OK;
}

/* ---- step -------------------------------------------------------------------- */

static void copy_box(fsop_state* s, double* dst, const double* src, int c) {
    int ni, nj, nk;
    box(s, c, &ni, &nj, &nk);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) P3(dst, i, j, k) = P3(src, i, j, k);
}
static void blend_box(fsop_state* s, double* dst, const double* a, double ca, double cb, int c) {
    int ni, nj, nk;
    box(s, c, &ni, &nj, &nk);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) P3(dst, i, j, k) = ca * P3(a, i, j, k) + cb * P3(dst, i, j, k);
}
static void average_box(fsop_state* s, double* dst, const double* a) {
    for (int k = 0; k < s->nz; k++)
        for (int j = 0; j < s->ny; j++)
            for (int i = 0; i < s->nx; i++) P3(dst, i, j, k) = (double)0.5 * (P3(a, i, j, k) + P3(dst, i, j, k));
}

/* SolveForRK, EULERIAN_FLUID_SOLVER_3D.h:1453-1476 */
static void solve_for_rk(fsop_state* s, double dt, fs_step_stats* st) {
    if (s->P.is_velocity_advection_active) fsop_velocity_advection(s, dt);
    if (s->P.is_external_force_active) fsop_apply_gravity(s, dt);
    if (s->P.is_sourcing_active) {
        s->visc_iters = 0;
        viscosity(s, dt);
        if (st) st->viscosity_iterations += s->visc_iters;
    }
    if (s->P.is_projection_active) fsop_projection(s, dt, st);
}

/* AdvanceOneTimeStepThread, orders 1-3 (EULERIAN_FLUID_SOLVER_3D.h:803-1111) */
int fsop_step(fsop_state* s, double dt, fs_step_stats* st) {
    if (st) memset(st, 0, sizeof(*st));
    const int order = s->P.order_for_time_advancing;
    if (order == 1 || order == 2) {
        if (s->P.is_levelset_advection_active) {
            copy_box(s, s->tmp, s->phi, 0);
            fsop_levelset_advection(s, dt);
            fsop_levelset_advection(s, dt);
            average_box(s, s->phi, s->tmp);
        }
        if (order == 1)
            solve_for_rk(s, dt, st);
        else {
            for (int c = 1; c < 4; c++) copy_box(s, s->rk[c], comp_field(s, c), c);
            solve_for_rk(s, dt, st);
            solve_for_rk(s, dt, st);
            for (int c = 1; c < 4; c++) blend_box(s, comp_field(s, c), s->rk[c], (double)0.5, (double)0.5, c);
        }
        copy_box(s, s->tmp, s->phi, 0);
        fsop_reinitialization(s, dt);
        fsop_reinitialization(s, dt);
        average_box(s, s->phi, s->tmp);
    } else {
        if (s->P.is_levelset_advection_active) fsop_levelset_advection(s, dt);
        for (int c = 1; c < 4; c++) copy_box(s, s->rk[c], comp_field(s, c), c);
        solve_for_rk(s, dt, st);
        solve_for_rk(s, dt, st);
        for (int c = 1; c < 4; c++) blend_box(s, comp_field(s, c), s->rk[c], (double)0.75, (double)0.25, c);
        solve_for_rk(s, dt, st);
        for (int c = 1; c < 4; c++) blend_box(s, comp_field(s, c), s->rk[c], (double)1 / (double)3, (double)2 / (double)3, c);
        copy_box(s, s->tmp, s->phi, 0);
        fsop_reinitialization(s, dt);
        fsop_reinitialization(s, dt);
        blend_box(s, s->phi, s->tmp, (double)0.75, (double)0.25, 0);
        fsop_reinitialization(s, dt);
        blend_box(s, s->phi, s->tmp, (double)1 / (double)3, (double)2 / (double)3, 0);
    }
    return FS_OK;
}

/* SIMULATION_WORLD::AdvanceOneFrame (SIMULATION_WORLD.h:320-330): the level
 * set's ComputeCurvatures, then the step */
int fsop_advance_frame(fsop_state* s, double dt, fs_step_stats* st) {
    curvatures(s);
    return fsop_step(s, dt, st);
}

/* CFLOneTimeStep, oil/water, dimensionless, semi-implicit
 * (EULERIAN_FLUID_SOLVER_3D.h:2673-2785) with the curvature maximum of the
 * current level set (the frame's ComputeCurvatures, SIMULATION_WORLD.h:324,
 * runs before DetermineTimeStep); dt = cfl / rate. */
int fsop_cfl_dt(fsop_state* s, double cfl, double* dt) {
    double mx[3] = {0, 0, 0};
    for (int c = 1; c < 4; c++) {
        int ni, nj, nk;
        box(s, c, &ni, &nj, &nk);
        const double* a = comp_field(s, c);
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++)
                    if (mx[c - 1] <= fabs(P3(a, i, j, k))) mx[c - 1] = fabs(P3(a, i, j, k));
    }
    memcpy(s->scr, s->phi, sizeof(double) * s->ntot);
    double* keep = s->phi;
    s->phi = s->scr;
    curvatures(s);
    s->phi = keep;
    double c_f = mx[0] * s->mac[1].x.inv + mx[1] * s->mac[2].y.inv + mx[2] * s->mac[3].z.inv;
    double g_f = 0;
    if (s->P.is_external_force_active) {
        double mn = s->Q.oil_density < s->P.water_density ? s->Q.oil_density : s->P.water_density;
        g_f = sqrt(s->R1 / (s->V0 * s->V0) * (fabs(s->P.gravity[1]) + fabs((double)1 / mn * s->f)) * s->mac[2].y.inv);
    }
    double rr = s->P.water_density / s->Q.oil_density, mind = fmin(fmin(s->mac[1].x.d, s->mac[2].y.d), s->mac[3].z.d);
    double s_f = sqrt((s->one_over_We * s->kmax) / ((rr < 1 ? rr : 1) * (mind * mind)));
    double v_f = 0;
    double cfl_rate = (double)0.5 * (c_f + v_f + sqrt((c_f + v_f) * (c_f + v_f) + (double)4 * (g_f * g_f) + (double)4 * (s_f * s_f)));
    *dt = cfl * ((double)1 / cfl_rate);
    return FS_OK;
}

int fsop_phase(fsop_state* s, int phase, double dt, fs_step_stats* st) {
    if (st) memset(st, 0, sizeof(*st));
    switch (phase) {
        case FS_PIPE_LEVELSET_ADVECTION: return fsop_levelset_advection(s, dt);
        case FS_PIPE_VELOCITY_ADVECTION: return fsop_velocity_advection(s, dt);
        case FS_PIPE_GRAVITY: return fsop_apply_gravity(s, dt);
        case FS_PIPE_SOURCING: return fsop_sourcing(s, dt, st);
        case FS_PIPE_PROJECTION: return fsop_projection(s, dt, st);
        case FS_PIPE_REINITIALIZATION: return fsop_reinitialization(s, dt);
    }
    return FS_ERR_BAD_ARG;
}

int fsop_get_array(fsop_state* s, int which, double* dst) {
    double* src = NULL;
    int c = 0;
    switch (which) {
        case FS_PIPE_ARR_PHI: src = s->phi; break;
        case FS_PIPE_ARR_SCALAR_GHOST: src = s->sfg; break;
        case FS_PIPE_ARR_GHOST_U: src = s->gx, c = 1; break;
        case FS_PIPE_ARR_GHOST_V: src = s->gy, c = 2; break;
        case FS_PIPE_ARR_GHOST_W: src = s->gz, c = 3; break;
        case FS_PIPE_ARR_ADV_U_Y: src = s->ag[0][0], c = 1; break;
        case FS_PIPE_ARR_ADV_U_Z: src = s->ag[0][1], c = 1; break;
        case FS_PIPE_ARR_ADV_V_Y: src = s->ag[1][0], c = 2; break;
        case FS_PIPE_ARR_ADV_V_Z: src = s->ag[1][1], c = 2; break;
        case FS_PIPE_ARR_ADV_W_Y: src = s->ag[2][0], c = 3; break;
        case FS_PIPE_ARR_ADV_W_Z: src = s->ag[2][1], c = 3; break;
        case FS_PIPE_ARR_VIS_U_Y: src = s->vg[0][0], c = 1; break;
        case FS_PIPE_ARR_VIS_U_Z: src = s->vg[0][1], c = 1; break;
        case FS_PIPE_ARR_VIS_V_Y: src = s->vg[1][0], c = 2; break;
        case FS_PIPE_ARR_VIS_V_Z: src = s->vg[1][1], c = 2; break;
        case FS_PIPE_ARR_VIS_W_Y: src = s->vg[2][0], c = 3; break;
        case FS_PIPE_ARR_VIS_W_Z: src = s->vg[2][1], c = 3; break;
        case FS_PIPE_ARR_BOUNDARY: src = s->bls; break;
        case FS_PIPE_ARR_DIVERGENCE: src = s->div; break;
        case FS_PIPE_ARR_JUMP: src = s->jc; break;
        case FS_PIPE_ARR_PRESSURE: src = s->p; break;
        default: return FS_ERR_BAD_ARG;
    }
    int ni, nj, nk;
    box(s, c, &ni, &nj, &nk);
    long q = 0;
    for (int k = -PG; k < nk + PG; k++)
        for (int j = -PG; j < nj + PG; j++)
            for (int i = -PG; i < ni + PG; i++) dst[q++] = P3(src, i, j, k);
    return FS_OK;
}
