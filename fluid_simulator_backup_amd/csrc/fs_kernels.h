// fs_kernels.h -- launchers of the step kernels (host-callable, internal).
#pragma once
#include <hip/hip_runtime.h>

#include "fs_bodies.h"  // FsBox, fs_box: box extents of a component (0 cell, 1..3 x-, y-, z-faces)
#include "fs_device.h"
#include "fs_layout.h"

// ---- ghost fills (FIELD_STRUCTURE_3D.h) ---------------------------------
void fs_copy_box(hipStream_t st, double* dst, const double* src, const FsLayout& L, FsBox b);
void fs_fill_zero_faces(hipStream_t st, double* dst, const FsLayout& L, FsBox b, int depth);
void fs_fill_const(hipStream_t st, double* dst, const double* src, const FsLayout& L, FsBox b, int depth);
void fs_fill_linear(hipStream_t st, double* dst, const double* src, const FsLayout& L, FsBox b, int depth, int edges);
// the face layers of fs_fill_linear with the box's values read from src and
// the ghost layers' from gsrc (fs_fill_linear: gsrc == src)
void fs_fill_linear_split(hipStream_t st, double* dst, const double* src, const double* gsrc, const FsLayout& L,
                          FsBox b, int depth);

// ---- advection ------------------------------------------------------------
// fast: the WENO5 derivatives in fs_weno5_fast's operation order (fast_weno)
void fs_weno5_stage(hipStream_t st, double* rho, const double* ghost, const double* gu, const double* gv,
                    const double* gw, const FsLayout& L, int comp, const FsGridConst& gc, double eps, double dt,
                    bool fast = false);
void fs_eno3_stage(hipStream_t st, double* rho, const double* ghost, const double* gu, const double* gv,
                   const double* gw, const FsLayout& L, int comp, const FsGridConst& gc, double dt);

// One WENO5 level-set stage with the level set extended by constant
// extrapolation and zero MAC ghosts, and its adjoint (accumulating).
void fs_weno5_transport(hipStream_t st, const double* phi, const double* u, const double* v, const double* w,
                        double* out, const FsLayout& L, const FsGridConst& gc, double eps, double dt);
void fs_adjoint_weno5_transport(hipStream_t st, const double* phi, const double* u, const double* v, const double* w,
                                const double* lam, double* lphi, double* lu, double* lv, double* lw, const FsLayout& L,
                                const FsGridConst& gc, double eps, double dt);

// ---- forces ---------------------------------------------------------------
void fs_gravity(hipStream_t st, double* u, double* v, double* w, const FsLayout& L, double gx, double gy, double gz);
void fs_viscosity(hipStream_t st, int comp, double* dst, const double* U, const double* V, const double* W,
                  const double* phi, const FsLayout& L, const FsMaterial& m, const FsGridConst& cell, double dt);

// lu, lv, lw += (the transpose of comp's explicit viscous update) src
void fs_adjoint_viscosity(hipStream_t st, int comp, const double* src, double* lu, double* lv, double* lw,
                          const double* phi, const FsLayout& L, const FsMaterial& m, const FsGridConst& cell,
                          double dt);

// ---- projection -----------------------------------------------------------
// Semi-implicit viscosity rows of MAC component comp (1..3): off-diagonal
// entries a[0..5] (i-1, i+1, j-1, j+1, k-1, k+1), diagonal and right-hand side.
void fs_visc_system(hipStream_t st, int comp, double* const a[6], double* diag, double* rhs, const double* U,
                    const double* V, const double* W, const double* self, const double* phi, const FsLayout& L,
                    const FsMaterial& m, const FsGridConst& cell, double dt);
void fs_density(hipStream_t st, const double* phi, double* rho, double* irho, const FsLayout& L, double light,
                double heavy);
void fs_divergence(hipStream_t st, const double* gu, const double* gv, const double* gw, double* div,
                   const FsLayout& L, const FsGridConst& cell, double one_over_dt);
// beta of every face between two interior cells (0 elsewhere), A(c,c) and
// b = -div - F with the jump terms of jc (jc = NULL: b = -div).
// pin[0..1]: the first cell's Neumann extras beta_b, beta_d.  axis_diag:
// the diagonal sums each face's beta over its own d^2 instead of the
// reference's 1/dx^2 for every face.
void fs_assemble(hipStream_t st, const double* phi, const double* irho, const double* div, const double* jc,
                 double* bx, double* by, double* bz, double* diag, double* rhs, double* pin, const FsLayout& L,
                 const FsGridConst& cell, int axis_diag);
// curvature of phi (its ghosts as left), jc = sigma * kappa when jc is given;
// kstat[0] = bits of max |kappa| near the interface, kstat[1] += cells whose
// gradient vanishes in both difference forms.  Neither slot is cleared here.
// exact: the corrected cross term and one-sided z difference.
void fs_curvature(hipStream_t st, const double* phi, double* kappa, double* jc, double sigma, const FsLayout& L,
                  const FsGridConst& cell, int exact, unsigned long long* kstat);
// unit normals of phi (ComputeNormals), nrm[0..2] = x, y, z components
void fs_normals(hipStream_t st, const double* phi, double* const* nrm, const FsLayout& L, const FsGridConst& cell);
// curvature by normals: kappa (jc = sigma * kappa when given), kstat[0] =
// bits of max |div n| over every cell (not cleared here)
void fs_curvature_normals(hipStream_t st, double* const* nrm, double* kappa, double* jc, double sigma,
                          const FsLayout& L, const FsGridConst& cell, unsigned long long* kstat);
// one spherical level-set source over the cells lo..hi (inclusive, inside
// the grid): mode 1 phi = min(d, phi), mode 2 phi = max(-d, phi) where
// d <= 3 dx; origin = (x_min, y_min, z_min) of the cell grid
void fs_source_sphere(hipStream_t st, double* phi, const FsLayout& L, const int lo[3], const int hi[3],
                      const double center[3], double radius, const FsGridConst& cell, const double origin[3],
                      int mode);
// b -= mean(b) with the per-plane row-order sum; scratch: nz+1 doubles.
void fs_remove_mean(hipStream_t st, double* rhs, double* scratch, const FsLayout& L);
void fs_gradient_update(hipStream_t st, int comp, int apply, double* vel, const double* p, const double* phi,
                        const double* rho, const double* jc, const FsLayout& L, const FsMaterial& m,
                        const FsGridConst& cell, double dt, unsigned long long* vmax_bits);
// Adjoint pieces of the projection at fixed phi: g += G^T lam (comp's faces
// to cells) and lam += scale * D^T z (cells to the faces of u, v, w).
void fs_adjoint_gradient(hipStream_t st, int comp, const double* lam, const double* phi, const double* rho, double* g,
                         const FsLayout& L, const FsMaterial& m, const FsGridConst& cell, double dt);
void fs_adjoint_divergence(hipStream_t st, double* lu, double* lv, double* lw, const double* z, const FsLayout& L,
                           const FsGridConst& cell, double scale);
// g += Gh^T lam for fs_gradient_update_beta's faces of comp
void fs_adjoint_gradient_beta(hipStream_t st, int comp, const double* lam, const double* beta, double* g,
                              const FsLayout& L, const FsGridConst& cell, double dt);
void fs_abs_max(hipStream_t st, const double* a, const FsLayout& L, FsBox b, unsigned long long* bits);
// fixed schedule: u_f -= dt beta_f (p_c - p_n + jump)/dx with the pressure
// matrix's face coefficients (beta = bx, by or bz of fs_assemble) and its
// jump terms (jc = NULL: none)
void fs_gradient_update_beta(hipStream_t st, int comp, double* vel, const double* p, const double* beta,
                             const double* phi, const double* jc, const FsLayout& L, const FsGridConst& cell,
                             double dt, unsigned long long* vmax_bits);
// zero the two normal wall planes of MAC component comp (1..3)
void fs_zero_walls(hipStream_t st, int comp, double* vel, const FsLayout& L);

// ---- reinitialization -----------------------------------------------------
void fs_sign(hipStream_t st, const double* phi, double* sgn, const FsLayout& L, double dx2);
// exact: the Euclidean gradient norm in every Godunov case (the reference's
// misplaced parenthesis of ADVECTION_METHOD_3D.h:1605 otherwise)
// One sweep from the ghost copy `ghost` (its box holds phi) into out's box;
// out may be phi or another field, not ghost.
void fs_reinit_sweep(hipStream_t st, double* out, const double* ghost, const double* sgn, const FsLayout& L,
                     const FsGridConst& cell, double eps, double dt, int exact, bool fast = false);

// ---- misc -----------------------------------------------------------------
void fs_average(hipStream_t st, double* phi, const double* tmp, const FsLayout& L);  // phi = 0.5*(tmp+phi)
// dst = ca*a + cb*dst on box b
void fs_combine(hipStream_t st, double* dst, const double* a, double ca, double cb, const FsLayout& L, FsBox b);
// box interior <- / -> an x-fastest array over the box grown by g ghost
// layers (g <= FS_G; gather includes the layers, scatter skips them)
void fs_scatter(hipStream_t st, double* dst, const double* lin, const FsLayout& L, FsBox b, int g = 0);
void fs_gather(hipStream_t st, double* lin, const double* src, const FsLayout& L, FsBox b, int g = 0);
void fs_fill_value(hipStream_t st, double* dst, long n, double v);
