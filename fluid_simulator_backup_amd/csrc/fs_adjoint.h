// fs_adjoint.h -- launchers of the reverse-mode kernels of the fixed-schedule
// step (host-callable, internal).  Adjoint fields use the padded layout of
// the forward fields; accumulators may receive contributions at ghost
// positions, which a fold then maps onto the interior values those ghosts
// were made from (and clears).
#pragma once
#include <hip/hip_runtime.h>

#include "fs_device.h"
#include "fs_kernels.h"
#include "fs_layout.h"

// Transpose of the level set's ghost refresh (copy + linear layers (g+1) f0
// - g f1 on the six faces, depth d): face-ghost values of a are added onto
// the two interior values each was extrapolated from, then cleared.
void fs_adj_fold_linear(hipStream_t st, double* a, const FsLayout& L, int d);
// Transpose of the constant extrapolation of cell-centred ghosts (faces,
// edges, vertices, depth d): every ghost value is added onto its clamped
// interior cell, then cleared.
void fs_adj_fold_const(hipStream_t st, double* a, const FsLayout& L, int d);
// a = 0 outside box b (the transpose of zero ghost faces: contributions to
// ghost positions belong to constants).
void fs_zero_outside(hipStream_t st, double* a, const FsLayout& L, FsBox b);

// Adjoint of one WENO5 upwind stage out = rho - dt (u X + v Y + w Z) of
// component comp (0 level set, 1..3 MAC), derivatives from ghost copy gh and
// sample velocity from gu, gv, gw: agh += (d out/d gh)^T lam and
// au, av, aw += (d out/d (gu, gv, gw))^T lam (atomics; ghost positions
// included).  The base term (d out/d rho = identity) is not added.
void fs_adj_weno5_stage(hipStream_t st, const double* lam, const double* gh, const double* gu, const double* gv,
                        const double* gw, double* agh, double* au, double* av, double* aw, const FsLayout& L,
                        int comp, const FsGridConst& gc, double eps, double dt);

// Level-set part of the adjoint of one explicit viscosity component update
// (k_viscosity): aphi += (d upd/d phi)^T lam at the velocities U, V, W that
// update read (atomics; phi's ghost positions included).
void fs_adj_viscosity_phi(hipStream_t st, int comp, const double* lam, const double* U, const double* V,
                          const double* W, const double* phi, double* aphi, const FsLayout& L, const FsMaterial& m,
                          const FsGridConst& cell, double dt);

// Level-set part of the adjoint of the fixed-schedule projection through the
// GFM face coefficients: aphi += sum over faces of d beta/d phi times
// [-dt od lam_f (p_c - p_n) - y^T (d A/d beta_f) p], with p the solved
// pressure and y = A^-1 Gh^T lam (interior cells, gather form).
void fs_adj_projection_phi(hipStream_t st, const double* phi, const double* irho, const double* p, const double* y,
                           const double* lu, const double* lv, const double* lw, double* aphi, const FsLayout& L,
                           const FsGridConst& cell, double dt);

// Adjoint of one Sussman sweep (k_reinit with the Euclidean norm) at the
// sweep's input phi (gh its fresh ghost copy): out += (d phi'/d phi)^T lam
// (atomics, ghost positions included) and lsg += (d phi'/d S)^T lam.
void fs_adj_reinit_sweep(hipStream_t st, const double* lam, const double* phi, const double* gh, const double* sgn,
                         double* out, double* lsg, const FsLayout& L, const FsGridConst& cell, double eps, double dt);
// lphi += lsg * dS/dphi with S = phi / sqrt(phi^2 + dx2)
void fs_adj_sign(hipStream_t st, const double* phi, const double* lsg, double* lphi, const FsLayout& L, double dx2);

// Sum over cells of (H_eps(phi) - H_eps(target))^2, H the projection's
// smoothed Heaviside; lphi (optional) = its gradient.  part: FS_LOSS_BLOCKS
// + 1 doubles of scratch; the loss lands in part[FS_LOSS_BLOCKS].
#define FS_LOSS_BLOCKS 1024
void fs_levelset_loss(hipStream_t st, const double* phi, const double* target, double* lphi, double* part,
                      const FsLayout& L, double eps);
