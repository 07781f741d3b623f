/*
 * fs_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, FP64) of the reference's per-timestep hot path
 * (EULERIAN_FLUID_SOLVER_3D::AdvanceOneTimeStepThread and everything under
 * it, air/water mode).  It is the checker for the HIP library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path never calls it.
 *
 * Entry points mirror include/fs_abi.h with the prefix fso_ and the same
 * argument meaning, so a parity test calls both sides the same way.
 */
#ifndef FS_ORACLE_H
#define FS_ORACLE_H
#include "../include/fs_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fso_state fso_state;

int  fso_create(const fs_params* params, const fs_grid* grid, fso_state** out);
void fso_destroy(fso_state* s);
int  fso_set_fields(fso_state* s, const double* phi, const double* u, const double* v,
                    const double* w, const double* p);
int  fso_get_fields(fso_state* s, double* phi, double* u, double* v, double* w, double* p);
/* LEVELSET_3D::Write / Read records (LEVELSET_3D.h:1132-1200); a checkpoint
 * is the records of phi, u, v, w, p back to back */
int  fso_write_levelset(fso_state* s, const char* path);
int  fso_read_levelset(fso_state* s, const char* path);
int  fso_write_checkpoint(fso_state* s, const char* path);
int  fso_read_checkpoint(fso_state* s, const char* path);

/* 1 = explicit viscosity updated in place in lexicographic order, as the
 * reference's single-thread sweep does (VISCOSITY_3D.h:483-559);
 * 0 = out-of-place (every face reads the pre-sweep component), the order the
 * HIP kernels use.  Default 1. */
void fso_set_viscosity_inplace(fso_state* s, int inplace);
/* Worker threads for the stencil loops (results do not depend on it). */
void fso_set_threads(int n);

int  fso_advance_frame(fso_state* s, double dt, fs_step_stats* st);
int  fso_step(fso_state* s, double dt, fs_step_stats* st);
int  fso_cfl_dt(fso_state* s, double cfl, int gravity_norm, double* dt);

int  fso_set_levelset_sources(fso_state* s, int n, const fs_levelset_source* src);
int  fso_set_frame(fso_state* s, int frame);
int  fso_apply_levelset_sources(fso_state* s);
int  fso_compute_curvatures(fso_state* s);
int  fso_get_curvature(fso_state* s, double* kappa, double* abs_max_curvature);
int  fso_levelset_advection(fso_state* s, double dt);
int  fso_velocity_advection(fso_state* s, double dt);
int  fso_apply_gravity(fso_state* s, double dt);
int  fso_sourcing(fso_state* s, double dt);
int  fso_projection(fso_state* s, double dt, fs_step_stats* st);
int  fso_reinitialization(fso_state* s, double dt);

int  fso_assemble_pressure_system(fso_state* s, double dt, double* diag, double* bx,
                                  double* by, double* bz, double* rhs);
int  fso_apply_pressure_operator(fso_state* s, const double* x, double* q);

/* Sparse view of the last assembly, for tests that check the operator:
 * n rows, nnz entries, CSR arrays copied into caller buffers when non-NULL. */
int  fso_csr_info(fso_state* s, int* n, int* nnz);
int  fso_csr_copy(fso_state* s, int* row_ptr, int* col, double* val, double* b);

/* Raw 1-D kernels for closed-form pins. */
/* WENO5-JS one-sided derivatives from 7 samples v[0..6] centred on v[3]
 * (ADVECTION_METHOD_3D.h:55-183): writes d_minus, d_plus. */
void fso_weno5_1d(const double* v, double one_over_dx, double epsilon,
                  double* d_minus, double* d_plus);
/* ENO3 one-sided derivatives from v[0..6] centred on v[3]
 * (ADVECTION_METHOD_3D.h:1653-1718). */
void fso_eno3_1d(const double* v, double dx, double one_over_dx, double one_over_2dx,
                 double one_over_3dx, double* d_minus, double* d_plus);

/* Test hooks for closed-form pins. */
/* One axis of the Godunov case tree (ADVECTION_METHOD_3D.h:908-1622): the
 * selected one-sided derivative in *val; returns the case (1 upwind m, 2
 * upwind p, 3 expansion, 4/5 shock taking m/p, 0 none). */
int  fso_godunov_axis(double sg, double m, double p, double* val);
/* |grad phi| - 1 from the three axes' one-sided derivatives (exact = 1: the
 * Euclidean norm in every case; 0: the reference's two leaves as written);
 * returns 0 when no case is met. */
int  fso_godunov_norm(double sg, double mx, double px, double my, double py, double mz, double pz,
                      int exact, double* g);
/* The projection's system div(beta grad p) = source with the jump
 * p(phi > 0) - p(phi <= 0) = jump (NULL: none) at the current level set,
 * solved from zero; p receives the solution. */
int  fso_solve_poisson(fso_state* s, const double* source, const double* jump, double* p, fs_step_stats* st);
/* The persistent ghost copies over their boxes grown by 3 (NULL skips). */
int  fso_get_ghost_copies(fso_state* s, double* sghost, double* gu, double* gv, double* gw);

/* ---- oil/water pipe mode (fs_oracle_pipe.c) ---------------------------------
 * The same entry points as the HIP library's fs_create_pipe handle
 * (fs_pipe_* in include/fs_abi.h), prefix fsop_. */
typedef struct fsop_state fsop_state;
int  fsop_create(const fs_params* params, const fs_pipe_params* pipe, const fs_grid* grid, fsop_state** out);
void fsop_destroy(fsop_state* s);
int  fsop_constants(fsop_state* s, double* out);
int  fsop_initial_fields(fsop_state* s);
int  fsop_set_fields(fsop_state* s, const double* phi, const double* u, const double* v, const double* w,
                     const double* p);
int  fsop_get_fields(fsop_state* s, double* phi, double* u, double* v, double* w, double* p);
int  fsop_phase(fsop_state* s, int phase, double dt, fs_step_stats* st);
int  fsop_step(fsop_state* s, double dt, fs_step_stats* st);
int  fsop_advance_frame(fsop_state* s, double dt, fs_step_stats* st);
int  fsop_compute_curvatures(fsop_state* s);
int  fsop_cfl_dt(fsop_state* s, double cfl, double* dt);
int  fsop_get_array(fsop_state* s, int which, double* dst);
int  fsop_levelset_advection(fsop_state* s, double dt);
int  fsop_velocity_advection(fsop_state* s, double dt);
int  fsop_apply_gravity(fsop_state* s, double dt);
int  fsop_sourcing(fsop_state* s, double dt, fs_step_stats* st);
int  fsop_projection(fsop_state* s, double dt, fs_step_stats* st);
int  fsop_reinitialization(fsop_state* s, double dt);
/* the pipe mode's sin polynomial (|x| <= pi) */
double fsop_sin(double x);

#ifdef __cplusplus
}
#endif
#endif
