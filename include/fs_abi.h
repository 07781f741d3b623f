/*
 * fs_abi.h -- C-ABI of the MI355X-native timestep hot path of a two-phase
 * level-set Navier-Stokes solver on a uniform MAC grid.
 *
 * The library (libfs_hip.so) replaces the per-timestep work of the
 * reference class EULERIAN_FLUID_SOLVER_3D (EULERIAN_FLUID_SOLVER_3D.h) with
 * hand-written HIP kernels for gfx950.  Every entry point below names the
 * reference function it stands in for.  Nothing here uses C++ or torch types:
 * plain structs, pointers and sizes only.
 *
 * Conventions
 *   - every function returns 0 (FS_OK) or a negative FS_ERR_* code;
 *     fs_last_error() gives a message for the last failure on a handle;
 *   - one host thread per handle; calls on one handle are not re-entrant;
 *   - all GPU work of a handle runs on one HIP stream owned by the handle and
 *     every call returns after that work has finished;
 *   - field buffers exchanged with the caller are interior-only and x-fastest:
 *       cell field (phi, p)   nx*ny*nz        index i + nx*(j + ny*k)
 *       x-faces (u)           (nx+1)*ny*nz    index i + (nx+1)*(j + ny*k)
 *       y-faces (v)           nx*(ny+1)*nz    index i + nx*(j + (ny+1)*k)
 *       z-faces (w)           nx*ny*(nz+1)    index i + nx*(j + ny*k)
 *     which is the k-j-i order of ARRAY_3D (ARRAY_3D.h:98) and of the
 *     level-set dump format (LEVELSET_3D.h:1132-1200).
 */
#ifndef FS_ABI_H
#define FS_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define FS_OK                 0
#define FS_ERR_BAD_ARG       -1   /* invalid argument (null, dt <= 0, bad size) */
#define FS_ERR_OOM           -2   /* device allocation failed or over the size cap */
#define FS_ERR_HIP           -3   /* a HIP runtime call failed */
#define FS_ERR_NOT_CONVERGED -4   /* PCG hit max_iteration; fields are still updated */
#define FS_ERR_NONFINITE     -5   /* NaN/Inf found in the velocity or level set */
#define FS_ERR_UNSUPPORTED   -6   /* option outside the implemented hot path */

/* ---- option values -------------------------------------------------------- */
/* Pressure solve precision. */
#define FS_FP64          0   /* Jacobi-PCG in FP64 (PCG_METHOD.cpp:4-126) */
#define FS_FP32_REFINE   1   /* FP32 Jacobi-PCG inside FP64 iterative refinement */
#define FS_FP32_MG       2   /* FP32 CG preconditioned by one geometric-multigrid V-cycle, inside FP64
                                iterative refinement.  No counterpart in the reference: MACROS.h:3
                                declares MULTIGRID and PROJECTION_3D.h:55 a multigrid smoother, but its
                                script parser never selects it (PROJECTION_3D.h:207-226).  Both FP32
                                paths stop on rel_tolerance, whatever stop_rule says. */

/* PCG stop rule. */
#define FS_STOP_ABS_PRECOND  0   /* r^T D^-1 r < tol^2, the reference rule (PCG_METHOD.cpp:85-90) */
#define FS_STOP_REL_RESIDUAL 1   /* ||b - A x||_2 / ||b||_2 < rel_tolerance */

/* Step schedule. */
#define FS_SCHEDULE_COMPAT   0   /* reference call order, ghost refresh points and quirks */
#define FS_SCHEDULE_FIXED    1   /* same kernels, quirks corrected: ghost copies refreshed
                                     before use, every face corrected with the
                                     matrix's coefficients, walls zeroed after
                                     forces, Euclidean Godunov norm (piecewise
                                     smooth: fs_adjoint_step differentiates it) */

/* ---- plain-data structs --------------------------------------------------- */

/* Grid of cell centres.  Spacing follows GRID_STRUCTURE_3D::Initialize
 * (GRID_STRUCTURE_3D.h:91-137): dx = (x_max - x_min)/(nx - 1), with
 * x_min/x_max the first/last cell centre. */
typedef struct fs_grid {
    int    nx, ny, nz;
    double x_min, y_min, z_min;
    double x_max, y_max, z_max;
} fs_grid;

/* The script keys of the hot path (EULERIAN_FLUID_SOLVER_3D.h:192-260,
 * ADVECTION_3D.h:102-113, PROJECTION_3D.h:178-227), air/water mode. */
/* fs_params.curvature_method */
#define FS_CURV_LEVELSET 0
#define FS_CURV_NORMALS  1
#define FS_CURV_BOTH     2
#define FS_CURV_NONE     3

typedef struct fs_params {
    /* phase labels: 1 = air_bubble_rising (phi <= 0 is the light phase),
     * 0 = water_drop (phi <= 0 is water); PROJECTION_3D.h:605-639 */
    int    air_bubble_rising;
    double water_density, air_density;
    double water_viscosity, air_viscosity;
    double gravity[3];
    double surface_tension;                 /* sigma of the jump p_water - p_air = sigma * curvature */
    int    use_delta_function_formulation;  /* explicit viscosity on/off */
    int    order_for_time_advancing;        /* 1 Euler, 2 RK2, 3 TVD-RK3 (EULERIAN_FLUID_SOLVER_3D.h:806-1109) */
    int    sussmanstyle_reinitialization;
    int    scaling_number_for_reinitialization;   /* fictitious dt = dx/scaling */
    int    iteration_number_for_reinitialization; /* sweeps per call */
    double levelset_epsilon;   /* LEVELSET_ADVECTION epsilon: WENO5 eps and reinit clamp */
    double velocity_epsilon;   /* VELOCITY_ADVECTION epsilon (WENO5 velocity) */
    int    use_5th_weno_v;     /* 1: WENO5 on MAC velocity, 0: ENO3 (use_3rd_eno_v) */
    /* PROJECTION_WATER */
    double tolerance;          /* absolute tolerance of the reference stop rule */
    int    max_iteration;
    int    use_previous_pressure;
    int    stop_rule;          /* FS_STOP_* */
    double rel_tolerance;      /* for FS_STOP_REL_RESIDUAL and the FP32 precisions */
    int    precision;          /* FS_FP64 | FS_FP32_REFINE | FS_FP32_MG */
    int    schedule;           /* FS_SCHEDULE_* */
    /* debugging switches, EULERIAN_FLUID_SOLVER_3D.h:103-107 */
    int    is_levelset_advection_active;
    int    is_velocity_advection_active;
    int    is_external_force_active;
    int    is_sourcing_active;
    int    is_projection_active;
    /* VISCOSITY_3D semi_implicit_approach (VISCOSITY_3D.h:227, :853-989):
     * per MAC component solve (I + sum_f c_f/d^2) u = u* + cross terms with
     * Dirichlet walls, Jacobi-PCG with the PROJECTION_WATER tolerance,
     * max_iteration and stop rule; the CFL drops its viscous term
     * (EULERIAN_FLUID_SOLVER_3D.h:2770-2773).  Needs
     * use_delta_function_formulation like the explicit path. */
    int    semi_implicit_viscosity;
    /* No counterpart in the reference: 1 takes the WENO5 one-sided
     * derivatives of the level-set stages, the WENO5 velocity stages and the
     * reinitialization sweeps in fewer operations (fused multiply-adds, the
     * weights in an algebraically equal product form, one division per cell
     * and axis), FP64 throughout; the results then differ from the
     * reference's operation order in the last bits of each derivative.
     * 0 keeps that order, bit for bit. */
    int    fast_weno;
    /* CSF_model (EULERIAN_FLUID_SOLVER_3D.h:210): Sourcing's
     * ApplySurfaceTension (:2089-2093, :2102-2196) computes the curvature
     * (with its ghost refresh of phi) and sets the projection's CSF flag,
     * which zeroes the jump condition from then on (PROJECTION_3D.h:754-757);
     * its force is applied only in oil/water mode (:2106), so in this
     * air/water path the model removes the surface-tension jump and adds no
     * force, as the reference does.  The CFL's surface-tension term stays. */
    int    csf_model;
    /* How LEVELSET_3D::ComputeCurvatures forms the curvature
     * (curvature_by_normal_vector / curvature_by_levelset,
     * EULERIAN_FLUID_SOLVER_3D.h:758-759, LEVELSET_3D.h:963-1082):
     * FS_CURV_LEVELSET (0, the shipped scripts' choice) from the level set's
     * derivatives; FS_CURV_NORMALS the divergence of the unit normals, clamped
     * to -1/min(dx,dy,dz) where |kappa| exceeds 1/min(dx,dy,dz) (its maximum
     * over all cells, unclamped, feeds the CFL); FS_CURV_BOTH the normals'
     * pass, then the level set's, which overwrites kappa (the maximum is over
     * both); FS_CURV_NONE neither (kappa and the jump stay zero).  In the
     * compatible schedule the normals are the ones the last projection's
     * ComputeNormals left (PROJECTION_3D.h:803; zero before the first); the
     * fixed schedule forms them from the current level set first. */
    int    curvature_method;
} fs_params;

/* Per-step (or per-phase) statistics. */
typedef struct fs_step_stats {
    int    pcg_iterations;       /* inner PCG iterations, summed over the step's projections */
    int    refinement_sweeps;    /* outer FP64 sweeps (FS_FP32_REFINE, FS_FP32_MG), else 0 */
    double residual_precond;     /* sqrt(r^T D^-1 r) at exit, PCG_METHOD.cpp:111 */
    double residual_rel;         /* ||b - A x||_2 / ||b||_2 at exit (recomputed in FP64) */
    double max_velocity[3];      /* max |u|,|v|,|w| after the gradient update, PROJECTION_3D.h:1236,1345,1427 */
    double projection_ms;        /* device time of the last projection */
    double step_ms;              /* device time of the last fs_advance_frame/fs_step */
    int    viscosity_iterations; /* PCG iterations of the semi-implicit viscosity solves */
} fs_step_stats;

typedef struct fs_solver fs_solver;

/* ---- lifecycle ------------------------------------------------------------ */

/* Replaces EULERIAN_FLUID_SOLVER_3D::InitializeFromScriptBlock
 * (EULERIAN_FLUID_SOLVER_3D.h:192-800): allocates the MAC fields, ghost
 * fields and level set on device.  device < 0 keeps the current device. */
int  fs_create(const fs_params* params, const fs_grid* grid, int device, fs_solver** out);
void fs_destroy(fs_solver* s);
const char* fs_last_error(const fs_solver* s);

/* Copy host fields in.  phi is the initial level set written by
 * SIMULATION_WORLD.cpp:116-129; u,v,w the MAC velocity
 * (EULERIAN_FLUID_SOLVER_3D.h:23-25).  Ghost cells are set as at the end of
 * InitializeFromScriptBlock (level-set ghosts by constant extrapolation,
 * EULERIAN_FLUID_SOLVER_3D.h:757; every other ghost and p zero).
 * Any pointer may be NULL to leave that field as it is. */
int  fs_set_fields(fs_solver* s, const double* phi, const double* u,
                   const double* v, const double* w, const double* p);
/* Copy fields out (any pointer may be NULL). */
int  fs_get_fields(fs_solver* s, double* phi, double* u, double* v, double* w, double* p);

/* fs_set_fields / fs_get_fields with a choice of layout and memory.
 * FS_LAYOUT_INTERIOR is the interior-only layout above.  FS_LAYOUT_GHOSTED
 * is the reference's own arrays, ARRAY_3D over the box grown by its ghost
 * layers, x fastest: 2 for phi and u, v, w (EULERIAN_FLUID_SOLVER_3D.h:
 * 407-459), 1 for p (PROJECTION_3D.h:362); on input the ghost values are
 * ignored (phi's are extrapolated as for fs_set_fields), on output they are
 * what the solver's ghost cells hold.  FS_MEM_DEVICE: the pointers are
 * device memory of the handle's GPU (a caller that keeps its state on the
 * GPU), copied with kernels on the handle's stream; FS_MEM_HOST: host
 * memory, as fs_set_fields. */
#define FS_LAYOUT_INTERIOR 0
#define FS_LAYOUT_GHOSTED  1
#define FS_MEM_HOST        0
#define FS_MEM_DEVICE      1
int  fs_set_fields_ex(fs_solver* s, const double* phi, const double* u, const double* v, const double* w,
                      const double* p, int layout, int memory);
int  fs_get_fields_ex(fs_solver* s, double* phi, double* u, double* v, double* w, double* p, int layout,
                      int memory);

/* Asynchronous steps (no counterpart in the reference, whose steps block
 * on their thread barriers).  With async set, fs_step and fs_advance_frame
 * return once the step's last host decision is taken (the projection's stop
 * tests still wait for the device), leaving the remaining kernels -- the
 * reinitialization, the last blends -- queued on the handle's stream, and
 * report step_ms as 0.  fs_synchronize waits for the stream; fs_get_stream
 * gives it (a hipStream_t) so a caller can order its own work after the
 * step.  Every other call still waits before it returns. */
int  fs_set_async(fs_solver* s, int async);
int  fs_synchronize(fs_solver* s);
int  fs_get_stream(fs_solver* s, void** stream);

/* ---- files ------------------------------------------------------------------
 * fs_write_levelset / fs_read_levelset: LEVELSET_3D::Write / Read
 * (LEVELSET_3D.h:1132-1200), the reference's one binary field format.  A
 * record is int i_start, j_start, k_start, i_res, j_res, k_res (starts 0);
 * double x_min, y_min, z_min, x_max, y_max, z_max (first and last sample
 * positions); int ghost_width; then i_res*j_res*k_res doubles, i fastest, in
 * native byte order.  Reading checks that the record's grid is the solver's
 * (FS_ERR_BAD_ARG otherwise) and sets phi as fs_set_fields does.
 * fs_write_checkpoint / fs_read_checkpoint: the state (phi, u, v, w, p) as
 * five such records back to back; u, v, w on their MAC grids (faces, the
 * cell extents widened by dx/2 along their axis,
 * EULERIAN_FLUID_SOLVER_3D.h:405-431), ghost widths 2, 2, 2, 2, 1.  Its first
 * record alone is a level-set file the reference can read.  The reference
 * writes only level sets (FIELD_STRUCTURE_3D::Write is declared, not
 * defined); the checkpoint extends its layout to the whole state. */
int  fs_write_levelset(fs_solver* s, const char* path);
int  fs_read_levelset(fs_solver* s, const char* path);
int  fs_write_checkpoint(fs_solver* s, const char* path);
int  fs_read_checkpoint(fs_solver* s, const char* path);

/* ---- the timestep -------------------------------------------------------- */

/* One frame as SIMULATION_WORLD::AdvanceOneFrame runs it
 * (SIMULATION_WORLD.h:320-330): the level-set ghost refresh of
 * LEVELSET_3D::ComputeCurvatures (LEVELSET_3D.h:998) followed by
 * EULERIAN_FLUID_SOLVER_3D::AdvanceOneTimeStepThread(dt)
 * (EULERIAN_FLUID_SOLVER_3D.h:803-1111).  stats may be NULL. */
int  fs_advance_frame(fs_solver* s, double dt, fs_step_stats* stats);

/* EULERIAN_FLUID_SOLVER_3D::AdvanceOneTimeStepThread alone. */
int  fs_step(fs_solver* s, double dt, fs_step_stats* stats);

/* CFL time step: EULERIAN_FLUID_SOLVER_3D::CFLOneTimeStep()
 * (EULERIAN_FLUID_SOLVER_3D.h:2673-2785) times cfl.  gravity_norm = 0 uses
 * |g_y| as the reference does (:2717); 1 uses |g|.  Returns in *dt. */
int  fs_cfl_dt(fs_solver* s, double cfl, int gravity_norm, double* dt);

/* fs_cfl_dt without a handle and without any allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * u, v, w are device pointers to interior MAC fields in the layout above;
 * dt is a device pointer that receives cfl / rate (NaN when a velocity is
 * not finite), the value fs_cfl_dt gives for the same fields.  The work is
 * enqueued on `stream` (a hipStream_t; NULL is the null stream) and the call
 * returns without waiting.  Partial maxima live in a fixed device array of
 * 3 x 2048 doubles per GPU, so each velocity component may have at most
 * 2048 * 2^17 faces (FS_ERR_OOM beyond); surface tension is not supported
 * (its term needs the curvature: FS_ERR_UNSUPPORTED).  Calls that overlap
 * on one GPU share that array and race. */
int  fs_cfl_dt_noalloc(const fs_params* params, const fs_grid* grid, const double* u, const double* v,
                       const double* w, double cfl, int gravity_norm, double* dt, void* stream);

/* ---- phases (EULERIAN_FLUID_SOLVER_3D.h:1590-1606, 2076-2237) ------------ */
int  fs_compute_curvatures(fs_solver* s);           /* LEVELSET_3D::ComputeCurvatures, LEVELSET_3D.h:963-1082 */
int  fs_levelset_advection(fs_solver* s, double dt); /* ADVECTION_3D::Solve_Levelset, ADVECTION_3D.h:204-242 */
int  fs_velocity_advection(fs_solver* s, double dt); /* ADVECTION_3D::Solve_Velocity, ADVECTION_3D.h:244-358 */
int  fs_apply_gravity(fs_solver* s, double dt);      /* ApplyGravity, EULERIAN_FLUID_SOLVER_3D.h:2237-2270 */
int  fs_sourcing(fs_solver* s, double dt);           /* Sourcing -> VISCOSITY_3D::ApplyViscosity, VISCOSITY_3D.h:448-563 */
int  fs_projection(fs_solver* s, double dt, fs_step_stats* stats); /* PROJECTION_3D::Solve, PROJECTION_3D.h:416-594 */
int  fs_reinitialization(fs_solver* s, double dt);   /* Reinitialization, EULERIAN_FLUID_SOLVER_3D.h:2198-2225 */

/* ---- level-set sources (objects of the scene) -------------------------------
 * EULERIAN_FLUID_SOLVER_3D::SourcingFromObjectList
 * (EULERIAN_FLUID_SOLVER_3D.h:2592-2671) for spherical objects
 * (SIMULATION_WORLD.cpp:355-364, SPHERE.h): at the end of the Sourcing phase
 * (:2076-2087), and whenever fs_apply_levelset_sources is called, each source
 * whose frame window holds the current frame visits the cells of its box
 * (the grid cells from the one holding center - half_width to the one holding
 * center + half_width per axis, clamped to the grid,
 * SIMULATION_OBJECT.h:149-160) and, where its signed distance
 * d = |c - center| - radius at the point c = (x_min + (i + 1/2) dx, ...)
 * (GRID_STRUCTURE_3D::CellCenter, GRID_STRUCTURE_3D.h:269-272) is at most
 * 3 dx, sets phi = min(d, phi) (mode FS_SOURCE_ADD, FLUID_CONTROL_OPTIONS
 * source_levelset) or phi = max(-d, phi) (FS_SOURCE_DELETE,
 * delete_water_levelset).  Sources act in the order given.  The velocity
 * sourcing of the reference is commented out (:2644-2657) and none is done. */
#define FS_SOURCE_NONE    0
#define FS_SOURCE_ADD     1
#define FS_SOURCE_DELETE  2
typedef struct fs_levelset_source {
    double center[3];
    double radius;
    double half_width;      /* box half width: radius + 3 dx of the world grid */
    int    mode;            /* FS_SOURCE_* */
    int    start_frame, stop_frame;  /* inclusive window (source_start_frame, source_stop_frame) */
} fs_levelset_source;
/* Replaces the solver's object list (n = 0 clears it; at most 64). */
int  fs_set_levelset_sources(fs_solver* s, int n, const fs_levelset_source* src);
/* The frame number the sources compare with (the reference's borrowed
 * frame counter, EULERIAN_FLUID_SOLVER_3D.h:154-162); 0 after fs_create. */
int  fs_set_frame(fs_solver* s, int frame);
/* SourcingFromObjectList once, outside a step (SIMULATION_WORLD.h:224-227). */
int  fs_apply_levelset_sources(fs_solver* s);

/* ---- surface tension ------------------------------------------------------
 * With surface_tension != 0 each projection computes the level set's
 * curvature (LEVELSET_3D::ComputeCurvatures, curvature_by_levelset,
 * LEVELSET_3D.h:998-1080) and, with use_delta_function_formulation, the jump
 * condition jc = surface_tension * curvature
 * (PROJECTION_3D::DetermineJumpConditionField, PROJECTION_3D.h:746-769); the
 * pressure right-hand side gains the jump terms of
 * POISSON_SOLVER_3D.h:1215-1328 and the gradient update the jump of
 * PROJECTION_3D.h:1178-1230, :1287-1339, :1396-1423.  fs_cfl_dt adds the
 * surface-tension rate s_f (EULERIAN_FLUID_SOLVER_3D.h:2740-2744) from the
 * curvature of the current level set.  A cell whose gradient vanishes in
 * central and one-sided differences, where the reference stops the program
 * (LEVELSET_3D.h:1040-1044), returns FS_ERR_NONFINITE.
 *
 * fs_get_curvature: ComputeCurvatures on the current level set (its ghost
 * refresh included) and the curvature out: kappa nx*ny*nz (may be NULL) and
 * abs_max_curvature, the maximum |curvature| over cells with |phi| < 0.1
 * (may be NULL).  FS_ERR_UNSUPPORTED when surface_tension is 0. */
int  fs_get_curvature(fs_solver* s, double* kappa, double* abs_max_curvature);

/* ---- inspection of the pressure system (per-kernel parity) ------------- */

/* Assemble the projection's linear system for the current fields without
 * solving it (PROJECTION_3D.h:596-833 up to POISSON_SOLVER_3D::Solve's
 * mean removal, POISSON_SOLVER_3D.h:225-246).  Outputs (host, may be NULL):
 *   diag  nx*ny*nz   A(c,c) of each cell          (POISSON_SOLVER_3D.h:1213)
 *   bx    (nx+1)*ny*nz  beta of x-faces between two full cells, else 0
 *   by, bz likewise for y- and z-faces            (POISSON_SOLVER_3D.h:1080-1085)
 *   rhs   nx*ny*nz   b after the Neumann mean removal */
int  fs_assemble_pressure_system(fs_solver* s, double dt, double* diag, double* bx,
                                 double* by, double* bz, double* rhs);

/* q = A x with the matrix-free operator of the last assembly
 * (CSR_MATRIX::Multiply, CSR_MATRIX.h:236-260).  x, q: nx*ny*nz host. */
int  fs_apply_pressure_operator(fs_solver* s, const double* x, double* q);

/* ---- measurement ---------------------------------------------------------- */

/* Times the PCG kernels of one iteration on the operator of the last
 * assembly: [0] the p update fused with SpMV + p.q, [1] the x/r/z update with
 * r.z and r.r; [2] is 0 (no separate p-update kernel).  `reps` back-to-back
 * launches each, timed with HIP events on the handle's stream.  precision:
 * FS_FP64 or FS_FP32_REFINE (needs a handle created with it).  avg_ms[0..2]
 * receives the mean duration per launch.  No reference
 * counterpart: it replaces the wall-clock LOG timers around PCGMethod
 * (PCG_METHOD.cpp:6, LOG.h:191-338) for roofline accounting. */
int  fs_time_pcg_kernels(fs_solver* s, int precision, int reps, double* avg_ms);

/* In-loop kernel sampling.  While enabled, ordinary solves bracket the PCG
 * kernels of one iteration per launch batch (16 iterations) with HIP events
 * on the handle's stream; fs_get_kernel_samples returns, per kernel (as in
 * fs_time_pcg_kernels) of the FP64 or the FP32 inner iteration, the mean
 * duration and the sample count since sampling was last enabled (enabling
 * clears them).  No reference counterpart: measurement hooks beside
 * fs_time_pcg_kernels. */
int  fs_set_kernel_sampling(fs_solver* s, int enable);
int  fs_get_kernel_samples(fs_solver* s, int precision, double* avg_ms, int* count);

/* Frame kernels sampled while fs_set_kernel_sampling is on: FS_KS_REINIT
 * times the first sweep of each reinitialization call (k_reinit) and counts
 * every sweep; FS_KS_MG_SMOOTH times the first finest-level smoother launch
 * of each V-cycle of the FS_FP32_MG solve (k_mg_rb_sweep, one red-black sweep)
 * and counts every finest-level smoother launch.  Returns the mean duration (ms), the number
 * of timed launches and the number of launches since sampling was enabled.
 * No reference counterpart: measurement hooks beside fs_time_pcg_kernels. */
#define FS_KS_REINIT 0
#define FS_KS_MG_SMOOTH 1
#define FS_KS_COUNT 2
int  fs_get_frame_kernel_samples(fs_solver* s, int kernel, double* avg_ms, long long* samples,
                                 long long* launches);

/* ---- adjoints (no reference counterpart) ---------------------------------
 * Reverse-mode pieces of the timestep for gradients through the GPU path.
 *
 * fs_adjoint_projection: the projection (fs_projection) at the current
 * level set is linear in the velocity, u' = (I - G A^-1 M B) u with
 * B u = -div(u)/dt, M the Neumann mean removal, A the GFM matrix and G the
 * gradient update (z faces, or all faces in FS_SCHEDULE_FIXED).  Given the
 * adjoint lam of u', v', w' (host face fields, interior only, in place) it
 * returns lam - B^T M A^-1 G^T lam: one more solve with the same symmetric
 * operator, FP64, with the stop rule of the forward solve.  The level set is
 * held fixed (its influence on the coefficients is not differentiated);
 * nothing of the forward state is changed. */
int  fs_adjoint_projection(fs_solver* s, double dt, double* lu, double* lv, double* lw, fs_step_stats* stats);

/* fs_adjoint_sourcing: the explicit viscous update of fs_sourcing (x, then
 * y reading the new u, then z reading the new u and v) is linear in the
 * velocity at a fixed level set; this applies its transpose (z, y, x) to the
 * adjoint lam of u, v, w (host face fields, in place).  FS_ERR_UNSUPPORTED
 * with semi_implicit_viscosity. */
int  fs_adjoint_sourcing(fs_solver* s, double dt, double* lu, double* lv, double* lw);

/* fs_weno5_transport: one level-set stage of ADVECTION_3D::Solve_Levelset
 * (ADVECTION_3D.h:204-242, ADVECTION_METHOD_3D.h:42-279) on explicit inputs,
 * with the level set's ghosts its constant extrapolation and the MAC ghosts
 * zero: out = phi - dt (u X + v Y + w Z), X the upwind WENO5 derivative.
 * fs_adjoint_weno5_transport: its vector-Jacobian product at (phi, u, v,
 * w): given lam (cell field) it returns lphi = (d out/d phi)^T lam, with the
 * derivatives of the WENO5 weights and smoothness indicators, and lu, lv, lw
 * = (d out/d velocity)^T lam; the upwind choice is piecewise constant and
 * not differentiated.  Host buffers, interior only. */
int  fs_weno5_transport(fs_solver* s, double dt, const double* phi, const double* u, const double* v,
                        const double* w, double* out);
int  fs_adjoint_weno5_transport(fs_solver* s, double dt, const double* phi, const double* u, const double* v,
                                const double* w, const double* lam, double* lphi, double* lu, double* lv, double* lw);

/* ---- adjoint time step -------------------------------------------------------
 * No reference counterpart (the reference has no adjoint): reverse-mode
 * derivatives of EULERIAN_FLUID_SOLVER_3D::AdvanceOneTimeStepThread
 * (EULERIAN_FLUID_SOLVER_3D.h:803-1111) as fs_step computes it with
 * schedule FS_SCHEDULE_FIXED, which is piecewise smooth: every ghost copy is
 * a function of the current fields, the pressure gradient uses the matrix's
 * face coefficients, the Godunov norm is Euclidean.  Every phase is
 * differentiated, the level set's influence on the GFM coefficients, the
 * smoothed densities and viscosities, and the reinitialization's sign
 * function included; the discrete choices (upwind side, Godunov case, +-eps
 * clamp, a cell's phase) are held at the forward's.  FS_ERR_UNSUPPORTED for
 * the compat schedule, semi_implicit_viscosity, or ENO3 velocity advection
 * (use_5th_weno_v = 0).  Orders 1, 2 and 3.  The adjoint pressure solves run
 * the FP64 Jacobi-PCG with the handle's stop rule.
 *
 * fs_adjoint_step: the step from the current fields is recomputed with its
 * tape; lphi, lu, lv, lw (host, interior layout as fs_set_fields) carry the
 * adjoint of the step's output in and of its input out.  The fields are left
 * as they were.
 *
 * fs_levelset_loss: sum over cells of (H(phi) - H(phi_target))^2 with H the
 * projection's smoothed Heaviside (PROJECTION_3D.h:1452-1472) at
 * eps = 1.5 dx, for the current level set.
 *
 * fs_adjoint_rollout: from the current fields, `steps` fs_step calls of
 * length dt, the loss above at the end, and its gradient with respect to the
 * initial phi, u, v, w (any output may be NULL), by the checkpointed reverse
 * sweep: the state at the start of every checkpoint_every-th step is kept on
 * the device, each segment is re-run from its checkpoint and its steps are
 * reversed one by one.  The fields are left at the initial state. */
int  fs_adjoint_step(fs_solver* s, double dt, double* lphi, double* lu, double* lv, double* lw,
                     fs_step_stats* stats);
int  fs_levelset_loss(fs_solver* s, const double* phi_target, double* loss);
int  fs_adjoint_rollout(fs_solver* s, int steps, double dt, int checkpoint_every, const double* phi_target,
                        double* loss, double* grad_phi, double* grad_u, double* grad_v, double* grad_w,
                        fs_step_stats* stats);

/* ---- batches of independent scenes ---------------------------------------
 * No reference counterpart: the reference advances one world per process
 * (SIMULATION_WORLD::AdvanceOneFrame, SIMULATION_WORLD.h:320-330).  A batch
 * serves many uneven scenes from one process: scene i is a full solver
 * (params[i], grids[i]) with its own stream; fs_batch_advance_frame runs one
 * fs_advance_frame on every scene, host threads (threads <= 0: up to 16)
 * taking scenes largest first, so the small scenes' kernels overlap on the
 * GPU.  Results are those of each scene advanced alone.  fs_batch_scene
 * returns scene i's handle (owned by the batch) for fields and phases.
 * rcs (may be NULL) receives each scene's return code; the function returns
 * the first failure or FS_OK. */
typedef struct fs_batch fs_batch;
int  fs_batch_create(int n, const fs_params* params, const fs_grid* grids, int device, int threads, fs_batch** out);
void fs_batch_destroy(fs_batch* b);
fs_solver* fs_batch_scene(fs_batch* b, int i);
int  fs_batch_advance_frame(fs_batch* b, const double* dt, fs_step_stats* stats, int* rcs);
const char* fs_batch_last_error(const fs_batch* b);

/* ---- oil/water pipe mode ------------------------------------------------------
 * EULERIAN_FLUID_SOLVER_3D in oil_water_simulation mode (PIPE_OPTIONS,
 * WORLD_DISCRETIZATION_3D.h:91-110): core-annular flow of oil (phi <= 0)
 * inside water in a cylinder of radius a (R2/R1 in units of the core radius
 * R1), the pipe along x (is_parallel: wall at sqrt(y^2 + z^2) = a, periodic
 * in x with FillGhostCellsPeriodicInXDirection, FIELD_STRUCTURE_3D.h:
 * 1088-1327) or along y (is_vertical: wall at sqrt(x^2 + z^2) = a, periodic
 * in y with FillGhostCellsPeriodicInYDirection, :1328-1710, whose masked form
 * zeroes the copy outside the wall and extrapolates the velocity linearly
 * beyond it), the node-periodic ghost copies, the wall by theta-extrapolated
 * ghost velocities
 * (ADVECTION_3D.h:380-993, VISCOSITY_3D.h:2048-2625, PROJECTION_3D.h:
 * 1521-1690), the centreline faces held fixed (ADVECTION_3D.h:279-307), the
 * dimensionless driving pressure and gravity (EULERIAN_FLUID_SOLVER_3D.h:
 * 2277-2433), the semi-implicit viscosity with its x-periodic rows
 * (VISCOSITY_3D.h:565-989, POISSON_SOLVER_3D.h:342-378, 3641-3828) and the
 * masked GFM projection with the curvature jump 1/We kappa (PROJECTION_3D.h:
 * 414-432, 642-848, 901-1097, 1101-1449; POISSON_SOLVER_3D.h:1360-1815).
 *
 * Scope: the shipped scripts' configuration (3D_script.sim parallel,
 * 3D_script_1/2/3.sim vertical) -- cylindrical boundary, Neumann pressure,
 * dimensionless form, semi-implicit viscosity, ENO3 velocity advection,
 * WENO5 level set, boundary_order_velocity 1, the "old_fashioned" pressure
 * assembly.  regular_boundary, the dimensional form, explicit viscosity, the
 * CSF model, purvis_style and second-order wall extrapolation return
 * FS_ERR_UNSUPPORTED.  The pipe must lie strictly inside the box: the first
 * and last two cells and faces of every line across the pipe outside the
 * wall (FS_ERR_BAD_ARG otherwise); every extent >= 8.
 *
 * fs_params keys used: water_density, water_viscosity, gravity,
 * surface_tension, order_for_time_advancing, the reinitialization keys,
 * levelset_epsilon, tolerance, max_iteration, use_previous_pressure,
 * stop_rule, rel_tolerance and the debugging switches; precision must be
 * FS_FP64 and schedule FS_SCHEDULE_COMPAT (the reference's order and quirks).
 */
typedef struct fs_pipe_params {
    int    is_parallel;          /* 1: pipe along x (is_parallel); 0: pipe along y (is_vertical) */
    double a;                    /* R2/R1 (SIMULATION_WORLD.h:124) */
    double R2;                   /* pipe radius (:125) */
    double A_0;                  /* interface amplitude as the script gives it; scaled by a/R2 (:126-127) */
    double alpha;                /* interface wave number (:128) */
    double m, eta, K, g;         /* FLUID_SOLVER_UNIFORM_OIL_WATER keys (EULERIAN_FLUID_SOLVER_3D.h:503-506) */
    double oil_density, oil_viscosity;
    int    dimensionless_form;   /* 1 only */
    int    cylindrical_boundary; /* 1 only (PROJECTION_WATER cylindrical_boundary) */
    int    boundary_order_velocity; /* 1 only (VELOCITY_ADVECTION boundary_order_velocity) */
} fs_pipe_params;

/* Replaces EULERIAN_FLUID_SOLVER_3D::InitializeFromScriptBlock in oil/water
 * mode (EULERIAN_FLUID_SOLVER_3D.h:192-800): the derived constants R1 = R2/a,
 * f, A, V0, J, Re, We of :496-520 and the device fields.  The handle then
 * serves fs_set_fields / fs_get_fields (interior layout), fs_step,
 * fs_advance_frame, fs_cfl_dt, fs_set_async / fs_synchronize /
 * fs_get_stream and fs_destroy; every other fs_* entry point refuses it with
 * FS_ERR_UNSUPPORTED. */
int  fs_create_pipe(const fs_params* params, const fs_pipe_params* pipe, const fs_grid* grid, int device,
                    fs_solver** out);
/* The reference's initial state of the scene: the core-annular velocity
 * profile on the x faces (EULERIAN_FLUID_SOLVER_3D.h:536-563), v = w = 0,
 * the perturbed interface phi = sqrt(y^2 + z^2) - (A_0 cos(alpha x) + 1)
 * (SIMULATION_WORLD.cpp:290-300), the pipe's level set sqrt(y^2 + z^2) - a
 * (:305-319), p = 0. */
int  fs_pipe_initial_fields(fs_solver* s);
/* Derived constants: out[0..6] = R1, f, A, V0, J, Re, We. */
int  fs_pipe_constants(fs_solver* s, double* out);

/* One phase of the step alone, for parity tests and callers that schedule
 * phases themselves: FS_PIPE_LEVELSET_ADVECTION (ADVECTION_3D::Solve_Levelset),
 * FS_PIPE_VELOCITY_ADVECTION (Solve_Velocity), FS_PIPE_GRAVITY (ApplyGravity),
 * FS_PIPE_SOURCING (Sourcing: the semi-implicit viscosity),
 * FS_PIPE_PROJECTION (PROJECTION_3D::Solve), FS_PIPE_REINITIALIZATION
 * (Reinitialization, one call of iteration_number sweeps). */
#define FS_PIPE_LEVELSET_ADVECTION  0
#define FS_PIPE_VELOCITY_ADVECTION  1
#define FS_PIPE_GRAVITY             2
#define FS_PIPE_SOURCING            3
#define FS_PIPE_PROJECTION          4
#define FS_PIPE_REINITIALIZATION    5
int  fs_pipe_phase(fs_solver* s, int phase, double dt, fs_step_stats* stats);

/* A persistent array of the step, copied over its whole box grown by 3 ghost
 * layers ((ni + 6) * (nj + 6) * (nk + 6) doubles, x fastest; ni = nx + 1 on
 * x faces, ...), for tests that check the reference's ghost copies. */
#define FS_PIPE_ARR_PHI         0   /* level set with its ghosts */
#define FS_PIPE_ARR_SCALAR_GHOST 1  /* scalar_field_ghost */
#define FS_PIPE_ARR_GHOST_U     2   /* vector_field_mac_ghost_x, _y, _z */
#define FS_PIPE_ARR_GHOST_V     3
#define FS_PIPE_ARR_GHOST_W     4
#define FS_PIPE_ARR_ADV_U_Y     5   /* ADVECTION_3D's velocity_field_mac_ghost_x_y, x_z, y_y, y_z, z_y, z_z
                                        (the vertical pipe: x_x, x_z, y_x, y_z, z_x, z_z) */
#define FS_PIPE_ARR_ADV_U_Z     6
#define FS_PIPE_ARR_ADV_V_Y     7
#define FS_PIPE_ARR_ADV_V_Z     8
#define FS_PIPE_ARR_ADV_W_Y     9
#define FS_PIPE_ARR_ADV_W_Z    10
#define FS_PIPE_ARR_VIS_U_Y    11   /* VISCOSITY_3D's copies of the same six */
#define FS_PIPE_ARR_VIS_U_Z    12
#define FS_PIPE_ARR_VIS_V_Y    13
#define FS_PIPE_ARR_VIS_V_Z    14
#define FS_PIPE_ARR_VIS_W_Y    15
#define FS_PIPE_ARR_VIS_W_Z    16
#define FS_PIPE_ARR_BOUNDARY   17   /* the pipe's level set (cell) */
#define FS_PIPE_ARR_DIVERGENCE 18
#define FS_PIPE_ARR_JUMP       19   /* jc_on_solution */
#define FS_PIPE_ARR_PRESSURE   20
#define FS_PIPE_ARR_COUNT      21
int  fs_pipe_get_array(fs_solver* s, int which, double* dst);

/* In-loop samples of the pipe solves' two loop kernels while
 * fs_set_kernel_sampling is on: one launch of each per launch batch of every
 * solve is timed with HIP events on the handle's stream (a launch that the
 * stop test ended early is dropped).  The three viscosity components are
 * solved together, one launch iterating each system still running.  Returns
 * the mean duration (ms), the mean algorithmic bytes of the timed launches
 * over the rows of the systems they iterated (FS_PIPE_KS_SPMV,
 * k_pcg_spmv_dir: 120 per row -- 7 ELL values and columns, the count, p, z
 * and the new p and q; FS_PIPE_KS_UPDATE, k_pcg_update2: 64 per row -- x, r
 * read and written, p, q, D read, z written), the number of timed launches
 * and the number of launches since sampling was enabled.  No reference counterpart:
 * measurement hooks beside fs_get_frame_kernel_samples. */
#define FS_PIPE_KS_SPMV   0
#define FS_PIPE_KS_UPDATE 1
int  fs_pipe_kernel_samples(fs_solver* s, int kernel, double* avg_ms, double* avg_bytes, long long* samples,
                            long long* launches);

#ifdef __cplusplus
}
#endif
#endif /* FS_ABI_H */
