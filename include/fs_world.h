/*
 * fs_world.h -- C-ABI of the scene layer above the timestep library: the
 * reference's SIMULATION_WORLD (SIMULATION_WORLD.h, SIMULATION_WORLD.cpp)
 * for air/water and oil/water pipe scenes, driven by its .sim scripts, and the script reader
 * (SCRIPT_READER.h) it parses them with.
 *
 * A world owns one fs_solver (fs_abi.h) and advances it frame by frame as
 * SIMULATION_WORLD::AdvanceOneFrame does.  Same conventions as fs_abi.h:
 * 0 or a negative FS_ERR_* code, one host thread per handle.
 */
#ifndef FS_WORLD_H
#define FS_WORLD_H

#include "fs_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scripts (SCRIPT_READER / SCRIPT_BLOCK, SCRIPT_READER.cpp) ---------------
 * fs_script_open reads and parses a script with the reference's rules:
 * comment removal (SCRIPT_READER.cpp:298-381), #INCLUDE <file> relative to
 * the script's directory (:232-296), NAME { ... } blocks (:383-421), values
 * by type (:155-214).  block_path names a block as "A.B.C" from the top
 * level (SCRIPT_READER::FindBlock, :482-507); "" is the script itself for
 * fs_script_child_*.  The getters behave as SCRIPT_BLOCK's (:842-987): the
 * default when the block or key is absent, atoi/atof of the value text,
 * true only for "true", vector components 0 when the value is not a
 * three-component vector.  fs_script_get_string copies at most len-1
 * characters and a terminating zero. */
typedef struct fs_script fs_script;
int  fs_script_open(const char* path, fs_script** out);
void fs_script_close(fs_script* s);
int  fs_script_get_int(const fs_script* s, const char* block_path, const char* key, int def, int* out);
int  fs_script_get_float(const fs_script* s, const char* block_path, const char* key, double def, double* out);
int  fs_script_get_bool(const fs_script* s, const char* block_path, const char* key, int def, int* out);
int  fs_script_get_vector3(const fs_script* s, const char* block_path, const char* key, const double def[3],
                           double out[3]);
int  fs_script_get_int3(const fs_script* s, const char* block_path, const char* key, const int def[3], int out[3]);
int  fs_script_get_string(const fs_script* s, const char* block_path, const char* key, const char* def,
                          char* out, int len);
/* number of child blocks of block_path, and the name of child i
 * (READ_SCRIPT_BLOCKS, SCRIPT_READER.h:10-11) */
int  fs_script_child_count(const fs_script* s, const char* block_path, int* n);
int  fs_script_child_name(const fs_script* s, const char* block_path, int i, char* out, int len);

/* ---- worlds ----------------------------------------------------------------
 * fs_world_create: SIMULATION_WORLD::InitializeFormScript
 * (SIMULATION_WORLD.h:90-240) for air_water_simulation scripts:
 *   - frame settings of block SIMULATION_WORLD: dt, max_dt, CFL, frame_rate,
 *     num_substeps, last_frame (:100-134);
 *   - the grid of WORLD_DISCRETIZATION_AIR_WATER.GRID_STRUCTURE_3D_LARGE or
 *     _SMALL (WORLD_DISCRETIZATION_3D.h:59-90, GRID_STRUCTURE_3D.h:201-219),
 *     scaled by the solver block's resolution_scale (GRID_STRUCTURE_3D.h:149-152);
 *   - the solver keys of block FLUID_SOLVER_UNIFORM_AIR_WATER and its
 *     LEVELSET_ADVECTION, VELOCITY_ADVECTION and PROJECTION_WATER blocks
 *     (EULERIAN_FLUID_SOLVER_3D.h:192-300, ADVECTION_3D.h:102-113,
 *     PROJECTION_3D.h:178-203, VISCOSITY_3D.h:186-231), with their defaults;
 *   - the initial level set of an OBJECT_LIST block of shape_type WATER: in
 *     air/water mode the sphere of radius 1/3 about the origin sampled at
 *     x_min + i dx (SIMULATION_WORLD.cpp:116-128, which overwrites the
 *     sea-surface, sphere-bubble and user-defined shapes written before it);
 *   - OBJECT_LIST blocks of shape_type SPHERE become level-set sources
 *     (fs_set_levelset_sources) with their FLUID_CONTROL_OPTIONS
 *     (FLUID_CONTROL_OPTIONS.h:45-61), applied once at frame 0
 *     (SIMULATION_WORLD.h:224-227) and in every Sourcing phase.
 * For oil_water_simulation scripts (the shipped 3D_script.sim) it sets up
 * the pipe instead (fs_create_pipe, fs_pipe_initial_fields): a, R2, A_0,
 * alpha of block SIMULATION_WORLD (SIMULATION_WORLD.h:122-129), the grid of
 * WORLD_DISCRETIZATION_OIL_WATER.GRID_STRUCTURE_3D_PARALLEL
 * (WORLD_DISCRETIZATION_3D.h:91-110), the keys of
 * FLUID_SOLVER_UNIFORM_OIL_WATER and its blocks (m, eta, K, g, oil_density,
 * oil_viscosity, boundary_order_velocity, cylindrical_boundary, ...), and the
 * interface and flow the scene starts from (SIMULATION_WORLD.cpp:130-319).
 * FS_ERR_UNSUPPORTED names what a script asks for that this build does not
 * run: the pipe mode outside fs_create_pipe's scope, rigid, cylinder or mesh objects, other Poisson
 * solvers or boundary conditions, curvature by normals, the CSF model, a
 * missing WATER object (the level set would keep its zeros).
 * opts may be NULL (FP64, the reference's stop rule, compatible schedule). */
typedef struct fs_world fs_world;
typedef struct fs_world_options {
    int    precision;       /* FS_FP64 | FS_FP32_REFINE */
    int    stop_rule;       /* FS_STOP_ABS_PRECOND: the script's tolerance */
    double rel_tolerance;   /* for FS_STOP_REL_RESIDUAL */
    int    schedule;        /* FS_SCHEDULE_* */
} fs_world_options;

typedef struct fs_world_info {
    fs_params params;       /* as passed to fs_create */
    fs_grid   grid;
    double dt, max_dt, cfl, frame_rate;
    double accu_dt;         /* sum of the CFL steps taken */
    int    num_substeps;
    int    frame;           /* frames advanced */
    int    last_frame;
    int    n_sources;
    int    oil_water;       /* 1: an oil_water_simulation script (a pipe handle, fs_create_pipe) */
    fs_pipe_params pipe;    /* its pipe keys (oil_water only) */
} fs_world_info;

int  fs_world_create(const char* script_path, const fs_world_options* opts, int device, fs_world** out);
void fs_world_destroy(fs_world* w);
const char* fs_world_last_error(const fs_world* w);
/* The world's solver handle (fields, files, statistics); owned by the world. */
fs_solver* fs_world_solver(fs_world* w);
int  fs_world_get_info(const fs_world* w, fs_world_info* info);

/* SIMULATION_WORLD::AdvanceOneFrame (SIMULATION_WORLD.h:320-330) with
 * AdvanceOneFrameThread (:273-318): the level set's ComputeCurvatures, then
 * with CFL or frame_rate zero num_substeps steps of the script's dt, else one
 * step of dt = CFL / (CFL rate) (DetermineTimeStep, :332-344, with the
 * rule's |g_y|) added to accu_dt.  stats: the last step's, with PCG
 * iterations and device time summed over the frame's steps. */
int  fs_world_advance_frame(fs_world* w, fs_step_stats* stats);

/* The script layer alone: what fs_world_create would set up, without a
 * device.  sources may be NULL; at most max_sources are copied
 * (info->n_sources counts all). */
int  fs_world_parse(const char* script_path, const fs_world_options* opts, fs_world_info* info,
                    fs_levelset_source* sources, int max_sources);

#ifdef __cplusplus
}
#endif
#endif
