// Scene layer: SIMULATION_WORLD for air/water and oil/water (pipe) scripts,
// over the C-ABI of the timestep library.  The script keys, their defaults and the order of the
// set-up follow the reference files cited at each step.
#include "../../include/fs_world.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "fs_script.h"

namespace {

struct WorldFail {
    int code;
    std::string msg;
};

[[noreturn]] void unsupported(const std::string& what) { throw WorldFail{FS_ERR_UNSUPPORTED, what}; }

// The set-up a script asks for: parameters, grid, frame settings, the
// initial level set's shape and the objects.
struct WorldSetup {
    fs_params P{};
    fs_grid G{};
    double dt = 0, max_dt = 0, cfl = 0, frame_rate = 0;
    int num_substeps = 1, last_frame = 0;
    bool water_object = false;
    std::vector<fs_levelset_source> sources;
    bool oil_water = false;  // the pipe mode (fs_create_pipe)
    fs_pipe_params Q{};
};

// GRID_STRUCTURE_3D::InitializeFromBlock (GRID_STRUCTURE_3D.h:201-219); the
// result's resolution scaled by `scale` as GRID_STRUCTURE_3D::Initialize(grid,
// scale) does (:149-152).  Returns the world grid's spacing along x in wdx.
fs_grid grid_from_block(const fs::ScriptBlock& b, double solver_scale, double* wdx) {
    const double rs = b.GetFloat("resolution_scale", 1.0);
    const int zero[3] = {0, 0, 0};
    const double dzero[3] = {0, 0, 0};
    int start[3], res[3];
    double mn[3], mx[3];
    b.GetInt3("start_indices", zero, start);
    b.GetInt3("base_grid_resolution", zero, res);
    b.GetVector3("base_grid_min", dzero, mn);
    b.GetVector3("base_grid_max", dzero, mx);
    if (start[0] != 0 || start[1] != 0 || start[2] != 0) unsupported("start_indices other than (0,0,0)");
    int wres[3], sres[3];
    for (int a = 0; a < 3; a++) {
        wres[a] = (int)((double)res[a] * rs);
        sres[a] = (int)((double)wres[a] * solver_scale);
    }
    if (wres[0] < 2) throw WorldFail{FS_ERR_BAD_ARG, "grid resolution below 2"};
    *wdx = (mx[0] - mn[0]) / (wres[0] - 1);  // GRID_STRUCTURE_3D.h:113-117
    fs_grid g;
    g.nx = sres[0];
    g.ny = sres[1];
    g.nz = sres[2];
    g.x_min = mn[0];
    g.y_min = mn[1];
    g.z_min = mn[2];
    g.x_max = mx[0];
    g.y_max = mx[1];
    g.z_max = mx[2];
    return g;
}

// the options of this build shared by both modes
void build_options(fs_params& P, const fs_world_options* opts) {
    P.stop_rule = opts ? opts->stop_rule : FS_STOP_ABS_PRECOND;
    P.rel_tolerance = opts ? opts->rel_tolerance : 1e-8;
    P.precision = opts ? opts->precision : FS_FP64;
    P.schedule = opts ? opts->schedule : FS_SCHEDULE_COMPAT;
}

// SIMULATION_WORLD::InitializeFormScript for oil_water_simulation scripts:
// the pipe (SIMULATION_WORLD.h:122-129), WORLD_DISCRETIZATION_OIL_WATER
// (WORLD_DISCRETIZATION_3D.h:91-110) and FLUID_SOLVER_UNIFORM_OIL_WATER
// (EULERIAN_FLUID_SOLVER_3D.h:192-260, 496-520, ADVECTION_3D.h:100-113,
// PROJECTION_3D.h:178-228, VISCOSITY_3D.h:184-231)
WorldSetup parse_oil_water(const fs::Script& sc, const fs::ScriptBlock& world, const fs_world_options* opts,
                           WorldSetup w) {
    w.oil_water = true;
    fs_pipe_params& Q = w.Q;
    Q.a = world.GetFloat("a", 1);
    Q.R2 = world.GetFloat("R2", 1);
    Q.A_0 = world.GetFloat("A_0", 1);  // scaled by a/R2 in fs_create_pipe, SIMULATION_WORLD.h:127
    Q.alpha = world.GetFloat("alpha", 1);
    const fs::ScriptBlock& wd = world.Find("WORLD_DISCRETIZATION_OIL_WATER");
    const fs::ScriptBlock& po = wd.Find("PIPE_OPTIONS");
    const bool vertical = po.GetBoolean("is_vertical", false), parallel = po.GetBoolean("is_parallel", false);
    if (vertical && parallel)
        unsupported("PIPE_OPTIONS sets both is_vertical and is_parallel (the reference would run both pipes' "
                    "boundary conditions on one grid)");
    if (!vertical && !parallel) unsupported("PIPE_OPTIONS sets neither is_vertical nor is_parallel");
    Q.is_parallel = parallel ? 1 : 0;
    const fs::ScriptBlock& fb = sc.Find("FLUID_SOLVER_UNIFORM_OIL_WATER");
    double wdx = 0;
    // WORLD_DISCRETIZATION_3D.h:96-109: each pipe reads its own grid block
    w.G = grid_from_block(wd.Find(vertical ? "GRID_STRUCTURE_3D_VERTICAL" : "GRID_STRUCTURE_3D_PARALLEL"),
                          fb.GetFloat("resolution_scale", 1.0), &wdx);
    if (w.G.nx < 8 || w.G.ny < 8 || w.G.nz < 8) throw WorldFail{FS_ERR_BAD_ARG, "pipe grid below 8 cells per axis"};
    fs_params& P = w.P;
    if (!fb.GetBoolean("use_water_solver", true)) unsupported("use_water_solver = false");
    if (fb.GetInteger("ghost_width", 3) != 3) unsupported("ghost_width other than 3");
    P.surface_tension = fb.GetFloat("surface_tension", 0);
    if (fb.GetBoolean("fastsweeping_reinitialization", false)) unsupported("fastsweeping_reinitialization");
    P.sussmanstyle_reinitialization = fb.GetBoolean("sussmanstyle_reinitialization", false);
    P.scaling_number_for_reinitialization = fb.GetInteger("scaling_number_for_reinitialization", 5);
    P.iteration_number_for_reinitialization = fb.GetInteger("iteration_number_for_reinitialization", 10);
    P.use_delta_function_formulation = fb.GetBoolean("use_delta_function_formulation", false);
    if (!P.use_delta_function_formulation)
        unsupported("use_delta_function_formulation = false in the pipe mode (no viscosity and no jump)");
    if (fb.GetBoolean("CSF_model", false)) unsupported("CSF_model in the pipe mode");
    if (fb.GetBoolean("curvature_by_normal_vector", false) || !fb.GetBoolean("curvature_by_levelset", false))
        unsupported("curvature other than by level set in the pipe mode");
    P.curvature_method = FS_CURV_LEVELSET;
    P.water_viscosity = fb.GetFloat("water_viscosity", 1);
    P.water_density = fb.GetFloat("water_density", 1);
    Q.oil_viscosity = fb.GetFloat("oil_viscosity", 0);
    Q.oil_density = fb.GetFloat("oil_density", 1);
    const double g0[3] = {0, -9.8, 0};
    fb.GetVector3("gravity", g0, P.gravity);
    P.order_for_time_advancing = fb.GetInteger("order_for_time_advancing", 1);
    P.is_levelset_advection_active = fb.GetBoolean("is_Levelset_advection_active", false);
    P.is_velocity_advection_active = fb.GetBoolean("is_Velocity_advection_active", false);
    P.is_external_force_active = fb.GetBoolean("is_external_force_active", false);
    P.is_sourcing_active = fb.GetBoolean("is_sourcing_active", false);
    P.is_projection_active = fb.GetBoolean("is_projection_active", false);
    if (!fb.GetBoolean("semi_implicit_approach", false))
        unsupported("the explicit viscosity of the pipe mode (semi_implicit_approach = false)");
    P.semi_implicit_viscosity = 1;
    Q.dimensionless_form = fb.GetBoolean("dimensionless_form", false) ? 1 : 0;
    if (!Q.dimensionless_form) unsupported("the dimensional form of the pipe mode (dimensionless_form = false)");
    if (!fb.GetBoolean("is_periodic", false)) unsupported("a pipe that is not periodic (is_periodic = false)");
    Q.m = fb.GetFloat("m", 1);
    Q.eta = fb.GetFloat("eta", 1);
    Q.K = fb.GetFloat("K", 1);
    Q.g = fb.GetFloat("g", 1);
    const fs::ScriptBlock& ls = fb.Find("LEVELSET_ADVECTION");
    if (!ls.GetBoolean("use_5th_weno", false)) unsupported("level-set advection other than use_5th_weno");
    P.levelset_epsilon = ls.GetFloat("epsilon", 10e-6);
    const fs::ScriptBlock& va = fb.Find("VELOCITY_ADVECTION");
    if (va.GetBoolean("use_5th_weno_v", false) || !va.GetBoolean("use_3rd_eno_v", false))
        unsupported("velocity advection other than use_3rd_eno_v in the pipe mode");
    P.use_5th_weno_v = 0;
    P.velocity_epsilon = va.GetFloat("epsilon", 10e-6);
    Q.boundary_order_velocity = va.GetInteger("boundary_order_velocity", 1);
    if (Q.boundary_order_velocity != 1) unsupported("boundary_order_velocity other than 1");
    const fs::ScriptBlock& pw = fb.Find("PROJECTION_WATER");
    P.tolerance = pw.GetFloat("tolerance", 1e-4);
    P.max_iteration = pw.GetInteger("max_iteration", 30);
    P.use_previous_pressure = pw.GetBoolean("use_previous_pressure", false);
    if (!P.use_previous_pressure) unsupported("use_previous_pressure = false in the pipe mode");
    if (pw.GetString("poisson_solver_type", "NULL") != "PCG")
        unsupported("poisson_solver_type other than \"PCG\" (CG, multigrid, hybrid, Gauss-Seidel)");
    if (!pw.GetBoolean("use_variable_density", false)) unsupported("use_variable_density = false");
    if (pw.GetBoolean("Dirichlet_Boundary_Condition", false) || !pw.GetBoolean("Neumann_Boundary_Condition", false))
        unsupported("pressure boundary conditions other than Neumann");
    if (pw.GetBoolean("regular_boundary", false) || !pw.GetBoolean("cylindrical_boundary", false))
        unsupported("a pipe boundary other than cylindrical_boundary");
    Q.cylindrical_boundary = 1;
    if (!pw.GetBoolean("old_fashioned", true) || pw.GetBoolean("purvis_style", false))
        unsupported("the Purvis-style jump discretisation");
    build_options(P, opts);
    // the WATER object seeds the interface (SIMULATION_WORLD.cpp:130-319);
    // other objects are outside this mode
    const fs::ScriptBlock& objs = sc.Find("OBJECT_LIST");
    for (const fs::ScriptBlock& o : objs.children) {
        const std::string shape = o.GetString("shape_type", "NONE"), type = o.GetString("object_type", "NONE");
        if (shape == "WATER") w.water_object = true;
        else if (type == "RIGID" || shape == "CYLINDER" || shape == "OBJ" || shape == "SPHERE")
            unsupported("objects other than the WATER block in the pipe mode (" + o.name + ")");
    }
    if (!w.water_object) unsupported("no OBJECT_LIST block of shape_type WATER (the interface would not be seeded)");
    return w;
}

WorldSetup parse(const std::string& path, const fs_world_options* opts) {
    fs::Script sc;
    std::string err;
    if (!sc.Load(path, &err)) throw WorldFail{FS_ERR_BAD_ARG, err};
    WorldSetup w;

    // SIMULATION_WORLD::InitializeFormScript, SIMULATION_WORLD.h:90-134
    const fs::ScriptBlock& world = sc.Find("SIMULATION_WORLD");
    w.dt = world.GetFloat("dt", 0.01);
    w.max_dt = world.GetFloat("max_dt", 100);
    w.cfl = world.GetFloat("CFL", 0);
    w.frame_rate = world.GetFloat("frame_rate", 24);
    w.num_substeps = world.GetInteger("num_substeps", 1);
    w.last_frame = world.GetInteger("last_frame");
    if (!world.GetBoolean("use_eulerian_solver")) unsupported("use_eulerian_solver = false (no solver to advance)");
    const bool air_water = world.GetBoolean("air_water_simulation", false);
    const bool oil_water = world.GetBoolean("oil_water_simulation", false);
    if (air_water && oil_water) unsupported("both air_water_simulation and oil_water_simulation are set");
    if (oil_water) return parse_oil_water(sc, world, opts, w);
    if (!air_water) unsupported("neither air_water_simulation nor oil_water_simulation is set");

    // the solver block, EULERIAN_FLUID_SOLVER_3D::InitializeFromScriptBlock,
    // EULERIAN_FLUID_SOLVER_3D.h:192-300
    const fs::ScriptBlock& fb = sc.Find("FLUID_SOLVER_UNIFORM_AIR_WATER");

    // WORLD_DISCRETIZATION_3D::Initialize, WORLD_DISCRETIZATION_3D.h:59-90:
    // the large grid, then the small one when both are set
    const fs::ScriptBlock& wd = world.Find("WORLD_DISCRETIZATION_AIR_WATER");
    const bool large = wd.GetBoolean("large_bubble", false), small = wd.GetBoolean("small_bubble", false);
    if (!large && !small) unsupported("WORLD_DISCRETIZATION_AIR_WATER sets neither large_bubble nor small_bubble");
    const double solver_scale = fb.GetFloat("resolution_scale", 1.0);  // EULERIAN_FLUID_SOLVER_3D.h:390
    double wdx = 0;
    w.G = grid_from_block(wd.Find(small ? "GRID_STRUCTURE_3D_SMALL" : "GRID_STRUCTURE_3D_LARGE"), solver_scale, &wdx);
    if (w.G.nx < 3 || w.G.ny < 3 || w.G.nz < 3) throw WorldFail{FS_ERR_BAD_ARG, "grid below 3 cells per axis"};

    fs_params& P = w.P;
    if (!fb.GetBoolean("use_water_solver", true)) unsupported("use_water_solver = false");
    if (fb.GetInteger("ghost_width", 3) != 3) unsupported("ghost_width other than 3");
    P.surface_tension = fb.GetFloat("surface_tension", 0);
    if (fb.GetBoolean("fastsweeping_reinitialization", false)) unsupported("fastsweeping_reinitialization");
    P.sussmanstyle_reinitialization = fb.GetBoolean("sussmanstyle_reinitialization", false);
    P.scaling_number_for_reinitialization = fb.GetInteger("scaling_number_for_reinitialization", 5);
    P.iteration_number_for_reinitialization = fb.GetInteger("iteration_number_for_reinitialization", 10);
    P.use_delta_function_formulation = fb.GetBoolean("use_delta_function_formulation", false);
    // EULERIAN_FLUID_SOLVER_3D.h:210; in air/water mode it zeroes the jump
    // condition and applies no force (fs_params.csf_model)
    P.csf_model = fb.GetBoolean("CSF_model", false) ? 1 : 0;
    // viscosities and densities: the solver's and VISCOSITY_3D's reads of
    // the same keys (VISCOSITY_3D.h:186-204) differ only in the air
    // viscosity's default (0 and 1), so that key must be given
    if (!fb.Value("air_viscosity")) unsupported("air_viscosity not given (its two defaults in the reference differ)");
    P.water_viscosity = fb.GetFloat("water_viscosity", 1);
    P.air_viscosity = fb.GetFloat("air_viscosity", 0);
    const double g0[3] = {0, -9.8, 0};
    fb.GetVector3("gravity", g0, P.gravity);
    P.water_density = fb.GetFloat("water_density", 1);
    P.air_density = fb.GetFloat("air_density", 1);
    P.order_for_time_advancing = fb.GetInteger("order_for_time_advancing", 1);
    P.is_levelset_advection_active = fb.GetBoolean("is_Levelset_advection_active", false);
    P.is_velocity_advection_active = fb.GetBoolean("is_Velocity_advection_active", false);
    P.is_external_force_active = fb.GetBoolean("is_external_force_active", false);
    P.is_sourcing_active = fb.GetBoolean("is_sourcing_active", false);
    P.is_projection_active = fb.GetBoolean("is_projection_active", false);
    // the semi-implicit viscosity runs only in oil/water mode
    // (VISCOSITY_3D.h:565-989), so an air/water script never reaches it
    P.semi_implicit_viscosity = 0;
    // the level set's two curvature flags (EULERIAN_FLUID_SOLVER_3D.h:758-759,
    // both false by default: no curvature at all)
    {
        const bool by_normals = fb.GetBoolean("curvature_by_normal_vector", false),
                   by_levelset = fb.GetBoolean("curvature_by_levelset", false);
        P.curvature_method = by_normals ? (by_levelset ? FS_CURV_BOTH : FS_CURV_NORMALS)
                                        : (by_levelset ? FS_CURV_LEVELSET : FS_CURV_NONE);
    }

    // ADVECTION_3D::InitializeFromBlock, ADVECTION_3D.h:100-113
    const fs::ScriptBlock& ls = fb.Find("LEVELSET_ADVECTION");
    if (!ls.GetBoolean("use_5th_weno", false)) unsupported("level-set advection other than use_5th_weno");
    if (ls.GetBoolean("use_csl", false)) unsupported("use_csl");
    P.levelset_epsilon = ls.GetFloat("epsilon", 10e-6);
    const fs::ScriptBlock& va = fb.Find("VELOCITY_ADVECTION");
    const bool weno_v = va.GetBoolean("use_5th_weno_v", false), eno_v = va.GetBoolean("use_3rd_eno_v", false);
    if (weno_v == eno_v) unsupported("velocity advection needs exactly one of use_5th_weno_v and use_3rd_eno_v");
    P.use_5th_weno_v = weno_v ? 1 : 0;
    P.velocity_epsilon = va.GetFloat("epsilon", 10e-6);

    // PROJECTION_3D::InitializeFromBlock, PROJECTION_3D.h:178-228
    const fs::ScriptBlock& pw = fb.Find("PROJECTION_WATER");
    P.tolerance = pw.GetFloat("tolerance", 1e-4);
    P.max_iteration = pw.GetInteger("max_iteration", 30);
    P.use_previous_pressure = pw.GetBoolean("use_previous_pressure", false);
    const bool bubble = pw.GetBoolean("air_bubble_rising", false), drop = pw.GetBoolean("water_drop", false);
    if (bubble == drop) unsupported("PROJECTION_WATER needs exactly one of air_bubble_rising and water_drop");
    P.air_bubble_rising = bubble ? 1 : 0;
    if (pw.GetString("poisson_solver_type", "NULL") != "PCG")
        unsupported("poisson_solver_type other than \"PCG\" (CG, multigrid, hybrid, Gauss-Seidel)");
    if (!pw.GetBoolean("use_variable_density", false)) unsupported("use_variable_density = false");
    if (pw.GetBoolean("Dirichlet_Boundary_Condition", false) || !pw.GetBoolean("Neumann_Boundary_Condition", false))
        unsupported("boundary conditions other than Neumann walls");
    if (!pw.GetBoolean("old_fashioned", true) || pw.GetBoolean("purvis_style", false))
        unsupported("the Purvis-style jump discretisation");

    build_options(P, opts);

    // SIMULATION_WORLD::InitializeObjectListFromScript, SIMULATION_WORLD.cpp:4-480
    const fs::ScriptBlock& objs = sc.Find("OBJECT_LIST");
    for (const fs::ScriptBlock& o : objs.children) {
        const std::string shape = o.GetString("shape_type", "NONE");
        if (shape == "WATER") w.water_object = true;  // :61-129
    }
    for (const fs::ScriptBlock& o : objs.children) {  // :331-478
        const std::string type = o.GetString("object_type", "NONE"), shape = o.GetString("shape_type", "NONE");
        if (type == "RIGID") unsupported("rigid objects (" + o.name + ")");
        if (shape == "CYLINDER" || shape == "OBJ") unsupported("object shape " + shape + " (" + o.name + ")");
        if (shape != "SPHERE") continue;  // WATER, WALL and unknown shapes add no object
        const double zero[3] = {0, 0, 0};
        fs_levelset_source s{};
        o.GetVector3("position", zero, s.center);
        s.radius = o.GetFloat("size");
        s.half_width = s.radius + 3.0 * wdx;  // :359, world_discretization.dx
        // FLUID_CONTROL_OPTIONS::InitializeFromScript, FLUID_CONTROL_OPTIONS.h:45-61
        const fs::ScriptBlock& fc = o.Find("FLUID_CONTROL_OPTIONS");
        s.mode = fc.GetBoolean("source_levelset", false)         ? FS_SOURCE_ADD
                 : fc.GetBoolean("delete_water_levelset", false) ? FS_SOURCE_DELETE
                                                                 : FS_SOURCE_NONE;
        s.start_frame = fc.GetInteger("source_start_frame", -INT_MAX);
        s.stop_frame = fc.GetInteger("source_stop_frame", INT_MAX);
        w.sources.push_back(s);
    }
    if (!w.water_object) unsupported("no OBJECT_LIST block of shape_type WATER (the level set would stay zero)");
    if (w.sources.size() > 64) unsupported("more than 64 sphere objects");
    return w;
}

void fill_info(const WorldSetup& w, fs_world_info* info) {
    std::memset(info, 0, sizeof(*info));
    info->params = w.P;
    info->grid = w.G;
    info->dt = w.dt;
    info->max_dt = w.max_dt;
    info->cfl = w.cfl;
    info->frame_rate = w.frame_rate;
    info->num_substeps = w.num_substeps;
    info->last_frame = w.last_frame;
    info->n_sources = (int)w.sources.size();
    info->oil_water = w.oil_water ? 1 : 0;
    info->pipe = w.Q;
}

}  // namespace

struct fs_script {
    fs::Script sc;
};

struct fs_world {
    WorldSetup setup;
    fs_solver* solver = nullptr;
    int frame = 0;
    double accu_dt = 0;
    std::string err;
};

// ---- scripts ---------------------------------------------------------------
namespace {
const fs::ScriptBlock* block_of(const fs_script* s, const char* path) {
    if (!s || !path) return nullptr;
    return &s->sc.Find(path);
}
void copy_str(const std::string& v, char* out, int len) {
    if (!out || len <= 0) return;
    size_t n = std::min((size_t)(len - 1), v.size());
    std::memcpy(out, v.data(), n);
    out[n] = 0;
}
}  // namespace

extern "C" {

int fs_script_open(const char* path, fs_script** out) {
    if (!path || !out) return FS_ERR_BAD_ARG;
    fs_script* s = new (std::nothrow) fs_script;
    if (!s) return FS_ERR_OOM;
    std::string err;
    if (!s->sc.Load(path, &err)) {
        delete s;
        return FS_ERR_BAD_ARG;
    }
    *out = s;
    return FS_OK;
}

void fs_script_close(fs_script* s) { delete s; }

int fs_script_get_int(const fs_script* s, const char* block_path, const char* key, int def, int* out) {
    const fs::ScriptBlock* b = block_of(s, block_path);
    if (!b || !key || !out) return FS_ERR_BAD_ARG;
    *out = b->GetInteger(key, def);
    return FS_OK;
}

int fs_script_get_float(const fs_script* s, const char* block_path, const char* key, double def, double* out) {
    const fs::ScriptBlock* b = block_of(s, block_path);
    if (!b || !key || !out) return FS_ERR_BAD_ARG;
    *out = b->GetFloat(key, def);
    return FS_OK;
}

int fs_script_get_bool(const fs_script* s, const char* block_path, const char* key, int def, int* out) {
    const fs::ScriptBlock* b = block_of(s, block_path);
    if (!b || !key || !out) return FS_ERR_BAD_ARG;
    *out = b->GetBoolean(key, def != 0) ? 1 : 0;
    return FS_OK;
}

int fs_script_get_vector3(const fs_script* s, const char* block_path, const char* key, const double def[3],
                          double out[3]) {
    const fs::ScriptBlock* b = block_of(s, block_path);
    if (!b || !key || !out || !def) return FS_ERR_BAD_ARG;
    b->GetVector3(key, def, out);
    return FS_OK;
}

int fs_script_get_int3(const fs_script* s, const char* block_path, const char* key, const int def[3], int out[3]) {
    const fs::ScriptBlock* b = block_of(s, block_path);
    if (!b || !key || !out || !def) return FS_ERR_BAD_ARG;
    b->GetInt3(key, def, out);
    return FS_OK;
}

int fs_script_get_string(const fs_script* s, const char* block_path, const char* key, const char* def, char* out,
                         int len) {
    const fs::ScriptBlock* b = block_of(s, block_path);
    if (!b || !key || !out || len <= 0) return FS_ERR_BAD_ARG;
    copy_str(b->GetString(key, def ? def : "Null"), out, len);
    return FS_OK;
}

int fs_script_child_count(const fs_script* s, const char* block_path, int* n) {
    if (!s || !block_path || !n) return FS_ERR_BAD_ARG;
    *n = block_path[0] ? (int)s->sc.Find(block_path).children.size() : (int)s->sc.Blocks().size();
    return FS_OK;
}

int fs_script_child_name(const fs_script* s, const char* block_path, int i, char* out, int len) {
    if (!s || !block_path || !out || len <= 0) return FS_ERR_BAD_ARG;
    const std::vector<fs::ScriptBlock>& v = block_path[0] ? s->sc.Find(block_path).children : s->sc.Blocks();
    if (i < 0 || i >= (int)v.size()) return FS_ERR_BAD_ARG;
    copy_str(v[i].name, out, len);
    return FS_OK;
}

// ---- worlds ----------------------------------------------------------------
int fs_world_parse(const char* script_path, const fs_world_options* opts, fs_world_info* info,
                   fs_levelset_source* sources, int max_sources) {
    if (!script_path || !info || max_sources < 0) return FS_ERR_BAD_ARG;
    try {
        WorldSetup w = parse(script_path, opts);
        fill_info(w, info);
        for (int i = 0; sources && i < (int)w.sources.size() && i < max_sources; i++) sources[i] = w.sources[i];
        return FS_OK;
    } catch (WorldFail& f) {
        return f.code;
    } catch (std::bad_alloc&) {
        return FS_ERR_OOM;
    }
}

int fs_world_create(const char* script_path, const fs_world_options* opts, int device, fs_world** out) {
    if (!script_path || !out) return FS_ERR_BAD_ARG;
    *out = nullptr;
    std::unique_ptr<fs_world> w(new (std::nothrow) fs_world);
    if (!w) return FS_ERR_OOM;
    try {
        w->setup = parse(script_path, opts);
    } catch (WorldFail& f) {
        return f.code;
    } catch (std::bad_alloc&) {
        return FS_ERR_OOM;
    }
    const WorldSetup& S = w->setup;
    int rc;
    if (S.oil_water) {
        // the pipe's level set, the interface and the core-annular flow
        // (EULERIAN_FLUID_SOLVER_3D.h:496-563, SIMULATION_WORLD.cpp:288-319)
        if ((rc = fs_create_pipe(&S.P, &S.Q, &S.G, device, &w->solver)) != FS_OK) return rc;
        if ((rc = fs_pipe_initial_fields(w->solver)) != FS_OK) return rc;
        *out = w.release();
        return FS_OK;
    }
    rc = fs_create(&S.P, &S.G, device, &w->solver);
    if (rc != FS_OK) return rc;
    // SIMULATION_WORLD.cpp:116-128: in air/water mode the WATER object's level
    // set is the sphere of radius 1/3 about the origin, sampled at the
    // reference's x_min + i dx (no half-cell offset), ghosts by constant
    // extrapolation (FillGhostCellsFromThreaded(.., false))
    const fs_grid& g = S.G;
    const double dx = (g.x_max - g.x_min) / (g.nx - 1), dy = (g.y_max - g.y_min) / (g.ny - 1),
                 dz = (g.z_max - g.z_min) / (g.nz - 1);
    std::vector<double> phi;
    try {
        phi.resize((size_t)g.nx * g.ny * g.nz);
    } catch (std::bad_alloc&) {
        return FS_ERR_OOM;
    }
    for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++) {
                double x = g.x_min + i * dx, y = g.y_min + j * dy, z = g.z_min + k * dz;
                phi[(size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * k)] =
                    std::sqrt(x * x + y * y + z * z) - (double)1 / 3;
            }
    if ((rc = fs_set_fields(w->solver, phi.data(), nullptr, nullptr, nullptr, nullptr)) != FS_OK) return rc;
    if (!S.sources.empty()) {
        // SIMULATION_WORLD.h:224-227: SourcingFromObjectList once, frame 0
        if ((rc = fs_set_levelset_sources(w->solver, (int)S.sources.size(), S.sources.data())) != FS_OK) return rc;
        if ((rc = fs_set_frame(w->solver, 0)) != FS_OK) return rc;
        if ((rc = fs_apply_levelset_sources(w->solver)) != FS_OK) return rc;
    }
    *out = w.release();
    return FS_OK;
}

void fs_world_destroy(fs_world* w) {
    if (!w) return;
    fs_destroy(w->solver);
    delete w;
}

const char* fs_world_last_error(const fs_world* w) {
    if (!w) return "null world";
    return w->err.empty() ? fs_last_error(w->solver) : w->err.c_str();
}

fs_solver* fs_world_solver(fs_world* w) { return w ? w->solver : nullptr; }

int fs_world_get_info(const fs_world* w, fs_world_info* info) {
    if (!w || !info) return FS_ERR_BAD_ARG;
    fill_info(w->setup, info);
    info->frame = w->frame;
    info->accu_dt = w->accu_dt;
    return FS_OK;
}

int fs_world_advance_frame(fs_world* w, fs_step_stats* stats) {
    if (!w) return FS_ERR_BAD_ARG;
    const WorldSetup& S = w->setup;
    int rc;
    // SIMULATION_WORLD::AdvanceOneFrame, SIMULATION_WORLD.h:320-330: the
    // level set's ComputeCurvatures first.  A pipe handle has no level-set
    // sources; its fs_advance_frame is ComputeCurvatures and a step, and its
    // CFL takes the curvature maximum of the current level set itself.
    bool curvature_pending = S.oil_water;
    if (!S.oil_water) {
        if ((rc = fs_set_frame(w->solver, w->frame)) != FS_OK) return rc;
        if ((rc = fs_compute_curvatures(w->solver)) != FS_OK) return rc;
    }
    fs_step_stats sum{}, st{};
    int result = FS_OK;
    auto step = [&](double dt) {
        int r = curvature_pending ? fs_advance_frame(w->solver, dt, &st) : fs_step(w->solver, dt, &st);
        curvature_pending = false;
        if (r != FS_OK && r != FS_ERR_NOT_CONVERGED) return r;
        if (r == FS_ERR_NOT_CONVERGED) result = r;
        int its = sum.pcg_iterations + st.pcg_iterations, vits = sum.viscosity_iterations + st.viscosity_iterations;
        int sweeps = sum.refinement_sweeps + st.refinement_sweeps;
        double pms = sum.projection_ms + st.projection_ms, sms = sum.step_ms + st.step_ms;
        sum = st;
        sum.pcg_iterations = its;
        sum.viscosity_iterations = vits;
        sum.refinement_sweeps = sweeps;
        sum.projection_ms = pms;
        sum.step_ms = sms;
        return FS_OK;
    };
    // AdvanceOneFrameThread, SIMULATION_WORLD.h:273-318
    if (S.cfl == 0 || S.frame_rate == 0) {
        for (int sub = 0; sub < S.num_substeps; sub++)
            if ((rc = step(S.dt)) != FS_OK) return rc;
    } else {
        double dt = 0;
        // DetermineTimeStep()*CFL with CFLOneTimeStep()'s |g_y| (:289, :332-344)
        if ((rc = fs_cfl_dt(w->solver, S.cfl, 0, &dt)) != FS_OK) return rc;
        w->accu_dt += dt;
        if ((rc = step(dt)) != FS_OK) return rc;
    }
    w->frame++;
    if (stats) *stats = sum;
    return result;
}

}  // extern "C"
