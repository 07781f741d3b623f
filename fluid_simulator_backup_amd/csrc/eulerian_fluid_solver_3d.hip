// eulerian_fluid_solver_3d.hip -- host class and C-ABI (include/fs_abi.h).
//
// EulerianFluidSolver3D keeps the call surface of the reference class
// EULERIAN_FLUID_SOLVER_3D (EULERIAN_FLUID_SOLVER_3D.h): the same phase
// functions, called in the same order, on device-resident fields.  All work
// of a handle runs on its own HIP stream; each public call returns after the
// stream has drained.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <atomic>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/fs_abi.h"
#include "fs_adjoint.h"
#include "fs_batch.h"
#include "fs_kernels.h"
#include "fs_layout.h"
#include "fs_mg.h"
#include "fs_probe.h"
#include "fs_cfl.h"
#include "fs_pipe.h"

#ifndef FS_FP32_RESTART
#define FS_FP32_RESTART 4000  // FS_FP32_REFINE: restart a recursion that dropped under 100x in this many iterations
#endif
#ifndef FS_MG_SWEEPS_LOW
#define FS_MG_SWEEPS_LOW 2   // FS_FP32_MG smoothing sweeps each way below 10:1 density contrast
#endif
#ifndef FS_MG_SWEEPS_HIGH
#define FS_MG_SWEEPS_HIGH 4  // and from 10:1
#endif
#ifndef FS_MG_INNER_RTOL
#define FS_MG_INNER_RTOL 1e-5  // FS_FP32_MG: relative residual of each sweep's FP32 solve
#endif
#include "fs_pcg.h"

namespace {

struct HipError {
    hipError_t e;
    const char* what;
};

#define HIPCHECK(x)                                  \
    do {                                             \
        hipError_t _e = (x);                         \
        if (_e != hipSuccess) throw HipError{_e, #x}; \
    } while (0)

struct FsFail {
    int code;
    std::string msg;
};

}  // namespace

class EulerianFluidSolver3D {
   public:
    fs_params P;
    fs_grid G;
    FsLayout L;
    FsGridConst cell, macx, macy, macz;
    FsMaterial mat;
    int device = 0;
    hipStream_t st = nullptr;
    // second stream: the step's reinitialization runs on it, on a copy of the
    // advected level set, while the projection runs on st (ConcurrentReinit)
    hipStream_t st2 = nullptr;
    hipEvent_t ev_ls = nullptr, ev_rn = nullptr;
    double* phi_r = nullptr;
    // The reinitialization's sweeps past `reinit_switch` run on st3, a
    // stream on every CU (after an event on st2): once SolveForRK is done,
    // the CUs st2's mask leaves to it would idle.  The switch point follows
    // the last concurrent step's timings: SolveForRK's device time over the
    // masked sweep's (ev_t: reinitialization start and switch on st2,
    // SolveForRK start and end on st).
    hipStream_t st3 = nullptr;
    hipEvent_t ev_sw = nullptr, ev_t[4] = {};
    int reinit_switch = -1, reinit_count = 0, last_switch = -1;
    bool timed_concurrent = false, switched = false;
    bool batch_member = false;  // set by fs_batch_create
    bool async = false;         // fs_set_async: steps return without the final wait
    hipEvent_t ev[4] = {};

    // persistent fields (padded layout, FP64)
    double *phi = nullptr, *sghost = nullptr, *sghost2 = nullptr, *u = nullptr, *v = nullptr, *w = nullptr;
    // one spare per component: each keeps its component's extents, so ghost
    // cells of every velocity buffer stay zero across the swaps
    double *spare_u = nullptr, *spare_v = nullptr, *spare_w = nullptr;
    // velocity at the start of a Runge-Kutta step (orders 2 and 3), the
    // reference's water_velocity_field_*_rk
    double *u_rk = nullptr, *v_rk = nullptr, *w_rk = nullptr;
    // semi-implicit viscosity rows (entries to i-1, i+1, j-1, j+1, k-1, k+1,
    // diagonal, right-hand side) and its PCG vectors
    double* va[6] = {};
    double *vdiag = nullptr, *vrhs = nullptr;
    FsPcgWork wv{};
    double *gu = nullptr, *gv = nullptr, *gw = nullptr, *p = nullptr, *div = nullptr, *rho = nullptr,
           *irho = nullptr;
    double *sgn = nullptr, *tmp = nullptr, *bx = nullptr, *by = nullptr, *bz = nullptr, *diag = nullptr,
           *rhs = nullptr;
    // PCG workspaces
    FsPcgWork w64{}, w32{};
    float *bx32 = nullptr, *by32 = nullptr, *bz32 = nullptr, *diag32 = nullptr, *e32 = nullptr, *r32 = nullptr;
    float* phi32 = nullptr;  // the level set of the last assembly, for the FP32 SpMV's coefficients
    // FS_FP32_MG: the multigrid hierarchy, the CG's partial sums and totals
    FsMg mg;
    double *mg_part = nullptr, *mg_mid = nullptr;
    FsMgCg *mg_cg = nullptr, *hmg_cg = nullptr;  // the CG's device scalars and their pinned host copy
    // the projection's operator forms its face coefficients from phi (set by
    // the forward assembly, whose phi and densities define them)
    bool op_phi = false;
    FsPcgState *S64 = nullptr, *S32 = nullptr, *hS = nullptr;  // hS: pinned host mirror
    double* scratch = nullptr;                                   // mean removal partials (nz+1) + pin (2)
    unsigned long long* vmax = nullptr;                          // 4 words
    unsigned long long* hvmax = nullptr;                         // pinned
    // surface tension (sigma != 0): curvature, jump condition sigma * kappa
    // (zero ghosts), and [bits of max |kappa| near the interface, count of
    // cells with a vanishing gradient] of the last curvature pass
    double *kappa = nullptr, *jc = nullptr;
    unsigned long long *kstat = nullptr, *hkstat = nullptr;
    // the scene's objects that source the level set, and the frame number
    // their windows compare with
    std::vector<fs_levelset_source> sources;
    int frame = 0;
    double* staging = nullptr;                                   // host <-> device exchange
    long staging_n = 0;
    bool assembled = false;
    std::string last_error;

    // last-projection statistics
    fs_step_stats stats{};
    double update_drop = 0.01;  // reliable update when ||r32||^2 fell by this factor
    double hbb = 0;             // host staging of ||b||^2 for the FP32 state
    bool debug = false;         // FS_DEBUG: trace refinement on stderr
    int dbg_lines = 0;
    int chunk = 16;
    // in-loop kernel sampling (measurement only): one iteration per launch
    // batch bracketed with events; sums per precision (0 FP64, 1 FP32) and
    // kernel (0 p-update + SpMV + dot, 1 update; 2 unused)
    FsKernelProbe probe{};
    bool sampling = false;
    double ksum[2][3] = {};
    int kcount[2][3] = {};
    // frame kernels sampled the same way: FS_KS_REINIT (the first sweep of
    // each reinitialization call), FS_KS_MG_SMOOTH (the first level-0
    // smoother launch of each V-cycle)
    FsLaunchProbe kprobe[FS_KS_COUNT];

    // ---- tape of one fixed-schedule step (what its reverse pass reads) ----
    // Fields are drawn from one pool per component box (0 cells, 1..3 MAC)
    // so a recycled field's ghosts stay zero.
    struct RkRecord {
        double *u = nullptr, *v = nullptr, *w = nullptr;     // velocity at the stage's start
        double *u2 = nullptr, *v2 = nullptr, *w2 = nullptr;  // at Sourcing's start
        double *ux = nullptr, *vy = nullptr;                 // after the x / y viscosity updates
        double *p = nullptr, *phi = nullptr;                 // solved pressure; the stage's level set
    };
    struct StepTape {
        std::vector<double*> pool[4];
        size_t used[4] = {};
        double *phi = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *p = nullptr;  // step input
        std::vector<double*> ls;                   // level set before each advection stage
        std::vector<RkRecord> rk;                  // one per SolveForRK
        double* phi_adv = nullptr;                 // level set before reinitialization
        std::vector<std::vector<double*>> reinit;  // per call: level set before each sweep
        void reset() {
            for (auto& u_ : used) u_ = 0;
            ls.clear();
            rk.clear();
            reinit.clear();
            phi = u = v = w = p = phi_adv = nullptr;
        }
    };
    StepTape tape;
    bool recording = false;
    // adjoint fields (lazily allocated) and saved states of a rollout
    struct AdjointWork {
        double *lphi = nullptr, *lu = nullptr, *lv = nullptr, *lw = nullptr;  // adjoint of the state
        double *ladv = nullptr, *a0 = nullptr, *a1 = nullptr, *lsg = nullptr, *ghw = nullptr, *y = nullptr;
        double *phiw = nullptr, *tgt = nullptr, *part = nullptr;
        double *aU = nullptr, *aV = nullptr, *aW = nullptr, *src = nullptr, *du = nullptr, *dv = nullptr,
               *dw = nullptr;
        bool ready = false;
    } A;
    struct SavedState {
        double *phi = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *p = nullptr;
    };
    std::vector<SavedState> saved;

    void* dalloc(size_t bytes) {
        void* ptr = nullptr;
        if (hipMalloc(&ptr, bytes) != hipSuccess) throw FsFail{FS_ERR_OOM, "hipMalloc failed"};
        HIPCHECK(hipMemsetAsync(ptr, 0, bytes, st));
        return ptr;
    }
    double* dfield() { return (double*)dalloc(sizeof(double) * (size_t)L.total); }
    float* ffield() { return (float*)dalloc(sizeof(float) * (size_t)L.total); }

    // EULERIAN_FLUID_SOLVER_3D::InitializeFromScriptBlock, :192-800
    void InitializeFromScriptBlock(const fs_params& params, const fs_grid& grid, int dev) {
        P = params;
        G = grid;
        if (G.nx < 3 || G.ny < 3 || G.nz < 3) throw FsFail{FS_ERR_BAD_ARG, "grid must be at least 3 cells per axis"};
        if ((long)G.nx * G.ny * G.nz > (1L << 31)) throw FsFail{FS_ERR_BAD_ARG, "grid too large"};
        if (!std::isfinite(P.surface_tension)) throw FsFail{FS_ERR_BAD_ARG, "surface_tension must be finite"};
        if (P.order_for_time_advancing < 1 || P.order_for_time_advancing > 3)
            throw FsFail{FS_ERR_BAD_ARG, "order_for_time_advancing must be 1, 2 or 3"};
        if (P.precision != FS_FP64 && P.precision != FS_FP32_REFINE && P.precision != FS_FP32_MG)
            throw FsFail{FS_ERR_BAD_ARG, "precision"};
        if (P.scaling_number_for_reinitialization <= 0) throw FsFail{FS_ERR_BAD_ARG, "scaling_number"};
        if (P.curvature_method < FS_CURV_LEVELSET || P.curvature_method > FS_CURV_NONE)
            throw FsFail{FS_ERR_BAD_ARG, "curvature_method"};
        debug = std::getenv("FS_DEBUG") != nullptr;
        // tuning override of the reliable-update trigger (residual^2 drop, default 0.01)
        if (const char* e = std::getenv("FS_FP32_UPDATE_DROP")) {
            double v = std::atof(e);
            if (v > 0 && v < 1) update_drop = v;
        }
        if (dev >= 0) HIPCHECK(hipSetDevice(dev));
        HIPCHECK(hipGetDevice(&device));
        HIPCHECK(hipStreamCreate(&st));
        for (auto& e : ev) HIPCHECK(hipEventCreate(&e));
        for (auto& e : probe.ev) HIPCHECK(hipEventCreate(&e));
        for (auto& k : kprobe) k.create();
        L = fs_make_layout(G.nx, G.ny, G.nz);
        cell.x = fs_make_axis(G.x_min, G.x_max, G.nx);
        cell.y = fs_make_axis(G.y_min, G.y_max, G.ny);
        cell.z = fs_make_axis(G.z_min, G.z_max, G.nz);
        // MAC grids, EULERIAN_FLUID_SOLVER_3D.h:407,415,423
        macx = cell;
        macx.x = fs_make_axis(G.x_min - (double)0.5 * cell.x.d, G.x_max + (double)0.5 * cell.x.d, G.nx + 1);
        macy = cell;
        macy.y = fs_make_axis(G.y_min - (double)0.5 * cell.y.d, G.y_max + (double)0.5 * cell.y.d, G.ny + 1);
        macz = cell;
        macz.z = fs_make_axis(G.z_min - (double)0.5 * cell.z.d, G.z_max + (double)0.5 * cell.z.d, G.nz + 1);
        mat = FsMaterial{P.air_bubble_rising, P.water_density, P.air_density, P.water_viscosity, P.air_viscosity};
        if (P.precision != FS_FP64 && !fixed() && !cubic())
            throw FsFail{FS_ERR_UNSUPPORTED,
                         "FP32 refinement on unequal spacings needs the fixed schedule: the compatible schedule's "
                         "matrix puts 1/dx^2 on the diagonal for every face, so constants are not in its null space"};

        double** f64[] = {&phi, &sghost, &sghost2, &u, &v, &w, &spare_u, &spare_v, &spare_w, &gu, &gv, &gw, &p, &div, &rho, &irho, &sgn, &tmp,
                          &bx, &by, &bz, &diag, &rhs};
        for (double** f : f64) *f = dfield();
        if (P.order_for_time_advancing > 1) {
            u_rk = dfield();
            v_rk = dfield();
            w_rk = dfield();
        }
        const size_t npart = sizeof(double) * FS_PCG_SLOTS * FS_PCG_MAX_BLOCKS;
        w64.r = dfield();
        w64.s = dfield();
        w64.p = dfield();
        w64.q = dfield();
        w64.p2 = dfield();
        w64.partials = (double*)dalloc(npart);
        if (P.precision == FS_FP32_REFINE || P.precision == FS_FP32_MG) {
            bx32 = ffield();
            by32 = ffield();
            bz32 = ffield();
            diag32 = ffield();
            phi32 = ffield();
            e32 = ffield();
            r32 = ffield();
            w32.r = ffield();
            w32.s = ffield();
            w32.p = ffield();
            w32.q = ffield();
            w32.p2 = ffield();
            w32.partials = (double*)dalloc(npart);
        }
        if (P.semi_implicit_viscosity) {
            for (auto& a : va) a = dfield();
            vdiag = dfield();
            vrhs = dfield();
            wv.r = dfield();
            wv.s = dfield();
            wv.p = dfield();
            wv.q = dfield();
            wv.p2 = dfield();
            wv.partials = (double*)dalloc(npart);
        }
        S64 = (FsPcgState*)dalloc(sizeof(FsPcgState));
        S32 = (FsPcgState*)dalloc(sizeof(FsPcgState));
        HIPCHECK(hipHostMalloc((void**)&hS, sizeof(FsPcgState)));
        scratch = (double*)dalloc(sizeof(double) * (G.nz + 8));
        vmax = (unsigned long long*)dalloc(sizeof(unsigned long long) * 4);
        HIPCHECK(hipHostMalloc((void**)&hvmax, sizeof(unsigned long long) * 4));
        if (tension()) {
            kappa = dfield();
            jc = dfield();
            kstat = (unsigned long long*)dalloc(sizeof(unsigned long long) * 2);
            HIPCHECK(hipHostMalloc((void**)&hkstat, sizeof(unsigned long long) * 2));
            hkstat[0] = hkstat[1] = 0;
        }
        staging_n = (long)(G.nx + 1) * (G.ny + 1) * (G.nz + 1);
        staging = (double*)dalloc(sizeof(double) * staging_n);
        HIPCHECK(hipStreamSynchronize(st));
    }

    ~EulerianFluidSolver3D() {
        // every stream drained before any buffer is freed
        for (hipStream_t s_ : {st, st2, st3})
            if (s_) (void)hipStreamSynchronize(s_);
        double* f64[] = {phi, phi_r, sghost, sghost2, u, v, w, spare_u, spare_v, spare_w, gu, gv, gw, p, div, rho, irho, sgn, tmp, bx, by, bz, diag, rhs,
                         scratch, staging, w64.partials, w32.partials, u_rk, v_rk, w_rk,
                         va[0], va[1], va[2], va[3], va[4], va[5], vdiag, vrhs, wv.partials, kappa, jc,
                         nrm[0], nrm[1], nrm[2]};
        for (double* f : f64)
            if (f) (void)hipFree(f);
        void* fv[] = {w64.r, w64.s, w64.p, w64.q, w32.r, w32.s, w32.p, w32.q, bx32, by32, bz32, diag32, e32, r32, phi32,
                      S64, S32, vmax, kstat, wv.r, wv.s, wv.p, wv.q, w64.p2, w32.p2, wv.p2};
        for (void* f : fv)
            if (f) (void)hipFree(f);
        for (auto& pl : tape.pool)
            for (double* f : pl) (void)hipFree(f);
        double* adj[] = {A.lphi, A.lu, A.lv, A.lw, A.ladv, A.a0, A.a1, A.lsg, A.ghw, A.y, A.phiw, A.tgt, A.part,
                         A.aU, A.aV, A.aW, A.src, A.du, A.dv, A.dw};
        for (double* f : adj)
            if (f) (void)hipFree(f);
        for (auto& sv : saved)
            for (double* f : {sv.phi, sv.u, sv.v, sv.w, sv.p})
                if (f) (void)hipFree(f);
        if (hS) (void)hipHostFree(hS);
        if (hvmax) (void)hipHostFree(hvmax);
        if (hkstat) (void)hipHostFree(hkstat);
        if (mg_part) (void)hipFree(mg_part);
        if (mg_mid) (void)hipFree(mg_mid);
        if (mg_cg) (void)hipFree(mg_cg);
        if (hmg_cg) (void)hipHostFree(hmg_cg);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : probe.ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& k : kprobe) k.destroy();
        for (hipEvent_t e : {ev_ls, ev_rn, ev_sw, ev_t[0], ev_t[1], ev_t[2], ev_t[3]})
            if (e) (void)hipEventDestroy(e);
        if (st3) (void)hipStreamDestroy(st3);
        if (st2) (void)hipStreamDestroy(st2);
        if (st) (void)hipStreamDestroy(st);
    }

    FsBox box(int comp) const { return fs_box(L, comp); }

    void sync() { HIPCHECK(hipStreamSynchronize(st)); }
    void check_launch() { HIPCHECK(hipGetLastError()); }

    // ---- field exchange ---------------------------------------------------
    // An x-fastest array over component comp's box grown by g ghost layers,
    // in host memory (through the staging buffer) or device memory.
    void put(double* dst, const double* src, int comp, int g = 0, bool device = false) {
        FsBox b = box(comp);
        const double* lin = src;
        if (!device) {
            size_t n = (size_t)(b.ni + 2 * g) * (b.nj + 2 * g) * (b.nk + 2 * g);
            EnsureStaging((long)n);
            HIPCHECK(hipMemcpyAsync(staging, src, n * sizeof(double), hipMemcpyHostToDevice, st));
            lin = staging;
        }
        fs_scatter(st, dst, lin, L, b, g);
        sync();
    }
    void get(const double* src, double* out, int comp, int g = 0, bool device = false) {
        FsBox b = box(comp);
        if (device) {
            fs_gather(st, out, src, L, b, g);
            sync();
            return;
        }
        size_t n = (size_t)(b.ni + 2 * g) * (b.nj + 2 * g) * (b.nk + 2 * g);
        EnsureStaging((long)n);
        fs_gather(st, staging, src, L, b, g);
        HIPCHECK(hipMemcpyAsync(out, staging, n * sizeof(double), hipMemcpyDeviceToHost, st));
        sync();
    }
    void EnsureStaging(long n) {
        if (n <= staging_n) return;
        sync();
        HIPCHECK(hipFree(staging));
        staging = nullptr;
        staging = (double*)dalloc(sizeof(double) * (size_t)n);
        staging_n = n;
    }
    // ghost layers of each field in the reference's arrays (fs_set_fields_ex
    // layout FS_LAYOUT_GHOSTED): phi and the MAC velocities 2
    // (EULERIAN_FLUID_SOLVER_3D.h:407-459), the pressure 1 (PROJECTION_3D.h:362)
    static int ref_ghost(int field) { return field == 4 ? 1 : 2; }

    void SetFields(const double* hphi, const double* hu, const double* hv, const double* hw, const double* hp,
                   int layout = FS_LAYOUT_INTERIOR, int memory = FS_MEM_HOST) {
        const bool dev = memory == FS_MEM_DEVICE;
        auto g = [&](int f) { return layout == FS_LAYOUT_GHOSTED ? ref_ghost(f) : 0; };
        if (hphi) {
            put(phi, hphi, 0, g(0), dev);
            // level-set ghosts by constant extrapolation, EULERIAN_FLUID_SOLVER_3D.h:757
            fs_fill_const(st, phi, phi, L, box(0), 2);
        }
        if (hu) put(u, hu, 1, g(1), dev);
        if (hv) put(v, hv, 2, g(2), dev);
        if (hw) put(w, hw, 3, g(3), dev);
        if (hp) put(p, hp, 0, g(4), dev);
        check_launch();
        sync();
    }
    void GetFields(double* hphi, double* hu, double* hv, double* hw, double* hp, int layout = FS_LAYOUT_INTERIOR,
                   int memory = FS_MEM_HOST) {
        const bool dev = memory == FS_MEM_DEVICE;
        auto g = [&](int f) { return layout == FS_LAYOUT_GHOSTED ? ref_ghost(f) : 0; };
        if (hphi) get(phi, hphi, 0, g(0), dev);
        if (hu) get(u, hu, 1, g(1), dev);
        if (hv) get(v, hv, 2, g(2), dev);
        if (hw) get(w, hw, 3, g(3), dev);
        if (hp) get(p, hp, 0, g(4), dev);
    }

    // ---- phases -----------------------------------------------------------
    // LEVELSET_3D::ComputeCurvatures, LEVELSET_3D.h:963-1082: phi's ghosts
    // re-extrapolated linearly (:998), then the curvature and its maximum
    // near the interface (the CFL rule's s_f).  With zero surface tension
    // nothing reads the curvature and only the ghost refresh is kept.  With_jc:
    // also jc = sigma * kappa (DetermineJumpConditionField, PROJECTION_3D.h:746-769).
    // curvature_method: by the normals first (fixed schedule: formed from
    // the current phi; compatible: those the last projection left), then by
    // the level set, which overwrites kappa; the maximum spans both passes.
    void ComputeCurvatures(bool with_jc = false) {
        fs_fill_linear(st, phi, phi, L, box(0), 2, 1);
        if (!tension()) return;
        HIPCHECK(hipMemsetAsync(kstat, 0, sizeof(unsigned long long) * 2, st));
        if (curv_normals()) {
            EnsureNormals();
            if (fixed()) fs_normals(st, phi, nrm, L, cell);
            fs_curvature_normals(st, nrm, kappa, with_jc && !curv_levelset() ? jc : nullptr, P.surface_tension, L,
                                 cell, kstat);
        }
        if (curv_levelset())
            fs_curvature(st, phi, kappa, with_jc ? jc : nullptr, P.surface_tension, L, cell, fixed() ? 1 : 0,
                         kstat);
        HIPCHECK(hipMemcpyAsync(hkstat, kstat, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, st));
    }
    bool curv_normals() const {
        return P.curvature_method == FS_CURV_NORMALS || P.curvature_method == FS_CURV_BOTH;
    }
    bool curv_levelset() const {
        return P.curvature_method == FS_CURV_LEVELSET || P.curvature_method == FS_CURV_BOTH;
    }
    // LEVELSET_3D::normal, zero until the first ComputeNormals
    double* nrm[3] = {nullptr, nullptr, nullptr};
    void EnsureNormals() {
        for (double*& f : nrm)
            if (!f) f = dfield();
    }
    // jump condition of the pressure solve, or NULL where it is zero: no
    // surface tension, or use_delta_function_formulation off (jc_on_solution
    // then keeps its initial zeros, PROJECTION_3D.h:358, :750)
    const double* jump() const {
        return tension() && P.use_delta_function_formulation && !csf_on && P.curvature_method != FS_CURV_NONE ? jc
                                                                                                             : nullptr;
    }
    // the projection's CSF_model flag, set by the first Sourcing with the
    // CSF model on and kept for the handle's life (as water_projection's)
    bool csf_on = false;
    // the curvature pass of a call reached a cell whose gradient vanishes in
    // both difference forms (the reference stops there, LEVELSET_3D.h:1041-1045);
    // call after a sync
    void check_curvature() const {
        if (tension() && hkstat[1] != 0)
            throw FsFail{FS_ERR_NONFINITE, "curvature: level-set gradient vanishes in both difference forms"};
    }

    bool fixed() const { return P.schedule == FS_SCHEDULE_FIXED; }
    // unequal spacings: the reference's diagonal (1/dx^2 for every face,
    // POISSON_SOLVER_3D.h:1213) differs from the operator's flux form
    bool cubic() const {
        double a = (double)1 / (cell.x.d * cell.x.d), b = (double)1 / (cell.y.d * cell.y.d),
               c = (double)1 / (cell.z.d * cell.z.d);
        return a == b && a == c;
    }
    // the fixed schedule's diagonal: each face's beta over its own d^2
    int axis_diag() const { return fixed() && !cubic() ? 1 : 0; }
    bool tension() const { return P.surface_tension != 0.0; }
    // WENO5 derivatives in the fewer-operation form (not bit for bit)
    bool fast() const { return P.fast_weno != 0; }

    // copy of field src (component box comp) onto the step tape
    double* Record(const double* src, int comp) {
        auto& pl = tape.pool[comp];
        size_t& n = tape.used[comp];
        if (n == pl.size()) pl.push_back(dfield());
        double* f = pl[n++];
        fs_copy_box(st, f, src, L, box(comp));
        return f;
    }

    // Ghost copies as functions of the current fields (fixed schedule): the
    // level set's by linear extrapolation of its own interior, layer g at
    // (g+1) phi_0 - g phi_1, faces only; the MAC velocities' with zero faces.
    void RefreshLevelsetGhost() {
        fs_copy_box(st, sghost, phi, L, box(0));
        fs_fill_linear(st, sghost, sghost, L, box(0), 3, 0);
    }
    void RefreshVelocityGhosts() {
        fs_copy_box(st, gu, u, L, box(1));
        fs_fill_zero_faces(st, gu, L, box(1), 3);
        fs_copy_box(st, gv, v, L, box(2));
        fs_fill_zero_faces(st, gv, L, box(2), 3);
        fs_copy_box(st, gw, w, L, box(3));
        fs_fill_zero_faces(st, gw, L, box(3), 3);
    }
    void ZeroWalls() {
        fs_zero_walls(st, 1, u, L);
        fs_zero_walls(st, 2, v, L);
        fs_zero_walls(st, 3, w, L);
    }

    // ADVECTION_3D::Solve_Levelset, ADVECTION_3D.h:204-242.  The compat
    // schedule reads the ghost copies the earlier phases left (the last
    // reinitialization sweep's, the last projection's); the fixed schedule
    // refreshes them first.
    void Levelset_Advection(double dt) {
        if (recording) tape.ls.push_back(Record(phi, 0));
        if (fixed()) {
            RefreshLevelsetGhost();
            RefreshVelocityGhosts();
        }
        fs_weno5_stage(st, phi, sghost, gu, gv, gw, L, 0, cell, P.levelset_epsilon, dt, fast());
    }

    // ADVECTION_3D::Solve_Velocity, ADVECTION_3D.h:244-358 (air/water).  The
    // reference's WENO5 branch passes the x ghost copy for all three
    // components (:314-316); the fixed schedule reads each component's own.
    void Velocity_Advection(double dt) {
        if (fixed()) RefreshVelocityGhosts();
        if (P.use_5th_weno_v) {
            bool own = fixed();
            fs_weno5_stage(st, u, gu, gu, gv, gw, L, 1, macx, P.velocity_epsilon, dt, fast());
            fs_weno5_stage(st, v, own ? gv : gu, gu, gv, gw, L, 2, macy, P.velocity_epsilon, dt, fast());
            fs_weno5_stage(st, w, own ? gw : gu, gu, gv, gw, L, 3, macz, P.velocity_epsilon, dt, fast());
        } else {
            fs_eno3_stage(st, u, gu, gu, gv, gw, L, 1, macx, dt);
            fs_eno3_stage(st, v, gv, gu, gv, gw, L, 2, macy, dt);
            fs_eno3_stage(st, w, gw, gu, gv, gw, L, 3, macz, dt);
        }
    }

    // ApplyGravity, EULERIAN_FLUID_SOLVER_3D.h:2237-2270
    // (the fixed schedule then zeroes the normal wall faces)
    void ApplyGravity(double dt) {
        fs_gravity(st, u, v, w, L, dt * P.gravity[0], dt * P.gravity[1], dt * P.gravity[2]);
        if (fixed()) ZeroWalls();
    }

    // Sourcing -> VISCOSITY_3D::ApplyViscosity (air/water explicit), each
    // component out of place into its spare buffer, then swapped in.  The
    // fixed schedule reads phi's ghosts as the constant extrapolation of the
    // current level set and zeroes each component's wall faces after its
    // update.
    // EULERIAN_FLUID_SOLVER_3D::Sourcing, :2076-2100: the viscosity, then the
    // objects' level-set sourcing when the scene has objects.
    int Sourcing(double dt) {
        int rc = ApplyViscosity(dt);
        if (!sources.empty()) SourcingFromObjectList();
        // CSF_model, :2089-2093: ApplySurfaceTension's curvature (and its
        // ghost refresh of phi); its force is oil/water-only (:2106-2195).
        // The projection's CSF flag it sets zeroes the jump condition from
        // the next projection on (PROJECTION_3D.h:754-757).
        if (P.csf_model) {
            ComputeCurvatures();
            csf_on = true;
        }
        return rc;
    }

    // SourcingFromObjectList, :2592-2671, spherical objects: each source in
    // its frame window over the cells of its box (SIMULATION_OBJECT::
    // BuildAABBGrid, SIMULATION_OBJECT.h:149-160: GRID_STRUCTURE_3D::ClampedCell
    // of the box corners, GRID_STRUCTURE_3D.h:310-319).
    void SourcingFromObjectList() {
        const FsAxis* ax[3] = {&cell.x, &cell.y, &cell.z};
        const double origin[3] = {G.x_min, G.y_min, G.z_min};
        const int res[3] = {G.nx, G.ny, G.nz};
        for (const fs_levelset_source& o : sources) {
            if (o.mode == FS_SOURCE_NONE || !(o.start_frame <= frame && frame <= o.stop_frame)) continue;
            int lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                double fl = std::floor((o.center[a] - o.half_width - origin[a]) * ax[a]->inv);
                double fh = std::floor((o.center[a] + o.half_width - origin[a]) * ax[a]->inv);
                lo[a] = (int)std::min(std::max(fl, 0.0), (double)(res[a] - 1));
                hi[a] = (int)std::min(std::max(fh, 0.0), (double)(res[a] - 1));
            }
            fs_source_sphere(st, phi, L, lo, hi, o.center, o.radius, cell, origin, o.mode);
        }
        check_launch();
    }

    // VISCOSITY_3D::ApplyViscosity (air/water explicit, or the semi-implicit
    // systems), VISCOSITY_3D.h:448-563
    int ApplyViscosity(double dt) {
        if (!P.use_delta_function_formulation) return FS_OK;
        if (fixed()) fs_fill_const(st, phi, phi, L, box(0), 2);
        if (P.semi_implicit_viscosity) {
            int rc = SemiImplicitViscosity(dt);
            if (fixed()) ZeroWalls();
            return rc;
        }
        if (recording) {
            RkRecord& r = tape.rk.back();
            r.u2 = Record(u, 1);
            r.v2 = Record(v, 2);
            r.w2 = Record(w, 3);
        }
        fs_viscosity(st, 1, spare_u, u, v, w, phi, L, mat, cell, dt);
        std::swap(u, spare_u);
        if (fixed()) fs_zero_walls(st, 1, u, L);
        if (recording) tape.rk.back().ux = Record(u, 1);
        fs_viscosity(st, 2, spare_v, u, v, w, phi, L, mat, cell, dt);
        std::swap(v, spare_v);
        if (fixed()) fs_zero_walls(st, 2, v, L);
        if (recording) tape.rk.back().vy = Record(v, 2);
        fs_viscosity(st, 3, spare_w, u, v, w, phi, L, mat, cell, dt);
        std::swap(w, spare_w);
        if (fixed()) fs_zero_walls(st, 3, w, L);
        return FS_OK;
    }

    // VISCOSITY_3D::ApplyViscosity, semi-implicit branch (VISCOSITY_3D.h:853-989):
    // the three systems are built from the velocities at entry (the spares
    // hold them) and solved in place one after the other with Jacobi-PCG
    // (POISSON_SOLVER_3D::SolveForViscosity, :254-266), warm-started from the
    // current component.  The unknowns are the component's faces, so the PCG
    // runs on a layout whose extents are that face box.
    int SemiImplicitViscosity(double dt) {
        fs_copy_box(st, spare_u, u, L, box(1));
        fs_copy_box(st, spare_v, v, L, box(2));
        fs_copy_box(st, spare_w, w, L, box(3));
        int rc = FS_OK;
        double* comps[3] = {u, v, w};
        for (int comp = 1; comp <= 3; comp++) {
            double* x = comps[comp - 1];
            fs_visc_system(st, comp, va, vdiag, vrhs, spare_u, spare_v, spare_w, x, phi, L, mat, cell, dt);
            FsBox b = box(comp);
            FsLayout Lf = L;
            Lf.nx = b.ni;
            Lf.ny = b.nj;
            Lf.nz = b.nk;
            // vectors of the previous component's box must not leave values
            // in this box's zero ghost ring
            for (void* f : {wv.r, wv.s, wv.p, wv.q, wv.p2})
                HIPCHECK(hipMemsetAsync(f, 0, sizeof(double) * (size_t)L.total, st));
            FsPcgOp o{};
            o.kind = FS_OP_ROWS;
            for (int e = 0; e < 6; e++) o.a[e] = va[e];
            o.dg = vdiag;
            int its = run_pcg_rows(Lf, o, x);
            stats.viscosity_iterations += its;
            if (hS->done != 1) {
                rc = FS_ERR_NOT_CONVERGED;
                last_error = "semi-implicit viscosity PCG reached max_iteration";
            }
        }
        return rc;
    }

    // Jacobi-PCG on an FS_OP_ROWS system with the PROJECTION_WATER stop rule.
    int run_pcg_rows(const FsLayout& Lf, const FsPcgOp& o, double* x) {
        init_state(S64, P.stop_rule, P.tolerance, P.rel_tolerance, P.max_iteration, false);
        fs_pcg64_begin(st, Lf, o, vrhs, vdiag, x, wv, S64);
        read_state(S64);
        int n = 0;
        while (!hS->done) {
            fs_pcg64_iterations(st, Lf, o, vdiag, x, wv, S64, n, chunk);
            n += chunk;
            check_launch();
            read_state(S64);
        }
        return hS->iter;
    }

    // PROJECTION_3D::Solve up to the linear system (:416-434, :596-834 and
    // POISSON_SOLVER_3D::Solve :165-246).
    void AssemblePressureSystem(double dt) {
        double light = P.air_bubble_rising ? P.air_density : P.water_density;
        double heavy = P.air_bubble_rising ? P.water_density : P.air_density;
        fs_density(st, phi, rho, irho, L, light, heavy);
        // DetermineDivergence: ghost copies of u, v, w with zero faces
        FsBox bu = box(1), bv = box(2), bw = box(3);
        fs_copy_box(st, gu, u, L, bu);
        fs_fill_zero_faces(st, gu, L, bu, 3);
        fs_copy_box(st, gv, v, L, bv);
        fs_fill_zero_faces(st, gv, L, bv, 3);
        fs_copy_box(st, gw, w, L, bw);
        fs_fill_zero_faces(st, gw, L, bw, 3);
        fs_divergence(st, gu, gv, gw, div, L, cell, 1 / dt);
        // DetermineJumpConditionField: ComputeCurvatures, then jc.  Without
        // surface tension its refresh of phi's ghosts is overwritten by
        // ComputeNormals below before anything reads it, so it is skipped.
        if (tension()) ComputeCurvatures(P.use_delta_function_formulation != 0 && !csf_on);
        if (!P.use_previous_pressure) fs_fill_value(st, p, L.total, 0.0);
        fs_fill_const(st, phi, phi, L, box(0), 2);   // ComputeNormals, LEVELSET_3D.h:931
        // ... and its normals, which the compatible schedule's next curvature
        // by normals reads (PROJECTION_3D.h:803)
        if (tension() && curv_normals() && !fixed()) {
            EnsureNormals();
            fs_normals(st, phi, nrm, L, cell);
        }
        fs_fill_const(st, irho, irho, L, box(0), 1); // PROJECTION_3D.h:813
        fs_assemble(st, phi, irho, div, jump(), bx, by, bz, diag, rhs, S64pin(), L, cell, axis_diag());
        op_phi = !std::getenv("FS_BETA_ARRAYS");
        fs_remove_mean(st, rhs, scratch, L);
        assembled = true;
    }

    FsPcgOp op64() const {
        FsPcgOp o{};
        o.bx = bx;
        o.by = by;
        o.bz = bz;
        o.idx2 = (double)1 / (cell.x.d * cell.x.d);
        o.axis_diag = fixed() ? 1 : 0;
        if (op_phi) {
            // 1/rho as k_density forms it: light where phi <= 0
            double light = P.air_bubble_rising ? P.air_density : P.water_density;
            double heavy = P.air_bubble_rising ? P.water_density : P.air_density;
            o.phi = phi;
            o.v_neg = 1 / light;
            o.v_pos = 1 / heavy;
        }
        o.idy2 = (double)1 / (cell.y.d * cell.y.d);
        o.idz2 = (double)1 / (cell.z.d * cell.z.d);
        return o;
    }

    // Host copy of the solver configuration into a device state, keeping the
    // pin extras computed on device.
    void init_state(FsPcgState* dS, int stop_rule, double tol, double rtol, int max_iter, bool with_pin = true) {
        FsPcgState h{};
        h.max_iter = max_iter;
        h.stop_rule = stop_rule;
        h.tol2 = tol * tol;
        h.rtol2 = rtol * rtol;
        HIPCHECK(hipMemcpyAsync(dS, &h, sizeof(h), hipMemcpyHostToDevice, st));
        // pin extras written by the diagonal kernel (fs_assemble)
        if (with_pin)
            HIPCHECK(hipMemcpyAsync(&dS->pin_b, &S64pin()[0], 2 * sizeof(double), hipMemcpyDeviceToDevice, st));
        sync();
    }

    // FP64 residual r = b - A p of the pinned system with its constant mode
    // settled in FP64 (fs_deflate_constant); returns r.r, sets bb.
    double true_residual(double& bb) {
        FsPcgOp o = op64();
        fs_residual64(st, L, o, rhs, p, (double*)w64.r, r32, w64, S64);
        read_state(S64);
        double pin = axis_diag() ? o.idy2 * hS->pin_b + o.idz2 * hS->pin_d : o.idx2 * (hS->pin_b + hS->pin_d);
        if (pin > 0) {
            fs_deflate_constant(st, L, p, (double*)w64.r, r32, S64, pin);
            read_state(S64);
        }
        bb = hS->bb;
        return hS->rr;
    }
    double* S64pin() { return scratch + G.nz + 2; }

    // After a launch batch and read_state: add the probed iteration's kernel
    // durations, if that iteration ran in full (it began before the stop).
    void harvest(int prec) {
        if (!sampling || !probe.recorded) return;
        probe.recorded = false;
        if (hS->done && hS->iter <= probe.iter) return;
        for (int k = 0; k < FS_PCG_KERNELS; k++) {
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, probe.ev[2 * k], probe.ev[2 * k + 1]));
            ksum[prec][k] += ms;
            kcount[prec][k]++;
        }
    }

    void read_state(FsPcgState* dS) {
        HIPCHECK(hipMemcpyAsync(hS, dS, sizeof(FsPcgState), hipMemcpyDeviceToHost, st));
        sync();
    }

    // The iteration kernels' tiling for this handle's grid (fs_pcg_tune),
    // once per precision: a concurrent reinitialization finishes first so
    // that the timings are the kernels' own.
    bool pcg_tuned[2] = {false, false};
    void TunePcg(const FsPcgOp& o, const void* dg, FsPcgWork& w, FsPcgState* dS, int prec) {
        if (pcg_tuned[prec]) return;
        if (st2) HIPCHECK(hipStreamSynchronize(st2));
        if (st3) HIPCHECK(hipStreamSynchronize(st3));
        fs_pcg_tune(st, L, o, dg, w, dS, prec);
        check_launch();
        pcg_tuned[prec] = true;
    }

    // Jacobi-PCG to completion; returns reference-style iteration count.
    int run_pcg64(int stop_rule, double tol, double rtol, int max_iter, double* x) {
        FsPcgOp o = op64();
        TunePcg(o, diag, w64, S64, 0);
        init_state(S64, stop_rule, tol, rtol, max_iter);
        fs_pcg64_begin(st, L, o, rhs, diag, x, w64, S64);
        read_state(S64);
        int n = 0;
        while (!hS->done) {
            fs_pcg64_iterations(st, L, o, diag, x, w64, S64, n, chunk, sampling ? &probe : nullptr);
            n += chunk;
            check_launch();
            read_state(S64);
            harvest(0);
        }
        stats.residual_precond = std::sqrt(hS->rs_last);
        return hS->iter;
    }

    FsPcgOp op32() const {
        FsPcgOp o = op64();
        if (op_phi) o.phi = phi32;
        o.bx = bx32;
        o.by = by32;
        o.bz = bz32;
        return o;
    }

    // FP32 Jacobi-PCG on the correction e with FP64 reliable updates: when
    // the FP32 recursive residual has dropped by `update_drop` since the last
    // update (or reached the target), e is folded into the FP64 pressure and
    // the residual b - A p is recomputed in FP64 and replaces the
    // recursion's residual between its r-update and p-update, so beta and the
    // new direction come from the refreshed residual and the Krylov history
    // is kept.  Returns FP32 iterations; `updates` counts FP64 residual
    // recomputations after the first.
    //
    // The reference's system is a Neumann operator L pinned at one cell,
    // A = L + pin e0 e0^T, with a zero-sum right-hand side.  A constant
    // shift of x changes the residual only in the pinned cell, so each FP64
    // residual first settles the constant (x += sum(r)/pin, leaving a
    // zero-sum residual) and the FP32 recursion solves the pin-free L on
    // zero-sum vectors: its corrections carry no large constant (which FP32
    // could not resolve the small variations on) and it converges faster
    // than the pinned system, whose tiny pin eigenvalue slows Jacobi-PCG.
    int run_pcg32_reliable(double& rr64, double& bb, int& updates) {
        FsPcgOp o32 = op32();
        TunePcg(o32, diag32, w32, S32, 1);
        const double rt2 = P.rel_tolerance * P.rel_tolerance;
        init_state(S64, FS_STOP_REL_RESIDUAL, 0, P.rel_tolerance, 0);
        rr64 = true_residual(bb);
        updates = 0;
        if (bb == 0 || rr64 <= rt2 * bb) return 0;
        init_state(S32, FS_STOP_REL_RESIDUAL, 0, P.rel_tolerance, P.max_iteration, false);
        HIPCHECK(hipMemsetAsync(e32, 0, sizeof(float) * L.total, st));
        fs_pcg32_begin(st, L, o32, r32, diag32, e32, w32, S32);
        // the recursion stops against the global target ||b||
        hbb = bb;
        HIPCHECK(hipMemcpyAsync(&S32->bb, &hbb, sizeof(double), hipMemcpyHostToDevice, st));
        double rr_ref = rr64;
        int n = 0, done_before = 0, n_update = 0;
        while (true) {
            fs_pcg32_chunk(st, L, o32, diag32, e32, w32, S32, n, chunk, sampling ? &probe : nullptr);
            check_launch();
            read_state(S32);
            harvest(1);
            if (hS->done == 2) {
                n = hS->iter;
                break;
            }
            // a recursion that has not dropped the residual 100x in
            // FS_FP32_RESTART iterations has stagnated at FP32's attainable
            // accuracy: fold e into the pressure and restart it from the
            // FP64 residual (restarted CG; the 64-bubble pool at 512^3 needs it)
            if (hS->done == 0 && n + chunk - n_update >= FS_FP32_RESTART && !(hS->rr < update_drop * rr_ref)) {
                n += chunk;
                fs_add_correction(st, L, p, e32);
                HIPCHECK(hipMemsetAsync(e32, 0, sizeof(float) * L.total, st));
                rr64 = true_residual(bb);
                updates++;
                done_before += n;
                if (debug) fprintf(stderr, "[fs] fp32 restart after %d rel64=%.3e\n", done_before, std::sqrt(rr64 / bb));
                if (rr64 <= rt2 * bb || updates > 400 || done_before >= P.max_iteration) {
                    n = 0;
                    break;
                }
                init_state(S32, FS_STOP_REL_RESIDUAL, 0, P.rel_tolerance, P.max_iteration - done_before, false);
                fs_pcg32_begin(st, L, o32, r32, diag32, e32, w32, S32);
                HIPCHECK(hipMemcpyAsync(&S32->bb, &hbb, sizeof(double), hipMemcpyHostToDevice, st));
                rr_ref = rr64;
                n = n_update = 0;
                continue;
            }
            bool stopped = hS->done == 1;
            if (debug && dbg_lines++ < 120)
                fprintf(stderr, "[fs]   chunk n0=%d done=%d iter=%d rel32=%.3e rz=%.3e,%.3e\n", n, hS->done, hS->iter,
                        std::sqrt(hS->rr / bb), hS->rz[0], hS->rz[1]);
            // the last iteration that ran: the one that met the target, or
            // the chunk's last; the next K1 continues from it
            int last = stopped ? hS->iter : n + chunk - 1;
            n = last + 1;
            double rr32 = hS->rr;
            if (stopped || rr32 < update_drop * rr_ref) {
                fs_add_correction(st, L, p, e32);
                HIPCHECK(hipMemsetAsync(e32, 0, sizeof(float) * L.total, st));
                rr64 = true_residual(bb);
                updates++;
                if (debug)
                    fprintf(stderr, "[fs] fp32 n=%d stopped=%d rel32=%.3e rel64=%.3e\n", n, (int)stopped,
                            std::sqrt(rr32 / bb), std::sqrt(rr64 / bb));
                if (rr64 <= rt2 * bb || updates > 400) break;
                fs_pcg32_reliable_update(st, L, r32, diag32, w32, S32);
                rr_ref = rr64;
                n_update = n;
            }
        }
        return done_before + n;
    }

    // ---- FS_FP32_MG ----------------------------------------------------------
    // CG on the FP32 pin-free operator with one multigrid V-cycle as the
    // preconditioner (fs_mg.h), from e = 0 for the zero-sum residual r32,
    // until ||r|| <= rtol ||r32|| or max_it.  Returns the iterations.
    int run_mgpcg32(double rtol, int max_it) {
        const FsLayout& Lz = L;
        const int nb = fs_mgc_blocks(Lz);
        if (!mg_part) {
            mg_part = (double*)dalloc(sizeof(double) * 2 * (size_t)nb);
            mg_mid = (double*)dalloc(sizeof(double) * 2 * FS_MG_MID);
            mg_cg = (FsMgCg*)dalloc(sizeof(FsMgCg));
            HIPCHECK(hipHostMalloc((void**)&hmg_cg, sizeof(FsMgCg)));
        }
        const float ix = (float)cell.x.inv_d2, iy = (float)cell.y.inv_d2, iz = (float)cell.z.inv_d2;
        float *x = e32, *r = (float*)w32.r, *z = (float*)w32.s, *p = (float*)w32.p, *q = (float*)w32.q;
        fs_mg_zero(st, x, L.total, nullptr);
        HIPCHECK(hipMemcpyAsync(r, r32, sizeof(float) * L.total, hipMemcpyDeviceToDevice, st));
        fs_mgc_start(st, mg_cg, rtol * rtol, max_it);
        mg.skip = nullptr;
        // the in-loop probe samples this V-cycle, which always runs; the
        // loop's finest-level launches are counted below
        FsLaunchProbe* kp = mg.probe;
        mg.Apply(st, r, z);
        mg.probe = nullptr;
        fs_mgc_dot2(st, Lz, r, z, mg_part, mg_cg);
        fs_mgc_total(st, mg_part, nb, 2, mg_mid, mg_cg, FS_MGC_INIT);
        HIPCHECK(hipMemcpyAsync(p, z, sizeof(float) * L.total, hipMemcpyDeviceToDevice, st));
        // Iterations are enqueued in batches with the scalars on the device;
        // the host reads the state after each batch.  A batch is as long as
        // the convergence rate seen so far says is left (2 at first, at most
        // 8), and iterations past the stop return at once (`done`, which
        // the V-cycle reads through mg.skip).
        mg.skip = &mg_cg->done;
        int enq = 0, batch = 2;
        hmg_cg->done = 0;
        hmg_cg->it = 0;
        while (enq < max_it) {
            const int k = std::min(batch, max_it - enq);
            for (int n = 0; n < k; n++) {
                fs_mgc_spmv_dot(st, Lz, bx32, by32, bz32, ix, iy, iz, p, q, mg_part, mg_cg);
                fs_mgc_total(st, mg_part, nb, 1, mg_mid, mg_cg, FS_MGC_PQ);
                fs_mgc_update(st, Lz, x, r, p, q, mg_part, mg_cg);
                fs_mgc_total(st, mg_part, nb, 1, mg_mid, mg_cg, FS_MGC_RR);
                mg.Apply(st, r, z);
                fs_mgc_dot2(st, Lz, r, z, mg_part, mg_cg);
                fs_mgc_total(st, mg_part, nb, 2, mg_mid, mg_cg, FS_MGC_RZ);
                fs_mgc_direction(st, Lz, z, p, mg_cg);
            }
            enq += k;
            check_launch();
            HIPCHECK(hipMemcpyAsync(hmg_cg, mg_cg, sizeof(FsMgCg), hipMemcpyDeviceToHost, st));
            sync();
            if (hmg_cg->done) break;
            // iterations left at the mean rate so far: rr falls by rate^it
            const double target = hmg_cg->rtol2 * hmg_cg->rr0;
            if (hmg_cg->it > 0 && hmg_cg->rr > 0 && hmg_cg->rr < hmg_cg->rr0 && target > 0) {
                const double per = std::log(hmg_cg->rr / hmg_cg->rr0) / hmg_cg->it;
                const double left = std::log(target / hmg_cg->rr) / per;
                batch = (int)std::min(8.0, std::max(1.0, std::ceil(left)));
            } else {
                batch = 2;
            }
        }
        mg.skip = nullptr;
        mg.probe = kp;
        if (kp) {
            // V-cycles the loop ran: one per update, but for the last when
            // the stop came right after it
            const int d = hmg_cg->done;
            kp->launches += (long)std::max(0, hmg_cg->it - (d == 1 || d == 3 ? 1 : 0)) * mg.FinestLaunchesPerCycle();
        }
        return hmg_cg->it;
    }

    // FP32 multigrid-preconditioned CG inside FP64 iterative refinement:
    // each sweep solves the FP32 correction to FS_MG_INNER_RTOL of its
    // right-hand side
    // and adds it to the FP64 pressure; the FP64 residual (its constant mode
    // settled through the pin, as in the Jacobi refinement) decides the stop.
    void SolvePressureMg() {
        mg.Init(L, cell.x.inv_d2, cell.y.inv_d2, cell.z.inv_d2);
        // two sweeps each way at low density contrast; four from 10:1, where
        // they halve the CG iterations (air/water dam break 107 -> 59, 64
        // bubbles 263 -> 133 per frame at 256^3) for twice the smoothing
        {
            const double hi = std::max(P.water_density, P.air_density), lo = std::min(P.water_density, P.air_density);
            int nu = lo > 0 && hi / lo < 10 ? FS_MG_SWEEPS_LOW : FS_MG_SWEEPS_HIGH;
            if (const char* e = std::getenv("FS_MG_SWEEPS")) nu = std::max(1, std::atoi(e));  // tuning override
            mg.SetSweeps(nu);
        }
        mg.probe = sampling ? &kprobe[FS_KS_MG_SMOOTH] : nullptr;
        mg.SetOperator(st, bx32, by32, bz32);
        if (!mg.Tuned()) {
            // time the sweeps with the GPU to themselves: let a concurrent
            // reinitialization finish first (once per handle)
            if (st2) HIPCHECK(hipStreamSynchronize(st2));
            if (st3) HIPCHECK(hipStreamSynchronize(st3));
            mg.Tune(st, (float*)w32.r, (float*)w32.s);
        }
        init_state(S64, FS_STOP_REL_RESIDUAL, 0, P.rel_tolerance, 0);
        const double rt2 = P.rel_tolerance * P.rel_tolerance;
        double bb = 0, rr = true_residual(bb);
        int sweeps = 0, its = 0;
        const bool fixed_inner = std::getenv("FS_MG_FIXED_INNER") != nullptr;
        while (bb > 0 && rr > rt2 * bb && sweeps < 40) {
            // Each sweep's FP32 solve asks for the reduction still missing
            // (half of it again, as a margin), between FS_MG_INNER_RTOL --
            // about what FP32 attains -- and 0.1: a sweep that starts close
            // to the target then takes one or two iterations, not a full
            // FS_MG_INNER_RTOL reduction past it.
            double inner = FS_MG_INNER_RTOL;
            if (!fixed_inner) {
                const double need = 0.5 * P.rel_tolerance / std::sqrt(rr / bb);
                inner = std::min(0.1, std::max((double)FS_MG_INNER_RTOL, need));
            }
            // restarted CG: at most 500 iterations per sweep
            its += run_mgpcg32(inner, std::min(P.max_iteration, 500));
            fs_add_correction(st, L, p, e32);
            rr = true_residual(bb);
            sweeps++;
            if (debug) fprintf(stderr, "[fs] mg sweep %d its %d rel64=%.3e\n", sweeps, its, std::sqrt(rr / bb));
        }
        stats.pcg_iterations += its;
        stats.refinement_sweeps += sweeps;
        stats.residual_rel = bb > 0 ? std::sqrt(rr / bb) : 0;
        stats.residual_precond = 0;
        if (bb > 0 && rr > rt2 * bb) throw FsFail{FS_ERR_NOT_CONVERGED, "multigrid refinement did not reach rel_tolerance"};
    }

    // Pressure solve on p (warm start), FP64 or FP32-inner / FP64-refined.
    void SolvePressure() {
        FsPcgOp o = op64();
        if (P.precision == FS_FP64) {
            // With the relative stop rule the recursion's residual can drift
            // from the true one over thousands of iterations (the 64-bubble
            // pool at 512^3 stopped at a true 1.8e-7 against 1e-8): PCG is
            // restarted from the true residual b - A p, warm-started at p,
            // until the true residual meets the target.  A restart that does
            // not halve the true residual has reached FP64's attainable
            // accuracy for this system; the solve then reports
            // FS_ERR_NOT_CONVERGED with that residual, as it does when the
            // iteration budget runs out.  The reference's absolute rule stops
            // on the recursion alone (PCG_METHOD.cpp:81-90).
            const bool rel = P.stop_rule == FS_STOP_REL_RESIDUAL;
            const double rt2 = P.rel_tolerance * P.rel_tolerance;
            int its = 0, restarts = 0;
            bool converged = false, stagnated = false;
            double prev = INFINITY;
            while (true) {
                its += run_pcg64(P.stop_rule, P.tolerance, P.rel_tolerance, P.max_iteration - its, p);
                converged = hS->done == 1;
                fs_residual64(st, L, o, rhs, p, (double*)w64.r, r32 ? r32 : (float*)w64.q, w64, S64);
                read_state(S64);
                const double rr = hS->rr, bb = hS->bb;
                stats.residual_rel = bb > 0 ? std::sqrt(rr / bb) : 0;
                if (!converged || !rel || bb == 0 || rr <= rt2 * bb) break;
                if (its >= P.max_iteration || !(rr < (double)0.25 * prev)) {
                    converged = false;
                    stagnated = its < P.max_iteration;
                    break;
                }
                prev = rr;
                restarts++;
                stats.refinement_sweeps++;  // one restart from the true residual
                if (debug) fprintf(stderr, "[fs] fp64 restart %d after %d its, true rel %.3e\n", restarts, its,
                                   stats.residual_rel);
            }
            stats.pcg_iterations += its;
            if (!converged) {
                char msg[160];
                if (stagnated)
                    snprintf(msg, sizeof msg,
                             "PCG stagnated at a true relative residual of %.3e, above rel_tolerance %.3e",
                             stats.residual_rel, P.rel_tolerance);
                else
                    snprintf(msg, sizeof msg, "PCG reached max_iteration (true relative residual %.3e)",
                             stats.residual_rel);
                throw FsFail{FS_ERR_NOT_CONVERGED, msg};
            }
            return;
        }
        // FP32 inner PCG inside FP64 iterative refinement: the outer residual
        // b - A x, the correction update and all dot products stay FP64.
        (void)o;
        fs_make_op32(st, L, bx, by, bz, diag, bx32, by32, bz32, diag32);
        if (P.precision == FS_FP32_MG) {
            SolvePressureMg();
            return;
        }
        if (op_phi) fs_to_float(st, phi, phi32, L.total);
        double rr = 0, bb = 0;
        int updates = 0;
        stats.pcg_iterations += run_pcg32_reliable(rr, bb, updates);
        stats.refinement_sweeps += updates;
        stats.residual_rel = bb > 0 ? std::sqrt(rr / bb) : 0;
        stats.residual_precond = 0;
        double rt2 = P.rel_tolerance * P.rel_tolerance;
        if (bb > 0 && rr > rt2 * bb) throw FsFail{FS_ERR_NOT_CONVERGED, "refinement did not reach rel_tolerance"};
    }

    // PROJECTION_3D::UpdateVelocity, :836-841 (z faces only in air/water
    // mode).  The fixed schedule corrects every face with the matrix's own
    // face coefficients.
    void UpdateVelocity(double dt) {
        HIPCHECK(hipMemsetAsync(vmax, 0, sizeof(unsigned long long) * 4, st));
        if (fixed()) {
            fs_gradient_update_beta(st, 1, u, p, bx, phi, jump(), L, cell, dt, vmax + 0);
            fs_gradient_update_beta(st, 2, v, p, by, phi, jump(), L, cell, dt, vmax + 1);
            fs_gradient_update_beta(st, 3, w, p, bz, phi, jump(), L, cell, dt, vmax + 2);
        } else {
            fs_gradient_update(st, 1, 0, u, p, phi, rho, jump(), L, mat, cell, dt, vmax + 0);
            fs_gradient_update(st, 2, 0, v, p, phi, rho, jump(), L, mat, cell, dt, vmax + 1);
            fs_gradient_update(st, 3, 1, w, p, phi, rho, jump(), L, mat, cell, dt, vmax + 2);
        }
        HIPCHECK(hipMemcpyAsync(hvmax, vmax, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost, st));
        sync();
        check_curvature();
        for (int c = 0; c < 3; c++) {
            double m;
            memcpy(&m, &hvmax[c], sizeof(double));
            stats.max_velocity[c] = m;
        }
    }

    // PROJECTION_3D::Solve, PROJECTION_3D.h:416-594
    int Projection(double dt) {
        HIPCHECK(hipEventRecord(ev[2], st));
        AssemblePressureSystem(dt);
        // pin extras of cell (0,0,0) for the operator kernels
        int rc = FS_OK;
        try {
            SolvePressure();
        } catch (FsFail& f) {
            if (f.code != FS_ERR_NOT_CONVERGED) throw;
            rc = f.code;
            last_error = f.msg;
        }
        UpdateVelocity(dt);
        HIPCHECK(hipEventRecord(ev[3], st));
        sync();
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, ev[2], ev[3]));
        stats.projection_ms += ms;
        for (int c = 0; c < 3; c++)
            if (!std::isfinite(stats.max_velocity[c])) throw FsFail{FS_ERR_NONFINITE, "non-finite velocity after projection"};
        return rc;
    }

    // Adjoint of Projection at the current interface (phi held fixed):
    // u' = u + Gh A^-1 M B u, so lam <- lam + B^T M A^-1 Gh^T lam on the face
    // fields lu, lv, lw (device,
    // zero ghosts).  The operator is rebuilt from phi exactly as Projection
    // builds it, without the persistent ghost copies the forward refreshes;
    // the solve is the FP64 Jacobi-PCG with the forward's stop rule.
    int AdjointProjection(double dt, double* lu, double* lv, double* lw) {
        double light = P.air_bubble_rising ? P.air_density : P.water_density;
        double heavy = P.air_bubble_rising ? P.water_density : P.air_density;
        fs_density(st, phi, rho, irho, L, light, heavy);
        fs_fill_const(st, phi, phi, L, box(0), 2);  // as AssemblePressureSystem
        fs_fill_const(st, irho, irho, L, box(0), 1);
        fs_fill_value(st, div, L.total, 0.0);
        fs_assemble(st, phi, irho, div, nullptr, bx, by, bz, diag, rhs, S64pin(), L, cell, axis_diag());
        op_phi = false;
        assembled = true;
        // rhs = G^T lam (the faces the forward corrects)
        fs_fill_value(st, rhs, L.total, 0.0);
        if (fixed()) {
            fs_adjoint_gradient_beta(st, 1, lu, bx, rhs, L, cell, dt);
            fs_adjoint_gradient_beta(st, 2, lv, by, rhs, L, cell, dt);
            fs_adjoint_gradient_beta(st, 3, lw, bz, rhs, L, cell, dt);
        } else {
            fs_adjoint_gradient(st, 3, lw, phi, rho, rhs, L, mat, cell, dt);
        }
        // y = A^-1 rhs (A symmetric), then M y
        fs_fill_value(st, tmp, L.total, 0.0);
        int its = run_pcg64(P.stop_rule, P.tolerance, P.rel_tolerance, P.max_iteration, tmp);
        bool converged = hS->done == 1;
        fs_remove_mean(st, tmp, scratch, L);
        // rhs holds Gh^T lam with Gh = dU/dp (the minus sign of the update
        // included), so lam <- lam + B^T M A^-1 Gh^T lam with B = -D/dt
        fs_adjoint_divergence(st, lu, lv, lw, tmp, L, cell, -1 / dt);
        check_launch();
        sync();
        stats.pcg_iterations = its;
        return converged ? FS_OK : FS_ERR_NOT_CONVERGED;
    }

    // Adjoint of the explicit Sourcing at the current level set: the x, y, z
    // updates run in that order, each reading the components already
    // updated, so their transposes run z, y, x; each scatters from a copy of
    // its component's adjoint.
    void AdjointSourcing(double dt, double* lu, double* lv, double* lw) {
        if (!P.use_delta_function_formulation) return;
        if (P.semi_implicit_viscosity)
            throw FsFail{FS_ERR_UNSUPPORTED, "adjoint of the semi-implicit viscosity is not built"};
        double* lam[3] = {lu, lv, lw};
        for (int comp = 3; comp >= 1; comp--) {
            fs_copy_box(st, tmp, lam[comp - 1], L, box(comp));
            fs_adjoint_viscosity(st, comp, tmp, lu, lv, lw, phi, L, mat, cell, dt);
        }
        check_launch();
        sync();
    }

    // ---- reverse pass of one fixed-schedule step ---------------------------
    // The step is differentiated as the fixed schedule computes it: every
    // ghost copy is a function of the current fields, the gradient update
    // uses the matrix's coefficients, the Godunov norm is Euclidean, so the
    // step is piecewise smooth.  Discrete choices (upwind side, Godunov
    // case, the +-eps clamp, the phase of a cell) are held at the forward's.
    void CheckAdjointSupport() const {
        if (!fixed()) throw FsFail{FS_ERR_UNSUPPORTED, "the adjoint step differentiates the fixed schedule"};
        if (P.use_delta_function_formulation && P.is_sourcing_active && P.semi_implicit_viscosity)
            throw FsFail{FS_ERR_UNSUPPORTED, "adjoint of the semi-implicit viscosity is not built"};
        if (P.is_velocity_advection_active && !P.use_5th_weno_v)
            throw FsFail{FS_ERR_UNSUPPORTED, "adjoint velocity advection needs use_5th_weno_v (ENO3's stencil choice "
                                             "is discontinuous)"};
        if (tension() && P.use_delta_function_formulation && P.is_projection_active)
            throw FsFail{FS_ERR_UNSUPPORTED, "adjoint of the surface-tension jump (curvature of phi) is not built"};
        if (!sources.empty() && P.is_sourcing_active)
            throw FsFail{FS_ERR_UNSUPPORTED, "adjoint of the level-set sources (a min/max of phi) is not built"};
        if (P.csf_model && P.is_sourcing_active)
            throw FsFail{FS_ERR_UNSUPPORTED, "adjoint with the CSF model is not built"};
    }

    void EnsureAdjoint() {
        if (A.ready) return;
        double** f[] = {&A.lphi, &A.lu, &A.lv, &A.lw, &A.ladv, &A.a0, &A.a1, &A.lsg, &A.ghw, &A.y, &A.phiw, &A.tgt,
                        &A.aU, &A.aV, &A.aW, &A.src, &A.du, &A.dv, &A.dw};
        for (double** x : f) *x = dfield();
        A.part = (double*)dalloc(sizeof(double) * (FS_LOSS_BLOCKS + 8));
        A.ready = true;
    }

    void Zero(double* a) { fs_fill_value(st, a, L.total, 0.0); }
    // dst = ca * a + cb * dst on component comp's box
    void Blend(double* dst, const double* a, double ca, double cb, int comp) {
        fs_combine(st, dst, a, ca, cb, L, box(comp));
    }
    void VelocityGhostsFrom(const double* su, const double* sv, const double* sw) {
        const double* s3[3] = {su, sv, sw};
        double* g3[3] = {gu, gv, gw};
        for (int c = 0; c < 3; c++) {
            fs_copy_box(st, g3[c], s3[c], L, box(c + 1));
            fs_fill_zero_faces(st, g3[c], L, box(c + 1), 3);
        }
    }

    // lam <- (dR/dphi)^T lam for the recorded reinitialization call `call`
    // (identity when reinitialization is off); lam and spare swap roles.
    void ReverseReinitCall(size_t call, double*& lam, double*& spare) {
        if (call >= tape.reinit.size()) return;
        const std::vector<double*>& sweeps = tape.reinit[call];
        if (sweeps.empty()) return;
        double fictious_dt = cell.x.d / P.scaling_number_for_reinitialization;
        double dx2 = cell.x.d * cell.x.d;
        fs_sign(st, sweeps[0], sgn, L, dx2);
        Zero(A.lsg);
        for (int s = (int)sweeps.size() - 1; s >= 0; s--) {
            fs_copy_box(st, A.ghw, sweeps[s], L, box(0));
            fs_fill_linear(st, A.ghw, A.ghw, L, box(0), 3, 0);
            Zero(spare);
            fs_adj_reinit_sweep(st, lam, sweeps[s], A.ghw, sgn, spare, A.lsg, L, cell, P.levelset_epsilon,
                                fictious_dt);
            fs_adj_fold_linear(st, spare, L, 3);
            std::swap(lam, spare);
        }
        fs_adj_sign(st, sweeps[0], A.lsg, lam, L, dx2);
    }

    // (A.lu, A.lv, A.lw) <- adjoint of one SolveForRK stage's input; the
    // level-set parts (viscosity and projection coefficients) add to A.ladv.
    int ReverseRk(size_t r, double dt) {
        const RkRecord& R = tape.rk.at(r);
        double* lam[3] = {A.lu, A.lv, A.lw};
        int rc = FS_OK;
        // the stage's level set with constant ghosts, as Sourcing and the
        // projection's assembly read it
        fs_copy_box(st, A.phiw, R.phi, L, box(0));
        fs_fill_const(st, A.phiw, A.phiw, L, box(0), 2);
        if (P.is_projection_active) {
            double light = P.air_bubble_rising ? P.air_density : P.water_density;
            double heavy = P.air_bubble_rising ? P.water_density : P.air_density;
            fs_density(st, A.phiw, rho, irho, L, light, heavy);
            fs_fill_const(st, irho, irho, L, box(0), 1);
            Zero(div);
            fs_assemble(st, A.phiw, irho, div, nullptr, bx, by, bz, diag, rhs, S64pin(), L, cell, axis_diag());
            op_phi = false;
            assembled = true;
            Zero(rhs);
            fs_adjoint_gradient_beta(st, 1, A.lu, bx, rhs, L, cell, dt);
            fs_adjoint_gradient_beta(st, 2, A.lv, by, rhs, L, cell, dt);
            fs_adjoint_gradient_beta(st, 3, A.lw, bz, rhs, L, cell, dt);
            Zero(A.y);
            stats.pcg_iterations += run_pcg64(P.stop_rule, P.tolerance, P.rel_tolerance, P.max_iteration, A.y);
            if (hS->done != 1) {
                rc = FS_ERR_NOT_CONVERGED;
                last_error = "adjoint pressure solve reached max_iteration";
            }
            fs_adj_projection_phi(st, A.phiw, irho, R.p, A.y, A.lu, A.lv, A.lw, A.ladv, L, cell, dt);
            fs_remove_mean(st, A.y, scratch, L);
            fs_adjoint_divergence(st, A.lu, A.lv, A.lw, A.y, L, cell, -1 / dt);
        }
        if (P.is_sourcing_active && P.use_delta_function_formulation) {
            Zero(A.a0);
            const double* U[3] = {R.u2, R.ux, R.ux};  // what the x, y, z updates read
            const double* V[3] = {R.v2, R.v2, R.vy};
            for (int comp = 3; comp >= 1; comp--) {
                fs_zero_walls(st, comp, lam[comp - 1], L);
                fs_copy_box(st, A.src, lam[comp - 1], L, box(comp));
                fs_adjoint_viscosity(st, comp, A.src, A.lu, A.lv, A.lw, A.phiw, L, mat, cell, dt);
                fs_adj_viscosity_phi(st, comp, A.src, U[comp - 1], V[comp - 1], R.w2, A.phiw, A.a0, L, mat, cell, dt);
            }
            for (int c = 1; c <= 3; c++) fs_zero_outside(st, lam[c - 1], L, box(c));
            fs_adj_fold_const(st, A.a0, L, 2);
            Blend(A.ladv, A.a0, 1, 1, 0);
        }
        if (P.is_external_force_active)
            for (int c = 1; c <= 3; c++) fs_zero_walls(st, c, lam[c - 1], L);
        if (P.is_velocity_advection_active) {
            VelocityGhostsFrom(R.u, R.v, R.w);
            double* acc[3] = {A.aU, A.aV, A.aW};
            const double* g3[3] = {gu, gv, gw};
            const FsGridConst* mg[3] = {&macx, &macy, &macz};
            for (double* a : acc) Zero(a);
            for (int comp = 1; comp <= 3; comp++)
                fs_adj_weno5_stage(st, lam[comp - 1], g3[comp - 1], gu, gv, gw, acc[comp - 1], A.aU, A.aV, A.aW, L,
                                   comp, *mg[comp - 1], P.velocity_epsilon, dt);
            for (int c = 1; c <= 3; c++) {
                fs_zero_outside(st, acc[c - 1], L, box(c));
                Blend(lam[c - 1], acc[c - 1], 1, 1, c);
            }
        }
        return rc;
    }

    // lam <- (d phi'/d phi)^T lam for a level-set stage from `phik`; the
    // sample velocity's part accumulates in A.aU, A.aV, A.aW (gu, gv, gw
    // hold the velocity ghosts the stage read).
    void ReverseLevelsetStage(const double* phik, double* lam, double dt) {
        fs_copy_box(st, A.ghw, phik, L, box(0));
        fs_fill_linear(st, A.ghw, A.ghw, L, box(0), 3, 0);
        Zero(A.a0);
        fs_adj_weno5_stage(st, lam, A.ghw, gu, gv, gw, A.a0, A.aU, A.aV, A.aW, L, 0, cell, P.levelset_epsilon, dt);
        fs_adj_fold_linear(st, A.a0, L, 3);
        Blend(lam, A.a0, 1, 1, 0);
    }

    // Reverse pass over the recorded step: on entry A.lphi, A.lu, A.lv, A.lw
    // hold the adjoint of the step's output, on return of its input.
    int ReverseStep(double dt) {
        const int order = P.order_for_time_advancing;
        int rc = FS_OK;
        // reinitialization and its Runge-Kutta blends with tmp = phi_adv
        double *lam = A.lphi, *spare = A.a1;
        Zero(A.ladv);
        if (order == 3) {
            Blend(A.ladv, lam, (double)1 / 3, 0, 0);
            Blend(lam, lam, 0, (double)2 / 3, 0);
            ReverseReinitCall(2, lam, spare);
            Blend(A.ladv, lam, (double)0.75, 1, 0);
            Blend(lam, lam, 0, (double)0.25, 0);
        } else {
            Blend(A.ladv, lam, (double)0.5, 0, 0);
            Blend(lam, lam, 0, (double)0.5, 0);
        }
        ReverseReinitCall(1, lam, spare);
        ReverseReinitCall(0, lam, spare);
        Blend(A.ladv, lam, 1, 1, 0);
        // velocity stages and their blends with the step's input velocity
        double* lv3[3] = {A.lu, A.lv, A.lw};
        double* d3[3] = {A.du, A.dv, A.dw};
        auto rk = [&](size_t r) {
            int e = ReverseRk(r, dt);
            if (e != FS_OK) rc = e;
        };
        if (order == 1) {
            rk(0);
        } else {
            for (int c = 0; c < 3; c++) {
                Blend(d3[c], lv3[c], order == 2 ? (double)0.5 : (double)1 / 3, 0, c + 1);
                Blend(lv3[c], lv3[c], 0, order == 2 ? (double)0.5 : (double)2 / 3, c + 1);
            }
            if (order == 3) {
                rk(2);
                for (int c = 0; c < 3; c++) {
                    Blend(d3[c], lv3[c], (double)0.75, 1, c + 1);
                    Blend(lv3[c], lv3[c], 0, (double)0.25, c + 1);
                }
            }
            rk(1);
            rk(0);
            for (int c = 0; c < 3; c++) Blend(lv3[c], d3[c], 1, 1, c + 1);
        }
        // level-set advection with the step's input velocity
        if (P.is_levelset_advection_active && !tape.ls.empty()) {
            VelocityGhostsFrom(tape.u, tape.v, tape.w);
            for (double* a : {A.aU, A.aV, A.aW}) Zero(a);
            if (order == 3) {
                ReverseLevelsetStage(tape.ls[0], A.ladv, dt);
            } else {
                Blend(A.a1, A.ladv, (double)0.5, 0, 0);
                Blend(A.ladv, A.ladv, 0, (double)0.5, 0);
                ReverseLevelsetStage(tape.ls[1], A.ladv, dt);
                ReverseLevelsetStage(tape.ls[0], A.ladv, dt);
                Blend(A.ladv, A.a1, 1, 1, 0);
            }
            double* acc[3] = {A.aU, A.aV, A.aW};
            for (int c = 1; c <= 3; c++) {
                fs_zero_outside(st, acc[c - 1], L, box(c));
                Blend(lv3[c - 1], acc[c - 1], 1, 1, c);
            }
        }
        fs_copy_box(st, A.lphi, A.ladv, L, box(0));
        check_launch();
        return rc;
    }

    // the recorded step, from the current fields
    int RecordStep(double dt) {
        recording = true;
        int rc;
        try {
            rc = AdvanceOneTimeStepThread(dt);
        } catch (...) {
            recording = false;
            throw;
        }
        recording = false;
        return rc;
    }

    SavedState& Saved(size_t n) {
        while (saved.size() <= n) {
            SavedState sv;
            sv.phi = dfield();
            sv.u = dfield();
            sv.v = dfield();
            sv.w = dfield();
            sv.p = dfield();
            saved.push_back(sv);
        }
        return saved[n];
    }
    void SaveState(SavedState& sv) {
        fs_copy_box(st, sv.phi, phi, L, box(0));
        fs_copy_box(st, sv.u, u, L, box(1));
        fs_copy_box(st, sv.v, v, L, box(2));
        fs_copy_box(st, sv.w, w, L, box(3));
        fs_copy_box(st, sv.p, p, L, box(0));
    }
    void LoadState(const double* sphi, const double* su, const double* sv, const double* sw, const double* sp) {
        fs_copy_box(st, phi, sphi, L, box(0));
        fs_fill_const(st, phi, phi, L, box(0), 2);
        fs_copy_box(st, u, su, L, box(1));
        fs_copy_box(st, v, sv, L, box(2));
        fs_copy_box(st, w, sw, L, box(3));
        fs_copy_box(st, p, sp, L, box(0));
    }
    void LoadState(const SavedState& s_) { LoadState(s_.phi, s_.u, s_.v, s_.w, s_.p); }

    // host adjoint fields in, through the spares (their ghosts hold scatter
    // leftovers afterwards and are cleared: the forward relies on zero ghosts)
    template <typename F>
    int WithAdjointFields(double* hu, double* hv, double* hw, F&& f) {
        put(spare_u, hu, 1);
        put(spare_v, hv, 2);
        put(spare_w, hw, 3);
        int rc = f(spare_u, spare_v, spare_w);
        get(spare_u, hu, 1);
        get(spare_v, hv, 2);
        get(spare_w, hw, 3);
        for (double* a : {spare_u, spare_v, spare_w}) fs_fill_value(st, a, L.total, 0.0);
        sync();
        return rc;
    }

    // EULERIAN_FLUID_SOLVER_3D::Reinitialization, :2198-2225.  The fixed
    // schedule extrapolates the ghost copy from phi's interior only and uses
    // the Euclidean norm in every Godunov case.
    void Reinitialization() {
        if (!P.sussmanstyle_reinitialization) return;
        double fictious_dt = cell.x.d / P.scaling_number_for_reinitialization;
        fs_sign(st, phi, sgn, L, cell.x.d * cell.x.d);
        if (recording) tape.reinit.emplace_back();
        const int n = P.iteration_number_for_reinitialization;
        // The sweeps ping-pong between two ghost copies: sweep `it` reads
        // `src` and writes the next one's box (phi's on the last sweep), whose
        // ghost layers are then filled; what each sweep reads is what the
        // reference's copy-then-fill gives (ADVECTION_3D::ReinitializationBySussman,
        // ADVECTION_3D.h:360-378).  The compatible schedule's fill reads the
        // deeper layers' sources from phi's own ghost layers, as the
        // reference's FillGhostCellsContinuousDerivativesFrom(phi) does.
        // A recorded step keeps phi current after every sweep for the tape.
        double *src = sghost, *nxt = sghost2;
        // the probe times the first sweep of each call: an event pair between
        // two later sweeps of the CU-masked stream slowed the overlapped frame
        // by 7% (27.9 -> 29.9 ms at 256^3), the first sweep's pair did not
        for (int it = 0; it < n; it++) {
            if (reinit_switch >= 0 && reinit_count++ == reinit_switch) SwitchReinitStream();
            if (recording) tape.reinit.back().push_back(Record(phi, 0));
            if (it == 0 || recording) {
                if (fixed()) {
                    RefreshLevelsetGhost();
                } else {
                    fs_copy_box(st, sghost, phi, L, box(0));
                    fs_fill_linear(st, sghost, phi, L, box(0), 3, 0);
                }
            }
            const bool last = it == n - 1 || recording;
            const bool timed = sampling && it == 0;
            if (timed) kprobe[FS_KS_REINIT].begin(st);
            fs_reinit_sweep(st, last ? phi : nxt, src, sgn, L, cell, P.levelset_epsilon, fictious_dt, fixed(),
                            fast());
            if (timed) kprobe[FS_KS_REINIT].end(st);
            if (!last) {
                if (fixed())
                    fs_fill_linear(st, nxt, nxt, L, box(0), 3, 0);
                else
                    fs_fill_linear_split(st, nxt, nxt, phi, L, box(0), 3);
                std::swap(src, nxt);
            }
        }
        // the compatible schedule's level-set advection reads the ghost copy
        // the last sweep read: keep it in sghost
        if (src != sghost) std::swap(sghost, sghost2);
        if (sampling) kprobe[FS_KS_REINIT].launches += n;
    }

    // EULERIAN_FLUID_SOLVER_3D::SolveForRK, :1453-1476; a non-converged
    // projection is remembered and the step goes on, as the reference does.
    void SolveForRK(double dt, int& rc) {
        if (recording) {
            RkRecord r;
            r.u = Record(u, 1);
            r.v = Record(v, 2);
            r.w = Record(w, 3);
            r.phi = Record(phi, 0);
            tape.rk.push_back(r);
        }
        if (P.is_velocity_advection_active) Velocity_Advection(dt);
        if (P.is_external_force_active) ApplyGravity(dt);
        if (P.is_sourcing_active) {
            int r = Sourcing(dt);
            if (r != FS_OK) rc = r;
        }
        if (P.is_projection_active) {
            int r = Projection(dt);
            if (r != FS_OK) rc = r;
            if (recording) tape.rk.back().p = Record(p, 0);
        }
    }

    // velocity = ca*velocity_rk + cb*velocity on every MAC component
    void BlendVelocity(double ca, double cb) {
        fs_combine(st, u, u_rk, ca, cb, L, box(1));
        fs_combine(st, v, v_rk, ca, cb, L, box(2));
        fs_combine(st, w, w_rk, ca, cb, L, box(3));
    }
    void SaveVelocity() {
        fs_copy_box(st, u_rk, u, L, box(1));
        fs_copy_box(st, v_rk, v, L, box(2));
        fs_copy_box(st, w_rk, w, L, box(3));
    }

    // EULERIAN_FLUID_SOLVER_3D::AdvanceOneTimeStepThread, :803-1111.  tmp
    // plays temp_for_levelset.
    int AdvanceOneTimeStepThread(double dt) {
        int rc = FS_OK;
        const int order = P.order_for_time_advancing;
        if (recording) {
            tape.reset();
            tape.phi = Record(phi, 0);
            tape.u = Record(u, 1);
            tape.v = Record(v, 2);
            tape.w = Record(w, 3);
            tape.p = Record(p, 0);
        }
        LevelsetStage(dt);
        // The reinitialization reads only the advected level set (and, in the
        // compatible schedule, its ghost layers).  SolveForRK changes neither
        // the box nor, beyond the constant fill replayed below, the ghosts,
        // unless surface tension refreshes them or level-set sources change
        // the box.  Then the reinitialization is enqueued first, on the
        // second stream and a full copy of phi, and overlaps the projection;
        // phi takes the result after SolveForRK.  The values are those of
        // the sequential order.
        const bool concurrent = ConcurrentReinit();
        if (concurrent) {
            EnsureSecondStream();
            const int sweeps = P.iteration_number_for_reinitialization * (order == 3 ? 3 : 2);
            const int at = ReinitSwitchPoint(sweeps);
            HIPCHECK(hipEventRecord(ev_ls, st));
            HIPCHECK(hipStreamWaitEvent(st2, ev_ls, 0));
            // enqueued on st2 (then st3) with phi and phi_r exchanged; the
            // guard gives st and phi back however the stage leaves
            ReinitScope scope(*this);
            st = st2;
            HIPCHECK(hipEventRecord(ev_t[0], st));
            HIPCHECK(hipMemcpyAsync(phi_r, phi, sizeof(double) * (size_t)L.total, hipMemcpyDeviceToDevice, st));
            std::swap(phi, phi_r);
            scope.swapped = true;
            // the ghost layers SolveForRK will leave on phi: the constant
            // fill of ComputeNormals (assembly) and of the fixed schedule's
            // viscosity, a function of the box, which SolveForRK leaves as is
            if (ReplaysConstFill()) fs_fill_const(st, phi, phi, L, box(0), 2);
            reinit_switch = at < sweeps ? at : -1;
            reinit_count = 0;
            switched = false;
            ReinitializationStage(order);
            reinit_switch = -1;
            if (!switched) HIPCHECK(hipEventRecord(ev_t[1], st));  // every sweep ran masked
            last_switch = switched ? at : sweeps;
            HIPCHECK(hipEventRecord(ev_rn, st));
            scope.release();
        }
        try {
            if (concurrent) HIPCHECK(hipEventRecord(ev_t[2], st));
            VelocityStages(dt, rc);
            if (concurrent) {
                HIPCHECK(hipEventRecord(ev_t[3], st));
                timed_concurrent = true;
            }
        } catch (...) {
            // the reinitialization still writes phi_r: wait for it and take
            // its result, so later calls see one consistent state
            if (concurrent) {
                (void)hipStreamWaitEvent(st, ev_rn, 0);
                std::swap(phi, phi_r);
            }
            throw;
        }
        if (concurrent) {
            HIPCHECK(hipStreamWaitEvent(st, ev_rn, 0));
            std::swap(phi, phi_r);
        } else {
            if (recording) tape.phi_adv = Record(phi, 0);
            ReinitializationStage(order);
        }
        check_launch();
        return rc;
    }

    // the level-set stages: RK2 (:810-832, :888-911) or, at order 3, one
    // Euler stage (:1010-1013)
    void LevelsetStage(double dt) {
        if (!P.is_levelset_advection_active) return;
        if (P.order_for_time_advancing == 3) {
            Levelset_Advection(dt);
        } else {
            fs_copy_box(st, tmp, phi, L, box(0));
            Levelset_Advection(dt);
            Levelset_Advection(dt);
            fs_average(st, phi, tmp, L);
        }
    }

    // SolveForRK and the velocity blends of the step's order
    void VelocityStages(double dt, int& rc) {
        const int order = P.order_for_time_advancing;
        if (order == 1) {
            SolveForRK(dt, rc);  // :834
            return;
        }
        SaveVelocity();  // :913-932, :1015-1034
        SolveForRK(dt, rc);
        SolveForRK(dt, rc);
        if (order == 2) {
            BlendVelocity((double)0.5, (double)0.5);  // :937-953
        } else {
            BlendVelocity((double)0.75, (double)0.25);  // :1039-1055
            SolveForRK(dt, rc);
            BlendVelocity((double)1 / (double)3, (double)2 / (double)3);  // :1059-1075
        }
    }

    // whether this scene's step may take its reinitialization from a batch's
    // scene-table stage: the conditions of the concurrent order, and the
    // constant fill the stage replays
    bool BatchReinitEligible() const {
        return P.sussmanstyle_reinitialization && !tension() && !P.csf_model && sources.empty() && !recording &&
               (P.iteration_number_for_reinitialization & 1) == 1;
    }
    bool ReplaysConstFill() const {
        return P.is_projection_active || (fixed() && P.is_sourcing_active && P.use_delta_function_formulation);
    }
    // the scene's entry of a batch's scene table (buffers 0/1 phi and its
    // copy, in the order phi, phi_r at the time of the call)
    FsReinitScene ReinitSceneEntry() {
        if (!phi_r) phi_r = dfield();
        FsReinitScene e{};
        e.L = L;
        e.ix = cell.x.inv;
        e.iy = cell.y.inv;
        e.iz = cell.z.inv;
        e.eps = P.levelset_epsilon;
        e.dt = cell.x.d / P.scaling_number_for_reinitialization;
        e.dx2 = cell.x.d * cell.x.d;
        e.exact = fixed() ? 1 : 0;
        e.fill_const = ReplaysConstFill() ? 1 : 0;
        e.fast = fast() ? 1 : 0;
        double* b[6] = {phi, phi_r, tmp, sgn, sghost, sghost2};
        for (int q = 0; q < 6; q++) e.buf[q] = b[q];
        return e;
    }

    // reinitialization, RK2 (:836-858, :958-980) or RK3 (:1080-1109)
    void ReinitializationStage(int order) {
        fs_copy_box(st, tmp, phi, L, box(0));
        Reinitialization();
        Reinitialization();
        if (order == 3) {
            fs_combine(st, phi, tmp, (double)0.75, (double)0.25, L, box(0));
            Reinitialization();
            fs_combine(st, phi, tmp, (double)1 / (double)3, (double)2 / (double)3, L, box(0));
        } else {
            fs_average(st, phi, tmp, L);
        }
    }

    // whether the step's reinitialization may overlap SolveForRK (above).  A
    // batch's scenes already fill the GPU from their own streams, so they
    // keep the sequential order.
    bool ConcurrentReinit() const {
        const bool serial = std::getenv("FS_SERIAL_REINIT") != nullptr;
        // (the CSF model refreshes phi's ghosts inside SolveForRK, as surface
        // tension does)
        return !serial && !batch_member && !recording && P.sussmanstyle_reinitialization && !tension() &&
               !P.csf_model && sources.empty();
    }
    // Restores the main stream and phi after the concurrent stage's
    // enqueue, also when an enqueue throws (phi then waits for what st2
    // already holds).
    struct ReinitScope {
        EulerianFluidSolver3D& S;
        hipStream_t main;
        bool swapped = false, active = true;
        explicit ReinitScope(EulerianFluidSolver3D& s) : S(s), main(s.st) {}
        void release() {
            if (swapped) std::swap(S.phi, S.phi_r);
            swapped = false;
            S.st = main;
            active = false;
        }
        ~ReinitScope() {
            if (!active) return;
            hipStream_t cur = S.st;
            S.st = main;
            S.reinit_switch = -1;
            if (swapped) {
                // what was enqueued on phi (now phi_r's buffer) finishes
                // before main touches it
                if (hipEventRecord(S.ev_rn, cur) == hipSuccess) (void)hipStreamWaitEvent(main, S.ev_rn, 0);
                std::swap(S.phi, S.phi_r);
            }
        }
    };

    // continue the reinitialization on the all-CU stream st3
    void SwitchReinitStream() {
        switched = true;
        HIPCHECK(hipEventRecord(ev_sw, st));
        HIPCHECK(hipEventRecord(ev_t[1], st));
        HIPCHECK(hipStreamWaitEvent(st3, ev_sw, 0));
        st = st3;
    }

    // Sweeps of this step's reinitialization to run on the masked stream:
    // as many as the last concurrent step's SolveForRK lasted in masked
    // sweeps (half of them at the first step); FS_REINIT_MASKED_SWEEPS
    // fixes the count (negative: all of them, the single-stream order).
    int ReinitSwitchPoint(int sweeps) {
        if (const char* e = std::getenv("FS_REINIT_MASKED_SWEEPS")) {
            int k = std::atoi(e);
            return k < 0 ? sweeps : std::min(k, sweeps);
        }
        int at = sweeps / 2;
        if (timed_concurrent && last_switch > 0) {
            float masked = 0, solve = 0;
            if (hipEventElapsedTime(&masked, ev_t[0], ev_t[1]) == hipSuccess &&
                hipEventElapsedTime(&solve, ev_t[2], ev_t[3]) == hipSuccess && masked > 0) {
                const double per = (double)masked / last_switch;
                at = (int)std::lround((double)solve / per);
            }
        }
        return std::max(0, std::min(at, sweeps));
    }

    // the second stream, its events and the level-set copy, on first use
    void EnsureSecondStream() {
        if (st2) return;
        if (!ev_sw) HIPCHECK(hipEventCreateWithFlags(&ev_sw, hipEventDisableTiming));
        for (auto& e : ev_t)
            if (!e) HIPCHECK(hipEventCreate(&e));
        if (!st3) HIPCHECK(hipStreamCreate(&st3));
        // events and the copy first: the stream, made last, marks completion
        if (!ev_ls) HIPCHECK(hipEventCreateWithFlags(&ev_ls, hipEventDisableTiming));
        if (!ev_rn) HIPCHECK(hipEventCreateWithFlags(&ev_rn, hipEventDisableTiming));
        if (!phi_r) phi_r = dfield();
        // The reinitialization stream runs on the lowest-indexed 3/4 of
        // the CUs, 5/8 with fast WENO5 (FS_REINIT_CU_FRAC overrides the
        // fraction; 0 or 1 lifts the mask).  The projection beside it is
        // latency-bound and keeps the remaining CUs to itself.  Measured at
        // 256^3 on an MI355X: reference-order sweeps 42.3 ms per frame
        // against 45.4 ms unmasked (5/8 of the CUs: 45.4, 7/8: 44.1); the
        // fast sweeps, half the instructions, 28.1 ms at 5/8 against 28.4
        // (1/2), 28.7 (3/4) and 28.7 (7/8).
        const char* fr = std::getenv("FS_REINIT_CU_FRAC");
        const double f = fr ? std::atof(fr) : (fast() ? 0.625 : 0.75);
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, device));
        const int n = prop.multiProcessorCount, keep = (int)std::lround(n * f);
        if (f > 0 && f < 1 && keep > 0 && n >= 8) {
            std::vector<uint32_t> mask((n + 31) / 32, 0u);
            for (int c = 0; c < keep; c++) mask[c / 32] |= 1u << (c % 32);
            // a device that refuses the mask gets an unmasked stream
            if (hipExtStreamCreateWithCUMask(&st2, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
                (void)hipGetLastError();
                st2 = nullptr;
            }
        }
        if (!st2) HIPCHECK(hipStreamCreate(&st2));
    }

    // EULERIAN_FLUID_SOLVER_3D::CFLOneTimeStep, :2673-2785
    double CFLOneTimeStep(int gravity_norm) {
        HIPCHECK(hipMemsetAsync(vmax, 0, sizeof(unsigned long long) * 4, st));
        fs_abs_max(st, u, L, box(1), vmax + 0);
        fs_abs_max(st, v, L, box(2), vmax + 1);
        fs_abs_max(st, w, L, box(3), vmax + 2);
        HIPCHECK(hipMemcpyAsync(hvmax, vmax, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost, st));
        sync();
        double m[3];
        for (int c = 0; c < 3; c++) memcpy(&m[c], &hvmax[c], sizeof(double));
        for (double x : m)
            if (!std::isfinite(x)) throw FsFail{FS_ERR_NONFINITE, "non-finite velocity"};
        FsCflTerms t = fs_cfl_terms(P, G, gravity_norm);
        // surface tension restriction (:2740-2744, air/water) from the
        // curvature of the current level set, as SIMULATION_WORLD::AdvanceOneFrame
        // computes it right before the CFL rule (SIMULATION_WORLD.h:324, :289)
        if (tension()) {
            ComputeCurvatures();
            sync();
            check_curvature();
            double kmax;
            memcpy(&kmax, &hkstat[0], sizeof(double));
            double dmin = std::min(std::min(macx.x.d, macy.y.d), macz.z.d);
            double rmin = P.water_density > P.air_density ? P.air_density : P.water_density;
            t.s_f = std::sqrt((P.surface_tension * kmax) / (rmin * (dmin * dmin)));
        }
        return (double)1 / fs_cfl_rate(t, m[0], m[1], m[2]);
    }
};

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

struct fs_solver {
    EulerianFluidSolver3D* impl = nullptr;
    FsPipe* pipe = nullptr;  // the oil/water pipe mode (fs_create_pipe)
    std::string err;
};

// ---- on-disk records (LEVELSET_3D::Write / Read, LEVELSET_3D.h:1132-1200) --
// One record: int i_start, j_start, k_start, i_res, j_res, k_res; double
// x_min, y_min, z_min, x_max, y_max, z_max; int ghost_width; then the
// i_res*j_res*k_res values in GRID_ITERATION_3D order (i fastest, k
// slowest), each field written on its own, native byte order.  A checkpoint
// is five records back to back: phi and p on the cell grid, u, v, w on their
// MAC grids (EULERIAN_FLUID_SOLVER_3D.h:405-431); its first record alone is
// a level-set file.
namespace {

struct FsRecordGrid {
    int res[3];
    double mn[3], mx[3];
    int ghost;
};

// grid of component comp (0 cells, 1..3 MAC) as the reference records it,
// with the ghost width of the reference's field: level set and MAC velocity
// 2, pressure 1 (EULERIAN_FLUID_SOLVER_3D.h:435-438, PROJECTION_3D.h:362)
FsRecordGrid record_grid(const EulerianFluidSolver3D& S, int comp, int ghost) {
    FsRecordGrid g;
    const FsGridConst* gc[4] = {&S.cell, &S.macx, &S.macy, &S.macz};
    const double mn[3] = {S.G.x_min, S.G.y_min, S.G.z_min}, mx[3] = {S.G.x_max, S.G.y_max, S.G.z_max};
    const double d[3] = {gc[0]->x.d, gc[0]->y.d, gc[0]->z.d};
    const int n[3] = {S.G.nx, S.G.ny, S.G.nz};
    for (int a = 0; a < 3; a++) {
        bool face = comp == a + 1;
        g.res[a] = n[a] + (face ? 1 : 0);
        g.mn[a] = face ? mn[a] - (double)0.5 * d[a] : mn[a];
        g.mx[a] = face ? mx[a] + (double)0.5 * d[a] : mx[a];
    }
    g.ghost = ghost;
    return g;
}

void write_record(FILE* f, const FsRecordGrid& g, const std::vector<double>& v) {
    const int start = 0;
    bool ok = true;
    for (int a = 0; a < 3; a++) ok &= fwrite(&start, sizeof(int), 1, f) == 1;
    for (int a = 0; a < 3; a++) ok &= fwrite(&g.res[a], sizeof(int), 1, f) == 1;
    for (int a = 0; a < 3; a++) ok &= fwrite(&g.mn[a], sizeof(double), 1, f) == 1;
    for (int a = 0; a < 3; a++) ok &= fwrite(&g.mx[a], sizeof(double), 1, f) == 1;
    ok &= fwrite(&g.ghost, sizeof(int), 1, f) == 1;
    ok &= fwrite(v.data(), sizeof(double), v.size(), f) == v.size();
    if (!ok) throw FsFail{FS_ERR_BAD_ARG, "short write"};
}

// reads one record and checks that its grid is `want` (extents to 1e-12 of
// the domain size); returns its values
std::vector<double> read_record(FILE* f, const FsRecordGrid& want, const char* what) {
    int start[3], res[3], ghost;
    double mn[3], mx[3];
    bool ok = fread(start, sizeof(int), 3, f) == 3 && fread(res, sizeof(int), 3, f) == 3 &&
              fread(mn, sizeof(double), 3, f) == 3 && fread(mx, sizeof(double), 3, f) == 3 &&
              fread(&ghost, sizeof(int), 1, f) == 1;
    if (!ok) throw FsFail{FS_ERR_BAD_ARG, std::string(what) + ": truncated header"};
    for (int a = 0; a < 3; a++) {
        double tol = 1e-12 * (std::fabs(want.mx[a] - want.mn[a]) + 1.0);
        if (res[a] != want.res[a] || std::fabs(mn[a] - want.mn[a]) > tol || std::fabs(mx[a] - want.mx[a]) > tol)
            throw FsFail{FS_ERR_BAD_ARG, std::string(what) + ": grid does not match the solver's"};
    }
    std::vector<double> v((size_t)res[0] * res[1] * res[2]);
    if (fread(v.data(), sizeof(double), v.size(), f) != v.size())
        throw FsFail{FS_ERR_BAD_ARG, std::string(what) + ": truncated values"};
    return v;
}

struct FileCloser {
    FILE* f;
    ~FileCloser() {
        if (f) fclose(f);
    }
};

}  // namespace

template <typename F>
static int guarded(fs_solver* s, F&& f) {
    if (s && s->pipe) {
        s->err = "not available for the oil/water pipe mode";
        return FS_ERR_UNSUPPORTED;
    }
    if (!s || !s->impl) return FS_ERR_BAD_ARG;
    try {
        int rc = f(*s->impl);
        if (rc != FS_OK) s->err = s->impl->last_error;
        return rc;
    } catch (FsFail& e) {
        s->err = e.msg;
        return e.code;
    } catch (HipError& e) {
        s->err = std::string(e.what) + ": " + hipGetErrorString(e.e);
        return FS_ERR_HIP;
    }
}

// a call on a pipe-mode handle: FsPipeError and HIP failures to codes
template <typename F>
static int pipe_guarded(fs_solver* s, F&& f) {
    try {
        return f(*s->pipe);
    } catch (FsPipeError& e) {
        s->err = e.msg;
        return e.code;
    } catch (HipError& e) {
        s->err = std::string(e.what) + ": " + hipGetErrorString(e.e);
        return FS_ERR_HIP;
    }
}

extern "C" {

int fs_create_pipe(const fs_params* params, const fs_pipe_params* pipe, const fs_grid* grid, int device,
                   fs_solver** out) {
    if (!params || !pipe || !grid || !out) return FS_ERR_BAD_ARG;
    *out = nullptr;
    try {
        FsPipe* p = fs_pipe_new(*params, *pipe, *grid, device);
        auto* s = new fs_solver();
        s->pipe = p;
        *out = s;
        return FS_OK;
    } catch (FsPipeError& e) {
        return e.code;
    }
}

int fs_pipe_initial_fields(fs_solver* s) {
    if (!s || !s->pipe) return FS_ERR_BAD_ARG;
    return pipe_guarded(s, [&](FsPipe& P) {
        fs_pipe_initial(&P);
        return FS_OK;
    });
}

int fs_pipe_constants(fs_solver* s, double* out) {
    if (!s || !s->pipe || !out) return FS_ERR_BAD_ARG;
    fs_pipe_consts(s->pipe, out);
    return FS_OK;
}

int fs_pipe_phase(fs_solver* s, int phase, double dt, fs_step_stats* stats) {
    if (!s || !s->pipe) return FS_ERR_BAD_ARG;
    if (!(dt > 0) || !std::isfinite(dt)) return FS_ERR_BAD_ARG;
    return pipe_guarded(s, [&](FsPipe& P) { return fs_pipe_run_phase(&P, phase, dt, stats); });
}

int fs_pipe_get_array(fs_solver* s, int which, double* dst) {
    if (!s || !s->pipe || !dst) return FS_ERR_BAD_ARG;
    return pipe_guarded(s, [&](FsPipe& P) {
        fs_pipe_array(&P, which, dst);
        return FS_OK;
    });
}

int fs_pipe_kernel_samples(fs_solver* s, int kernel, double* avg_ms, double* avg_bytes, long long* samples,
                           long long* launches) {
    if (!s || !s->pipe || !avg_ms || !avg_bytes || !samples || !launches) return FS_ERR_BAD_ARG;
    return pipe_guarded(s, [&](FsPipe& P) {
        fs_pipe_samples(&P, kernel, avg_ms, avg_bytes, samples, launches);
        return FS_OK;
    });
}

int fs_create(const fs_params* params, const fs_grid* grid, int device, fs_solver** out) {
    if (!params || !grid || !out) return FS_ERR_BAD_ARG;
    *out = nullptr;
    auto* s = new fs_solver();
    s->impl = new EulerianFluidSolver3D();
    try {
        s->impl->InitializeFromScriptBlock(*params, *grid, device);
    } catch (FsFail& e) {
        delete s->impl;
        delete s;
        return e.code;
    } catch (HipError& e) {
        delete s->impl;
        delete s;
        return e.e == hipErrorOutOfMemory ? FS_ERR_OOM : FS_ERR_HIP;
    }
    *out = s;
    return FS_OK;
}

void fs_destroy(fs_solver* s) {
    if (!s) return;
    if (s->pipe) fs_pipe_delete(s->pipe);
    delete s->impl;
    delete s;
}

const char* fs_last_error(const fs_solver* s) { return s ? s->err.c_str() : "null handle"; }

int fs_set_fields(fs_solver* s, const double* phi, const double* u, const double* v, const double* w,
                  const double* p) {
    if (s && s->pipe)
        return pipe_guarded(s, [&](FsPipe& P) {
            fs_pipe_set(&P, phi, u, v, w, p);
            return FS_OK;
        });
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.SetFields(phi, u, v, w, p);
        return FS_OK;
    });
}

int fs_set_fields_ex(fs_solver* s, const double* phi, const double* u, const double* v, const double* w,
                     const double* p, int layout, int memory) {
    if ((layout != FS_LAYOUT_INTERIOR && layout != FS_LAYOUT_GHOSTED) || (memory != FS_MEM_HOST && memory != FS_MEM_DEVICE))
        return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.SetFields(phi, u, v, w, p, layout, memory);
        return FS_OK;
    });
}

int fs_get_fields_ex(fs_solver* s, double* phi, double* u, double* v, double* w, double* p, int layout, int memory) {
    if ((layout != FS_LAYOUT_INTERIOR && layout != FS_LAYOUT_GHOSTED) || (memory != FS_MEM_HOST && memory != FS_MEM_DEVICE))
        return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.GetFields(phi, u, v, w, p, layout, memory);
        return FS_OK;
    });
}

int fs_set_async(fs_solver* s, int async) {
    if (s && s->pipe) {
        fs_pipe_async(s->pipe, async != 0);
        return FS_OK;
    }
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.async = async != 0;
        return FS_OK;
    });
}

int fs_synchronize(fs_solver* s) {
    if (s && s->pipe)
        return pipe_guarded(s, [&](FsPipe& P) {
            fs_pipe_sync(&P);
            return FS_OK;
        });
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.sync();
        S.check_launch();
        return FS_OK;
    });
}

int fs_get_stream(fs_solver* s, void** stream) {
    if (!stream) return FS_ERR_BAD_ARG;
    if (s && s->pipe) {
        *stream = (void*)fs_pipe_stream(s->pipe);
        return FS_OK;
    }
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        *stream = (void*)S.st;
        return FS_OK;
    });
}

int fs_get_fields(fs_solver* s, double* phi, double* u, double* v, double* w, double* p) {
    if (s && s->pipe)
        return pipe_guarded(s, [&](FsPipe& P) {
            fs_pipe_get(&P, phi, u, v, w, p);
            return FS_OK;
        });
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.GetFields(phi, u, v, w, p);
        return FS_OK;
    });
}

static int step_impl(EulerianFluidSolver3D& S, double dt, fs_step_stats* stats, bool frame) {
    if (!(dt > 0) || !std::isfinite(dt)) throw FsFail{FS_ERR_BAD_ARG, "dt must be positive and finite"};
    S.stats = fs_step_stats{};
    HIPCHECK(hipEventRecord(S.ev[0], S.st));
    if (frame) S.ComputeCurvatures();
    int rc = S.AdvanceOneTimeStepThread(dt);
    HIPCHECK(hipEventRecord(S.ev[1], S.st));
    if (!S.async) {
        S.sync();
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
        S.stats.step_ms = ms;
    }
    if (stats) *stats = S.stats;
    return rc;
}

// ---- batches of independent scenes ------------------------------------------

struct fs_batch {
    std::vector<fs_solver*> scenes;
    std::vector<long> order;  // scenes by decreasing cell count (largest first)
    int device = 0;
    int threads = 1;
    std::string err;
    // the scene-table reinitialization (fs_batch.h): built at the first frame
    // when every scene qualifies, run on its own stream
    FsBatchReinit table;
    int table_state = 0;  // 0 not tried, 1 built, -1 scenes do not qualify
    hipStream_t bst = nullptr;
    hipEvent_t ev_brn = nullptr;
    std::vector<hipEvent_t> ev_ls;
    std::vector<double*> buf0, buf1;  // each scene's table buffers 0 and 1 (its phi and phi_r at the build)
    ~fs_batch() {
        if (bst) (void)hipStreamSynchronize(bst);
        for (hipEvent_t e : ev_ls)
            if (e) (void)hipEventDestroy(e);
        if (ev_brn) (void)hipEventDestroy(ev_brn);
        if (bst) (void)hipStreamDestroy(bst);
    }
};

// Whether the batch's scenes take their reinitialization from the scene
// table, building it on first use: FS_BATCH_TABLE is set, and every scene
// qualifies (the concurrent order's conditions, an odd sweep count) with the
// same sweep count and calls per stage.  Opt-in: on 48 uneven scenes of
// 12^3..64^3 the table stage (193 ms/frame) trails the per-scene stages on
// their own streams (182 ms/frame), whose reinitialization overlaps the
// other scenes' projections where the table has to wait for every scene's
// advected level set.
static bool batch_table(fs_batch* b) {
    if (!std::getenv("FS_BATCH_TABLE")) return false;
    if (b->table_state != 0) {
        if (b->table_state < 0) return false;
        // sources added since, or a scene's phi outside the two table buffers
        for (size_t i = 0; i < b->scenes.size(); i++) {
            EulerianFluidSolver3D& S = *b->scenes[i]->impl;
            if (!S.BatchReinitEligible()) return false;
        }
        return true;
    }
    b->table_state = -1;
    const EulerianFluidSolver3D& S0 = *b->scenes[0]->impl;
    const int sweeps = S0.P.iteration_number_for_reinitialization, calls = S0.P.order_for_time_advancing == 3 ? 3 : 2;
    for (fs_solver* s : b->scenes) {
        const EulerianFluidSolver3D& S = *s->impl;
        if (!S.BatchReinitEligible() || S.P.iteration_number_for_reinitialization != sweeps ||
            (S.P.order_for_time_advancing == 3 ? 3 : 2) != calls)
            return false;
    }
    std::vector<FsReinitScene> entries;
    for (fs_solver* s : b->scenes) {
        EulerianFluidSolver3D& S = *s->impl;
        entries.push_back(S.ReinitSceneEntry());
        b->buf0.push_back(S.phi);
        b->buf1.push_back(S.phi_r);
        hipEvent_t e = nullptr;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->ev_ls.push_back(e);
        S.sync();
    }
    HIPCHECK(hipEventCreateWithFlags(&b->ev_brn, hipEventDisableTiming));
    HIPCHECK(hipStreamCreate(&b->bst));
    b->table.Build(entries, sweeps, calls);
    b->table_state = 1;
    return true;
}

// A frame of every scene with the reinitialization from the scene table:
// each scene's level-set stages on its stream; the table's stage for all
// scenes on the batch stream once every scene's advected level set is there
// (it works on the copies); each scene's SolveForRK on its stream beside it;
// then each scene takes its reinitialized copy as phi.
static int batch_frame_table(fs_batch* b, const double* dt, fs_step_stats* stats, std::vector<int>& rc) {
    const int n = (int)b->scenes.size();
    int parity = -1;
    for (int i = 0; i < n; i++) {
        const EulerianFluidSolver3D& S = *b->scenes[i]->impl;
        const int q = S.phi == b->buf0[i] && S.phi_r == b->buf1[i] ? 0 : (S.phi == b->buf1[i] && S.phi_r == b->buf0[i] ? 1 : -1);
        if (q < 0 || (parity >= 0 && q != parity)) {
            b->err = "batched reinitialization: scene " + std::to_string(i) + " is out of step with the table";
            return FS_ERR_UNSUPPORTED;
        }
        parity = q;
    }
    std::atomic<int> next{0};
    auto pool = [&](auto&& body) {
        next = 0;
        auto work = [&]() {
            if (hipSetDevice(b->device) != hipSuccess) return;
            for (int q = next++; q < n; q = next++) {
                int i = (int)b->order[q];
                if (rc[i] != FS_OK) continue;
                rc[i] = guarded(b->scenes[i], [&](EulerianFluidSolver3D& S) { return body(S, i); });
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < b->threads; t++) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
    };
    // 1. level-set stages
    pool([&](EulerianFluidSolver3D& S, int i) {
        if (!(dt[i] > 0) || !std::isfinite(dt[i])) throw FsFail{FS_ERR_BAD_ARG, "dt must be positive and finite"};
        S.stats = fs_step_stats{};
        HIPCHECK(hipEventRecord(S.ev[0], S.st));
        S.ComputeCurvatures();
        S.LevelsetStage(dt[i]);
        HIPCHECK(hipEventRecord(b->ev_ls[i], S.st));
        return FS_OK;
    });
    for (int i = 0; i < n; i++)
        if (rc[i] != FS_OK) {
            for (fs_solver* s : b->scenes) (void)hipStreamSynchronize(s->impl->st);
            return rc[i];
        }
    // 2. the reinitialization of every scene's copy
    try {
        for (int i = 0; i < n; i++) HIPCHECK(hipStreamWaitEvent(b->bst, b->ev_ls[i], 0));
        b->table.Run(b->bst, parity);
        HIPCHECK(hipEventRecord(b->ev_brn, b->bst));
    } catch (const std::exception& e) {
        b->err = std::string("batched reinitialization: ") + e.what();
        return FS_ERR_HIP;
    } catch (const HipError& e) {
        b->err = std::string("batched reinitialization: ") + hipGetErrorString(e.e);
        return FS_ERR_HIP;
    }
    // 3. SolveForRK beside it, then phi takes the copy
    pool([&](EulerianFluidSolver3D& S, int i) {
        int r = FS_OK;
        try {
            S.VelocityStages(dt[i], r);
        } catch (...) {
            (void)hipStreamWaitEvent(S.st, b->ev_brn, 0);
            std::swap(S.phi, S.phi_r);
            throw;
        }
        HIPCHECK(hipStreamWaitEvent(S.st, b->ev_brn, 0));
        std::swap(S.phi, S.phi_r);
        S.check_launch();
        HIPCHECK(hipEventRecord(S.ev[1], S.st));
        S.sync();
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
        S.stats.step_ms = ms;
        if (stats) stats[i] = S.stats;
        return r;
    });
    return FS_OK;
}

int fs_batch_create(int n, const fs_params* params, const fs_grid* grids, int device, int threads, fs_batch** out) {
    if (n <= 0 || !params || !grids || !out) return FS_ERR_BAD_ARG;
    *out = nullptr;
    auto* b = new fs_batch();
    for (int i = 0; i < n; i++) {
        fs_solver* s = nullptr;
        int rc = fs_create(&params[i], &grids[i], device, &s);
        if (rc != FS_OK) {
            for (fs_solver* t : b->scenes) fs_destroy(t);
            delete b;
            return rc;
        }
        s->impl->batch_member = true;
        b->scenes.push_back(s);
        b->order.push_back(i);
    }
    std::stable_sort(b->order.begin(), b->order.end(), [&](long a, long c) {
        auto cells = [&](long k) { return (long)grids[k].nx * grids[k].ny * grids[k].nz; };
        return cells(a) > cells(c);
    });
    (void)hipGetDevice(&b->device);
    b->threads = threads > 0 ? std::min(threads, n) : std::min(n, 16);
    *out = b;
    return FS_OK;
}

void fs_batch_destroy(fs_batch* b) {
    if (!b) return;
    for (fs_solver* s : b->scenes) fs_destroy(s);
    delete b;
}

fs_solver* fs_batch_scene(fs_batch* b, int i) {
    if (!b || i < 0 || i >= (int)b->scenes.size()) return nullptr;
    return b->scenes[i];
}

const char* fs_batch_last_error(const fs_batch* b) { return b ? b->err.c_str() : "null handle"; }

// Every scene advances one frame.  Host threads take scenes from a shared
// counter, largest first; each scene's work runs on its own stream, so the
// kernels of small scenes overlap on the device.
int fs_batch_advance_frame(fs_batch* b, const double* dt, fs_step_stats* stats, int* rcs) {
    if (!b || !dt) return FS_ERR_BAD_ARG;
    const int n = (int)b->scenes.size();
    std::vector<int> rc(n, FS_OK);
    bool table = false;
    try {
        table = batch_table(b);
    } catch (const HipError& e) {
        b->err = std::string("batch table: ") + hipGetErrorString(e.e);
        return FS_ERR_HIP;
    } catch (const FsFail& f) {
        b->err = "batch table: " + f.msg;
        return f.code;
    } catch (const std::exception& e) {
        b->err = std::string("batch table: ") + e.what();
        return FS_ERR_HIP;
    }
    if (table) {
        int r = batch_frame_table(b, dt, stats, rc);
        if (r != FS_OK) {
            if (b->err.empty()) b->err = "batched frame failed";
            for (int i = 0; i < n; i++)
                if (rcs) rcs[i] = rc[i];
            return r;
        }
        int first = FS_OK;
        for (int i = 0; i < n; i++) {
            if (rcs) rcs[i] = rc[i];
            if (rc[i] != FS_OK && first == FS_OK) {
                first = rc[i];
                b->err = "scene " + std::to_string(i) + ": " + b->scenes[i]->err;
            }
        }
        return first;
    }
    std::atomic<int> next{0};
    auto work = [&]() {
        if (hipSetDevice(b->device) != hipSuccess) return;
        for (int q = next++; q < n; q = next++) {
            int i = (int)b->order[q];
            rc[i] = guarded(b->scenes[i], [&](EulerianFluidSolver3D& S) {
                return step_impl(S, dt[i], stats ? &stats[i] : nullptr, true);
            });
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < b->threads; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    int first = FS_OK;
    for (int i = 0; i < n; i++) {
        if (rcs) rcs[i] = rc[i];
        if (rc[i] != FS_OK && first == FS_OK) {
            first = rc[i];
            b->err = "scene " + std::to_string(i) + ": " + b->scenes[i]->err;
        }
    }
    return first;
}

int fs_advance_frame(fs_solver* s, double dt, fs_step_stats* stats) {
    if (s && s->pipe) return pipe_guarded(s, [&](FsPipe& P) { return fs_pipe_advance(&P, dt, stats, true); });
    return guarded(s, [&](EulerianFluidSolver3D& S) { return step_impl(S, dt, stats, true); });
}

int fs_step(fs_solver* s, double dt, fs_step_stats* stats) {
    if (s && s->pipe) return pipe_guarded(s, [&](FsPipe& P) { return fs_pipe_advance(&P, dt, stats, false); });
    return guarded(s, [&](EulerianFluidSolver3D& S) { return step_impl(S, dt, stats, false); });
}

int fs_cfl_dt(fs_solver* s, double cfl, int gravity_norm, double* dt) {
    if (!dt || !(cfl > 0)) return FS_ERR_BAD_ARG;
    // the pipe mode's CFL uses g_y and its dimensionless forcing as the
    // reference does (EULERIAN_FLUID_SOLVER_3D.h:2719-2733); gravity_norm is ignored
    if (s && s->pipe)
        return pipe_guarded(s, [&](FsPipe& P) {
            *dt = fs_pipe_cfl(&P, cfl);
            return FS_OK;
        });
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        *dt = cfl * S.CFLOneTimeStep(gravity_norm);
        return FS_OK;
    });
}

int fs_compute_curvatures(fs_solver* s) {
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.ComputeCurvatures();
        S.check_launch();
        S.sync();
        S.check_curvature();
        return FS_OK;
    });
}

int fs_set_levelset_sources(fs_solver* s, int n, const fs_levelset_source* src) {
    if (n < 0 || n > 64 || (n > 0 && !src)) return FS_ERR_BAD_ARG;
    for (int i = 0; i < n; i++) {
        const fs_levelset_source& o = src[i];
        bool finite = std::isfinite(o.radius) && std::isfinite(o.half_width);
        for (double c : o.center) finite = finite && std::isfinite(c);
        if (!finite || o.mode < FS_SOURCE_NONE || o.mode > FS_SOURCE_DELETE) return FS_ERR_BAD_ARG;
    }
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.sources.assign(src, src + n);
        return FS_OK;
    });
}

int fs_set_frame(fs_solver* s, int frame) {
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.frame = frame;
        return FS_OK;
    });
}

int fs_apply_levelset_sources(fs_solver* s) {
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.SourcingFromObjectList();
        S.sync();
        return FS_OK;
    });
}

int fs_get_curvature(fs_solver* s, double* kappa, double* abs_max_curvature) {
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        if (!S.tension()) throw FsFail{FS_ERR_UNSUPPORTED, "the curvature is computed only with surface tension"};
        S.ComputeCurvatures();
        S.check_launch();
        S.sync();
        S.check_curvature();
        if (kappa) S.get(S.kappa, kappa, 0);
        if (abs_max_curvature) memcpy(abs_max_curvature, &S.hkstat[0], sizeof(double));
        return FS_OK;
    });
}

#define PHASE(name, call)                                                              \
    int name(fs_solver* s, double dt) {                                                \
        if (!(dt > 0)) return FS_ERR_BAD_ARG;                                          \
        return guarded(s, [&](EulerianFluidSolver3D& S) {                              \
            call;                                                                      \
            S.check_launch();                                                          \
            S.sync();                                                                  \
            return FS_OK;                                                              \
        });                                                                            \
    }

PHASE(fs_levelset_advection, S.Levelset_Advection(dt))
PHASE(fs_velocity_advection, S.Velocity_Advection(dt))
PHASE(fs_apply_gravity, S.ApplyGravity(dt))

int fs_sourcing(fs_solver* s, double dt) {
    if (!(dt > 0)) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.stats = fs_step_stats{};
        int rc = S.Sourcing(dt);
        S.check_launch();
        S.sync();
        return rc;
    });
}
PHASE(fs_reinitialization, S.Reinitialization())

int fs_projection(fs_solver* s, double dt, fs_step_stats* stats) {
    if (!(dt > 0)) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.stats = fs_step_stats{};
        int rc = S.Projection(dt);
        if (stats) *stats = S.stats;
        return rc;
    });
}

int fs_adjoint_projection(fs_solver* s, double dt, double* lu, double* lv, double* lw, fs_step_stats* stats) {
    if (!(dt > 0) || !lu || !lv || !lw) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.stats = fs_step_stats{};
        int rc = S.WithAdjointFields(lu, lv, lw, [&](double* a, double* b, double* c) {
            return S.AdjointProjection(dt, a, b, c);
        });
        if (stats) *stats = S.stats;
        return rc;
    });
}

// The stage's inputs and outputs travel through scratch fields that no
// later phase reads before rewriting them: tmp (phi), the spares (u, v, w;
// cleared afterwards), div (output or lam), rhs and bx/by/bz (adjoints; the
// next projection reassembles them).
int fs_weno5_transport(fs_solver* s, double dt, const double* phi, const double* u, const double* v, const double* w,
                       double* out) {
    if (!(dt > 0) || !phi || !u || !v || !w || !out) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.put(S.tmp, phi, 0);
        S.put(S.spare_u, u, 1);
        S.put(S.spare_v, v, 2);
        S.put(S.spare_w, w, 3);
        fs_weno5_transport(S.st, S.tmp, S.spare_u, S.spare_v, S.spare_w, S.div, S.L, S.cell, S.P.levelset_epsilon,
                           dt);
        S.check_launch();
        S.get(S.div, out, 0);
        for (double* a : {S.spare_u, S.spare_v, S.spare_w}) fs_fill_value(S.st, a, S.L.total, 0.0);
        S.sync();
        return FS_OK;
    });
}

int fs_adjoint_weno5_transport(fs_solver* s, double dt, const double* phi, const double* u, const double* v,
                               const double* w, const double* lam, double* lphi, double* lu, double* lv, double* lw) {
    if (!(dt > 0) || !phi || !u || !v || !w || !lam || !lphi || !lu || !lv || !lw) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.assembled = false;
        S.put(S.tmp, phi, 0);
        S.put(S.spare_u, u, 1);
        S.put(S.spare_v, v, 2);
        S.put(S.spare_w, w, 3);
        S.put(S.div, lam, 0);
        for (double* a : {S.rhs, S.bx, S.by, S.bz}) fs_fill_value(S.st, a, S.L.total, 0.0);
        fs_adjoint_weno5_transport(S.st, S.tmp, S.spare_u, S.spare_v, S.spare_w, S.div, S.rhs, S.bx, S.by, S.bz, S.L,
                                   S.cell, S.P.levelset_epsilon, dt);
        S.check_launch();
        S.get(S.rhs, lphi, 0);
        S.get(S.bx, lu, 1);
        S.get(S.by, lv, 2);
        S.get(S.bz, lw, 3);
        for (double* a : {S.spare_u, S.spare_v, S.spare_w}) fs_fill_value(S.st, a, S.L.total, 0.0);
        S.sync();
        return FS_OK;
    });
}

int fs_adjoint_sourcing(fs_solver* s, double dt, double* lu, double* lv, double* lw) {
    if (!(dt > 0) || !lu || !lv || !lw) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        return S.WithAdjointFields(lu, lv, lw, [&](double* a, double* b, double* c) {
            S.AdjointSourcing(dt, a, b, c);
            return FS_OK;
        });
    });
}

int fs_adjoint_step(fs_solver* s, double dt, double* lphi, double* lu, double* lv, double* lw,
                    fs_step_stats* stats) {
    if (!(dt > 0) || !std::isfinite(dt) || !lphi || !lu || !lv || !lw) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.CheckAdjointSupport();
        S.EnsureAdjoint();
        S.stats = fs_step_stats{};
        int rc = S.RecordStep(dt);
        for (double* a : {S.A.lphi, S.A.lu, S.A.lv, S.A.lw}) S.Zero(a);
        S.put(S.A.lphi, lphi, 0);
        S.put(S.A.lu, lu, 1);
        S.put(S.A.lv, lv, 2);
        S.put(S.A.lw, lw, 3);
        int rr = S.ReverseStep(dt);
        S.get(S.A.lphi, lphi, 0);
        S.get(S.A.lu, lu, 1);
        S.get(S.A.lv, lv, 2);
        S.get(S.A.lw, lw, 3);
        S.LoadState(S.tape.phi, S.tape.u, S.tape.v, S.tape.w, S.tape.p);
        S.sync();
        if (stats) *stats = S.stats;
        return rc != FS_OK ? rc : rr;
    });
}

static double read_loss(EulerianFluidSolver3D& S) {
    double h = 0;
    HIPCHECK(hipMemcpyAsync(&h, S.A.part + FS_LOSS_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, S.st));
    S.sync();
    return h;
}

int fs_levelset_loss(fs_solver* s, const double* phi_target, double* loss) {
    if (!phi_target || !loss) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.EnsureAdjoint();
        S.put(S.A.tgt, phi_target, 0);
        fs_levelset_loss(S.st, S.phi, S.A.tgt, nullptr, S.A.part, S.L, 1.5 * S.cell.x.d);
        S.check_launch();
        *loss = read_loss(S);
        return FS_OK;
    });
}

// Checkpointed reverse sweep: the forward keeps the state at the start of
// every `checkpoint_every`-th step; the reverse pass walks the segments last
// to first, re-running a segment from its checkpoint to keep each step's
// input, then re-records and reverses its steps one by one.
int fs_adjoint_rollout(fs_solver* s, int steps, double dt, int checkpoint_every, const double* phi_target,
                       double* loss, double* gphi, double* gu, double* gv, double* gw, fs_step_stats* stats) {
    if (steps < 1 || checkpoint_every < 1 || !(dt > 0) || !std::isfinite(dt) || !phi_target || !loss)
        return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.CheckAdjointSupport();
        S.EnsureAdjoint();
        S.stats = fs_step_stats{};
        const int every = checkpoint_every;
        const int nck = (steps + every - 1) / every;
        S.Saved((size_t)(nck + std::min(every, steps)) - 1);  // checkpoints, then one segment's states
        int rc = FS_OK;
        auto note = [&](int r) {
            if (r != FS_OK && rc == FS_OK) rc = r;
        };
        S.put(S.A.tgt, phi_target, 0);
        for (int k = 0; k < steps; k++) {
            if (k % every == 0) S.SaveState(S.saved[k / every]);
            note(S.AdvanceOneTimeStepThread(dt));
        }
        for (double* a : {S.A.lphi, S.A.lu, S.A.lv, S.A.lw}) S.Zero(a);
        fs_levelset_loss(S.st, S.phi, S.A.tgt, S.A.lphi, S.A.part, S.L, 1.5 * S.cell.x.d);
        S.check_launch();
        *loss = read_loss(S);
        for (int seg = nck - 1; seg >= 0; seg--) {
            const int k0 = seg * every, k1 = std::min(steps, k0 + every);
            S.LoadState(S.saved[seg]);
            for (int k = k0; k < k1; k++) {
                S.SaveState(S.saved[nck + (k - k0)]);
                if (k + 1 < k1) note(S.AdvanceOneTimeStepThread(dt));
            }
            for (int k = k1 - 1; k >= k0; k--) {
                S.LoadState(S.saved[nck + (k - k0)]);
                note(S.RecordStep(dt));
                note(S.ReverseStep(dt));
            }
        }
        S.LoadState(S.saved[0]);
        if (gphi) S.get(S.A.lphi, gphi, 0);
        if (gu) S.get(S.A.lu, gu, 1);
        if (gv) S.get(S.A.lv, gv, 2);
        if (gw) S.get(S.A.lw, gw, 3);
        S.sync();
        if (stats) *stats = S.stats;
        return rc;
    });
}

static size_t box_size(const EulerianFluidSolver3D& S, int comp) {
    FsBox b = S.box(comp);
    return (size_t)b.ni * b.nj * b.nk;
}

int fs_write_levelset(fs_solver* s, const char* path) {
    if (!path) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        std::vector<double> v(box_size(S, 0));
        S.GetFields(v.data(), nullptr, nullptr, nullptr, nullptr);
        FileCloser fc{fopen(path, "wb")};
        if (!fc.f) throw FsFail{FS_ERR_BAD_ARG, std::string("cannot open ") + path};
        write_record(fc.f, record_grid(S, 0, 2), v);
        return FS_OK;
    });
}

int fs_read_levelset(fs_solver* s, const char* path) {
    if (!path) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        FileCloser fc{fopen(path, "rb")};
        if (!fc.f) throw FsFail{FS_ERR_BAD_ARG, std::string("cannot open ") + path};
        std::vector<double> v = read_record(fc.f, record_grid(S, 0, 2), "level set");
        S.SetFields(v.data(), nullptr, nullptr, nullptr, nullptr);
        return FS_OK;
    });
}

int fs_write_checkpoint(fs_solver* s, const char* path) {
    if (!path) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        std::vector<double> f[5];
        const int comp[5] = {0, 1, 2, 3, 0}, ghost[5] = {2, 2, 2, 2, 1};
        for (int q = 0; q < 5; q++) f[q].resize(box_size(S, comp[q]));
        S.GetFields(f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data());
        FileCloser fc{fopen(path, "wb")};
        if (!fc.f) throw FsFail{FS_ERR_BAD_ARG, std::string("cannot open ") + path};
        for (int q = 0; q < 5; q++) write_record(fc.f, record_grid(S, comp[q], ghost[q]), f[q]);
        return FS_OK;
    });
}

int fs_read_checkpoint(fs_solver* s, const char* path) {
    if (!path) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        FileCloser fc{fopen(path, "rb")};
        if (!fc.f) throw FsFail{FS_ERR_BAD_ARG, std::string("cannot open ") + path};
        const int comp[5] = {0, 1, 2, 3, 0}, ghost[5] = {2, 2, 2, 2, 1};
        const char* what[5] = {"phi", "u", "v", "w", "p"};
        std::vector<double> f[5];
        for (int q = 0; q < 5; q++) f[q] = read_record(fc.f, record_grid(S, comp[q], ghost[q]), what[q]);
        S.SetFields(f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data());
        return FS_OK;
    });
}

int fs_assemble_pressure_system(fs_solver* s, double dt, double* diag, double* bxh, double* byh, double* bzh,
                                double* rhs) {
    if (!(dt > 0)) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.AssemblePressureSystem(dt);
        S.check_launch();
        if (diag) S.get(S.diag, diag, 0);
        if (bxh) S.get(S.bx, bxh, 1);
        if (byh) S.get(S.by, byh, 2);
        if (bzh) S.get(S.bz, bzh, 3);
        if (rhs) S.get(S.rhs, rhs, 0);
        S.sync();
        S.check_curvature();
        return FS_OK;
    });
}

int fs_apply_pressure_operator(fs_solver* s, const double* x, double* q) {
    if (!x || !q) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        if (!S.assembled) throw FsFail{FS_ERR_BAD_ARG, "assemble the pressure system first"};
        FsBox b = S.box(0);
        double* xd = (double*)S.w64.r;
        double* qd = (double*)S.w64.q;
        S.put(xd, x, 0);
        S.init_state(S.S64, 0, 0, 0, 0);
        fs_apply_op64(S.st, S.L, S.op64(), xd, qd, S.S64);
        S.check_launch();
        S.get(qd, q, 0);
        (void)b;
        return FS_OK;
    });
}

int fs_time_pcg_kernels(fs_solver* s, int precision, int reps, double* avg_ms) {
    if (!avg_ms || reps <= 0 || (precision != FS_FP64 && precision != FS_FP32_REFINE)) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        if (!S.assembled) throw FsFail{FS_ERR_BAD_ARG, "no assembled pressure system"};
        if (precision == FS_FP32_REFINE && !S.bx32) throw FsFail{FS_ERR_BAD_ARG, "handle has no FP32 workspace"};
        // a state that never stops, with the iteration's scalars finite
        FsPcgState* dS = precision == FS_FP64 ? S.S64 : S.S32;
        S.init_state(dS, FS_STOP_ABS_PRECOND, 0, 0, 1 << 30);
        FsPcgState h{};
        HIPCHECK(hipMemcpy(&h, dS, sizeof(h), hipMemcpyDeviceToHost));
        h.rz[0] = 1;
        h.rz[1] = 1;
        h.done = 0;
        HIPCHECK(hipMemcpy(dS, &h, sizeof(h), hipMemcpyHostToDevice));
        FsPcgOp o = S.op64();
        FsPcgWork* w = &S.w64;
        const void* dg = S.diag;
        void* x = S.tmp;  // scratch x (tmp only lives inside a step): the pressure is left untouched
        if (precision == FS_FP32_REFINE) {
            o = S.op32();
            w = &S.w32;
            dg = S.diag32;
            x = S.e32;
        }
        avg_ms[2] = 0;  // the p-update runs inside kernel 0
        for (int which = 0; which < FS_PCG_KERNELS; which++) {
            // as iteration 1: kernel 0 forms the direction on the fly
            fs_pcg_time_kernels(S.st, S.L, o, dg, x, *w, dS, precision == FS_FP64 ? 0 : 1, which, 1);  // warm
            HIPCHECK(hipEventRecord(S.ev[0], S.st));
            for (int r = 0; r < reps; r++)
                fs_pcg_time_kernels(S.st, S.L, o, dg, x, *w, dS, precision == FS_FP64 ? 0 : 1, which, 1);
            HIPCHECK(hipEventRecord(S.ev[1], S.st));
            S.check_launch();
            S.sync();
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
            avg_ms[which] = ms / reps;
        }
        return FS_OK;
    });
}

int fs_set_kernel_sampling(fs_solver* s, int enable) {
    if (s && s->pipe)
        return pipe_guarded(s, [&](FsPipe& P) {
            fs_pipe_sampling(&P, enable != 0);
            return FS_OK;
        });
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        S.sampling = enable != 0;
        S.probe.recorded = false;
        HIPCHECK(hipStreamSynchronize(S.st));
        for (auto& k : S.kprobe) k.reset();
        for (int p = 0; p < 2; p++)
            for (int k = 0; k < 3; k++) {
                S.ksum[p][k] = 0;
                S.kcount[p][k] = 0;
            }
        return FS_OK;
    });
}

int fs_get_frame_kernel_samples(fs_solver* s, int kernel, double* avg_ms, long long* samples,
                                long long* launches) {
    if (!avg_ms || !samples || !launches || kernel < 0 || kernel >= FS_KS_COUNT) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        FsLaunchProbe& k = S.kprobe[kernel];
        k.harvest();
        *avg_ms = k.samples ? k.sum_ms / k.samples : 0;
        *samples = k.samples;
        *launches = k.launches;
        return FS_OK;
    });
}

int fs_get_kernel_samples(fs_solver* s, int precision, double* avg_ms, int* count) {
    if (!avg_ms || !count || (precision != FS_FP64 && precision != FS_FP32_REFINE)) return FS_ERR_BAD_ARG;
    return guarded(s, [&](EulerianFluidSolver3D& S) {
        int p = precision == FS_FP64 ? 0 : 1;
        for (int k = 0; k < 3; k++) {
            count[k] = S.kcount[p][k];
            avg_ms[k] = S.kcount[p][k] ? S.ksum[p][k] / S.kcount[p][k] : 0;
        }
        return FS_OK;
    });
}

}  // extern "C"
