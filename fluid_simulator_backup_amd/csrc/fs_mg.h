// fs_mg.h -- FP32 geometric multigrid V-cycle on the pressure operator, the
// preconditioner of the FS_FP32_MG inner solve, and that solve's CG.
//
// The operator on every level is in flux form, sum over faces of
// beta_f (x_c - x_n) / d_f^2, with zero coefficients on wall faces (the
// Neumann walls) and no pin: the inner solve works on zero-sum residuals of
// the pin-free operator, as the FP32 Jacobi inner solve does.
//
// Hierarchy: cell-centred, each level halves every axis (ceil); a coarse face
// coefficient is the mean of the fine face coefficients it covers (with the
// spacing doubled, this keeps the finite-volume fluxes), the residual is
// restricted with the transpose of the prolongation (divided by 8), and the
// prolongation is trilinear with clamped (zero-gradient) indices at the walls.
// Smoother: red-black Gauss-Seidel (two or four sweeps each way), red first
// before the coarse correction and black first after it, so the V-cycle is
// a symmetric operator.  Each sweep is one launch that updates both colours (the first
// colour's values held in LDS for three planes), out of place between x and
// the level's residual buffer.  The coarsest level (at most 512 cells) is
// swept in LDS by one workgroup in a single launch.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <vector>

#include "fs_layout.h"
#include "fs_probe.h"

struct FsMgLevel {
    FsLayout L;
    float idx2, idy2, idz2;         // 1/d^2 per axis on this level
    float *bx, *by, *bz, *diag;     // level 0: coefficients borrowed from the FP32 operator
    float *x, *b, *r;               // level 0: x and b are the caller's vectors
    int kc;                         // planes per workgroup of the fused sweep (0 until tuned)
};

class FsMg {
   public:
    FsMg() = default;
    ~FsMg();
    FsMg(const FsMg&) = delete;
    FsMg& operator=(const FsMg&) = delete;

    // Allocates the hierarchy for fine layout L0 (idempotent for one layout).
    void Init(const FsLayout& L0, double idx2, double idy2, double idz2);
    // Takes the fine coefficients (padded FP32 face fields, zero wall faces)
    // and forms every level's coefficients and diagonals.
    void SetOperator(hipStream_t st, const float* bx, const float* by, const float* bz);
    // z = M r: one V-cycle from z = 0.  r and z are padded FP32 cell fields
    // (zero ghosts).  The first call also tunes each level's sweep: the
    // planes per workgroup (2, 4, 8 or 16) that time fastest on this GPU,
    // which changes no value.
    void Apply(hipStream_t st, const float* r, float* z);
    int Levels() const { return (int)lv_.size(); }
    // finest-level smoother launches of one V-cycle
    int FinestLaunchesPerCycle() const { return (split_ ? 4 : 2) * nu_; }
    bool Tuned() const { return tuned_ || split_; }
    // tunes the fused sweep now, on the finest level with b = r and x = z
    // (read, and the level's own residual buffer written); Apply tunes on its
    // first call otherwise.  A caller with other streams drains them first so
    // the timings are the sweep's.
    void Tune(hipStream_t st, const float* r, float* z);
    // red-black sweeps before and after the coarse correction (an odd count
    // ends the fused sweeps' ping-pong in the residual buffer, copied back)
    void SetSweeps(int nu) {
        if (nu < 1) throw std::invalid_argument("multigrid sweeps must be at least 1");
        nu_ = nu;
    }
    // When set, the first finest-level smoother launch (one red-black sweep)
    // of every V-cycle is timed and every finest-level smoother launch
    // counted (measurement only).
    FsLaunchProbe* probe = nullptr;
    // When set, every V-cycle launch returns at once while *skip != 0 (the
    // CG's `done` flag, so iterations enqueued past convergence cost only
    // their launches).
    const int* skip = nullptr;

   private:
    void Cycle(hipStream_t st, int l);
    void Smooth(hipStream_t st, int l, int c1);
    void Release();
    bool tuned_ = false, built_ = false;
    std::vector<FsMgLevel> lv_;
    bool split_ = false;
    int nu_ = 2;
    float* diag0_ = nullptr;
};

// ---- the FS_FP32_MG conjugate gradient, scalars on the device -----------------
// The CG's scalars live in device memory (FsMgCg): every vector kernel reads
// alpha, beta and `done` there, and single-workgroup kernels form them from
// the partial sums, so a whole iteration (and its V-cycle, through
// FsMg::skip) is enqueued without the host; after `done` is set the
// remaining launches return at once.  The host reads the state every few
// iterations.  Partial sums: one double per workgroup of a z-marching
// 64 x 4 x FS_MGC_KZ tiling (fs_mgc_blocks), summed in a fixed order.
struct FsMgCg {
    double rz, rr0, rtol2;  // r.z of the current iteration, r.r at the start, rtol^2
    double pq, rr;          // p.q and r.r of the last iteration
    float alpha, beta;
    int it, max_it;
    int done;  // 0 running; 1 converged; 2 p.q <= 0; 3 max_it reached; 4 zero right-hand side
};
#define FS_MGC_KZ 8
#define FS_MG_MID 128
int fs_mgc_blocks(const FsLayout& L);
// state := start of a solve (it 0, done 0, rtol^2, max_it)
void fs_mgc_start(hipStream_t st, FsMgCg* S, double rtol2, int max_it);
// partials of a.b and a.a (two runs of nb); skipped once done
void fs_mgc_dot2(hipStream_t st, const FsLayout& L, const float* a, const float* b, double* part, const FsMgCg* S);
// q = A p (flux form, pin-free), partial p.q; skipped once done
void fs_mgc_spmv_dot(hipStream_t st, const FsLayout& L, const float* bx, const float* by, const float* bz, float idx2,
                     float idy2, float idz2, const float* p, float* q, double* part, const FsMgCg* S);
// x += alpha p, r -= alpha q, partial r.r; skipped once done
void fs_mgc_update(hipStream_t st, const FsLayout& L, float* x, float* r, const float* p, const float* q, double* part,
                   const FsMgCg* S);
// p = z + beta p; skipped once done
void fs_mgc_direction(hipStream_t st, const FsLayout& L, const float* z, float* p, const FsMgCg* S);
// Sum `slots` runs of nb partials in a fixed order (mid: slots * FS_MG_MID
// doubles) and update the state: FS_MGC_INIT (rz, rr0 from a.b, a.a; done 4
// if rr0 is not positive), FS_MGC_PQ (alpha = rz / p.q; done 2 if p.q is not
// positive), FS_MGC_RR (it += 1; done 1 at r.r <= rtol^2 rr0, else 3 at
// max_it), FS_MGC_RZ (beta = rz_new / rz, rz = rz_new).
#define FS_MGC_INIT 0
#define FS_MGC_PQ 1
#define FS_MGC_RR 2
#define FS_MGC_RZ 3
void fs_mgc_total(hipStream_t st, const double* part, int nb, int slots, double* mid, FsMgCg* S, int kind);
// v = 0 over a whole padded field (n elements); skipped while *skip != 0
void fs_mg_zero(hipStream_t st, float* v, long n, const int* skip);
