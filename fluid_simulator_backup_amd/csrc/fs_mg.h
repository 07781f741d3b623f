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

// ---- vector kernels of the FS_FP32_MG conjugate gradient ---------------------
// Partial sums: one double per workgroup of the cell grid (fs_mg_blocks), each
// kernel writes its own; fs_mg_total sums them in a fixed order into *out.
int fs_mg_blocks(const FsLayout& L);
// q = A p (flux form, pin-free); partial p.q
void fs_mg_spmv_dot(hipStream_t st, const FsLayout& L, const float* bx, const float* by, const float* bz, float idx2,
                    float idy2, float idz2, const float* p, float* q, double* part);
// x += alpha p, r -= alpha q; partial r.r
void fs_mg_update(hipStream_t st, const FsLayout& L, float* x, float* r, const float* p, const float* q, float alpha,
                  double* part);
// partials of a.b and b.b (two slots: part[0..nb), part[nb..2nb))
void fs_mg_dot2(hipStream_t st, const FsLayout& L, const float* a, const float* b, double* part);
// p = z + beta p
void fs_mg_direction(hipStream_t st, const FsLayout& L, const float* z, float* p, float beta);
// out[s] = sum of part[s*nb .. (s+1)*nb) in a fixed order, for s < slots;
// mid holds slots * FS_MG_MID doubles of intermediate sums
#define FS_MG_MID 128
void fs_mg_total(hipStream_t st, const double* part, int nb, int slots, double* mid, double* out);
