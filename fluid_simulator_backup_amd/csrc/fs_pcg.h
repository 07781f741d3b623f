// fs_pcg.h -- matrix-free Jacobi-PCG for the variable-density pressure
// system (PCG_METHOD.cpp:4-126 with the 7-point GFM operator of
// POISSON_SOLVER_3D.h:1030-1213 applied on the fly).
#pragma once
#include <hip/hip_runtime.h>

#include "fs_layout.h"

// Device-resident solver scalars.  Kernels read `done` first and return at
// once when it is set, so the host may enqueue iterations past convergence.
struct FsPcgState {
    double rz, rz_old, pq, alpha, beta, rr, bb;
    double pin_b, pin_d;  // Neumann pin extras of cell (0,0,0), POISSON_SOLVER_3D.h:1145-1179
    int iter, done;       // done: 1 converged, 2 max_iteration reached
    int max_iter, stop_rule;
    double tol2, rtol2;   // squared absolute / relative tolerances
    unsigned int counter[4];
};

struct FsPcgOp {
    const void* bx;
    const void* by;
    const void* bz;
    double idx2, idy2, idz2;  // 1/dx^2, 1/dy^2, 1/dz^2 (POISSON_SOLVER_3D.h:1076)
};

// Workspace of one precision.
struct FsPcgWork {
    void *r, *s, *p, *q;    // padded-layout vectors (ghosts stay zero)
    void* diag;             // A(c,c)
    double* partials;       // 3 * nblocks
    int nblocks;
};

int fs_pcg_blocks(const FsLayout& L);

// FP64 solve in place on x (warm start).  Returns iterations; fills state.
// rhs, diag, x are FP64 padded arrays.
void fs_pcg64_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* rhs, const double* diag,
                    double* x, FsPcgWork& w, FsPcgState* dstate);
void fs_pcg64_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* diag, double* x,
                         FsPcgWork& w, FsPcgState* dstate, int count);

// FP32 inner solve of A e = r (e starts at zero).
void fs_pcg32_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* rhs, const float* diag,
                    float* x, FsPcgWork& w, FsPcgState* dstate);
void fs_pcg32_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* diag, float* x,
                         FsPcgWork& w, FsPcgState* dstate, int count);

// q = A x (FP64, exact CSR row order) -- also used for the outer residual.
void fs_apply_op64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* x, double* q,
                   const FsPcgState* dstate);
// r = b - A x in FP64; r32 = (float) r; partial sums of r.r into state->rr
// and of b.b into state->bb (last block).
void fs_residual64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* b, const double* x, double* r,
                   float* r32, FsPcgWork& w, FsPcgState* dstate);
// One launch of each iteration kernel (for timing), ignoring `done`.
void fs_pcg_time_kernels(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const void* diag, void* x,
                         FsPcgWork& w, FsPcgState* dstate, int precision, int which);
// x += (double) e
void fs_add_correction(hipStream_t st, const FsLayout& L, double* x, const float* e);
// FP32 copies of the face coefficients and the FP32 Jacobi diagonal.
void fs_make_op32(hipStream_t st, const FsLayout& L, const double* bx, const double* by, const double* bz,
                  const double* diag, float* bx32, float* by32, float* bz32, float* diag32);
