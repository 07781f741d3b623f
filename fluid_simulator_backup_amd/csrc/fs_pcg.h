// fs_pcg.h -- matrix-free Jacobi-PCG for the variable-density pressure
// system (PCG_METHOD.cpp:4-126 with the 7-point GFM operator of
// POISSON_SOLVER_3D.h:1030-1213 applied on the fly).
#pragma once
#include <hip/hip_runtime.h>

#include "fs_layout.h"

// Device-resident solver scalars.  Written by one thread (begin-finalize,
// workgroup 0 of the fused K1, the loop-tail kernel); every kernel of
// iteration n reads rz and `done` at entry and returns at once when
// done != 0, so the host may enqueue iterations past convergence.
struct FsPcgState {
    double rz[2];         // r.z of iteration n lives in rz[n & 1]
    double rr, bb;        // r.r and b.b of the last iteration / of the start
    double rsum;          // sum of the FP64 residual (fs_residual64)
    double rs_last;       // r.z at exit (the reference's residual^2, PCG_METHOD.cpp:111)
    double pin_b, pin_d;  // Neumann pin extras of cell (0,0,0), POISSON_SOLVER_3D.h:1145-1179
    int iter, done;       // done: 1 converged, 2 max_iteration reached
    int max_iter, stop_rule;
    double tol2, rtol2;   // squared absolute / relative tolerances
};

// kind FS_OP_GFM: the pressure operator from face coefficients bx, by, bz
// (POISSON_SOLVER_3D.h:1030-1213).  kind FS_OP_ROWS (FP64 only): rows with
// explicit entries a[0..5] to the i-1, i+1, j-1, j+1, k-1, k+1 neighbours
// (0 where the neighbour is not an unknown) and diagonal dg, summed in that
// order as CSR_MATRIX::Multiply sums the rows the reference assigns in it
// (the semi-implicit viscosity system, POISSON_SOLVER_3D.h:3359-3496).
#define FS_OP_GFM 0
#define FS_OP_ROWS 1
struct FsPcgOp {
    const void* bx;
    const void* by;
    const void* bz;
    double idx2, idy2, idz2;  // 1/dx^2, 1/dy^2, 1/dz^2 (POISSON_SOLVER_3D.h:1076)
    // 0: the reference's diagonal, 1/dx^2 times the sum of the face betas
    // (POISSON_SOLVER_3D.h:1213); 1: each face's beta over its own d^2 (the
    // two agree on cubic cells)
    int axis_diag;
    // face coefficients computed in the fused SpMV from the level set phi
    // (cell field, the operator's precision, ghosts as at assembly) instead
    // of read from bx/by/bz: beta_f = gfm(v_c, v_n, phi_n, phi_c) with
    // v = v_neg where phi <= 0, else v_pos (1/rho), 0 on wall faces -- the
    // values fs_assemble stores, bit for bit in FP64
    const void* phi;
    double v_neg, v_pos;
    int kind;
    const double* a[6];
    const double* dg;
};

// Workspace of one precision.  The direction of iteration n lives in p
// (n even) or p2 (n odd).
struct FsPcgWork {
    void *r, *s, *p, *q;    // padded-layout vectors (ghosts stay zero)
    double* partials;       // FS_PCG_SLOTS * FS_PCG_MAX_BLOCKS per-workgroup partial sums
    void* p2;
    // tiling of the fused SpMV (0) and the update (1): z-planes per tile and
    // the cap on workgroups; 0 takes the fixed rule (fs_pcg_tune sets them)
    int zc[2], nbmax[2];
};

// Kernels per PCG iteration: the fused p-update + SpMV + p.q, and the
// x/r/s update with r.s and r.r.
#define FS_PCG_KERNELS 2

#define FS_PCG_MAX_BLOCKS 2048
#define FS_PCG_SLOTS 8

// Measurement only: HIP events recorded on the solve's stream around the
// kernels of the first iteration of a launch batch, ev[2w] before and
// ev[2w+1] after kernel w (0 p-update + SpMV + dot, 1 update).  `iter` is
// that iteration; `recorded` is set when the events were enqueued.
struct FsKernelProbe {
    hipEvent_t ev[2 * FS_PCG_KERNELS];
    int iter;
    bool recorded;
};

// FP64 / FP32 solves on x (warm start).  begin: r = b - A x, z = D^-1 r,
// p = z and the scalars; iterations: n0 .. n0+count-1, then the loop tail of
// the last (the host sees its stop without another batch).
void fs_pcg64_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* rhs, const double* diag,
                    double* x, FsPcgWork& w, FsPcgState* dstate);
void fs_pcg64_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* diag, double* x,
                         FsPcgWork& w, FsPcgState* dstate, int n0, int count, FsKernelProbe* probe = nullptr);
void fs_pcg32_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* rhs, const float* diag,
                    float* x, FsPcgWork& w, FsPcgState* dstate);
void fs_pcg32_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* diag, float* x,
                         FsPcgWork& w, FsPcgState* dstate, int n0, int count);

// q = A x (FP64, exact CSR row order).
void fs_apply_op64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* x, double* q,
                   const FsPcgState* dstate);
// The constant mode of the singly-pinned Neumann system, kept in FP64: with
// the residual r of x just computed (state->rsum = sum(r)), shift
// x += rsum / pin everywhere, which changes r only in the pinned cell
// (r0 -= rsum, exactly for the flux form of the operator), and update r32
// and state->rr to match.  pin = the pinned cell's extra diagonal.
void fs_deflate_constant(hipStream_t st, const FsLayout& L, double* x, double* r, float* r32, FsPcgState* dstate,
                         double pin);
// r = b - A x in FP64; r32 = (float) r; state->rr = r.r, state->bb = b.b,
// state->rsum = sum(r).
void fs_residual64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* b, const double* x, double* r,
                   float* r32, FsPcgWork& w, FsPcgState* dstate);
// FP32 iterations n0 .. n0+count-1 and the loop tail of the last (its r.r
// into state->rr, done = 1 if it met the target).  The probe, if any,
// brackets iteration n0.
void fs_pcg32_chunk(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* diag, float* x, FsPcgWork& w,
                    FsPcgState* dstate, int n0, int count, FsKernelProbe* probe = nullptr);
// Reliable update between an iteration's K2 and the next K1, with r32
// holding the rounded FP64 residual: the recursion's r = r32, s = r32/D and
// the r.s / r.r partials the next K1 reduces; clears `done`.
void fs_pcg32_reliable_update(hipStream_t st, const FsLayout& L, const float* r32, const float* diag, FsPcgWork& w,
                              FsPcgState* dstate);
// x += (double) e
void fs_add_correction(hipStream_t st, const FsLayout& L, double* x, const float* e);
// FP32 copies of the face coefficients and the FP32 Jacobi diagonal.
// dst[i] = (float)src[i] for i < n (a whole padded field)
void fs_to_float(hipStream_t st, const double* src, float* dst, long n);
void fs_make_op32(hipStream_t st, const FsLayout& L, const double* bx, const double* by, const double* bz,
                  const double* diag, float* bx32, float* by32, float* bz32, float* diag32);
// One launch of iteration kernel `which` (0 p-update + SpMV + dot, 1
// update) as iteration n, for timing.
void fs_pcg_time_kernels(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const void* diag, void* x,
                         FsPcgWork& w, FsPcgState* dstate, int precision, int which, int n);

// Tiling of the two iteration kernels for the workspace's shape (GFM
// operator): on the first call for a (grid, precision) in the process, each
// kernel runs every candidate of z-planes per tile {8, 16, 32, 64} and
// workgroup cap {256, 512, 1024, 2048} on the solve's own buffers, with a
// scratch solution and a state that never stops, and the fastest of each is
// kept; later calls (any handle of that shape) reuse the choice.  The values
// of a solve then depend on that choice through the order of its partial
// sums (deterministic within the process).  FS_PCG_TUNE=0 keeps the fixed
// rule; FS_PCG_TILING="zc0,nb0,zc1,nb1" sets one.  Call before begin: it leaves r, s, p, q and the partials undefined.
void fs_pcg_tune(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const void* diag, FsPcgWork& w,
                 FsPcgState* dstate, int precision);
