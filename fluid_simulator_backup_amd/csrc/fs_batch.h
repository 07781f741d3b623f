// fs_batch.h -- the reinitialization stage of many scenes in one launch per
// kernel (a scene table), captured once into a HIP graph per parity.
//
// Every scene of a batch runs its step's reinitialization on a copy of its
// advected level set (as the one-scene concurrent order does) while its own
// stream runs SolveForRK.  Here the copies of all scenes are reinitialized
// together: each kernel of the stage -- the copy, the constant ghost fill
// the assembly applies, the sign function, the 25 (or n) Sussman sweeps of
// each call with their ghost fills, the Runge-Kutta blends -- is one launch
// over every scene's tiles.  A fixed grid of workgroups walks the flattened
// tiles of all scenes in XCD-contiguous chunks; each tile finds its scene by
// binary search in the table's tile offsets.  The per-cell work is that of
// the one-scene kernels (fs_bodies.h), so the values are the same bit for
// bit.  The stage is the same every frame but for which of a scene's two
// level-set buffers is current, so one graph per parity is captured on
// first use and replayed afterwards.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "fs_bodies.h"

// tile spaces of the stage's kernels
enum { FS_BT_FULL = 0, FS_BT_CELL = 1, FS_BT_CONST = 2, FS_BT_FACE = 3, FS_BT_KINDS = 4 };

struct FsReinitScene {
    FsLayout L;
    double ix, iy, iz;      // 1/dx, 1/dy, 1/dz of the cell grid
    double eps, dt, dx2;    // WENO5 / clamp epsilon, fictitious time step, dx^2 of the sign function
    int exact;              // fixed schedule: ghost copies from the copy's own interior, Euclidean norm
    int fill_const;         // replay the constant ghost fill SolveForRK applies to phi
    int fast;               // fast_weno
    // 0 and 1: the level set and its copy (the graph of parity q reads buffer
    // q and leaves its result in 1 - q); 2 temp_for_levelset; 3 the sign
    // field; 4 and 5 the two ghost copies the sweeps ping-pong between
    double* buf[6];
};

class FsBatchReinit {
   public:
    FsBatchReinit() = default;
    ~FsBatchReinit();
    FsBatchReinit(const FsBatchReinit&) = delete;
    FsBatchReinit& operator=(const FsBatchReinit&) = delete;

    // Scenes of one batch: every one with `sweeps` (odd) Sussman sweeps per
    // call and `calls` calls per stage (2: orders 1 and 2, the blend an
    // average; 3: order 3, the TVD-RK3 blends).
    void Build(const std::vector<FsReinitScene>& scenes, int sweeps, int calls);
    // The stage on stream st for the scenes whose current level set is
    // buffer `parity`; replays the graph (captured at the first call of a
    // parity) unless graphs are off (FS_BATCH_NO_GRAPH).
    void Run(hipStream_t st, int parity);
    int Scenes() const { return (int)host_.size(); }

   private:
    void Enqueue(hipStream_t st, int parity);
    std::vector<FsReinitScene> host_;
    FsReinitScene* dev_ = nullptr;
    int* starts_ = nullptr;  // FS_BT_KINDS rows of (scenes + 1) tile offsets
    int total_[FS_BT_KINDS] = {};
    int sweeps_ = 0, calls_ = 0;
    hipGraphExec_t exec_[2] = {};
    hipGraph_t graph_[2] = {};
};
