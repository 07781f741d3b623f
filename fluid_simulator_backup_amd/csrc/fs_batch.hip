// fs_batch.hip -- scene-table kernels of a batch's reinitialization stage and
// their graph (fs_batch.h).
#include "fs_batch.h"

#include <algorithm>
#include <cstdlib>
#include <stdexcept>

#define BTX 64
#define BTY 4
#define BT_GRID 2048       // workgroups of every stage launch (a multiple of 8: one share per XCD)
#define BT_FULL_TILE 2048  // elements per tile of the whole-allocation copy
#define BT_CONST_D 2       // layers of the constant fill (ComputeNormals, LEVELSET_3D.h:931)
#define BT_FACE_D 3        // layers of the sweeps' linear fills

namespace {

void check(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// tiles of a scene in each tile space
int scene_tiles(const FsLayout& L, int kind) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    switch (kind) {
        case FS_BT_FULL: return ceil_div(L.total, BT_FULL_TILE);
        case FS_BT_CELL: return ceil_div(nx, BTX) * ceil_div(ny, BTY) * nz;
        case FS_BT_CONST: return ceil_div(ny + 2 * BT_CONST_D, BTX) * ceil_div(nz + 2 * BT_CONST_D, BTY);
        default: {
            const int na = std::max(nx, ny), nc = std::max(ny, nz);
            return 3 * ceil_div(na, BTX) * ceil_div(nc, BTY);
        }
    }
}

// Workgroup b walks its XCD-contiguous share of the T tiles of a kind:
// share q = (b % 8) * (G / 8) + b / 8 of G shares, so that the workgroups of
// one XCD (b % 8 fixed) take adjacent tiles and share their halo rows
// through one L2.  f(scene, tile within the scene) for each.
template <class F>
__device__ __forceinline__ void for_tiles(const FsReinitScene* __restrict__ sc, const int* __restrict__ starts,
                                          int ns, int T, F&& f) {
    const int G = gridDim.x, b = blockIdx.x;
    const int q = (b & 7) * (G >> 3) + (b >> 3);
    const int t0 = (int)((long)q * T / G), t1 = (int)((long)(q + 1) * T / G);
    if (t0 >= t1) return;
    // the scene of tile t0: the last s with starts[s] <= t0
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= t0) lo = mid;
        else hi = mid - 1;
    }
    int s = lo;
    for (int t = t0; t < t1; t++) {
        while (t >= starts[s + 1]) s++;
        f(sc[s], t - starts[s]);
    }
}

__device__ __forceinline__ void cell_of(const FsLayout& L, int t, int& i, int& j, int& k) {
    const int ntx = (L.nx + BTX - 1) / BTX, nty = (L.ny + BTY - 1) / BTY;
    i = (t % ntx) * BTX + threadIdx.x;
    j = ((t / ntx) % nty) * BTY + threadIdx.y;
    k = t / (ntx * nty);
}

// whole allocation, src -> dst (the copy of the level set, its ghosts included)
__global__ void __launch_bounds__(BTX* BTY) k_bt_copy_full(const FsReinitScene* sc, const int* starts, int ns, int T,
                                                           int src, int dst) {
    for_tiles(sc, starts, ns, T, [&](const FsReinitScene& s, int t) {
        const double* a = s.buf[src];
        double* d = s.buf[dst];
        const long base = (long)t * BT_FULL_TILE, tid = threadIdx.y * BTX + threadIdx.x;
        for (int m = 0; m < BT_FULL_TILE / (BTX * BTY); m++) {
            const long e = base + m * (BTX * BTY) + tid;
            if (e < s.L.total) d[e] = a[e];
        }
    });
}

// constant fill of buffer id's ghost layers (only scenes that replay it)
__global__ void __launch_bounds__(BTX* BTY) k_bt_fill_const(const FsReinitScene* sc, const int* starts, int ns, int T,
                                                            int id) {
    for_tiles(sc, starts, ns, T, [&](const FsReinitScene& s, int t) {
        if (!s.fill_const) return;
        const FsBox b = fs_box(s.L, 0);
        const int ntj = (b.nj + 2 * BT_CONST_D + BTX - 1) / BTX;
        const int jj = (t % ntj) * BTX + threadIdx.x - BT_CONST_D, kk = (t / ntj) * BTY + threadIdx.y - BT_CONST_D;
        if (jj >= b.nj + BT_CONST_D || kk >= b.nk + BT_CONST_D) return;
        fs_fill_const_row(s.buf[id], s.buf[id], s.L, b, BT_CONST_D, jj, kk);
    });
}

// cell box of src -> dst
__global__ void __launch_bounds__(BTX* BTY) k_bt_copy_box(const FsReinitScene* sc, const int* starts, int ns, int T,
                                                          int src, int dst) {
    for_tiles(sc, starts, ns, T, [&](const FsReinitScene& s, int t) {
        int i, j, k;
        cell_of(s.L, t, i, j, k);
        if (i >= s.L.nx || j >= s.L.ny) return;
        const long c = s.L.idx(i, j, k);
        s.buf[dst][c] = s.buf[src][c];
    });
}

__global__ void __launch_bounds__(BTX* BTY) k_bt_sign(const FsReinitScene* sc, const int* starts, int ns, int T,
                                                      int phi) {
    for_tiles(sc, starts, ns, T, [&](const FsReinitScene& s, int t) {
        int i, j, k;
        cell_of(s.L, t, i, j, k);
        if (i >= s.L.nx || j >= s.L.ny) return;
        fs_sign_cell(s.buf[phi], s.buf[3], s.L.idx(i, j, k), s.dx2);
    });
}

// Linear ghost fill of buffer dst (BT_FACE_D layers, faces) from its box
// `src` and, for the deeper layers' sources, gsrc; mode 0: src = gsrc =
// phi (the first sweep's copy, compatible schedule); 1: src = gsrc = dst
// (the fixed schedule, and the next copy in both after the compatible
// schedule's gsrc choice below); in mode 1 a compatible-schedule scene reads
// its deeper sources from phi (fs_fill_linear_split(nxt, nxt, phi)).
__global__ void __launch_bounds__(BTX* BTY) k_bt_linear_faces(const FsReinitScene* sc, const int* starts, int ns,
                                                              int T, int dst, int phi, int first) {
    for_tiles(sc, starts, ns, T, [&](const FsReinitScene& s, int t) {
        const FsBox b = fs_box(s.L, 0);
        const int na = max(b.ni, b.nj), nc = max(b.nj, b.nk);
        const int nta = (na + BTX - 1) / BTX, ntc = (nc + BTY - 1) / BTY, per = nta * ntc;
        const int axis = t / per, r = t % per;
        const int a = (r % nta) * BTX + threadIdx.x, c = (r / nta) * BTY + threadIdx.y;
        const double *src, *gsrc;
        if (first) {
            // the first sweep's copy: compatible schedule fs_fill_linear(sghost,
            // phi): box and deeper sources from phi; fixed schedule: from the
            // copy itself (RefreshLevelsetGhost)
            src = gsrc = s.exact ? s.buf[dst] : s.buf[phi];
        } else {
            // the next copy: its own box; deeper sources from phi's ghost
            // layers in the compatible schedule, its own in the fixed one
            src = s.buf[dst];
            gsrc = s.exact ? s.buf[dst] : s.buf[phi];
        }
        fs_linear_faces_row(s.buf[dst], src, gsrc, s.L, b, BT_FACE_D, axis, a, c);
    });
}

__global__ void __launch_bounds__(BTX* BTY) k_bt_reinit(const FsReinitScene* sc, const int* starts, int ns, int T,
                                                        int out, int gh) {
    for_tiles(sc, starts, ns, T, [&](const FsReinitScene& s, int t) {
        int i, j, k;
        cell_of(s.L, t, i, j, k);
        if (i >= s.L.nx || j >= s.L.ny) return;
        const long c = s.L.idx(i, j, k);
        if (s.fast)
            fs_reinit_cell<true>(s.buf[out], s.buf[gh], s.buf[3], s.L, s.ix, s.iy, s.iz, s.eps, s.dt, s.exact, c);
        else
            fs_reinit_cell<false>(s.buf[out], s.buf[gh], s.buf[3], s.L, s.ix, s.iy, s.iz, s.eps, s.dt, s.exact, c);
    });
}

// phi = ca * tmp + cb * phi (k_combine's order), or the average (k_average's)
__global__ void __launch_bounds__(BTX* BTY) k_bt_blend(const FsReinitScene* sc, const int* starts, int ns, int T,
                                                       int phi, int average, double ca, double cb) {
    for_tiles(sc, starts, ns, T, [&](const FsReinitScene& s, int t) {
        int i, j, k;
        cell_of(s.L, t, i, j, k);
        if (i >= s.L.nx || j >= s.L.ny) return;
        const long c = s.L.idx(i, j, k);
        double* p = s.buf[phi];
        const double* tmp = s.buf[2];
        if (average) {
            p[c] = (double)0.5 * (tmp[c] + p[c]);
        } else {
            double x = ca * tmp[c];
            double y = cb * p[c];
            p[c] = x + y;
        }
    });
}

}  // namespace

FsBatchReinit::~FsBatchReinit() {
    for (int q = 0; q < 2; q++) {
        if (exec_[q]) (void)hipGraphExecDestroy(exec_[q]);
        if (graph_[q]) (void)hipGraphDestroy(graph_[q]);
    }
    if (dev_) (void)hipFree(dev_);
    if (starts_) (void)hipFree(starts_);
}

void FsBatchReinit::Build(const std::vector<FsReinitScene>& scenes, int sweeps, int calls) {
    if (scenes.empty() || sweeps < 1 || (sweeps & 1) == 0 || (calls != 2 && calls != 3))
        throw std::invalid_argument("batched reinitialization: an odd sweep count and 2 or 3 calls");
    host_ = scenes;
    sweeps_ = sweeps;
    calls_ = calls;
    const int ns = (int)scenes.size();
    std::vector<int> st((size_t)FS_BT_KINDS * (ns + 1));
    for (int kind = 0; kind < FS_BT_KINDS; kind++) {
        long acc = 0;
        for (int s = 0; s < ns; s++) {
            st[(size_t)kind * (ns + 1) + s] = (int)acc;
            acc += scene_tiles(scenes[s].L, kind);
        }
        if (acc > (1L << 30)) throw std::invalid_argument("batched reinitialization: too many tiles");
        st[(size_t)kind * (ns + 1) + ns] = (int)acc;
        total_[kind] = (int)acc;
    }
    check(hipMalloc(&dev_, sizeof(FsReinitScene) * ns));
    check(hipMalloc(&starts_, sizeof(int) * st.size()));
    check(hipMemcpy(dev_, scenes.data(), sizeof(FsReinitScene) * ns, hipMemcpyHostToDevice));
    check(hipMemcpy(starts_, st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice));
}

void FsBatchReinit::Enqueue(hipStream_t s, int parity) {
    const int ns = (int)host_.size();
    const int A = parity, B = 1 - parity;  // the current level set, and the copy that is reinitialized
    auto row = [&](int kind) { return starts_ + (size_t)kind * (ns + 1); };
    const dim3 g(BT_GRID), t(BTX, BTY);
    const int TF = total_[FS_BT_FULL], TC = total_[FS_BT_CELL], TK = total_[FS_BT_CONST], TE = total_[FS_BT_FACE];
    // the copy, with the constant ghost fill SolveForRK leaves on phi
    hipLaunchKernelGGL(k_bt_copy_full, g, t, 0, s, dev_, row(FS_BT_FULL), ns, TF, A, B);
    hipLaunchKernelGGL(k_bt_fill_const, g, t, 0, s, dev_, row(FS_BT_CONST), ns, TK, B);
    // ReinitializationStage: temp_for_levelset = phi, the calls, the blends
    hipLaunchKernelGGL(k_bt_copy_box, g, t, 0, s, dev_, row(FS_BT_CELL), ns, TC, B, 2);
    for (int call = 0; call < calls_; call++) {
        if (call == 2) hipLaunchKernelGGL(k_bt_blend, g, t, 0, s, dev_, row(FS_BT_CELL), ns, TC, B, 0, 0.75, 0.25);
        hipLaunchKernelGGL(k_bt_sign, g, t, 0, s, dev_, row(FS_BT_CELL), ns, TC, B);
        int src = 4, nxt = 5;
        for (int it = 0; it < sweeps_; it++) {
            if (it == 0) {
                hipLaunchKernelGGL(k_bt_copy_box, g, t, 0, s, dev_, row(FS_BT_CELL), ns, TC, B, 4);
                hipLaunchKernelGGL(k_bt_linear_faces, g, t, 0, s, dev_, row(FS_BT_FACE), ns, TE, 4, B, 1);
            }
            const bool last = it == sweeps_ - 1;
            hipLaunchKernelGGL(k_bt_reinit, g, t, 0, s, dev_, row(FS_BT_CELL), ns, TC, last ? B : nxt, src);
            if (!last) {
                hipLaunchKernelGGL(k_bt_linear_faces, g, t, 0, s, dev_, row(FS_BT_FACE), ns, TE, nxt, B, 0);
                std::swap(src, nxt);
            }
        }
    }
    if (calls_ == 3)
        hipLaunchKernelGGL(k_bt_blend, g, t, 0, s, dev_, row(FS_BT_CELL), ns, TC, B, 0, (double)1 / (double)3,
                           (double)2 / (double)3);
    else
        hipLaunchKernelGGL(k_bt_blend, g, t, 0, s, dev_, row(FS_BT_CELL), ns, TC, B, 1, 0.0, 0.0);
}

void FsBatchReinit::Run(hipStream_t s, int parity) {
    if (std::getenv("FS_BATCH_NO_GRAPH")) {
        Enqueue(s, parity);
        check(hipGetLastError());
        return;
    }
    if (!exec_[parity]) {
        check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        Enqueue(s, parity);
        hipGraph_t gr = nullptr;
        check(hipStreamEndCapture(s, &gr));
        graph_[parity] = gr;
        check(hipGraphInstantiate(&exec_[parity], gr, nullptr, nullptr, 0));
    }
    check(hipGraphLaunch(exec_[parity], s));
}
