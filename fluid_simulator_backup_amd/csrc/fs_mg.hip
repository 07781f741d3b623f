// fs_mg.hip -- FP32 multigrid V-cycle and the vector kernels of the
// FS_FP32_MG inner conjugate gradient (the hierarchy is described in
// fs_mg.h).
#include "fs_mg.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#define MBX 64
#define MBY 4
#define MNT (MBX * MBY)
#ifndef MG_COARSE_CELLS
#define MG_COARSE_CELLS 512  // coarsen while a level has more cells
#endif
#ifndef MG_COARSE_SWEEPS
#define MG_COARSE_SWEEPS 12  // red-black sweeps of the coarsest level, each way (24: no fewer iterations; 6: more)
#endif

namespace {

void check(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

dim3 cells_grid(const FsLayout& L) { return dim3((L.nx + MBX - 1) / MBX, (L.ny + MBY - 1) / MBY, L.nz); }

// sum over faces of beta_f (x_c - x_n) / d_f^2 at cell c
__device__ __forceinline__ float flux_row(const FsLayout& L, const float* __restrict__ bx,
                                          const float* __restrict__ by, const float* __restrict__ bz, float ix,
                                          float iy, float iz, const float* __restrict__ x, long c) {
    const float xc = x[c];
    return bx[c] * (ix * (xc - x[c - 1])) + bx[c + 1] * (ix * (xc - x[c + 1])) + by[c] * (iy * (xc - x[c - L.sx])) +
           by[c + L.sx] * (iy * (xc - x[c + L.sx])) + bz[c] * (iz * (xc - x[c - L.sy])) +
           bz[c + L.sy] * (iz * (xc - x[c + L.sy]));
}

// workgroup sum of one double per thread into part[block], in a fixed order:
// a butterfly within each 64-lane wave (MBX = 64, one wave per row of the
// block), then the MBY wave sums in row order
__device__ __forceinline__ void block_partial(double v, double* part) {
    static_assert(MBX == 64, "one wave per block row");
    __shared__ double sh[MBY];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (threadIdx.x == 0) sh[threadIdx.y] = v;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double s = sh[0];
        for (int w = 1; w < MBY; w++) s += sh[w];
        part[((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
    }
    __syncthreads();  // sh is reused by a second call in the same kernel
}

// 1-D prolongation weight of coarse cell C for fine cell f (trilinear,
// indices clamped at the walls): 3/4 from f/2, 1/4 from its other neighbour
__device__ __forceinline__ float pweight(int f, int C, int nc) {
    int c0 = f >> 1;
    int c1 = (f & 1) ? c0 + 1 : c0 - 1;
    c1 = c1 < 0 ? 0 : (c1 > nc - 1 ? nc - 1 : c1);
    float w = 0;
    if (c0 == C) w += 0.75f;
    if (c1 == C) w += 0.25f;
    return w;
}

__global__ void __launch_bounds__(MNT) k_mg_diag(FsLayout L, const float* __restrict__ bx,
                                                 const float* __restrict__ by, const float* __restrict__ bz,
                                                 float ix, float iy, float iz, float* __restrict__ diag) {
    int i = blockIdx.x * MBX + threadIdx.x, j = blockIdx.y * MBY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    diag[c] = ix * (bx[c] + bx[c + 1]) + iy * (by[c] + by[c + L.sx]) + iz * (bz[c] + bz[c + L.sy]);
}

// conductance of the fine faces f-1, f, f+1 (stride st) in series with
// weights 1/4, 1/2, 1/4 (the path between two coarse cell centres): 0 if any
// is 0; faces outside [0, nf] are left out
__device__ __forceinline__ float series3(const float* __restrict__ b, long c, long st, int f, int nf) {
    float res = 0, w = 0;
    for (int d = -1; d <= 1; d++) {
        int g = f + d;
        if (g < 0 || g > nf) continue;
        float bv = b[c + d * st], wd = d == 0 ? 0.5f : 0.25f;
        if (bv == 0.f) return 0.f;
        res += wd / bv;
        w += wd;
    }
    return w / res;
}

// coarse face coefficients: over the fine face lines a coarse face covers,
// the mean of each line's series conductance across the faces the segment
// between the two coarse cell centres passes (harmonic along the normal,
// arithmetic across it, as conductances in series and in parallel combine)
__global__ void __launch_bounds__(MNT) k_mg_coarsen(FsLayout Lf, const float* __restrict__ fx,
                                                    const float* __restrict__ fy, const float* __restrict__ fz,
                                                    FsLayout Lc, float* __restrict__ cx, float* __restrict__ cy,
                                                    float* __restrict__ cz) {
    int I = blockIdx.x * MBX + threadIdx.x, J = blockIdx.y * MBY + threadIdx.y, K = blockIdx.z;
    if (I > Lc.nx || J > Lc.ny) return;
    long C = Lc.idx(I, J, K);
    // x-faces: fine face index min(2I, nx), over fine rows 2J, 2J+1 and planes 2K, 2K+1
    if (J < Lc.ny && K < Lc.nz) {
        int fi = min(2 * I, Lf.nx);
        float s = 0;
        int n = 0;
        for (int dk = 0; dk < 2; dk++)
            for (int dj = 0; dj < 2; dj++) {
                int j = 2 * J + dj, k = 2 * K + dk;
                if (j < Lf.ny && k < Lf.nz) {
                    s += series3(fx, Lf.idx(fi, j, k), 1, fi, Lf.nx);
                    n++;
                }
            }
        cx[C] = n ? s / (float)n : 0.f;
    }
    if (I < Lc.nx && K < Lc.nz) {
        int fj = min(2 * J, Lf.ny);
        float s = 0;
        int n = 0;
        for (int dk = 0; dk < 2; dk++)
            for (int di = 0; di < 2; di++) {
                int i = 2 * I + di, k = 2 * K + dk;
                if (i < Lf.nx && k < Lf.nz) {
                    s += series3(fy, Lf.idx(i, fj, k), Lf.sx, fj, Lf.ny);
                    n++;
                }
            }
        cy[C] = n ? s / (float)n : 0.f;
    }
    if (I < Lc.nx && J < Lc.ny) {
        int fk = min(2 * K, Lf.nz);
        float s = 0;
        int n = 0;
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                int i = 2 * I + di, j = 2 * J + dj;
                if (i < Lf.nx && j < Lf.ny) {
                    s += series3(fz, Lf.idx(i, j, fk), Lf.sy, fk, Lf.nz);
                    n++;
                }
            }
        cz[C] = n ? s / (float)n : 0.f;
    }
}

// red-black Gauss-Seidel half sweep: cells with (i+j+k) % 2 == color.
// Every smoother divides by the diagonal through the hardware reciprocal
// (within 1 ulp): the V-cycle is a preconditioner, and all its smoothers use
// the same form, so the fused and split sweeps still agree bit for bit.
__global__ void __launch_bounds__(MNT) k_mg_smooth(FsLayout L, const float* __restrict__ bx,
                                                   const float* __restrict__ by, const float* __restrict__ bz,
                                                   float ix, float iy, float iz, const float* __restrict__ diag,
                                                   const float* __restrict__ b, float* __restrict__ x, int color,
                                                   const int* __restrict__ skip) {
    if (skip && *skip) return;
    int i = blockIdx.x * MBX + threadIdx.x, j = blockIdx.y * MBY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny || ((i + j + k) & 1) != color) return;
    long c = L.idx(i, j, k);
    float d = diag[c];
    if (d == 0.f) return;
    float s = b[c] + bx[c] * (ix * x[c - 1]) + bx[c + 1] * (ix * x[c + 1]) + by[c] * (iy * x[c - L.sx]) +
              by[c + L.sx] * (iy * x[c + L.sx]) + bz[c] * (iz * x[c - L.sy]) + bz[c + L.sy] * (iz * x[c + L.sy]);
    x[c] = s * __builtin_amdgcn_rcpf(d);
}

__global__ void __launch_bounds__(MNT) k_mg_residual(FsLayout L, const float* __restrict__ bx,
                                                     const float* __restrict__ by, const float* __restrict__ bz,
                                                     float ix, float iy, float iz, const float* __restrict__ b,
                                                     const float* __restrict__ x, float* __restrict__ r,
                                                     const int* __restrict__ skip) {
    if (skip && *skip) return;
    int i = blockIdx.x * MBX + threadIdx.x, j = blockIdx.y * MBY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    r[c] = b[c] - flux_row(L, bx, by, bz, ix, iy, iz, x, c);
}

// coarse right-hand side: transpose of the prolongation over the 4x4x4 fine
// block around the coarse cell, divided by 8; coarse x set to 0
__global__ void __launch_bounds__(MNT) k_mg_restrict(FsLayout Lf, const float* __restrict__ rf, FsLayout Lc,
                                                     float* __restrict__ bc, float* __restrict__ xc,
                                                     const int* __restrict__ skip) {
    if (skip && *skip) return;
    int I = blockIdx.x * MBX + threadIdx.x, J = blockIdx.y * MBY + threadIdx.y, K = blockIdx.z;
    if (I >= Lc.nx || J >= Lc.ny) return;
    float s = 0;
    for (int dk = -1; dk <= 2; dk++) {
        int k = 2 * K + dk;
        if (k < 0 || k >= Lf.nz) continue;
        float wk = pweight(k, K, Lc.nz);
        if (wk == 0.f) continue;
        for (int dj = -1; dj <= 2; dj++) {
            int j = 2 * J + dj;
            if (j < 0 || j >= Lf.ny) continue;
            float wj = pweight(j, J, Lc.ny);
            if (wj == 0.f) continue;
            float row = 0;
            for (int di = -1; di <= 2; di++) {
                int i = 2 * I + di;
                if (i < 0 || i >= Lf.nx) continue;
                row += pweight(i, I, Lc.nx) * rf[Lf.idx(i, j, k)];
            }
            s += wk * wj * row;
        }
    }
    long C = Lc.idx(I, J, K);
    bc[C] = 0.125f * s;
    xc[C] = 0.f;
}

// x_fine += P x_coarse (trilinear, clamped)
__global__ void __launch_bounds__(MNT) k_mg_prolong(FsLayout Lc, const float* __restrict__ xc, FsLayout Lf,
                                                    float* __restrict__ xf, const int* __restrict__ skip) {
    if (skip && *skip) return;
    int i = blockIdx.x * MBX + threadIdx.x, j = blockIdx.y * MBY + threadIdx.y, k = blockIdx.z;
    if (i >= Lf.nx || j >= Lf.ny) return;
    auto pair = [](int f, int nc, int& c0, int& c1) {
        c0 = f >> 1;
        c1 = (f & 1) ? c0 + 1 : c0 - 1;
        c1 = c1 < 0 ? 0 : (c1 > nc - 1 ? nc - 1 : c1);
    };
    int i0, i1, j0, j1, k0, k1;
    pair(i, Lc.nx, i0, i1);
    pair(j, Lc.ny, j0, j1);
    pair(k, Lc.nz, k0, k1);
    const int ii[2] = {i0, i1}, jj[2] = {j0, j1}, kk[2] = {k0, k1};
    const float w[2] = {0.75f, 0.25f};
    float s = 0;
    for (int c = 0; c < 2; c++)
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) s += w[a] * w[b] * w[c] * xc[Lc.idx(ii[a], jj[b], kk[c])];
    xf[Lf.idx(i, j, k)] += s;
}

// coarsest level (at most MG_COARSE_CELLS cells): red-black sweeps (red
// first) then the mirrored sweeps, one workgroup, x from 0.  The level is
// copied into LDS in a compact layout with one ghost layer (ghost x zero,
// wall faces' coefficients zero), swept there and x written back.
#define MG_CPAD 2400  // (nx+2)(ny+2)(nz+2) of the coarsest level at most (Init coarsens until it fits)
__global__ void __launch_bounds__(MNT) k_mg_coarse(FsLayout L, const float* __restrict__ bx,
                                                   const float* __restrict__ by, const float* __restrict__ bz,
                                                   float ix, float iy, float iz, const float* __restrict__ diag,
                                                   const float* __restrict__ b, float* __restrict__ x, int sweeps,
                                                   const int* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ float sx[MG_CPAD], sbx[MG_CPAD], sby[MG_CPAD], sbz[MG_CPAD], sb[MG_CPAD], sd[MG_CPAD];
    const int t = threadIdx.y * MBX + threadIdx.x;
    const int px = L.nx + 2, py = L.ny + 2, pxy = px * py, ptot = pxy * (L.nz + 2);
    for (int m = t; m < ptot; m += MNT) {
        sx[m] = 0.f;
        sbx[m] = sby[m] = sbz[m] = sb[m] = sd[m] = 0.f;
    }
    __syncthreads();
    const int n = L.nx * L.ny * L.nz;
    for (int m = t; m < n; m += MNT) {
        int i = m % L.nx, j = (m / L.nx) % L.ny, k = m / (L.nx * L.ny);
        long c = L.idx(i, j, k);
        int q = (i + 1) + px * (j + 1) + pxy * (k + 1);
        sbx[q] = bx[c];
        sby[q] = by[c];
        sbz[q] = bz[c];
        sb[q] = b[c];
        sd[q] = diag[c];
        // the far wall faces belong to the ghost cell beyond
        if (i == L.nx - 1) sbx[q + 1] = bx[c + 1];
        if (j == L.ny - 1) sby[q + px] = by[c + L.sx];
        if (k == L.nz - 1) sbz[q + pxy] = bz[c + L.sy];
    }
    __syncthreads();
    // thread t owns the t-th cell of each colour in index order (at most
    // MNT cells of a colour: n <= MG_COARSE_CELLS = 2 MNT); -1 if none
    __shared__ int qcol[2][MNT];
    for (int m = t; m < 2 * MNT; m += MNT) qcol[m / MNT][m % MNT] = -1;
    __syncthreads();
    if (t == 0) {
        int cnt[2] = {0, 0};
        for (int k = 0; k < L.nz; k++)
            for (int j = 0; j < L.ny; j++)
                for (int i = 0; i < L.nx; i++) {
                    int h = (i + j + k) & 1;
                    qcol[h][cnt[h]++] = (i + 1) + px * (j + 1) + pxy * (k + 1);
                }
    }
    __syncthreads();
    const int qr = qcol[0][t], qb = qcol[1][t];
    for (int s = 0; s < 4 * sweeps; s++) {
        // forward sweeps red, black; then backward black, red
        const int half = s < 2 * sweeps ? (s & 1) : 1 - (s & 1);
        const int q = half ? qb : qr;
        if (q >= 0) {
            float d = sd[q];
            if (d != 0.f) {
                float v = sb[q] + sbx[q] * (ix * sx[q - 1]) + sbx[q + 1] * (ix * sx[q + 1]) +
                          sby[q] * (iy * sx[q - px]) + sby[q + px] * (iy * sx[q + px]) + sbz[q] * (iz * sx[q - pxy]) +
                          sbz[q + pxy] * (iz * sx[q + pxy]);
                sx[q] = v * __builtin_amdgcn_rcpf(d);
            }
        }
        __syncthreads();
    }
    for (int m = t; m < n; m += MNT) {
        int i = m % L.nx, j = (m / L.nx) % L.ny, k = m / (L.nx * L.ny);
        x[L.idx(i, j, k)] = sx[(i + 1) + px * (j + 1) + pxy * (k + 1)];
    }
}

// One symmetric-order red-black Gauss-Seidel sweep, out of place: the cells
// of colour c1 ((i+j+k) % 2 == c1) from x_in, then the other colour from
// those, written with them to x_out; the same arithmetic per cell as two
// k_mg_smooth launches, the diagonal formed from the coefficients as
// k_mg_diag forms it.  A workgroup owns a 64 x 16 column tile and marches
// up kc planes, two per step; the first colour's new values on the tile and
// a ring of one cell around it go to LDS (planes k-1 .. k+2), so the
// coefficients serve both colours.
#define SBX 64
#ifndef SBY
#define SBY 16
#endif
#define SRX (SBX + 2)
#define SRY (SBY + 2)
__global__ void __launch_bounds__(SBX* SBY) k_mg_rb_sweep(FsLayout L, const float* __restrict__ bx,
                                                       const float* __restrict__ by, const float* __restrict__ bz,
                                                       float ix, float iy, float iz, const float* __restrict__ b,
                                                       const float* __restrict__ x_in, float* __restrict__ x_out,
                                                       int c1, int kc, int gx, int gy, const int* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ float R[4][SRY * SRX];
    const int tx = threadIdx.x, ty = threadIdx.y, t = ty * SBX + tx;
    // workgroup b runs on XCD b % 8; with a multiple of 8 tiles, each XCD
    // takes a contiguous range of tiles (x fastest), so the rings that
    // neighbouring tiles share are read through one L2
    const int nt = gridDim.x, wg = blockIdx.x;
    const int tile = (nt & 7) == 0 ? (wg & 7) * (nt >> 3) + (wg >> 3) : wg;
    const int i0 = (tile % gx) * SBX, j0 = ((tile / gx) % gy) * SBY, k0 = (tile / (gx * gy)) * kc;
    const int k1 = min(L.nz, k0 + kc);
    // first colour at plane kk on the tile and its ring: one thread per cell
    // of that colour (half of each ring row, SRX even), every operand loaded
    // before any is used (addresses clamped into the box, 0 outside it).
    // The second colour's update reads R only at first-colour cells.
    static_assert(SRX % 2 == 0 && (SRX / 2) * SRY <= SBX * SBY, "one first-colour ring cell per thread");
    auto stage = [&](int kk) {
        if (t >= (SRX / 2) * SRY) return;
        float* Rp = R[(kk + 4) % 4];
        const int lj = t / (SRX / 2), j = j0 - 1 + lj;
        const int li = ((c1 - (i0 - 1 + j + kk)) & 1) + 2 * (t % (SRX / 2)), i = i0 - 1 + li;
        const bool in = i >= 0 && i < L.nx && j >= 0 && j < L.ny && kk >= 0 && kk < L.nz;
        const long c = L.idx(min(max(i, 0), L.nx - 1), min(max(j, 0), L.ny - 1), min(max(kk, 0), L.nz - 1));
        const float x0 = x_in[c], xw = x_in[c - 1], xe = x_in[c + 1], xs = x_in[c - L.sx], xn = x_in[c + L.sx],
                    xb = x_in[c - L.sy], xt = x_in[c + L.sy];
        const float cw = bx[c], ce = bx[c + 1], cs = by[c], cn = by[c + L.sx], cb = bz[c], ct = bz[c + L.sy];
        const float bc = b[c];
        const float d = ix * (cw + ce) + iy * (cs + cn) + iz * (cb + ct);
        const float v = bc + cw * (ix * xw) + ce * (ix * xe) + cs * (iy * xs) + cn * (iy * xn) + cb * (iz * xb) +
                        ct * (iz * xt);
        const float r = v * __builtin_amdgcn_rcpf(d);
        Rp[lj * SRX + li] = in ? (d != 0.f ? r : x0) : 0.f;
    };
    stage(k0 - 1);
    stage(k0);
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i < L.nx && j < L.ny;
    const int q = (ty + 1) * SRX + tx + 1;
    // two planes per step: the first colour staged for planes k+1 and k+2,
    // then the second colour of planes k and k+1 (R holds planes k-1..k+2)
    for (int k = k0; k < k1; k += 2) {
        // the second colour's operands, loaded before the staging so all the
        // step's loads are in flight together
        float x0[2], bc[2], cw[2], ce[2], cs[2], cn[2], cb[2], ct[2];
        long c[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            c[h] = L.idx(min(i, L.nx - 1), min(j, L.ny - 1), min(k + h, L.nz - 1));
            x0[h] = x_in[c[h]];
            bc[h] = b[c[h]];
            cw[h] = bx[c[h]];
            ce[h] = bx[c[h] + 1];
            cs[h] = by[c[h]];
            cn[h] = by[c[h] + L.sx];
            cb[h] = bz[c[h]];
            ct[h] = bz[c[h] + L.sy];
        }
        stage(k + 1);
        stage(k + 2);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int kk = k + h;
            if (own && kk < k1) {
                const float *Rm = R[(kk + 3) % 4], *R0 = R[kk % 4], *Rp = R[(kk + 1) % 4];
                const float d = ix * (cw[h] + ce[h]) + iy * (cs[h] + cn[h]) + iz * (cb[h] + ct[h]);
                const float v = bc[h] + cw[h] * (ix * R0[q - 1]) + ce[h] * (ix * R0[q + 1]) +
                                cs[h] * (iy * R0[q - SRX]) + cn[h] * (iy * R0[q + SRX]) + cb[h] * (iz * Rm[q]) +
                                ct[h] * (iz * Rp[q]);
                const float r = v * __builtin_amdgcn_rcpf(d);
                x_out[c[h]] = ((i + j + kk) & 1) == c1 ? R0[q] : (d != 0.f ? r : x0[h]);
            }
        }
        __syncthreads();
    }
}

// ---- CG vector kernels (scalars in FsMgCg) ----------------------------------
// A workgroup owns a 64 x 4 column tile and marches up FS_MGC_KZ planes,
// each thread keeping its column's running partial; one partial per
// workgroup.
dim3 cg_grid(const FsLayout& L) {
    return dim3((L.nx + MBX - 1) / MBX, (L.ny + MBY - 1) / MBY, (L.nz + FS_MGC_KZ - 1) / FS_MGC_KZ);
}

__global__ void __launch_bounds__(MNT) k_mgc_spmv_dot(FsLayout L, const float* __restrict__ bx,
                                                      const float* __restrict__ by, const float* __restrict__ bz,
                                                      float ix, float iy, float iz, const float* __restrict__ p,
                                                      float* __restrict__ q, double* part, const FsMgCg* S) {
    if (S->done) return;
    const int i = blockIdx.x * MBX + threadIdx.x, j = blockIdx.y * MBY + threadIdx.y, k0 = blockIdx.z * FS_MGC_KZ;
    const int k1 = min(L.nz, k0 + FS_MGC_KZ);
    double acc = 0;
    if (i < L.nx && j < L.ny) {
        long c = L.idx(i, j, k0);
        float pm = p[c - L.sy], pc = p[c];
        for (int k = k0; k < k1; k++, c += L.sy) {
            const float pp = p[c + L.sy];
            // flux_row's terms in its order, the z neighbours from the march
            const float v = bx[c] * (ix * (pc - p[c - 1])) + bx[c + 1] * (ix * (pc - p[c + 1])) +
                            by[c] * (iy * (pc - p[c - L.sx])) + by[c + L.sx] * (iy * (pc - p[c + L.sx])) +
                            bz[c] * (iz * (pc - pm)) + bz[c + L.sy] * (iz * (pc - pp));
            q[c] = v;
            acc += (double)pc * (double)v;
            pm = pc;
            pc = pp;
        }
    }
    block_partial(acc, part);
}

__global__ void __launch_bounds__(MNT) k_mgc_update(FsLayout L, float* __restrict__ x, float* __restrict__ r,
                                                    const float* __restrict__ p, const float* __restrict__ q,
                                                    double* part, const FsMgCg* S) {
    if (S->done) return;
    const float alpha = S->alpha;
    const int i = blockIdx.x * MBX + threadIdx.x, j = blockIdx.y * MBY + threadIdx.y, k0 = blockIdx.z * FS_MGC_KZ;
    const int k1 = min(L.nz, k0 + FS_MGC_KZ);
    double acc = 0;
    if (i < L.nx && j < L.ny) {
        long c = L.idx(i, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy) {
            x[c] += alpha * p[c];
            const float rv = r[c] - alpha * q[c];
            r[c] = rv;
            acc += (double)rv * (double)rv;
        }
    }
    block_partial(acc, part);
}

__global__ void __launch_bounds__(MNT) k_mgc_dot2(FsLayout L, const float* __restrict__ a,
                                                  const float* __restrict__ b, double* part, int nb,
                                                  const FsMgCg* S) {
    if (S->done) return;
    const int i = blockIdx.x * MBX + threadIdx.x, j = blockIdx.y * MBY + threadIdx.y, k0 = blockIdx.z * FS_MGC_KZ;
    const int k1 = min(L.nz, k0 + FS_MGC_KZ);
    double ab = 0, aa = 0;
    if (i < L.nx && j < L.ny) {
        long c = L.idx(i, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy) {
            const double av = a[c];
            ab += av * (double)b[c];
            aa += av * av;
        }
    }
    block_partial(ab, part);
    block_partial(aa, part + nb);
}

__global__ void __launch_bounds__(MNT) k_mgc_direction(FsLayout L, const float* __restrict__ z, float* __restrict__ p,
                                                       const FsMgCg* S) {
    if (S->done) return;
    const float beta = S->beta;
    const int i = blockIdx.x * MBX + threadIdx.x, j = blockIdx.y * MBY + threadIdx.y, k0 = blockIdx.z * FS_MGC_KZ;
    if (i >= L.nx || j >= L.ny) return;
    const int k1 = min(L.nz, k0 + FS_MGC_KZ);
    long c = L.idx(i, j, k0);
    for (int k = k0; k < k1; k++, c += L.sy) p[c] = z[c] + beta * p[c];
}

__global__ void k_mgc_start(FsMgCg* S, double rtol2, int max_it) {
    FsMgCg s{};
    s.rtol2 = rtol2;
    s.max_it = max_it;
    *S = s;
}

// Sums of `slots` runs of nb partials in two launches, each in a fixed
// order: workgroup (b, s) sums its contiguous share of run s into
// mid[s * FS_MG_MID + b]; then one workgroup sums each run's mids and
// updates the state.
__global__ void __launch_bounds__(MNT) k_mgc_total_mid(const double* part, int nb, double* mid, const FsMgCg* S,
                                                       int always) {
    if (!always && S->done) return;
    __shared__ double sh[MNT];
    const int t = threadIdx.x, b = blockIdx.x;
    const int share = (nb + FS_MG_MID - 1) / FS_MG_MID;
    const double* src = part + (long)blockIdx.y * nb;
    const int m0 = b * share, m1 = min(nb, m0 + share);
    double s = 0;
    for (int m = m0 + t; m < m1; m += MNT) s += src[m];
    sh[t] = s;
    __syncthreads();
    for (int w = MNT / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) mid[blockIdx.y * FS_MG_MID + b] = sh[0];
}

__global__ void __launch_bounds__(MNT) k_mgc_finish(const double* mid, int slots, FsMgCg* S, int kind) {
    if (kind != FS_MGC_INIT && S->done) return;
    __shared__ double sh[MNT];
    __shared__ double tot[2];
    const int t = threadIdx.x;
    for (int s = 0; s < slots; s++) {
        sh[t] = t < FS_MG_MID ? mid[s * FS_MG_MID + t] : 0.0;
        __syncthreads();
        for (int w = MNT / 2; w > 0; w >>= 1) {
            if (t < w) sh[t] += sh[t + w];
            __syncthreads();
        }
        if (t == 0) tot[s] = sh[0];
        __syncthreads();
    }
    if (t != 0) return;
    FsMgCg s = *S;
    const double v = tot[0];
    if (kind == FS_MGC_INIT) {
        s.rz = v;
        s.rr0 = tot[1];
        s.done = s.rr0 > 0 ? 0 : 4;
    } else if (kind == FS_MGC_PQ) {
        s.pq = v;
        if (!(v > 0))
            s.done = 2;
        else
            s.alpha = (float)(s.rz / v);
    } else if (kind == FS_MGC_RR) {
        s.rr = v;
        s.it++;
        if (v <= s.rtol2 * s.rr0)
            s.done = 1;
        else if (s.it >= s.max_it)
            s.done = 3;
    } else {
        s.beta = (float)(v / s.rz);
        s.rz = v;
    }
    *S = s;
}

__global__ void __launch_bounds__(MNT) k_mg_zero(float* __restrict__ v, long n, const int* skip) {
    if (skip && *skip) return;
    for (long m = (long)blockIdx.x * MNT + threadIdx.x; m < n; m += (long)gridDim.x * MNT) v[m] = 0.f;
}

}  // namespace

// ---- hierarchy ---------------------------------------------------------------
FsMg::~FsMg() { Release(); }

void FsMg::Release() {
    for (size_t l = 0; l < lv_.size(); l++) {
        FsMgLevel& v = lv_[l];
        if (l > 0)
            for (float* f : {v.bx, v.by, v.bz, v.diag, v.x, v.b})
                if (f) (void)hipFree(f);
        if (v.r) (void)hipFree(v.r);
    }
    lv_.clear();
    if (diag0_) (void)hipFree(diag0_);
    diag0_ = nullptr;
    built_ = tuned_ = false;
}

void FsMg::Init(const FsLayout& L0, double idx2, double idy2, double idz2) {
    if (built_ && lv_[0].L.nx == L0.nx && lv_[0].L.ny == L0.ny && lv_[0].L.nz == L0.nz) return;
    if (built_) throw std::runtime_error("multigrid hierarchy built for another grid");
    Release();  // what an earlier, failed Init left
    auto field = [](const FsLayout& L) {
        float* p = nullptr;
        check(hipMalloc(&p, sizeof(float) * (size_t)L.total));
        check(hipMemset(p, 0, sizeof(float) * (size_t)L.total));
        return p;
    };
    // every level is pushed before its buffers are allocated, so a failed
    // allocation leaves nothing the destructor does not free
    FsMgLevel v{};
    v.L = L0;
    v.idx2 = (float)idx2;
    v.idy2 = (float)idy2;
    v.idz2 = (float)idz2;
    lv_.push_back(v);
    lv_[0].r = field(L0);
    diag0_ = field(L0);
    lv_[0].diag = diag0_;
    auto too_big = [](const FsLayout& L) {
        return (long)L.nx * L.ny * L.nz > MG_COARSE_CELLS || (long)(L.nx + 2) * (L.ny + 2) * (L.nz + 2) > MG_CPAD;
    };
    split_ = std::getenv("FS_MG_SPLIT_SMOOTH") != nullptr;
    while (too_big(lv_.back().L)) {
        const FsMgLevel& f = lv_.back();
        FsMgLevel c{};
        c.L = fs_make_layout((f.L.nx + 1) / 2, (f.L.ny + 1) / 2, (f.L.nz + 1) / 2);
        // the spacing doubles along an axis that halves; an axis of one cell stays
        c.idx2 = f.L.nx > 1 ? f.idx2 * 0.25f : f.idx2;
        c.idy2 = f.L.ny > 1 ? f.idy2 * 0.25f : f.idy2;
        c.idz2 = f.L.nz > 1 ? f.idz2 * 0.25f : f.idz2;
        lv_.push_back(c);
        FsMgLevel& n = lv_.back();
        for (float** buf : {&n.bx, &n.by, &n.bz, &n.diag, &n.x, &n.b, &n.r}) *buf = field(n.L);
    }
    built_ = true;
}

void FsMg::SetOperator(hipStream_t st, const float* bx, const float* by, const float* bz) {
    lv_[0].bx = (float*)bx;
    lv_[0].by = (float*)by;
    lv_[0].bz = (float*)bz;
    for (size_t l = 0; l < lv_.size(); l++) {
        FsMgLevel& v = lv_[l];
        if (l > 0) {
            const FsMgLevel& f = lv_[l - 1];
            dim3 g((v.L.nx + 1 + MBX - 1) / MBX, (v.L.ny + 1 + MBY - 1) / MBY, v.L.nz + 1);
            hipLaunchKernelGGL(k_mg_coarsen, g, dim3(MBX, MBY), 0, st, f.L, f.bx, f.by, f.bz, v.L, v.bx, v.by, v.bz);
        }
        hipLaunchKernelGGL(k_mg_diag, cells_grid(v.L), dim3(MBX, MBY), 0, st, v.L, v.bx, v.by, v.bz, v.idx2, v.idy2,
                           v.idz2, v.diag);
    }
}

// nu_ red-black sweeps on level l, colour c1 first in each: fused sweeps
// ping-pong between x and the level's residual buffer (free until the
// residual is formed, and again after the restriction), ending in x; with
// FS_MG_SPLIT_SMOOTH set, two in-place half sweeps per sweep instead (the
// same values).
void FsMg::Smooth(hipStream_t st, int l, int c1) {
    FsMgLevel& v = lv_[l];
    const bool timed = probe && l == 0 && c1 == 0;
    if (split_) {
        const dim3 g = cells_grid(v.L), t(MBX, MBY);
        for (int s = 0; s < 2 * nu_; s++) {
            if (timed && s == 0) probe->begin(st);
            hipLaunchKernelGGL(k_mg_smooth, g, t, 0, st, v.L, v.bx, v.by, v.bz, v.idx2, v.idy2, v.idz2, v.diag, v.b,
                               v.x, (s & 1) ^ c1, skip);
            if (timed && s == 0) probe->end(st);
        }
        if (probe && l == 0) probe->launches += 2 * nu_;
        return;
    }
    const int gx = (v.L.nx + SBX - 1) / SBX, gy = (v.L.ny + SBY - 1) / SBY;
    // planes per workgroup: as tuned, else 16 on large levels, fewer where
    // that leaves under 1024 workgroups, at least 4 (or nz)
    int kc = v.kc;
    if (kc <= 0) {
        kc = (int)std::min<long>(16, std::max<long>(4, (long)v.L.nz * gx * gy / 1024));
        kc = std::min(kc, v.L.nz);
    }
    const dim3 g(gx * gy * ((v.L.nz + kc - 1) / kc)), t(SBX, SBY);
    float *in = v.x, *out = v.r;
    for (int s = 0; s < nu_; s++) {
        if (timed && s == 0) probe->begin(st);
        hipLaunchKernelGGL(k_mg_rb_sweep, g, t, 0, st, v.L, v.bx, v.by, v.bz, v.idx2, v.idy2, v.idz2, v.b, in, out, c1,
                           kc, gx, gy, skip);
        if (timed && s == 0) probe->end(st);
        std::swap(in, out);
    }
    if (in != v.x) check(hipMemcpyAsync(v.x, in, sizeof(float) * (size_t)v.L.total, hipMemcpyDeviceToDevice, st));
    if (probe && l == 0) probe->launches += nu_;
}

void FsMg::Cycle(hipStream_t st, int l) {
    FsMgLevel& v = lv_[l];
    const dim3 g = cells_grid(v.L), t(MBX, MBY);
    if (l + 1 == (int)lv_.size()) {
        hipLaunchKernelGGL(k_mg_coarse, dim3(1), t, 0, st, v.L, v.bx, v.by, v.bz, v.idx2, v.idy2, v.idz2, v.diag,
                           v.b, v.x, MG_COARSE_SWEEPS, skip);
        return;
    }
    Smooth(st, l, 0);  // red, black, ...
    hipLaunchKernelGGL(k_mg_residual, g, t, 0, st, v.L, v.bx, v.by, v.bz, v.idx2, v.idy2, v.idz2, v.b, v.x, v.r, skip);
    FsMgLevel& c = lv_[l + 1];
    hipLaunchKernelGGL(k_mg_restrict, cells_grid(c.L), t, 0, st, v.L, v.r, c.L, c.b, c.x, skip);
    Cycle(st, l + 1);
    hipLaunchKernelGGL(k_mg_prolong, g, t, 0, st, c.L, c.x, v.L, v.x, skip);
    Smooth(st, l, 1);  // ..., black, red
}

// Autotuner of the fused sweep: on each level but the coarsest, time one
// sweep (best of three after a warm-up) for 2, 4, 8 and 16 planes per
// workgroup and keep the fastest.  The sweeps read x and b and write the
// level's residual buffer, which holds nothing yet; each cell's arithmetic
// does not depend on the chunking, so the choice changes no value.
void FsMg::Tune(hipStream_t st, const float* r, float* z) {
    lv_[0].b = (float*)r;
    lv_[0].x = z;
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0));
    check(hipEventCreate(&e1));
    for (size_t l = 0; l + 1 < lv_.size(); l++) {
        FsMgLevel& v = lv_[l];
        const int gx = (v.L.nx + SBX - 1) / SBX, gy = (v.L.ny + SBY - 1) / SBY;
        float best = 0;
        int best_kc = 0;
        for (int kc : {2, 4, 8, 16}) {
            if (kc > v.L.nz && kc != 2) break;
            const int k = std::min(kc, v.L.nz);
            const dim3 g(gx * gy * ((v.L.nz + k - 1) / k)), t(SBX, SBY);
            float ms_min = 0;
            for (int rep = 0; rep < 4; rep++) {
                check(hipEventRecord(e0, st));
                hipLaunchKernelGGL(k_mg_rb_sweep, g, t, 0, st, v.L, v.bx, v.by, v.bz, v.idx2, v.idy2, v.idz2, v.b, v.x,
                                   v.r, 0, k, gx, gy, (const int*)nullptr);
                check(hipEventRecord(e1, st));
                check(hipEventSynchronize(e1));
                float ms = 0;
                check(hipEventElapsedTime(&ms, e0, e1));
                if (rep == 1 || (rep > 1 && ms < ms_min)) ms_min = ms;  // rep 0 warms up
            }
            if (best_kc == 0 || ms_min < best) {
                best = ms_min;
                best_kc = k;
            }
        }
        v.kc = best_kc;
        if (std::getenv("FS_DEBUG")) fprintf(stderr, "[fs] mg level %zu (%d x %d x %d): %d planes per workgroup, %.1f us\n",
                                             l, v.L.nx, v.L.ny, v.L.nz, best_kc, best * 1e3f);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tuned_ = true;
}

void FsMg::Apply(hipStream_t st, const float* r, float* z) {
    FsMgLevel& v = lv_[0];
    v.b = (float*)r;
    v.x = z;
    if (!tuned_ && !split_) Tune(st, r, z);
    fs_mg_zero(st, z, v.L.total, skip);
    Cycle(st, 0);
}

// ---- CG vector kernels -------------------------------------------------------
int fs_mgc_blocks(const FsLayout& L) {
    dim3 g = cg_grid(L);
    return (int)(g.x * g.y * g.z);
}

void fs_mgc_start(hipStream_t st, FsMgCg* S, double rtol2, int max_it) {
    hipLaunchKernelGGL(k_mgc_start, dim3(1), dim3(1), 0, st, S, rtol2, max_it);
}

void fs_mgc_dot2(hipStream_t st, const FsLayout& L, const float* a, const float* b, double* part, const FsMgCg* S) {
    hipLaunchKernelGGL(k_mgc_dot2, cg_grid(L), dim3(MBX, MBY), 0, st, L, a, b, part, fs_mgc_blocks(L), S);
}

void fs_mgc_spmv_dot(hipStream_t st, const FsLayout& L, const float* bx, const float* by, const float* bz, float idx2,
                     float idy2, float idz2, const float* p, float* q, double* part, const FsMgCg* S) {
    hipLaunchKernelGGL(k_mgc_spmv_dot, cg_grid(L), dim3(MBX, MBY), 0, st, L, bx, by, bz, idx2, idy2, idz2, p, q,
                       part, S);
}

void fs_mgc_update(hipStream_t st, const FsLayout& L, float* x, float* r, const float* p, const float* q, double* part,
                   const FsMgCg* S) {
    hipLaunchKernelGGL(k_mgc_update, cg_grid(L), dim3(MBX, MBY), 0, st, L, x, r, p, q, part, S);
}

void fs_mgc_direction(hipStream_t st, const FsLayout& L, const float* z, float* p, const FsMgCg* S) {
    hipLaunchKernelGGL(k_mgc_direction, cg_grid(L), dim3(MBX, MBY), 0, st, L, z, p, S);
}

void fs_mgc_total(hipStream_t st, const double* part, int nb, int slots, double* mid, FsMgCg* S, int kind) {
    hipLaunchKernelGGL(k_mgc_total_mid, dim3(FS_MG_MID, slots), dim3(MNT), 0, st, part, nb, mid, S,
                       kind == FS_MGC_INIT ? 1 : 0);
    hipLaunchKernelGGL(k_mgc_finish, dim3(1), dim3(MNT), 0, st, mid, slots, S, kind);
}

void fs_mg_zero(hipStream_t st, float* v, long n, const int* skip) {
    const long nb = std::min<long>((n + MNT - 1) / MNT, 4096);
    hipLaunchKernelGGL(k_mg_zero, dim3((unsigned)nb), dim3(MNT), 0, st, v, n, skip);
}
