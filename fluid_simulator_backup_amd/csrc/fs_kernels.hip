// fs_kernels.hip -- step kernels for gfx950 (FP64).
//
// One thread per cell/face, 64-wide x rows (one wavefront per x row segment,
// coalesced loads), 4 rows per workgroup.  Stencils read the padded layout
// directly; neighbours along y and z come through L2.
#include <cstdlib>

#include "fs_kernels.h"

#define BX 64
#define BY 4

static inline dim3 grid_for(FsBox b) { return dim3((b.ni + BX - 1) / BX, (b.nj + BY - 1) / BY, b.nk); }

// ---------------------------------------------------------------------------
// ghost fills
// ---------------------------------------------------------------------------

__global__ void k_copy_box(double* __restrict__ dst, const double* __restrict__ src, FsLayout L, FsBox b) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    long c = L.idx(i, j, k);
    dst[c] = src[c];
}

void fs_copy_box(hipStream_t st, double* dst, const double* src, const FsLayout& L, FsBox b) {
    hipLaunchKernelGGL(k_copy_box, grid_for(b), dim3(BX, BY), 0, st, dst, src, L, b);
}

// FillGhostCellsZero face blocks, FIELD_STRUCTURE_3D.h:567-660.
__global__ void k_zero_faces(double* __restrict__ dst, FsLayout L, FsBox b, int d, int axis) {
    int a = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
    if (axis == 0) {
        if (a >= b.nj) return;
        for (int g = 1; g <= d; g++) {
            dst[L.idx(-g, a, c)] = 0;
            dst[L.idx(b.ni - 1 + g, a, c)] = 0;
        }
    } else if (axis == 1) {
        if (a >= b.ni) return;
        for (int g = 1; g <= d; g++) {
            dst[L.idx(a, -g, c)] = 0;
            dst[L.idx(a, b.nj - 1 + g, c)] = 0;
        }
    } else {
        if (a >= b.ni) return;
        for (int g = 1; g <= d; g++) {
            dst[L.idx(a, c, -g)] = 0;
            dst[L.idx(a, c, b.nk - 1 + g)] = 0;
        }
    }
}

void fs_fill_zero_faces(hipStream_t st, double* dst, const FsLayout& L, FsBox b, int depth) {
    hipLaunchKernelGGL(k_zero_faces, dim3((b.nj + 63) / 64, b.nk), dim3(64), 0, st, dst, L, b, depth, 0);
    hipLaunchKernelGGL(k_zero_faces, dim3((b.ni + 63) / 64, b.nk), dim3(64), 0, st, dst, L, b, depth, 1);
    hipLaunchKernelGGL(k_zero_faces, dim3((b.ni + 63) / 64, b.nj), dim3(64), 0, st, dst, L, b, depth, 2);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return fs_clampi(v, lo, hi); }

// FillGhostCellsFrom (fs_fill_const_row): one thread per row of the grown box.
__global__ void k_fill_const(double* dst, const double* src, FsLayout L, FsBox b, int d) {
    int jj = blockIdx.x * blockDim.x + threadIdx.x - d, kk = blockIdx.y - d;
    if (jj >= b.nj + d) return;
    fs_fill_const_row(dst, src, L, b, d, jj, kk);
}

void fs_fill_const(hipStream_t st, double* dst, const double* src, const FsLayout& L, FsBox b, int depth) {
    hipLaunchKernelGGL(k_fill_const, dim3((b.nj + 2 * depth + 63) / 64, b.nk + 2 * depth), dim3(64), 0, st, dst, src,
                       L, b, depth);
}

// FillGhostCellsContinuousDerivativesFrom face blocks (fs_linear_faces_row).
// The three axes in one launch (blockIdx.z = axis): each writes only its own
// face layers and reads only its own row, so they are independent.
__global__ void k_linear_faces(double* dst, const double* src, const double* gsrc, FsLayout L, FsBox b, int d) {
    fs_linear_faces_row(dst, src, gsrc, L, b, d, blockIdx.z, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// The three filled edge blocks normal to z (FIELD_STRUCTURE_3D.h:1896-1946),
// in the reference's loop order; one thread per z-plane.
__global__ void k_linear_edges(double* dst, const double* src, FsLayout L, FsBox b, int d) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.nk) return;
    for (int j = -1; j >= -d; j--)
        for (int i = -1; i >= -d; i--)
            dst[L.idx(i, j, k)] = (double)0.5 * (((double)2 * src[L.idx(i + 1, j, k)] - src[L.idx(i + 2, j, k)]) +
                                                 ((double)2 * src[L.idx(i, j + 1, k)] - src[L.idx(i, j + 2, k)]));
    for (int j = b.nj; j <= b.nj - 1 + d; j++)
        for (int i = b.ni; i <= b.ni - 1 + d; i++)
            dst[L.idx(i, j, k)] = (double)0.5 * (((double)2 * src[L.idx(i - 1, j, k)] - src[L.idx(i - 2, j, k)]) +
                                                 ((double)2 * src[L.idx(i, j - 1, k)] - src[L.idx(i, j - 2, k)]));
    for (int j = b.nj; j <= b.nj - 1 + d; j++)
        for (int i = -1; i >= -d; i--)
            dst[L.idx(i, j, k)] = (double)0.5 * (((double)2 * src[L.idx(i + 1, j, k)] - src[L.idx(i + 2, j, k)]) +
                                                 ((double)2 * src[L.idx(i, j - 1, k)] - src[L.idx(i, j - 2, k)]));
}

void fs_fill_linear(hipStream_t st, double* dst, const double* src, const FsLayout& L, FsBox b, int depth,
                    int edges) {
    fs_fill_linear_split(st, dst, src, src, L, b, depth);
    if (edges) hipLaunchKernelGGL(k_linear_edges, dim3((b.nk + 63) / 64), dim3(64), 0, st, dst, src, L, b, depth);
}

void fs_fill_linear_split(hipStream_t st, double* dst, const double* src, const double* gsrc, const FsLayout& L,
                          FsBox b, int depth) {
    const int na = b.ni > b.nj ? b.ni : b.nj, nc = b.nj > b.nk ? b.nj : b.nk;
    hipLaunchKernelGGL(k_linear_faces, dim3((na + 63) / 64, nc, 3), dim3(64), 0, st, dst, src, gsrc, L, b, depth);
}

// ---------------------------------------------------------------------------
// advection (ADVECTION_METHOD_3D.h)
// ---------------------------------------------------------------------------



__device__ __forceinline__ double upwind(double rho, double dt, double uu, double vv, double ww, double mx, double px,
                                         double my, double py, double mz, double pz) {
    double X = uu > 0 ? mx : px;
    double Y = vv > 0 ? my : py;
    double Z = ww > 0 ? mz : pz;
    return rho - dt * (uu * X + vv * Y + ww * Z);
}

template <bool FAST>
__device__ __forceinline__ void weno5(const double* c, double inv, double eps, double& dmo, double& dpo) {
    fs_weno5_of<FAST>(c, inv, eps, dmo, dpo);
}

template <bool FAST>
__global__ void __launch_bounds__(BX* BY) k_weno5_stage(double* __restrict__ rho, const double* __restrict__ gh,
                                                        const double* __restrict__ gu, const double* __restrict__ gv,
                                                        const double* __restrict__ gw, FsLayout L, FsBox b, int comp,
                                                        double ix, double iy, double iz, double eps, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    long c = L.idx(i, j, k);
    double f[7], mx, px, my, py, mz, pz, uu, vv, ww;
    gather7(gh, c, 1, f);
    weno5<FAST>(f + 3, ix, eps, mx, px);
    gather7(gh, c, L.sx, f);
    weno5<FAST>(f + 3, iy, eps, my, py);
    gather7(gh, c, L.sy, f);
    weno5<FAST>(f + 3, iz, eps, mz, pz);
    vel_at(gu, gv, gw, L, comp, i, j, k, uu, vv, ww);
    rho[c] = upwind(rho[c], dt, uu, vv, ww, mx, px, my, py, mz, pz);
}

static const FsGridConst& gc_dummy(const FsGridConst& g) { return g; }

// ---- WENO5 transport and its adjoint -----------------------------------------
// One level-set stage as k_weno5_stage computes it, on a level set whose
// ghosts are its constant extrapolation (read by clamping the index) and
// MAC velocities with zero ghosts: phi' = phi - dt (u X + v Y + w Z), X the
// upwind WENO5 derivative.  The adjoint differentiates the WENO5 weights
// (smoothness indicators included); the upwind choice is piecewise constant.


// the seven samples of the extended level set along one axis
__device__ __forceinline__ void gather7_clamped(const double* __restrict__ phi, const FsLayout& L, int i, int j, int k,
                                                int axis, double f[7], long cl[7]) {
    for (int o = -3; o <= 3; o++) {
        int ii = i, jj = j, kk = k;
        if (axis == 0) ii = min(max(i + o, 0), L.nx - 1);
        else if (axis == 1) jj = min(max(j + o, 0), L.ny - 1);
        else kk = min(max(k + o, 0), L.nz - 1);
        cl[o + 3] = L.idx(ii, jj, kk);
        f[o + 3] = phi[cl[o + 3]];
    }
}

__global__ void __launch_bounds__(BX* BY) k_weno5_transport(const double* __restrict__ phi, const double* __restrict__ u,
                                                            const double* __restrict__ v, const double* __restrict__ w,
                                                            double* __restrict__ out, FsLayout L, double ix, double iy,
                                                            double iz, double eps, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k), cl[7];
    double f[7], mx, px, my, py, mz, pz, uu, vv, ww;
    gather7_clamped(phi, L, i, j, k, 0, f, cl);
    fs_weno5(f + 3, ix, eps, mx, px);
    gather7_clamped(phi, L, i, j, k, 1, f, cl);
    fs_weno5(f + 3, iy, eps, my, py);
    gather7_clamped(phi, L, i, j, k, 2, f, cl);
    fs_weno5(f + 3, iz, eps, mz, pz);
    vel_at(u, v, w, L, 0, i, j, k, uu, vv, ww);
    out[c] = upwind(phi[c], dt, uu, vv, ww, mx, px, my, py, mz, pz);
}

// lphi += (d phi'/d phi)^T lam, lu/lv/lw += (d phi'/d velocity)^T lam.
__global__ void __launch_bounds__(BX* BY) k_adj_weno5_transport(const double* __restrict__ phi,
                                                                const double* __restrict__ u, const double* __restrict__ v,
                                                                const double* __restrict__ w,
                                                                const double* __restrict__ lam, double* lphi, double* lu,
                                                                double* lv, double* lw, FsLayout L, double ix, double iy,
                                                                double iz, double eps, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    double l = lam[c];
    atomicAdd(&lphi[c], l);  // phi' = phi - ...
    if (l == 0) return;
    double uu, vv, ww;
    vel_at(u, v, w, L, 0, i, j, k, uu, vv, ww);
    const double vel[3] = {uu, vv, ww}, inv[3] = {ix, iy, iz};
    const long st[3] = {1, L.sx, L.sy};
    double* lface[3] = {lu, lv, lw};
    for (int axis = 0; axis < 3; axis++) {
        double f[7];
        long cl[7];
        gather7_clamped(phi, L, i, j, k, axis, f, cl);
        double dm, dp;
        fs_weno5(f + 3, inv[axis], eps, dm, dp);
        bool minus = vel[axis] > 0;
        // the selected one-sided derivative and its five differences
        double d[5], g[5];
        if (minus) {
            for (int q = 0; q < 5; q++) d[q] = inv[axis] * (f[q + 1] - f[q]);  // m1..m5 over f[0..5]
        } else {
            d[0] = inv[axis] * (f[6] - f[5]);  // p1
            for (int q = 1; q < 5; q++) d[q] = inv[axis] * (f[6 - q] - f[5 - q]);  // p2..p5 = m5..m2
        }
        weno5_grad(d, eps, g);
        double sc = -dt * vel[axis] * l * inv[axis];
        if (minus) {
            for (int q = 0; q < 5; q++) {  // m_{q+1} = inv (f[q+1] - f[q])
                atomicAdd(&lphi[cl[q + 1]], sc * g[q]);
                atomicAdd(&lphi[cl[q]], -sc * g[q]);
            }
        } else {
            atomicAdd(&lphi[cl[6]], sc * g[0]);
            atomicAdd(&lphi[cl[5]], -sc * g[0]);
            for (int q = 1; q < 5; q++) {
                atomicAdd(&lphi[cl[6 - q]], sc * g[q]);
                atomicAdd(&lphi[cl[5 - q]], -sc * g[q]);
            }
        }
        // velocity: the cell value is the mean of the two faces along the axis
        double lvel = -dt * l * (minus ? dm : dp);
        atomicAdd(&lface[axis][c], (double)0.5 * lvel);
        atomicAdd(&lface[axis][c + st[axis]], (double)0.5 * lvel);
    }
}

void fs_weno5_transport(hipStream_t st, const double* phi, const double* u, const double* v, const double* w,
                        double* out, const FsLayout& L, const FsGridConst& gc, double eps, double dt) {
    hipLaunchKernelGGL(k_weno5_transport, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, phi, u, v, w, out, L, gc.x.inv,
                       gc.y.inv, gc.z.inv, eps, dt);
}

void fs_adjoint_weno5_transport(hipStream_t st, const double* phi, const double* u, const double* v, const double* w,
                                const double* lam, double* lphi, double* lu, double* lv, double* lw, const FsLayout& L,
                                const FsGridConst& gc, double eps, double dt) {
    hipLaunchKernelGGL(k_adj_weno5_transport, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, phi, u, v, w, lam, lphi,
                       lu, lv, lw, L, gc.x.inv, gc.y.inv, gc.z.inv, eps, dt);
}

void fs_weno5_stage(hipStream_t st, double* rho, const double* ghost, const double* gu, const double* gv,
                    const double* gw, const FsLayout& L, int comp, const FsGridConst& gc, double eps, double dt,
                    bool fast) {
    FsBox b = fs_box(L, comp);
    if (fast)
        hipLaunchKernelGGL(k_weno5_stage<true>, grid_for(b), dim3(BX, BY), 0, st, rho, ghost, gu, gv, gw, L, b, comp,
                           gc.x.inv, gc.y.inv, gc.z.inv, eps, dt);
    else
        hipLaunchKernelGGL(k_weno5_stage<false>, grid_for(b), dim3(BX, BY), 0, st, rho, ghost, gu, gv, gw, L, b, comp,
                           gc.x.inv, gc.y.inv, gc.z.inv, eps, dt);
    (void)gc_dummy;
}

__global__ void __launch_bounds__(BX* BY) k_eno3_stage(double* __restrict__ rho, const double* __restrict__ gh,
                                                       const double* __restrict__ gu, const double* __restrict__ gv,
                                                       const double* __restrict__ gw, FsLayout L, FsBox b, int comp,
                                                       FsGridConst gc, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    long c = L.idx(i, j, k);
    // one_over_3dz is built from one_over_dx (ADVECTION_METHOD_3D.h:1649)
    double inv3x = gc.x.inv * (double)1 / 3, inv3y = gc.y.inv * (double)1 / 3, inv3z = gc.x.inv * (double)1 / 3;
    double f[7], mx, px, my, py, mz, pz, uu, vv, ww;
    gather7(gh, c, 1, f);
    fs_eno3(f + 3, gc.x.d, gc.x.inv, gc.x.inv2d, inv3x, mx, px);
    gather7(gh, c, L.sx, f);
    fs_eno3(f + 3, gc.y.d, gc.y.inv, gc.y.inv2d, inv3y, my, py);
    gather7(gh, c, L.sy, f);
    fs_eno3(f + 3, gc.z.d, gc.z.inv, gc.z.inv2d, inv3z, mz, pz);
    vel_at(gu, gv, gw, L, comp, i, j, k, uu, vv, ww);
    rho[c] = upwind(rho[c], dt, uu, vv, ww, mx, px, my, py, mz, pz);
}

void fs_eno3_stage(hipStream_t st, double* rho, const double* ghost, const double* gu, const double* gv,
                   const double* gw, const FsLayout& L, int comp, const FsGridConst& gc, double dt) {
    FsBox b = fs_box(L, comp);
    hipLaunchKernelGGL(k_eno3_stage, grid_for(b), dim3(BX, BY), 0, st, rho, ghost, gu, gv, gw, L, b, comp, gc, dt);
}

// ---------------------------------------------------------------------------
// forces
// ---------------------------------------------------------------------------

// ApplyGravity, EULERIAN_FLUID_SOLVER_3D.h:2237-2270: every face.
__global__ void k_gravity(double* u, double* v, double* w, FsLayout L, double gx, double gy, double gz) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i > L.nx || j > L.ny) return;
    long c = L.idx(i, j, k);
    if (j < L.ny && k < L.nz) u[c] += gx;
    if (i < L.nx && k < L.nz) v[c] += gy;
    if (i < L.nx && j < L.ny) w[c] += gz;
}

void fs_gravity(hipStream_t st, double* u, double* v, double* w, const FsLayout& L, double gx, double gy, double gz) {
    FsBox b{L.nx + 1, L.ny + 1, L.nz + 1};
    hipLaunchKernelGGL(k_gravity, grid_for(b), dim3(BX, BY), 0, st, u, v, w, L, gx, gy, gz);
}

// VISCOSITY_3D::ApplyViscosity, air/water explicit branch
// (VISCOSITY_3D.h:452-563), one component, out of place: dst = src + update,
// all reads from U, V, W (the component being updated is read from its
// pre-sweep copy).  The smoothed density/viscosity fields of the reference
// (:1258-1347) are evaluated at the face on the fly.
__global__ void __launch_bounds__(BX* BY) k_viscosity(int comp, double* __restrict__ dst, const double* __restrict__ U,
                                                      const double* __restrict__ V, const double* __restrict__ W,
                                                      const double* __restrict__ f, FsLayout L, FsBox b, FsMaterial m,
                                                      double eps, double odx, double ody, double odz, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    const long c = L.idx(i, j, k), sx = L.sx, sy = L.sy;
#define MU(x) m.viscosity(fs_heaviside_visc((x), eps))
    double upd, self;
    if (comp == 1) {
        double lh = (double)0.5 * (f[c - 1] + f[c]);
        double lcu = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 + sx] + f[c + sx]);
        double lcd = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 - sx] + f[c - sx]);
        double lct = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 + sy] + f[c + sy]);
        double lcb = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 - sy] + f[c - sy]);
        double rho_h = m.density(fs_heaviside_visc(lh, eps));
        double mcu = MU(lcu), mcd = MU(lcd), mr = MU(f[c]), mc = MU(f[c - 1]), mct = MU(lct), mcb = MU(lcb);
        double coef = dt * ((double)1 / rho_h);
        double first = (double)2 * coef * odx * (mr * odx * (U[c + 1] - U[c]) - mc * ody * (U[c] - U[c - 1]));
        double second = coef * ody * (mcu * (ody * (U[c + sx] - U[c]) + odx * (V[c + sx] - V[c - 1 + sx])) -
                                      mcd * (ody * (U[c] - U[c - sx]) + odx * (V[c] - V[c - 1])));
        double third = coef * odz * (mct * (odz * (U[c + sy] - U[c]) + odx * (W[c + sy] - W[c - 1 + sy])) -
                                     mcb * (odz * (U[c] - U[c - sy]) + odx * (W[c] - W[c - 1])));
        upd = first + second + third;
        self = U[c];
    } else if (comp == 2) {
        double lh = (double)0.5 * (f[c - sx] + f[c]);
        double lcu = (double)0.25 * (f[c - sx] + f[c + 1 - sx] + f[c] + f[c + 1]);
        double lcl = (double)0.25 * (f[c - 1 - sx] + f[c - sx] + f[c - 1] + f[c]);
        double lct = (double)0.25 * (f[c - sx + sy] + f[c - sx] + f[c + sy] + f[c]);
        double lcb = (double)0.25 * (f[c - sx - sy] + f[c - sx] + f[c - sy] + f[c]);
        double rho_h = m.density(fs_heaviside_visc(lh, eps));
        double mcu = MU(lcu), mcl = MU(lcl), mu_ = MU(f[c]), mc = MU(f[c - sx]), mct = MU(lct), mcb = MU(lcb);
        double coef = dt * ((double)1 / rho_h);
        double first = coef * odx * (mcu * (ody * (U[c + 1] - U[c + 1 - sx]) + odx * (V[c + 1] - V[c])) -
                                     mcl * (ody * (U[c] - U[c - sx]) + odx * (V[c] - V[c - 1])));
        double second = (double)2 * coef * ody * (mu_ * ody * (V[c + sx] - V[c]) - mc * ody * (V[c] - V[c - sx]));
        double third = coef * odz * (mct * (odz * (V[c + sy] - V[c]) + ody * (W[c] - W[c - sx])) -
                                     mcb * (odz * (V[c] - V[c - sy]) + ody * (W[c - sy] - W[c - sx - sy])));
        upd = first + second + third;
        self = V[c];
    } else {
        double lh = (double)0.5 * (f[c - sy] + f[c]);
        double lct = (double)0.25 * (f[c - sy] + f[c + 1 - sy] + f[c] + f[c + 1]);
        double lcb = (double)0.25 * (f[c - sy] + f[c - 1 - sy] + f[c] + f[c - 1]);
        double lcu = (double)0.25 * (f[c - sy] + f[c + sx - sy] + f[c] + f[c + sx]);
        double lcd = (double)0.25 * (f[c - sy] + f[c - sx - sy] + f[c] + f[c - sx]);
        double rho_h = m.density(fs_heaviside_visc(lh, eps));
        double mct = MU(lct), mcb = MU(lcb), mcu = MU(lcu), mcd = MU(lcd), mt = MU(f[c]), mc = MU(f[c - sy]);
        double coef = dt * ((double)1 / rho_h);
        double first = coef * odx * (mct * (odz * (U[c] - U[c - sy]) + odx * (W[c + 1] - W[c])) -
                                     mcb * (odz * (U[c - 1] - U[c - 1 - sy]) + odx * (W[c] - W[c - 1])));
        double second = coef * ody * (mcu * (odz * (V[c] - V[c - sy]) + ody * (W[c + sx] - W[c])) -
                                      mcd * (odz * (V[c - sx] - V[c - sx - sy]) + ody * (W[c] - W[c - sx])));
        double third = (double)2 * coef * odz * (mt * odz * (W[c + sy] - W[c]) - mc * odz * (W[c] - W[c - sy]));
        upd = first + second + third;
        self = W[c];
    }
#undef MU
    dst[c] = self + upd;
}

void fs_viscosity(hipStream_t st, int comp, double* dst, const double* U, const double* V, const double* W,
                  const double* phi, const FsLayout& L, const FsMaterial& m, const FsGridConst& cell, double dt) {
    FsBox b = fs_box(L, comp);
    double eps = 1.5 * cell.x.d;  // epsilon_for_mollification, EULERIAN_FLUID_SOLVER_3D.h:393
    hipLaunchKernelGGL(k_viscosity, grid_for(b), dim3(BX, BY), 0, st, comp, dst, U, V, W, phi, L, b, m, eps,
                       cell.x.inv, cell.y.inv, cell.z.inv, dt);
}

// Transpose of one component's explicit viscous update (k_viscosity): the
// update X_comp += sum of coefficients times U, V, W values is linear for a
// fixed phi, so its adjoint scatters coefficient * lam_comp(face) into the
// adjoints of every value the face read (atomics; ghost targets collect what
// belongs to the constant zero ghosts and are never read back).  src is the
// adjoint of the component after the update; lu, lv, lw accumulate.
__global__ void __launch_bounds__(BX* BY) k_adj_viscosity(int comp, const double* __restrict__ src, double* lu,
                                                          double* lv, double* lw, const double* __restrict__ f,
                                                          FsLayout L, FsBox b, FsMaterial m, double eps, double odx,
                                                          double ody, double odz, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    const long c = L.idx(i, j, k), sx = L.sx, sy = L.sy;
    const double l = src[c];
    if (l == 0) return;
#define MU(x) m.viscosity(fs_heaviside_visc((x), eps))
#define ADD(arr, off, w) atomicAdd(&(arr)[c + (off)], (w) * l)
    if (comp == 1) {
        double lh = (double)0.5 * (f[c - 1] + f[c]);
        double lcu = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 + sx] + f[c + sx]);
        double lcd = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 - sx] + f[c - sx]);
        double lct = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 + sy] + f[c + sy]);
        double lcb = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 - sy] + f[c - sy]);
        double coef = dt * ((double)1 / m.density(fs_heaviside_visc(lh, eps)));
        double mcu = MU(lcu), mcd = MU(lcd), mr = MU(f[c]), mc = MU(f[c - 1]), mct = MU(lct), mcb = MU(lcb);
        double a1 = 2 * coef * odx * mr * odx, a2 = 2 * coef * odx * mc * ody;
        double bu = coef * ody * mcu, bd = coef * ody * mcd, tt = coef * odz * mct, tb = coef * odz * mcb;
        ADD(lu, 1, a1);
        ADD(lu, 0, -a1 - a2 - bu * ody - bd * ody - tt * odz - tb * odz);
        ADD(lu, -1, a2);
        ADD(lu, sx, bu * ody);
        ADD(lu, -sx, bd * ody);
        ADD(lu, sy, tt * odz);
        ADD(lu, -sy, tb * odz);
        ADD(lv, sx, bu * odx);
        ADD(lv, -1 + sx, -bu * odx);
        ADD(lv, 0, -bd * odx);
        ADD(lv, -1, bd * odx);
        ADD(lw, sy, tt * odx);
        ADD(lw, -1 + sy, -tt * odx);
        ADD(lw, 0, -tb * odx);
        ADD(lw, -1, tb * odx);
    } else if (comp == 2) {
        double lh = (double)0.5 * (f[c - sx] + f[c]);
        double lcu = (double)0.25 * (f[c - sx] + f[c + 1 - sx] + f[c] + f[c + 1]);
        double lcl = (double)0.25 * (f[c - 1 - sx] + f[c - sx] + f[c - 1] + f[c]);
        double lct = (double)0.25 * (f[c - sx + sy] + f[c - sx] + f[c + sy] + f[c]);
        double lcb = (double)0.25 * (f[c - sx - sy] + f[c - sx] + f[c - sy] + f[c]);
        double coef = dt * ((double)1 / m.density(fs_heaviside_visc(lh, eps)));
        double mcu = MU(lcu), mcl = MU(lcl), mu_ = MU(f[c]), mc = MU(f[c - sx]), mct = MU(lct), mcb = MU(lcb);
        double au = coef * odx * mcu, al = coef * odx * mcl;
        double s1 = 2 * coef * ody * mu_ * ody, s2 = 2 * coef * ody * mc * ody;
        double tt = coef * odz * mct, tb = coef * odz * mcb;
        ADD(lu, 1, au * ody);
        ADD(lu, 1 - sx, -au * ody);
        ADD(lu, 0, -al * ody);
        ADD(lu, -sx, al * ody);
        ADD(lv, 1, au * odx);
        ADD(lv, 0, -au * odx - al * odx - s1 - s2 - tt * odz - tb * odz);
        ADD(lv, -1, al * odx);
        ADD(lv, sx, s1);
        ADD(lv, -sx, s2);
        ADD(lv, sy, tt * odz);
        ADD(lv, -sy, tb * odz);
        ADD(lw, 0, tt * ody);
        ADD(lw, -sx, -tt * ody);
        ADD(lw, -sy, -tb * ody);
        ADD(lw, -sx - sy, tb * ody);
    } else {
        double lh = (double)0.5 * (f[c - sy] + f[c]);
        double lct = (double)0.25 * (f[c - sy] + f[c + 1 - sy] + f[c] + f[c + 1]);
        double lcb = (double)0.25 * (f[c - sy] + f[c - 1 - sy] + f[c] + f[c - 1]);
        double lcu = (double)0.25 * (f[c - sy] + f[c + sx - sy] + f[c] + f[c + sx]);
        double lcd = (double)0.25 * (f[c - sy] + f[c - sx - sy] + f[c] + f[c - sx]);
        double coef = dt * ((double)1 / m.density(fs_heaviside_visc(lh, eps)));
        double mct = MU(lct), mcb = MU(lcb), mcu = MU(lcu), mcd = MU(lcd), mt = MU(f[c]), mc = MU(f[c - sy]);
        double at = coef * odx * mct, ab = coef * odx * mcb, bu = coef * ody * mcu, bd = coef * ody * mcd;
        double s1 = 2 * coef * odz * mt * odz, s2 = 2 * coef * odz * mc * odz;
        ADD(lu, 0, at * odz);
        ADD(lu, -sy, -at * odz);
        ADD(lu, -1, -ab * odz);
        ADD(lu, -1 - sy, ab * odz);
        ADD(lv, 0, bu * odz);
        ADD(lv, -sy, -bu * odz);
        ADD(lv, -sx, -bd * odz);
        ADD(lv, -sx - sy, bd * odz);
        ADD(lw, 1, at * odx);
        ADD(lw, 0, -at * odx - ab * odx - bu * ody - bd * ody - s1 - s2);
        ADD(lw, -1, ab * odx);
        ADD(lw, sx, bu * ody);
        ADD(lw, -sx, bd * ody);
        ADD(lw, sy, s1);
        ADD(lw, -sy, s2);
    }
#undef ADD
#undef MU
}

void fs_adjoint_viscosity(hipStream_t st, int comp, const double* src, double* lu, double* lv, double* lw,
                          const double* phi, const FsLayout& L, const FsMaterial& m, const FsGridConst& cell,
                          double dt) {
    FsBox b = fs_box(L, comp);
    double eps = 1.5 * cell.x.d;  // epsilon_for_mollification, EULERIAN_FLUID_SOLVER_3D.h:393
    hipLaunchKernelGGL(k_adj_viscosity, grid_for(b), dim3(BX, BY), 0, st, comp, src, lu, lv, lw, phi, L, b, m, eps,
                       cell.x.inv, cell.y.inv, cell.z.inv, dt);
}

// Semi-implicit viscosity system of one MAC component (VISCOSITY_3D.h:853-989
// coefficients and explicit terms; POISSON_SOLVER_3D.h:3359-3496 rows), one
// row per face of the component's box.  Unknown faces are the box; the ring
// beyond it is Dirichlet with the zero ghost velocity
// (SetupBoundaryConditionsForVelocity, regular_boundary, VISCOSITY_3D.h:1403-1417).
//   c[0..5]: the coefficients to the i+1, i-1, j+1, j-1, k+1, k-1 faces
//            (coef_1..coef_6), 2 dt mu/rho_half across the normal pair and
//            dt mu/rho_half across the tangential ones;
//   a[0..5]: matrix entries to i-1, i+1, j-1, j+1, k-1, k+1 (-c/d^2, 0 to a
//            Dirichlet face), diag = 1 + inv_dxdx * (full coefs, then
//            Dirichlet coefs, in that order; 1 if zero) as :3402-3468 sums it;
//   rhs:     the explicit term plus the Dirichlet contributions (:3472-3495).
// U, V, W are the velocities at entry (the reference reads ghost copies made
// before the first solve); self is the live component (its ghosts, the
// Dirichlet values, are zero).  The cross terms of x and z are those of
// VISCOSITY_3D.h:887,975; the y term's z part follows the explicit update
// (:522) where :930 differences w across x.
__global__ void __launch_bounds__(BX* BY) k_visc_system(int comp, double* __restrict__ a0, double* __restrict__ a1,
                                                        double* __restrict__ a2, double* __restrict__ a3,
                                                        double* __restrict__ a4, double* __restrict__ a5,
                                                        double* __restrict__ diag, double* __restrict__ rhs,
                                                        const double* __restrict__ U, const double* __restrict__ V,
                                                        const double* __restrict__ W, const double* __restrict__ self,
                                                        const double* __restrict__ f, FsLayout L, FsBox b,
                                                        FsMaterial m, double eps, double odx, double ody, double odz,
                                                        double idxx, double idyy, double idzz, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    const long c = L.idx(i, j, k), sx = L.sx, sy = L.sy;
#define MU(x) m.viscosity(fs_heaviside_visc((x), eps))
    double cf[6];  // coef_1 .. coef_6
    double ex;
    if (comp == 1) {
        double lh = (double)0.5 * (f[c - 1] + f[c]);
        double lcu = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 + sx] + f[c + sx]);
        double lcd = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 - sx] + f[c - sx]);
        double lct = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 + sy] + f[c + sy]);
        double lcb = (double)0.25 * (f[c - 1] + f[c] + f[c - 1 - sy] + f[c - sy]);
        double rh = m.density(fs_heaviside_visc(lh, eps));
        double mcu = MU(lcu), mcd = MU(lcd), mr = MU(f[c]), mc = MU(f[c - 1]), mct = MU(lct), mcb = MU(lcb);
        cf[0] = (double)2 * dt / rh * mr;
        cf[1] = (double)2 * dt / rh * mc;
        cf[2] = dt / rh * mcu;
        cf[3] = dt / rh * mcd;
        cf[4] = dt / rh * mct;
        cf[5] = dt / rh * mcb;
        ex = U[c] + dt / rh * (ody * (mcu * odx * (V[c + sx] - V[c - 1 + sx]) - mcd * odx * (V[c] - V[c - 1]))) +
             dt / rh * (odz * (mct * odx * (W[c + sy] - W[c - 1 + sy]) - mcb * odx * (W[c] - W[c - 1])));
    } else if (comp == 2) {
        double lh = (double)0.5 * (f[c - sx] + f[c]);
        double lcu = (double)0.25 * (f[c - sx] + f[c + 1 - sx] + f[c] + f[c + 1]);
        double lcl = (double)0.25 * (f[c - 1 - sx] + f[c - sx] + f[c - 1] + f[c]);
        double lct = (double)0.25 * (f[c - sx + sy] + f[c - sx] + f[c + sy] + f[c]);
        double lcb = (double)0.25 * (f[c - sx - sy] + f[c - sx] + f[c - sy] + f[c]);
        double rh = m.density(fs_heaviside_visc(lh, eps));
        double mcu = MU(lcu), mcl = MU(lcl), mu_ = MU(f[c]), mc = MU(f[c - sx]), mct = MU(lct), mcb = MU(lcb);
        cf[0] = dt / rh * mcu;
        cf[1] = dt / rh * mcl;
        cf[2] = (double)2 * dt / rh * mu_;
        cf[3] = (double)2 * dt / rh * mc;
        cf[4] = dt / rh * mct;
        cf[5] = dt / rh * mcb;
        ex = V[c] + dt / rh * (odx * (mcu * ody * (U[c + 1] - U[c + 1 - sx]) - mcl * ody * (U[c] - U[c - sx]))) +
             dt / rh * (odz * (mct * ody * (W[c] - W[c - sx]) - mcb * ody * (W[c - sy] - W[c - sx - sy])));
    } else {
        double lh = (double)0.5 * (f[c - sy] + f[c]);
        double lct = (double)0.25 * (f[c - sy] + f[c + 1 - sy] + f[c] + f[c + 1]);
        double lcb = (double)0.25 * (f[c - sy] + f[c - 1 - sy] + f[c] + f[c - 1]);
        double lcu = (double)0.25 * (f[c - sy] + f[c + sx - sy] + f[c] + f[c + sx]);
        double lcd = (double)0.25 * (f[c - sy] + f[c - sx - sy] + f[c] + f[c - sx]);
        double rh = m.density(fs_heaviside_visc(lh, eps));
        double mct = MU(lct), mcb = MU(lcb), mcu = MU(lcu), mcd = MU(lcd), mt = MU(f[c]), mc = MU(f[c - sy]);
        cf[0] = dt / rh * mct;
        cf[1] = dt / rh * mcb;
        cf[2] = dt / rh * mcu;
        cf[3] = dt / rh * mcd;
        cf[4] = (double)2 * dt / rh * mt;
        cf[5] = (double)2 * dt / rh * mc;
        ex = self[c] + dt / rh * (odx * odz * (mct * (U[c] - U[c - sy]) - mcb * (U[c - 1] - U[c - 1 - sy]))) +
             dt / rh * (ody * odz * (mcu * (V[c] - V[c - sy]) - mcd * (V[c - sx] - V[c - sx - sy])));
    }
#undef MU
    // neighbours in assignment order i-1, i+1, j-1, j+1, k-1, k+1 with their
    // coefficient, inverse square spacing and face offset
    const double nc[6] = {cf[1], cf[0], cf[3], cf[2], cf[5], cf[4]};
    const double ni[6] = {idxx, idxx, idyy, idyy, idzz, idzz};
    const long no[6] = {-1, 1, -sx, sx, -sy, sy};
    const bool full[6] = {i > 0, i < b.ni - 1, j > 0, j < b.nj - 1, k > 0, k < b.nk - 1};
    double* const out[6] = {a0, a1, a2, a3, a4, a5};
    double coef = 0;
    for (int e = 0; e < 6; e++) {
        if (full[e]) coef += nc[e];
        out[e][c] = full[e] ? -nc[e] * ni[e] : (double)0;
    }
    for (int e = 0; e < 6; e++)
        if (!full[e]) coef += nc[e];
    if (coef == 0) coef = 1;
    diag[c] = (double)1 + idxx * coef;
    double bv = ex;
    for (int e = 0; e < 6; e++)
        if (!full[e]) bv += nc[e] * ni[e] * self[c + no[e]];
    rhs[c] = bv;
}

void fs_visc_system(hipStream_t st, int comp, double* const a[6], double* diag, double* rhs, const double* U,
                    const double* V, const double* W, const double* self, const double* phi, const FsLayout& L,
                    const FsMaterial& m, const FsGridConst& cell, double dt) {
    FsBox b = fs_box(L, comp);
    double eps = 1.5 * cell.x.d;  // epsilon_for_mollification, EULERIAN_FLUID_SOLVER_3D.h:393
    double idxx = (double)1 / (cell.x.d * cell.x.d), idyy = (double)1 / (cell.y.d * cell.y.d),
           idzz = (double)1 / (cell.z.d * cell.z.d);
    hipLaunchKernelGGL(k_visc_system, grid_for(b), dim3(BX, BY), 0, st, comp, a[0], a[1], a[2], a[3], a[4], a[5],
                       diag, rhs, U, V, W, self, phi, L, b, m, eps, cell.x.inv, cell.y.inv, cell.z.inv, idxx, idyy,
                       idzz, dt);
}

// ---------------------------------------------------------------------------
// projection (PROJECTION_3D.h, POISSON_SOLVER_3D.h)
// ---------------------------------------------------------------------------

// DetermineProjectionDensity, PROJECTION_3D.h:596-679 (bit-exact sign test).
__global__ void k_density(const double* __restrict__ phi, double* __restrict__ rho, double* __restrict__ irho,
                          FsLayout L, double light, double heavy) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    double f = phi[c], r = rho[c];
    if (f <= 0) r = light;
    else if (f > 0) r = heavy;
    rho[c] = r;
    irho[c] = 1 / r;
}

void fs_density(hipStream_t st, const double* phi, double* rho, double* irho, const FsLayout& L, double light,
                double heavy) {
    hipLaunchKernelGGL(k_density, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, phi, rho, irho, L, light, heavy);
}

// DetermineDivergence, PROJECTION_3D.h:690-725.
__global__ void k_divergence(const double* __restrict__ gu, const double* __restrict__ gv,
                             const double* __restrict__ gw, double* __restrict__ div, FsLayout L, double odx,
                             double ody, double odz, double one_over_dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    double d = odx * (gu[c + 1] - gu[c]) + ody * (gv[c + L.sx] - gv[c]) + odz * (gw[c + L.sy] - gw[c]);
    div[c] = d * one_over_dt;
}

void fs_divergence(hipStream_t st, const double* gu, const double* gv, const double* gw, double* div,
                   const FsLayout& L, const FsGridConst& cell, double one_over_dt) {
    hipLaunchKernelGGL(k_divergence, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, gu, gv, gw, div, L, cell.x.inv,
                       cell.y.inv, cell.z.inv, one_over_dt);
}


// Faces between two interior (full) cells get beta, every other face 0.
__global__ void k_beta(const double* __restrict__ f, const double* __restrict__ v, double* __restrict__ bx,
                       double* __restrict__ by, double* __restrict__ bz, FsLayout L) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i > L.nx || j > L.ny) return;
    long c = L.idx(i, j, k);
    bool in_c = i < L.nx && j < L.ny && k < L.nz;
    if (j < L.ny && k < L.nz) bx[c] = (in_c && i > 0) ? gfm_beta(v[c], v[c - 1], f[c - 1], f[c]) : 0.0;
    if (i < L.nx && k < L.nz) by[c] = (in_c && j > 0) ? gfm_beta(v[c], v[c - L.sx], f[c - L.sx], f[c]) : 0.0;
    if (i < L.nx && j < L.ny) bz[c] = (in_c && k > 0) ? gfm_beta(v[c], v[c - L.sy], f[c - L.sy], f[c]) : 0.0;
}

// A(c,c) = inv_dxdx * coef (:1119-1213) with the Neumann pin on the first
// full cell (:1145-1179: + beta_b + beta_d, + beta_u if its top is Neumann)
// and b = -div - F_X - F_Y - F_Z (:1215-1328): on a face between the two
// phases F = +-(1/d^2) a beta with a = (jc_c |phi_n| + jc_n |phi_c|) /
// (|phi_c| + |phi_n|), + when phi_c <= 0; the jc_on_derivative terms are
// zero.  Faces to ghost cells join cells of one phase (constant ghosts).
__device__ __forceinline__ bool same_phase(double fc, double fn) { return (fc <= 0 && fn <= 0) || (fc > 0 && fn > 0); }
__device__ __forceinline__ double jump_a(double fc, double fn, double jcc, double jcn) {
    return (jcc * fabs(fn) + jcn * fabs(fc)) / (fabs(fc) + fabs(fn));
}
__device__ __forceinline__ double jump_face(double fc, double fn, double jcc, double jcn, double od2, double beta) {
    if (same_phase(fc, fn)) return 0;
    double a = jump_a(fc, fn, jcc, jcn);
    if (fc <= 0) return od2 * a * beta;
    return -od2 * a * beta;
}

__global__ void k_diag_rhs(const double* __restrict__ f, const double* __restrict__ v, const double* __restrict__ bx,
                           const double* __restrict__ by, const double* __restrict__ bz, const double* __restrict__ div,
                           const double* __restrict__ jc, double* __restrict__ diag, double* __restrict__ rhs,
                           double* __restrict__ pin, FsLayout L, double inv_dxdx, double odx2, double ody2,
                           double odz2, int axis) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    const long sx = L.sx, sy = L.sy;
    double coef = 0;
    coef += bx[c];
    coef += bx[c + 1];
    coef += by[c];
    coef += by[c + sx];
    coef += bz[c];
    coef += bz[c + sy];
    if (i == 0 && j == 0 && k == 0) {
        double vc = v[c], fc = f[c];
        double pb = gfm_beta(vc, v[c - sx], f[c - sx], fc);  // beta_b, face towards j-1
        double pd = gfm_beta(vc, v[c - sy], f[c - sy], fc);  // beta_d, face towards k-1
        coef += pb;
        coef += pd;
        pin[0] = pb;
        pin[1] = pd;
    }
    if (axis) {
        // each face over its own spacing (the fixed schedule on non-cubic
        // cells); pin extras on their own axes
        double cx = 0, cy = 0, cz = 0;
        cx += bx[c];
        cx += bx[c + 1];
        cy += by[c];
        cy += by[c + sx];
        cz += bz[c];
        cz += bz[c + sy];
        if (i == 0 && j == 0 && k == 0) {
            cy += pin[0];
            cz += pin[1];
        }
        diag[c] = cx + cy + cz == 0 ? odx2 : odx2 * cx + ody2 * cy + odz2 * cz;
    } else {
        if (coef == 0) coef = 1;
        diag[c] = inv_dxdx * coef;
    }
    if (!jc) {
        rhs[c] = -div[c];
        return;
    }
    const double fc = f[c], jcc = jc[c];
    double FL = jump_face(fc, f[c - 1], jcc, jc[c - 1], odx2, bx[c]);
    double FR = jump_face(fc, f[c + 1], jcc, jc[c + 1], odx2, bx[c + 1]);
    double FB = jump_face(fc, f[c - sx], jcc, jc[c - sx], ody2, by[c]);
    double FT = jump_face(fc, f[c + sx], jcc, jc[c + sx], ody2, by[c + sx]);
    double FD = jump_face(fc, f[c - sy], jcc, jc[c - sy], odz2, bz[c]);
    double FU = jump_face(fc, f[c + sy], jcc, jc[c + sy], odz2, bz[c + sy]);
    double FX = FL + FR, FY = FB + FT, FZ = FD + FU;
    rhs[c] = -div[c] - FX - FY - FZ;
}

void fs_assemble(hipStream_t st, const double* phi, const double* irho, const double* div, const double* jc,
                 double* bx, double* by, double* bz, double* diag, double* rhs, double* pin, const FsLayout& L,
                 const FsGridConst& cell, int axis_diag) {
    FsBox bb{L.nx + 1, L.ny + 1, L.nz + 1};
    hipLaunchKernelGGL(k_beta, grid_for(bb), dim3(BX, BY), 0, st, phi, irho, bx, by, bz, L);
    double inv_dxdx = (double)1 / (cell.x.d * cell.x.d);
    hipLaunchKernelGGL(k_diag_rhs, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, phi, irho, bx, by, bz, div, jc, diag,
                       rhs, pin, L, inv_dxdx, cell.x.inv_d2, cell.y.inv_d2, cell.z.inv_d2, axis_diag);
}

// One spherical object of EULERIAN_FLUID_SOLVER_3D::SourcingFromObjectList
// (EULERIAN_FLUID_SOLVER_3D.h:2592-2671) over the cells [lo, hi] of its box:
// d = |CellCenter(i,j,k) - center| - radius with CellCenter's half-cell
// offset (GRID_STRUCTURE_3D.h:269-272); where d <= 3 dx, phi = min(d, phi)
// (mode 1) or max(-d, phi) (mode 2), as MIN/MAX write them (MACROS.h:17,24).
__global__ void k_source_sphere(double* __restrict__ phi, FsLayout L, int i0, int j0, int k0, int ni, int nj,
                                double cx, double cy, double cz, double radius, double x0, double y0, double z0,
                                double dx, double dy, double dz, double lim, int mode) {
    int a = blockIdx.x * BX + threadIdx.x, b = blockIdx.y * BY + threadIdx.y;
    if (a >= ni || b >= nj) return;
    int i = i0 + a, j = j0 + b, k = k0 + blockIdx.z;
    double px = x0 + ((double)0.5 + (double)i) * dx - cx;
    double py = y0 + ((double)0.5 + (double)j) * dy - cy;
    double pz = z0 + ((double)0.5 + (double)k) * dz - cz;
    double d = sqrt(px * px + py * py + pz * pz) - radius;
    if (d <= lim) {
        long c = L.idx(i, j, k);
        double f = phi[c];
        if (mode == 1) phi[c] = d > f ? f : d;
        else phi[c] = -d > f ? -d : f;
    }
}

void fs_source_sphere(hipStream_t st, double* phi, const FsLayout& L, const int lo[3], const int hi[3],
                      const double center[3], double radius, const FsGridConst& cell, const double origin[3],
                      int mode) {
    int ni = hi[0] - lo[0] + 1, nj = hi[1] - lo[1] + 1, nk = hi[2] - lo[2] + 1;
    if (ni <= 0 || nj <= 0 || nk <= 0) return;
    dim3 g((ni + BX - 1) / BX, (nj + BY - 1) / BY, nk);
    hipLaunchKernelGGL(k_source_sphere, g, dim3(BX, BY), 0, st, phi, L, lo[0], lo[1], lo[2], ni, nj, center[0],
                       center[1], center[2], radius, origin[0], origin[1], origin[2], cell.x.d, cell.y.d, cell.z.d,
                       cell.x.d * (double)3, mode);
}

// Neumann mean removal, POISSON_SOLVER_3D.h:225-246.
// One workgroup per plane: each thread adds whole rows in x order (its row's
// cache lines stream through L1), the row sums go to LDS, and one thread adds
// them in row order.  The order (rows, then the plane's rows, then planes) is
// the oracle's; the reference's own order depends on its thread count
// (per-thread slabs, END_GRID_ITERATION_SUM).
#define PS_ROWS 4096
__global__ void __launch_bounds__(256) k_plane_sums(const double* __restrict__ rhs, double* __restrict__ part,
                                                    FsLayout L) {
    __shared__ double rows[PS_ROWS];
    const int k = blockIdx.x;
    double s = 0;
    for (int j0 = 0; j0 < L.ny; j0 += PS_ROWS) {
        const int nr = min(PS_ROWS, L.ny - j0);
        __syncthreads();
        for (int jj = threadIdx.x; jj < nr; jj += blockDim.x) {
            const double* row = rhs + L.idx(0, j0 + jj, k);
            double r = 0;
            for (int i = 0; i < L.nx; i++) r += row[i];
            rows[jj] = r;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int jj = 0; jj < nr; jj++) s += rows[jj];
    }
    if (threadIdx.x == 0) part[k] = s;
}
__global__ void k_total(double* part, int nz) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0;
    for (int k = 0; k < nz; k++) s += part[k];
    part[nz] = s;
}
__global__ void k_sub_mean(double* __restrict__ rhs, const double* __restrict__ part, FsLayout L, double inv_n) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    rhs[c] -= inv_n * part[L.nz];
}

void fs_remove_mean(hipStream_t st, double* rhs, double* scratch, const FsLayout& L) {
    hipLaunchKernelGGL(k_plane_sums, dim3(L.nz), dim3(256), 0, st, rhs, scratch, L);
    hipLaunchKernelGGL(k_total, dim3(1), dim3(64), 0, st, scratch, L.nz);
    double inv_n = 1 / (double)((long)L.nx * L.ny * L.nz);
    hipLaunchKernelGGL(k_sub_mean, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, rhs, scratch, L, inv_n);
}

// max that keeps a NaN (so a non-finite field is reported, not skipped)
__device__ __forceinline__ double nanmax(double a, double b) { return (a != a || a > b) ? a : b; }

// Workgroup max of non-negative a (BX x BY threads), one atomic per
// workgroup; non-negative doubles (and NaN above +inf) order like their bits.
__device__ __forceinline__ void wg_max_atomic(double a, unsigned long long* bits) {
    __shared__ double lds[BY];
    for (int o = 32; o > 0; o >>= 1) a = nanmax(a, __shfl_xor(a, o));
    if (threadIdx.x == 0) lds[threadIdx.y] = a;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double m = lds[0];
        for (int w = 1; w < BY; w++) m = nanmax(m, lds[w]);
        atomicMax(bits, (unsigned long long)__double_as_longlong(m));
    }
}

#define KZ 16  // z-planes per workgroup in the max-reducing kernels

// LEVELSET_3D::ComputeCurvatures, curvature_by_levelset branch
// (LEVELSET_3D.h:963-1082), on phi with its ghosts as the caller left them:
// central first, second and cross derivatives, curvature
// -(px^2 pyy - 2 px py pxy + py^2 pxx + px^2 pzz - 2 px pz pxz + pz^2 pxx +
// py^2 pzz - c py pz pyz + pz^2 pyy) / |grad phi|^3 with c = 1 as the
// reference writes it (:1033) or 2 (`exact`), one-sided differences when the
// central gradient vanishes (:1037, whose z term multiplies by dz unless
// `exact`; a still vanishing gradient gives 0 and counts in kstat[1], where
// the reference stops the program, :1040-1044), clamped below at
// -1/min(dx,dy,dz) (:1061-1068).  kappa[c] = the curvature; jc[c] = sigma * kappa
// (PROJECTION_3D.h:767) when jc is given; kstat[0] = bits of max |kappa| over
// cells with |phi| < 0.1 (:1070-1077; NaN skipped as the reference's MAX
// skips it).
__global__ void __launch_bounds__(BX* BY) k_curvature(const double* __restrict__ phi, double* __restrict__ kappa,
                                                      double* __restrict__ jc, double sigma, FsLayout L, FsGridConst g,
                                                      double tol, int exact, unsigned long long* kstat) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double amax = 0;
    for (int k = blockIdx.z * KZ; k < min(blockIdx.z * KZ + KZ, L.nz); k++) {
        if (i >= L.nx || j >= L.ny) continue;
        const long c = L.idx(i, j, k), sx = L.sx, sy = L.sy;
        const double ox = g.x.inv2d, oy = g.y.inv2d, oz = g.z.inv2d;
        double px = (phi[c + 1] - phi[c - 1]) * ox, py = (phi[c + sx] - phi[c - sx]) * oy,
               pz = (phi[c + sy] - phi[c - sy]) * oz;
        double pxx = (phi[c + 1] - 2 * phi[c] + phi[c - 1]) * g.x.inv_d2;
        double pyy = (phi[c + sx] - 2 * phi[c] + phi[c - sx]) * g.y.inv_d2;
        double pzz = (phi[c + sy] - 2 * phi[c] + phi[c - sy]) * g.z.inv_d2;
        double pxy = (phi[c + 1 + sx] - phi[c + 1 - sx] - phi[c - 1 + sx] + phi[c - 1 - sx]) * ox * oy;
        double pxz = (phi[c + 1 + sy] - phi[c + 1 - sy] - phi[c - 1 + sy] + phi[c - 1 - sy]) * ox * oz;
        double pyz = (phi[c + sx + sy] - phi[c + sx - sy] - phi[c - sx + sy] + phi[c - sx - sy]) * oy * oz;
        double mag = sqrt(px * px + py * py + pz * pz);
        double deno = mag * mag * mag;
        double curv = 0;
        bool ok = true;
        if (deno == 0) {
            px = (phi[c + 1] - phi[c]) * g.x.inv;
            py = (phi[c + sx] - phi[c]) * g.y.inv;
            pz = (phi[c + sy] - phi[c]) * (exact ? g.z.inv : g.z.d);
            mag = sqrt(px * px + py * py + pz * pz);
            deno = mag * mag * mag;
            ok = deno != 0;
            if (!ok) atomicAdd(&kstat[1], 1ull);
        }
        if (ok) {
            double one_over_deno = (double)1 / deno;
            double cyz = exact ? (double)2 * py * pz * pyz : py * pz * pyz;
            curv = -(px * px * pyy - (double)2 * px * py * pxy + py * py * pxx + px * px * pzz -
                     (double)2 * px * pz * pxz + pz * pz * pxx + py * py * pzz - cyz + pz * pz * pyy) *
                   one_over_deno;
        }
        double kv = curv < -tol ? -tol : curv;
        kappa[c] = kv;
        if (jc) jc[c] = sigma * kv;
        if (fabs(phi[c]) < 1e-1 && fabs(kv) > amax) amax = fabs(kv);
    }
    wg_max_atomic(amax, kstat);
}

// LEVELSET_3D::ComputeNormals (LEVELSET_3D.h:933-955): unit normals of phi,
// its ghosts read as their constant fill (:931): central differences,
// forward ones where the central gradient vanishes, each vector times the
// reciprocal of its magnitude unless that is zero (VECTOR_3D.h:153-165).
__global__ void __launch_bounds__(BX* BY) k_normals(const double* __restrict__ phi, double* __restrict__ n0,
                                                    double* __restrict__ n1, double* __restrict__ n2, FsLayout L,
                                                    FsGridConst g) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    auto P = [&](int a, int b, int c) {
        return phi[L.idx(fs_clampi(a, 0, L.nx - 1), fs_clampi(b, 0, L.ny - 1), fs_clampi(c, 0, L.nz - 1))];
    };
    double x = (P(i + 1, j, k) - P(i - 1, j, k)) * g.x.inv2d;
    double y = (P(i, j + 1, k) - P(i, j - 1, k)) * g.y.inv2d;
    double z = (P(i, j, k + 1) - P(i, j, k - 1)) * g.z.inv2d;
    double mag = sqrt(x * x + y * y + z * z);
    if (!(mag != 0)) {
        x = (P(i + 1, j, k) - P(i, j, k)) * g.x.inv;
        y = (P(i, j + 1, k) - P(i, j, k)) * g.y.inv;
        z = (P(i, j, k + 1) - P(i, j, k)) * g.z.inv;
    }
    double m = sqrt(x * x + y * y + z * z);
    if (m != 0) {
        double r = (double)1 / m;
        x *= r;
        y *= r;
        z *= r;
    }
    const long c = L.idx(i, j, k);
    n0[c] = x;
    n1[c] = y;
    n2[c] = z;
}

void fs_normals(hipStream_t st, const double* phi, double* const* nrm, const FsLayout& L, const FsGridConst& cell) {
    hipLaunchKernelGGL(k_normals, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, phi, nrm[0], nrm[1], nrm[2], L, cell);
}

// Curvature by normals (LEVELSET_3D.h:974-996): the central divergence of
// the normal field (its one ghost layer the constant fill, :976), |curv|
// above tol taken as -tol; kstat[0] = bits of max |curv| (unclamped) over
// every cell.  jc[c] = sigma * kappa when jc is given.
__global__ void __launch_bounds__(BX* BY) k_curvature_normals(const double* __restrict__ n0,
                                                              const double* __restrict__ n1,
                                                              const double* __restrict__ n2, double* __restrict__ kappa,
                                                              double* __restrict__ jc, double sigma, FsLayout L,
                                                              FsGridConst g, double tol, unsigned long long* kstat) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double amax = 0;
    for (int k = blockIdx.z * KZ; k < min(blockIdx.z * KZ + KZ, L.nz); k++) {
        if (i >= L.nx || j >= L.ny) continue;
        auto N = [&](const double* f, int a, int b, int c) {
            return f[L.idx(fs_clampi(a, 0, L.nx - 1), fs_clampi(b, 0, L.ny - 1), fs_clampi(c, 0, L.nz - 1))];
        };
        double curv = (N(n0, i + 1, j, k) - N(n0, i - 1, j, k)) * g.x.inv2d;
        curv += (N(n1, i, j + 1, k) - N(n1, i, j - 1, k)) * g.y.inv2d;
        curv += (N(n2, i, j, k + 1) - N(n2, i, j, k - 1)) * g.z.inv2d;
        const double kv = fabs(curv) <= tol ? curv : -tol;
        const long c = L.idx(i, j, k);
        kappa[c] = kv;
        if (jc) jc[c] = sigma * kv;
        if (fabs(curv) > amax) amax = fabs(curv);
    }
    wg_max_atomic(amax, kstat);
}

void fs_curvature_normals(hipStream_t st, double* const* nrm, double* kappa, double* jc, double sigma,
                          const FsLayout& L, const FsGridConst& cell, unsigned long long* kstat) {
    FsBox b = fs_box(L, 0);
    double tol = (double)1 / std::min(std::min(cell.x.d, cell.y.d), cell.z.d);
    dim3 g((b.ni + BX - 1) / BX, (b.nj + BY - 1) / BY, (b.nk + KZ - 1) / KZ);
    hipLaunchKernelGGL(k_curvature_normals, g, dim3(BX, BY), 0, st, nrm[0], nrm[1], nrm[2], kappa, jc, sigma, L, cell,
                       tol, kstat);
}

void fs_curvature(hipStream_t st, const double* phi, double* kappa, double* jc, double sigma, const FsLayout& L,
                  const FsGridConst& cell, int exact, unsigned long long* kstat) {
    FsBox b = fs_box(L, 0);
    double tol = (double)1 / std::min(std::min(cell.x.d, cell.y.d), cell.z.d);
    dim3 g((b.ni + BX - 1) / BX, (b.nj + BY - 1) / BY, (b.nk + KZ - 1) / KZ);
    hipLaunchKernelGGL(k_curvature, g, dim3(BX, BY), 0, st, phi, kappa, jc, sigma, L, cell, tol, exact, kstat);
}


// UpdateVelocityByPressureGradientVariableDensity, PROJECTION_3D.h:1099-1449,
// one component (comp 1,2,3).  Faces whose + cell is not full are skipped
// (:1165-1168); the bc product test (:1191) lets a face next to a Neumann
// ghost through only for the cell numbered 0.  The face level set comes from
// the projection's ghost copy (linear extrapolation at depth 1, :1115).
__global__ void __launch_bounds__(BX* BY) k_gradient(int comp, int apply, double* __restrict__ vel,
                                                     const double* __restrict__ p, const double* __restrict__ phi,
                                                     const double* __restrict__ rho, const double* __restrict__ jc,
                                                     FsLayout L, FsBox b, FsMaterial m, double eps, double od, double dt,
                                                     unsigned long long* vmax) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double a = 0;
    for (int k = blockIdx.z * KZ; k < min(blockIdx.z * KZ + KZ, b.nk); k++) {
        bool valid = i < b.ni && j < b.nj && i < L.nx && j < L.ny && k < L.nz;
        if (!valid) continue;
        long c = L.idx(i, j, k);
        long st = comp == 1 ? 1 : comp == 2 ? L.sx : L.sy;
        int nidx = comp == 1 ? i - 1 : comp == 2 ? j - 1 : k - 1;
        bool neighbour_full = nidx >= 0;
        double lc = phi[c], ln = phi[c - st];
        // projection ghost copy: interior values, or 2*phi(0) - phi(1) below 0
        double gn = neighbour_full ? ln : (double)2 * phi[c] - phi[c + st];
        double lface = (double)0.5 * (lc + gn);
        double H = fs_heaviside_proj(lface, eps);
        double density_half = m.density(H);
        double one_over_density_ijk = (double)1 / rho[c];
        // jc's ghosts are zero (PROJECTION_3D.h:358 fills them once)
        double jcc = jc ? jc[c] : (double)0, jcn = jc ? jc[c - st] : (double)0;
        double jump_gamma = (jcc * fabs(ln) + jcn * fabs(lc)) / (fabs(ln) + fabs(lc));
        double one_over_density_half = 1 / density_half;
        bool first_cell = (i == 0 && j == 0 && k == 0);
        double pc = p[c], pn = p[c - st];
        double vr = vel[c];
        if (apply && (neighbour_full || first_cell)) {
            if (lc * ln > 0) {
                vr -= dt * (pc - pn) * one_over_density_ijk * od;
            } else {
                if (lc <= 0) vr -= dt * (pc - pn + jump_gamma) * one_over_density_half * od;
                else if (lc > 0) vr -= dt * (pc - pn - jump_gamma) * one_over_density_half * od;
            }
            vel[c] = vr;
        }
        a = nanmax(a, fabs(vr));
    }
    wg_max_atomic(a, vmax);
}

void fs_gradient_update(hipStream_t st, int comp, int apply, double* vel, const double* p, const double* phi,
                        const double* rho, const double* jc, const FsLayout& L, const FsMaterial& m,
                        const FsGridConst& cell, double dt, unsigned long long* vmax_bits) {
    FsBox b = fs_box(L, comp);
    double dg = comp == 1 ? cell.x.d : comp == 2 ? cell.y.d : cell.z.d;
    double od = comp == 1 ? cell.x.inv : comp == 2 ? cell.y.inv : cell.z.inv;
    dim3 g((b.ni + BX - 1) / BX, (b.nj + BY - 1) / BY, (b.nk + KZ - 1) / KZ);
    hipLaunchKernelGGL(k_gradient, g, dim3(BX, BY), 0, st, comp, apply, vel, p, phi, rho, jc, L, b, m,
                       (double)1.5 * dg, od, dt, vmax_bits);
}

// Gradient update of the fixed schedule: every face takes the pressure
// matrix's own GFM coefficient, u_f -= dt beta_f (p_c - p_n + s_c a_f) / dx,
// with the right-hand side's jump term on faces between the phases (s_c = +1
// when phi_c <= 0, else -1; a_f as in k_diag_rhs), so the corrected velocity
// is discretely divergence-free (beta is 0 on wall faces, which are left
// alone).  jc = NULL: no jump.
__global__ void __launch_bounds__(BX* BY) k_gradient_beta(double* __restrict__ vel, const double* __restrict__ p,
                                                          const double* __restrict__ beta,
                                                          const double* __restrict__ phi,
                                                          const double* __restrict__ jc, FsLayout L, FsBox b,
                                                          long st, double od, double dt, unsigned long long* vmax) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double a = 0;
    for (int k = blockIdx.z * KZ; k < min(blockIdx.z * KZ + KZ, b.nk); k++) {
        if (i >= b.ni || j >= b.nj) continue;
        long c = L.idx(i, j, k);
        double bf = beta[c], vr = vel[c];
        if (bf != 0) {
            double dp = p[c] - p[c - st];
            if (jc) {
                double fc = phi[c], fn = phi[c - st];
                if (!same_phase(fc, fn)) {
                    double ja = jump_a(fc, fn, jc[c], jc[c - st]);
                    dp = fc <= 0 ? dp + ja : dp - ja;
                }
            }
            vr -= dt * bf * dp * od;
            vel[c] = vr;
        }
        a = nanmax(a, fabs(vr));
    }
    wg_max_atomic(a, vmax);
}

void fs_gradient_update_beta(hipStream_t st, int comp, double* vel, const double* p, const double* beta,
                             const double* phi, const double* jc, const FsLayout& L, const FsGridConst& cell,
                             double dt, unsigned long long* vmax_bits) {
    FsBox b = fs_box(L, comp);
    double od = comp == 1 ? cell.x.inv : comp == 2 ? cell.y.inv : cell.z.inv;
    long s = comp == 1 ? 1 : comp == 2 ? L.sx : L.sy;
    dim3 g((b.ni + BX - 1) / BX, (b.nj + BY - 1) / BY, (b.nk + KZ - 1) / KZ);
    hipLaunchKernelGGL(k_gradient_beta, g, dim3(BX, BY), 0, st, vel, p, beta, phi, jc, L, b, s, od, dt, vmax_bits);
}

// Normal wall faces of component comp set to zero (the fixed schedule's
// no-penetration walls after forces).
__global__ void k_zero_walls(double* __restrict__ vel, FsLayout L, FsBox b, int comp) {
    int a = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
    if (comp == 1) {
        if (a >= b.nj) return;
        vel[L.idx(0, a, c)] = 0;
        vel[L.idx(b.ni - 1, a, c)] = 0;
    } else if (comp == 2) {
        if (a >= b.ni) return;
        vel[L.idx(a, 0, c)] = 0;
        vel[L.idx(a, b.nj - 1, c)] = 0;
    } else {
        if (a >= b.ni) return;
        vel[L.idx(a, c, 0)] = 0;
        vel[L.idx(a, c, b.nk - 1)] = 0;
    }
}

void fs_zero_walls(hipStream_t st, int comp, double* vel, const FsLayout& L) {
    FsBox b = fs_box(L, comp);
    int n = comp == 1 ? b.nj : b.ni, m = comp == 3 ? b.nj : b.nk;
    hipLaunchKernelGGL(k_zero_walls, dim3((n + 63) / 64, m), dim3(64), 0, st, vel, L, b, comp);
}

// ---- adjoint of the projection (fixed interface) ---------------------------
// The projection at fixed phi is linear in the velocity:
//   u' = u + Gh A^-1 M B u,  B u = -div(u)/dt,  M = mean removal,
// A the GFM pressure matrix (symmetric), Gh = du/dp of the gradient update
// above (its minus sign included).  Its adjoint is
// lambda + B^T M A^-1 Gh^T lambda: these kernels are Gh^T (faces to cells)
// and D^T (cells to faces); the solve uses the same operator.

// Coefficient of face f (the face below cell f along comp's axis, neighbour
// cell f - st) in k_gradient's update u_f -= dt (p_f - p_n) coef; returns
// false for a face k_gradient leaves alone.  phi's ghost next to the first
// cell holds its own value (constant extrapolation of ComputeNormals).
__device__ __forceinline__ bool grad_face_coef(const double* __restrict__ phi, const double* __restrict__ rho,
                                               const FsLayout& L, int comp, int i, int j, int k, long st, FsMaterial m,
                                               double eps, double od, double& coef) {
    int nidx = comp == 1 ? i - 1 : comp == 2 ? j - 1 : k - 1;
    bool neighbour_full = nidx >= 0;
    bool first_cell = (i == 0 && j == 0 && k == 0);
    if (!neighbour_full && !first_cell) return false;
    long c = L.idx(i, j, k);
    double lc = phi[c], ln = neighbour_full ? phi[c - st] : lc;
    if (lc * ln > 0) {
        coef = (double)1 / rho[c] * od;
    } else {
        double gn = neighbour_full ? ln : (double)2 * phi[c] - phi[c + st];
        double H = fs_heaviside_proj((double)0.5 * (lc + gn), eps);
        coef = (double)1 / m.density(H) * od;
    }
    return true;
}

// g_c += sum over comp's faces f touching cell c of (d u_f / d p_c) lam_f
// (Gh^T lam).
__global__ void __launch_bounds__(BX* BY) k_adj_gradient(int comp, const double* __restrict__ lam,
                                                         const double* __restrict__ phi, const double* __restrict__ rho,
                                                         double* __restrict__ g, FsLayout L, FsMaterial m, double eps,
                                                         double od, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    long st = comp == 1 ? 1 : comp == 2 ? L.sx : L.sy;
    int idx = comp == 1 ? i : comp == 2 ? j : k;
    int n = comp == 1 ? L.nx : comp == 2 ? L.ny : L.nz;
    double acc = 0, coef;
    // the face below c: u_f -= dt (p_c - p_n) coef
    if (grad_face_coef(phi, rho, L, comp, i, j, k, st, m, eps, od, coef)) acc -= dt * coef * lam[c];
    // the face above c, i.e. below the next cell (whose neighbour c is full)
    if (idx + 1 < n) {
        int i2 = i + (comp == 1), j2 = j + (comp == 2), k2 = k + (comp == 3);
        if (grad_face_coef(phi, rho, L, comp, i2, j2, k2, st, m, eps, od, coef)) acc += dt * coef * lam[c + st];
    }
    g[c] += acc;
}

void fs_adjoint_gradient(hipStream_t st, int comp, const double* lam, const double* phi, const double* rho, double* g,
                         const FsLayout& L, const FsMaterial& m, const FsGridConst& cell, double dt) {
    double dg = comp == 1 ? cell.x.d : comp == 2 ? cell.y.d : cell.z.d;
    double od = comp == 1 ? cell.x.inv : comp == 2 ? cell.y.inv : cell.z.inv;
    hipLaunchKernelGGL(k_adj_gradient, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, comp, lam, phi, rho, g, L, m,
                       (double)1.5 * dg, od, dt);
}

// g_c += (Gh^T lam)_c for k_gradient_beta's update u_f -= dt beta_f (p_c -
// p_n) od: the face below c contributes -dt od beta lam, the face above +.
__global__ void __launch_bounds__(BX* BY) k_adj_gradient_beta(const double* __restrict__ lam,
                                                              const double* __restrict__ beta, double* __restrict__ g,
                                                              FsLayout L, long st, double od, double dt) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    g[c] += dt * od * (beta[c + st] * lam[c + st] - beta[c] * lam[c]);
}

void fs_adjoint_gradient_beta(hipStream_t st, int comp, const double* lam, const double* beta, double* g,
                              const FsLayout& L, const FsGridConst& cell, double dt) {
    double od = comp == 1 ? cell.x.inv : comp == 2 ? cell.y.inv : cell.z.inv;
    long s = comp == 1 ? 1 : comp == 2 ? L.sx : L.sy;
    hipLaunchKernelGGL(k_adj_gradient_beta, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, lam, beta, g, L, s, od, dt);
}

// lam_face += scale * D^T z: the divergence reads u(i+1) - u(i) per cell, so
// face i of each component gets od (z(i-1) - z(i)), z = 0 outside the cells.
__global__ void __launch_bounds__(BX* BY) k_adj_divergence(int comp, double* __restrict__ lam,
                                                           const double* __restrict__ z, FsLayout L, FsBox b, double od,
                                                           double scale) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    long c = L.idx(i, j, k);
    long st = comp == 1 ? 1 : comp == 2 ? L.sx : L.sy;
    int idx = comp == 1 ? i : comp == 2 ? j : k;
    int n = comp == 1 ? L.nx : comp == 2 ? L.ny : L.nz;
    double below = idx > 0 ? z[c - st] : 0.0, above = idx < n ? z[c] : 0.0;
    lam[c] += scale * (od * (below - above));
}

void fs_adjoint_divergence(hipStream_t st, double* lu, double* lv, double* lw, const double* z, const FsLayout& L,
                           const FsGridConst& cell, double scale) {
    double* lam[3] = {lu, lv, lw};
    double od[3] = {cell.x.inv, cell.y.inv, cell.z.inv};
    for (int comp = 1; comp <= 3; comp++) {
        FsBox b = fs_box(L, comp);
        hipLaunchKernelGGL(k_adj_divergence, grid_for(b), dim3(BX, BY), 0, st, comp, lam[comp - 1], z, L, b,
                           od[comp - 1], scale);
    }
}

__global__ void __launch_bounds__(BX* BY) k_abs_max(const double* __restrict__ a, FsLayout L, FsBox b,
                                                    unsigned long long* bits) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double v = 0;
    if (i < b.ni && j < b.nj)
        for (int k = blockIdx.z * KZ; k < min(blockIdx.z * KZ + KZ, b.nk); k++) v = nanmax(v, fabs(a[L.idx(i, j, k)]));
    wg_max_atomic(v, bits);
}

void fs_abs_max(hipStream_t st, const double* a, const FsLayout& L, FsBox b, unsigned long long* bits) {
    dim3 g((b.ni + BX - 1) / BX, (b.nj + BY - 1) / BY, (b.nk + KZ - 1) / KZ);
    hipLaunchKernelGGL(k_abs_max, g, dim3(BX, BY), 0, st, a, L, b, bits);
}

// ---------------------------------------------------------------------------
// reinitialization
// ---------------------------------------------------------------------------

// Sign function S = phi/sqrt(phi^2 + dx^2), EULERIAN_FLUID_SOLVER_3D.h:2211-2215.
__global__ void k_sign(const double* __restrict__ phi, double* __restrict__ sgn, FsLayout L, double dx2) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    fs_sign_cell(phi, sgn, L.idx(i, j, k), dx2);
}

void fs_sign(hipStream_t st, const double* phi, double* sgn, const FsLayout& L, double dx2) {
    hipLaunchKernelGGL(k_sign, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, phi, sgn, L, dx2);
}

// WENO5thReinitialization, ADVECTION_METHOD_3D.h:745-1637 (fs_reinit_cell).
template <bool FAST>
__global__ void __launch_bounds__(BX* BY) k_reinit(double* __restrict__ out, const double* __restrict__ gh,
                                                   const double* __restrict__ sgn, FsLayout L, double ix, double iy,
                                                   double iz, double eps, double dt, int exact) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    fs_reinit_cell<FAST>(out, gh, sgn, L, ix, iy, iz, eps, dt, exact, L.idx(i, j, k));
}

// The fast-WENO5 sweep marching z: each thread owns an (i, j) column and KC
// consecutive planes.  Along z the seven values of the stencil, their six
// first differences and the four triples' smoothness terms (fs_w5_tri) slide
// with k, so each cell loads one new value and forms one new triple along z
// instead of seven loads and four triples; x and y take the per-cell form.
// The same operations on the same values as fs_reinit_cell<true>, so the
// same bits.
#define FS_REINIT_KC 8
__global__ void __launch_bounds__(BX* BY) k_reinit_zm(double* __restrict__ out, const double* __restrict__ gh,
                                                      const double* __restrict__ sgn, FsLayout L, double ix,
                                                      double iy, double iz, double eps, double dt, int exact) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k0 = blockIdx.z * FS_REINIT_KC;
    if (i >= L.nx || j >= L.ny) return;
    const int k1 = min(k0 + FS_REINIT_KC, L.nz);
    const long sy = L.sy;
    long c = L.idx(i, j, k0);
    // z window at cell k: phi(k+2) (next difference's left end), d_{k-3}..d_{k+1}, T_{k-3}..T_{k-1}
    double d[6];
    FsW5Tri t[4];
    double prev = gh[c - 3 * sy];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const double nx_ = gh[c + (q - 2) * sy];
        d[q] = iz * (nx_ - prev);
        prev = nx_;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) t[q] = fs_w5_tri(d[q], d[q + 1], d[q + 2], eps);
    for (int k = k0; k < k1; k++, c += sy) {
        const double nx_ = gh[c + 3 * sy];
        d[5] = iz * (nx_ - prev);
        prev = nx_;
        t[3] = fs_w5_tri(d[3], d[4], d[5], eps);
        double f[7], mx, px, my, py, mz, pz;
        fs_w5_combine(d, t, mz, pz);
        gather7(gh, c, 1, f);
        fs_weno5_fast(f + 3, ix, eps, mx, px);
        gather7(gh, c, L.sx, f);
        fs_weno5_fast(f + 3, iy, eps, my, py);
        fs_reinit_finish(out, gh, sgn, mx, px, my, py, mz, pz, eps, dt, exact, c);
#pragma unroll
        for (int q = 0; q < 5; q++) d[q] = d[q + 1];
#pragma unroll
        for (int q = 0; q < 3; q++) t[q] = t[q + 1];
    }
}

void fs_reinit_sweep(hipStream_t st, double* out, const double* ghost, const double* sgn, const FsLayout& L,
                     const FsGridConst& cell, double eps, double dt, int exact, bool fast) {
    const FsBox b = fs_box(L, 0);
    if (fast && !std::getenv("FS_REINIT_PER_CELL"))
        hipLaunchKernelGGL(k_reinit_zm, dim3((b.ni + BX - 1) / BX, (b.nj + BY - 1) / BY, (b.nk + FS_REINIT_KC - 1) / FS_REINIT_KC),
                           dim3(BX, BY), 0, st, out, ghost, sgn, L, cell.x.inv, cell.y.inv, cell.z.inv, eps, dt, exact);
    else if (fast)
        hipLaunchKernelGGL(k_reinit<true>, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, out, ghost, sgn, L,
                           cell.x.inv, cell.y.inv, cell.z.inv, eps, dt, exact);
    else
        hipLaunchKernelGGL(k_reinit<false>, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, out, ghost, sgn, L,
                           cell.x.inv, cell.y.inv, cell.z.inv, eps, dt, exact);
}

// ---------------------------------------------------------------------------
// misc
// ---------------------------------------------------------------------------

__global__ void k_average(double* __restrict__ phi, const double* __restrict__ tmp, FsLayout L) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    phi[c] = (double)0.5 * (tmp[c] + phi[c]);
}

void fs_average(hipStream_t st, double* phi, const double* tmp, const FsLayout& L) {
    hipLaunchKernelGGL(k_average, grid_for(fs_box(L, 0)), dim3(BX, BY), 0, st, phi, tmp, L);
}

// dst = ca*a + cb*dst over a box: the Runge-Kutta stage blends of
// EULERIAN_FLUID_SOLVER_3D.h:939-951 (0.5/0.5), :1041-1053 (0.75/0.25),
// :1061-1073 ((1/3)/(2/3)), :1099 and :1107, each product rounded before
// the sum as written there.
__global__ void k_combine(double* __restrict__ dst, const double* __restrict__ a, double ca, double cb, FsLayout L,
                          FsBox b) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    long c = L.idx(i, j, k);
    double x = ca * a[c];
    double y = cb * dst[c];
    dst[c] = x + y;
}

void fs_combine(hipStream_t st, double* dst, const double* a, double ca, double cb, const FsLayout& L, FsBox b) {
    hipLaunchKernelGGL(k_combine, grid_for(b), dim3(BX, BY), 0, st, dst, a, ca, cb, L, b);
}

// lin is x-fastest over the box grown by g layers on every side (g = 0: the
// box itself); scatter writes the box's interior from it, gather reads the
// grown box (ghost layers included) into it
__global__ void k_scatter(double* __restrict__ dst, const double* __restrict__ lin, FsLayout L, FsBox b, int g) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    if (i >= b.ni || j >= b.nj) return;
    const long gi = b.ni + 2 * g, gj = b.nj + 2 * g;
    dst[L.idx(i, j, k)] = lin[(i + g) + gi * ((j + g) + gj * (k + g))];
}
__global__ void k_gather(double* __restrict__ lin, const double* __restrict__ src, FsLayout L, FsBox b, int g) {
    int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, k = blockIdx.z;
    const int gi = b.ni + 2 * g, gj = b.nj + 2 * g;
    if (i >= gi || j >= gj) return;
    lin[i + (long)gi * (j + (long)gj * k)] = src[L.idx(i - g, j - g, k - g)];
}

void fs_scatter(hipStream_t st, double* dst, const double* lin, const FsLayout& L, FsBox b, int g) {
    hipLaunchKernelGGL(k_scatter, grid_for(b), dim3(BX, BY), 0, st, dst, lin, L, b, g);
}
void fs_gather(hipStream_t st, double* lin, const double* src, const FsLayout& L, FsBox b, int g) {
    FsBox gb{b.ni + 2 * g, b.nj + 2 * g, b.nk + 2 * g};
    hipLaunchKernelGGL(k_gather, grid_for(gb), dim3(BX, BY), 0, st, lin, src, L, b, g);
}

__global__ void k_fill_value(double* dst, long n, double v) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n; t += (long)gridDim.x * blockDim.x) dst[t] = v;
}
void fs_fill_value(hipStream_t st, double* dst, long n, double v) {
    hipLaunchKernelGGL(k_fill_value, dim3(1024), dim3(256), 0, st, dst, n, v);
}
