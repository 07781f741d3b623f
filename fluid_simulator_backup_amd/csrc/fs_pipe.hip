// fs_pipe.hip -- the oil/water pipe mode on gfx950: core-annular flow in a
// cylinder along x (is_parallel) or y (is_vertical) -- EULERIAN_FLUID_SOLVER_3D
// in oil_water_simulation mode, cylindrical_boundary, dimensionless_form,
// semi_implicit_approach: the configuration of the reference's shipped
// scripts (3D_script.sim:32-99, 195-290 parallel; 3D_script_1/2/3.sim
// vertical).  The vertical pipe is the parallel one's axis mirror except
// where the reference's vertical branches differ, which each kernel notes.
//
// The reference's modules keep persistent ghosted arrays: the solver's
// scalar_field_ghost and vector_field_mac_ghost_x/y/z are written by
// ADVECTION_3D, VISCOSITY_3D and PROJECTION_3D in turn, and each module has
// its own per-axis copies velocity_field_mac_ghost_<c>_<axis>.  The cells
// outside the pipe and the ghost layers keep what the last writer left, and
// later stencils read them.  Every such array lives here in HBM with the same
// refresh points, in the padded layout of fs_layout.h (ghost width 3), zeroed
// at creation as ARRAY_3D::Initialize(..., true) zeroes them.  The arrays the
// reference computes only on their interior (boundary_phi_<c>, the pipe's
// level set r - a at the faces) are static and computed once;
// their ghost cells read 0, as the reference's do.
//
// Every kernel keeps the reference's operation order (-ffp-contract=off).
// The linear systems (pressure and the three semi-implicit viscosity
// systems) have a static pattern: the pipe wall fixes which cells are
// unknowns, so their bc fields and row numbering (k, j, i order over the FULL
// cells, POISSON_SOLVER_3D.h:527-545) are built once on the host.  Each solve
// assembles its values row by row on the device into an ELL matrix whose
// entries keep the reference's CSR order, and runs the Jacobi-PCG of
// PCG_METHOD.cpp:4-160 with device-resident scalars; dot products use one
// fixed order (blocks of 256 rows summed by a binary tree, block sums in
// order), which the CPU checker uses too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fs_kernels.h"
#include "fs_layout.h"
#include "fs_pipe.h"

#define PBX 64
#define PBY 4
#define RBLK 256  // rows per reduction block (the fixed order of the pipe's dot products)

#define BC_FULL 0    // MACROS.h:6
#define BC_DIR (-1)  // :7
#define BC_NEUM (-4) // :10
#define BC_PER (-5)  // :11
#define BC_OTHER (-6)

namespace {

#define PCHECK(x)                                                                                   \
    do {                                                                                            \
        hipError_t _e = (x);                                                                        \
        if (_e != hipSuccess) throw FsPipeError{FS_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(_e)}; \
    } while (0)

inline dim3 box_grid(FsBox b, int lo = 0) {
    return dim3((b.ni + 2 * lo + PBX - 1) / PBX, (b.nj + PBY - 1) / PBY, b.nk);
}

// the centreline faces ADVECTION_3D::Solve_Velocity holds fixed
// (ADVECTION_3D.h:248-307: along y for the vertical pipe, along x for the
// parallel one); set from creation: every phase that reads them runs after
// that call in a step
__host__ __device__ inline bool fixed_face(int c, int i, int j, int k, int nx, int ny, int nz, int vert) {
    if (vert) {
        if (c == 1) return (i == nx / 2 || i == nx / 2 + 1) && k == (nz - 1) / 2;
        if (c == 2) return i == (nx - 1) / 2 && k == (nz - 1) / 2;
        if (c == 3) return i == (nx - 1) / 2 && (k == nz / 2 || k == nz / 2 + 1);
        return false;
    }
    if (c == 1) return j == (ny - 1) / 2 && k == (nz - 1) / 2;
    if (c == 2) return (j == ny / 2 || j == ny / 2 + 1) && k == (nz - 1) / 2;
    if (c == 3) return j == (ny - 1) / 2 && (k == nz / 2 || k == nz / 2 + 1);
    return false;
}

__device__ __forceinline__ double theta_of(double nearv, double far) { return fabs(nearv) / (fabs(nearv) + fabs(far)); }

// ---- x-periodic ghost copies, FIELD_STRUCTURE_3D.h:1088-1327 ---------------
// One thread per cell of the row segment i in [-gw, ni-1+gw].  The first
// cell (cell fields) or the two end faces (x faces) and the ghosts beyond
// take the node-periodic image; with a mask (boundary_phi of the
// component) only where the mask at the written cell is < 1e-8, and the
// interior copy only where it is <= 0.  The reads and the writes never meet
// for nx >= 8, so dst == src works in one pass.
__global__ void k_periodic_x(double* dst, const double* src, const double* __restrict__ mask, FsLayout L, FsBox b,
                             int c, int gw, int copy) {
    int i = blockIdx.x * PBX + threadIdx.x - gw, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i > b.ni - 1 + gw || j >= b.nj) return;
    const int ie = b.ni - 1;
    const long o = L.idx(i, j, k);
    const bool left = i <= 0, right = c == 1 ? i >= ie : i >= ie + 1;
    if (left || right) {
        if (mask && !(mask[o] < (double)1e-8)) return;
        int si = left ? (c == 1 ? ie + (i - 1) : ie + i) : (c == 1 ? 0 + (i - (ie - 1)) : 0 + (i - ie));
        dst[o] = src[L.idx(si, j, k)];
    } else if (copy) {
        if (mask && !(mask[o] <= 0)) return;
        dst[o] = src[o];
    }
}

// ---- y-periodic ghost copies (the vertical pipe), FIELD_STRUCTURE_3D.h:1328-1710
// One thread per cell of the column segment j in [-gw, nj-1+gw] of every
// interior (i, k).  The interior copy (masked: src where the mask is < 1e-8,
// zero elsewhere), then the y ghosts -- for y faces the two end faces and
// beyond, for other fields j < 0 and j > je only -- from the node-periodic
// image, with a mask only where it is < 1e-8 at the written cell.  Reads
// and writes never meet for ny >= 8, so dst == src works in one pass.
__global__ void k_periodic_y(double* dst, const double* src, const double* __restrict__ mask, FsLayout L, FsBox b,
                             int c, int gw, int copy) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y - gw, k = blockIdx.z;
    if (i >= b.ni || j > b.nj - 1 + gw) return;
    const int je = b.nj - 1;
    const long o = L.idx(i, j, k);
    const bool left = c == 2 ? j <= 0 : j < 0, right = c == 2 ? j >= je : j > je;
    if (left || right) {
        if (!mask || mask[o] < (double)1e-8) {
            int sj = left ? (c == 2 ? je + (j - (0 + 1)) : je + (j - 0)) : (c == 2 ? 0 + (j - (je - 1)) : 0 + (j - je));
            dst[o] = src[L.idx(i, sj, k)];
            return;
        }
    }
    if (copy && j >= 0 && j <= je) {
        if (!mask || mask[o] < (double)1e-8)
            dst[o] = src[o];
        else
            dst[o] = 0;
    }
}

// The masked fill's four extrapolation blocks (FIELD_STRUCTURE_3D.h:1568-1670),
// run on one thread in the order x-right, x-left, z-top, z-bottom: from each
// cell outside the wall whose neighbour below (above) is inside, every cell
// to the box's edge, ghosts included, takes 2 src(t -/+ 1) - src(t -/+ 2).
// The value does not depend on which crossing wrote it, so a cell takes the
// last block that reaches it: xr[j, k] (xl[j, k]) is the first (last)
// right (left) crossing of the x row, zt[i, j] (zb[i, j]) the first (last)
// top (bottom) crossing of the z column, or +-2^30 where there is none.
__global__ void k_periodic_y_ext(double* __restrict__ dst, const double* __restrict__ src,
                                 const int* __restrict__ xr, const int* __restrict__ xl, const int* __restrict__ zt,
                                 const int* __restrict__ zb, FsLayout L, FsBox b, int gw) {
    int i = blockIdx.x * PBX + threadIdx.x - gw, j = blockIdx.y * PBY + threadIdx.y, k = (int)blockIdx.z - gw;
    if (i > b.ni - 1 + gw || j >= b.nj) return;
    const long o = L.idx(i, j, k), sy = L.sy;
    const bool icol = i >= 0 && i < b.ni, krow = k >= 0 && k < b.nk;
    if (icol && k <= zb[j * b.ni + i])
        dst[o] = (double)2 * src[o + sy] - src[o + 2 * sy];
    else if (icol && zt[j * b.ni + i] <= k)
        dst[o] = (double)2 * src[o - sy] - src[o - 2 * sy];
    else if (krow && i <= xl[k * b.nj + j])
        dst[o] = (double)2 * src[o + 1] - src[o + 2];
    else if (krow && xr[k * b.nj + j] <= i)
        dst[o] = (double)2 * src[o - 1] - src[o - 2];
}

// ---- wall conditions of the velocity ghost copies --------------------------

// ADVECTION_3D::SetupBoundaryConditionForVelocity, level-set branch,
// is_parallel, y faces (ADVECTION_3D.h:510-531): outside the pipe v = 0 and
// the ghost copy is zeroed (all four y/z neighbours outside) or extrapolated;
// the second test repeats `phi(j-1) < 0` and reads the j+1 face, whose value
// is the one before the loop (gin: a copy taken before it).
__global__ void k_lsbc_y(double* __restrict__ v, double* __restrict__ g, const double* __restrict__ gin,
                         const double* __restrict__ b, FsLayout L, FsBox bx) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= bx.ni || j >= bx.nj) return;
    const long o = L.idx(i, j, k), sx = L.sx, sy = L.sy;
    const double b0 = b[o];
    if (!(b0 > 0)) return;
    v[o] = 0;
    double val = 0;
    bool wr = false;
    if (b[o - sx] > 0 && b[o + sx] > 0 && b[o - sy] > 0 && b[o + sy] > 0) val = 0, wr = true;
    if (b[o - sx] < 0) {
        double th = theta_of(b[o - sx], b0);
        val = (th - 1) / th * gin[o - sx];
        wr = true;
    }
    if (b[o - sx] < 0) {
        double th = theta_of(b[o + sx], b0);
        val = (th - 1) / th * gin[o + sx];
        wr = true;
    }
    if (wr) g[o] = val;
}

// the vertical pipe's x faces (ADVECTION_3D.h:445-466): outside the pipe
// u = 0 and the copy zeroed (all four x/z neighbours outside) or
// extrapolated from the inside neighbour left, then right (inside faces are
// never written here, so the loop reads no value it wrote)
__global__ void k_lsbc_x(double* __restrict__ u, double* __restrict__ g, const double* __restrict__ b, FsLayout L,
                         FsBox bx) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= bx.ni || j >= bx.nj) return;
    const long o = L.idx(i, j, k), sy = L.sy;
    const double b0 = b[o];
    if (!(b0 > 0)) return;
    u[o] = 0;
    double val = 0;
    bool wr = false;
    if (b[o - 1] > 0 && b[o + 1] > 0 && b[o - sy] > 0 && b[o + sy] > 0) val = 0, wr = true;
    if (b[o - 1] < 0) {
        double th = theta_of(b[o - 1], b0);
        val = (th - 1) / th * g[o - 1];
        wr = true;
    }
    if (b[o + 1] < 0) {
        double th = theta_of(b[o + 1], b0);
        val = (th - 1) / th * g[o + 1];
        wr = true;
    }
    if (wr) g[o] = val;
}

// the same for z faces (ADVECTION_3D.h:533-550, and :468-489 for the vertical pipe, s1 = 1): w = 0 outside, the copy
// zeroed or extrapolated from the inside neighbour below or above (inside
// faces are never written here, so the loop reads no value it wrote)
__global__ void k_lsbc_z(double* __restrict__ w, double* __restrict__ g, const double* __restrict__ b, FsLayout L,
                         FsBox bx, long s1) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= bx.ni || j >= bx.nj) return;
    const long o = L.idx(i, j, k), sy = L.sy;
    const double b0 = b[o];
    if (!(b0 > 0)) return;
    w[o] = 0;
    if (b[o - s1] > 0 && b[o + s1] > 0 && b[o - sy] > 0 && b[o + sy] > 0) g[o] = 0;
    if (b[o - sy] < 0) {
        double th = theta_of(b[o - sy], b0);
        g[o] = (th - 1) / th * g[o - sy];
    }
    if (b[o + sy] < 0) {
        double th = theta_of(b[o + sy], b0);
        g[o] = (th - 1) / th * g[o + sy];
    }
}

// ADVECTION_3D::SetupBoundaryConditionForVelocity, velocity branch,
// boundary_order_velocity 1 (ADVECTION_3D.h:553-851): at every crossing
// inside -> outside along the wall's first axis (y for the parallel pipe,
// x for the vertical one: stride sx) or z, the first outside value of the
// axis' copy is (theta - 1)/theta times the inside value and the next three
// are pow(10 t, t) = 400, 27000, 2560000.  Each written cell belongs to one
// crossing (a line across the convex pipe meets it in one interval).
__global__ void k_velbc_cross(double* __restrict__ ay, double* __restrict__ az, const double* __restrict__ g,
                              const double* __restrict__ b, FsLayout L, FsBox bx, long sx) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= bx.ni || j >= bx.nj) return;
    const long o = L.idx(i, j, k), sy = L.sy;
    const double b0 = b[o];
    if (!(b0 < 0)) return;
    const double sent[3] = {400.0, 27000.0, 2560000.0};
    const double gv = g[o];
    if (b[o + sx] > 0) {
        double th = theta_of(b0, b[o + sx]);
        ay[o + sx] = (th - 1) / th * gv;
        for (int t = 2; t <= 4; t++) ay[o + t * sx] = sent[t - 2];
    }
    if (b[o - sx] > 0) {
        double th = theta_of(b0, b[o - sx]);
        ay[o - sx] = (th - 1) / th * gv;
        for (int t = 2; t <= 4; t++) ay[o - t * sx] = sent[t - 2];
    }
    if (b[o + sy] > 0) {
        double th = theta_of(b0, b[o + sy]);
        az[o + sy] = (th - 1) / th * gv;
        for (int t = 2; t <= 4; t++) az[o + t * sy] = sent[t - 2];
    }
    if (b[o - sy] > 0) {
        double th = theta_of(b0, b[o - sy]);
        az[o - sy] = (th - 1) / th * gv;
        for (int t = 2; t <= 4; t++) az[o - t * sy] = sent[t - 2];
    }
}

// VISCOSITY_3D::SetupBoundaryConditionsForVelocity(thread_id),
// boundary_order_velocity 1 (VISCOSITY_3D.h:2255-2352): each face outside
// the pipe takes, in the reference's else-if order, zero in both copies
// (all four neighbours outside), the extrapolation from the first inside
// neighbour j-1, j+1 (y copy), k-1, k+1 (z copy), or zero in both.
// With the vertical pipe (:2120-2232) the wall's first axis is x (sx = 1,
// ay the x copy), and the y faces extrapolate from v itself (g = v), which
// equals its shared copy on the inside faces read.
__global__ void k_viscbc(double* __restrict__ ay, double* __restrict__ az, const double* __restrict__ g,
                         const double* __restrict__ b, FsLayout L, FsBox bx, long sx) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= bx.ni || j >= bx.nj) return;
    const long o = L.idx(i, j, k), sy = L.sy;
    const double b0 = b[o];
    if (!(b0 > 0)) return;
    if (b[o - sx] > 0 && b[o + sx] > 0 && b[o - sy] > 0 && b[o + sy] > 0) {
        ay[o] = 0;
        az[o] = 0;
    } else if (b[o - sx] < -0) {
        double th = theta_of(b[o - sx], b0);
        ay[o] = (th - 1) / th * g[o - sx];
    } else if (b[o + sx] < -0) {
        double th = theta_of(b[o + sx], b0);
        ay[o] = (th - 1) / th * g[o + sx];
    } else if (b[o - sy] < -0) {
        double th = theta_of(b[o - sy], b0);
        az[o] = (th - 1) / th * g[o - sy];
    } else if (b[o + sy] < -0) {
        double th = theta_of(b[o + sy], b0);
        az[o] = (th - 1) / th * g[o + sy];
    } else {
        ay[o] = 0;
        az[o] = 0;
    }
}

// PROJECTION_3D::SetupBoundaryConditionsForVelocityDivergence
// (PROJECTION_3D.h:1566-1685), one face component: zero (all four neighbours
// across the pipe outside: y and z for the parallel pipe, s1 = sx; x and z
// for the vertical one, s1 = 1), the extrapolation from the inside
// neighbour along the face's own axis (stride st), or zero.  Inside faces
// are not written, so reading them in place is safe.
__global__ void k_divbc(double* __restrict__ g, const double* __restrict__ b, FsLayout L, FsBox bx, long st, long sx) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= bx.ni || j >= bx.nj) return;
    const long o = L.idx(i, j, k), sy = L.sy;
    const double b0 = b[o];
    if (!(b0 > 0)) return;
    if (b[o - sx] > 0 && b[o + sx] > 0 && b[o - sy] > 0 && b[o + sy] > 0)
        g[o] = 0;
    else if (b[o - st] < 0) {
        double th = theta_of(b[o - st], b0);
        g[o] = (th - (double)1) / th * g[o - st];
    } else if (b[o + st] < 0) {
        double th = theta_of(b[o + st], b0);
        g[o] = (th - (double)1) / th * g[o + st];
    } else
        g[o] = 0;
}

// ---- advection ----------------------------------------------------------------

__device__ __forceinline__ double upwind_p(double rho, double dt, double uu, double vv, double ww, double mx,
                                           double px, double my, double py, double mz, double pz) {
    double X = uu > 0 ? mx : px, Y = vv > 0 ? my : py, Z = ww > 0 ? mz : pz;
    return rho - dt * (uu * X + vv * Y + ww * Z);
}

// ADVECTION_METHOD_3D::WENO5th with the pipe's level set
// (ADVECTION_METHOD_3D.h:281-523): cells outside the pipe skipped
__global__ void __launch_bounds__(PBX* PBY) k_weno5_pipe(double* __restrict__ phi, const double* __restrict__ gh,
                                                         const double* __restrict__ gu, const double* __restrict__ gv,
                                                         const double* __restrict__ gw,
                                                         const double* __restrict__ bls, FsLayout L, double ix,
                                                         double iy, double iz, double eps, double dt) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const long c = L.idx(i, j, k);
    if (bls[c] > 0) return;
    double f[7], mx, px, my, py, mz, pz, uu, vv, ww;
    gather7(gh, c, 1, f);
    fs_weno5(f + 3, ix, eps, mx, px);
    gather7(gh, c, L.sx, f);
    fs_weno5(f + 3, iy, eps, my, py);
    gather7(gh, c, L.sy, f);
    fs_weno5(f + 3, iz, eps, mz, pz);
    vel_at(gu, gv, gw, L, 0, i, j, k, uu, vv, ww);
    phi[c] = upwind_p(phi[c], dt, uu, vv, ww, mx, px, my, py, mz, pz);
}

// ADVECTION_METHOD_3D::ENO3rd with the pipe (ADVECTION_METHOD_3D.h:
// 1936-2243): fixed faces and faces with boundary_phi >= 0 keep their
// value; x derivatives from sxa, y from sya, z from sza
__global__ void __launch_bounds__(PBX* PBY) k_eno3_pipe(double* __restrict__ rho, const double* __restrict__ sxa,
                                                        const double* __restrict__ sya, const double* __restrict__ sza,
                                                        const double* __restrict__ gu, const double* __restrict__ gv,
                                                        const double* __restrict__ gw, const double* __restrict__ b,
                                                        FsLayout L, FsBox bx, int comp, FsGridConst gc, double dt,
                                                        int vert) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= bx.ni || j >= bx.nj) return;
    if (fixed_face(comp, i, j, k, L.nx, L.ny, L.nz, vert)) return;
    const long c = L.idx(i, j, k);
    if (!(b[c] < (double)0)) return;
    const double inv3x = gc.x.inv * (double)1 / 3, inv3y = gc.y.inv * (double)1 / 3, inv3z = gc.z.inv * (double)1 / 3;
    double f[7], mx, px, my, py, mz, pz, uu, vv, ww;
    gather7(sxa, c, 1, f);
    fs_eno3(f + 3, gc.x.d, gc.x.inv, gc.x.inv2d, inv3x, mx, px);
    gather7(sya, c, L.sx, f);
    fs_eno3(f + 3, gc.y.d, gc.y.inv, gc.y.inv2d, inv3y, my, py);
    gather7(sza, c, L.sy, f);
    fs_eno3(f + 3, gc.z.d, gc.z.inv, gc.z.inv2d, inv3z, mz, pz);
    vel_at(gu, gv, gw, L, comp, i, j, k, uu, vv, ww);
    rho[c] = upwind_p(rho[c], dt, uu, vv, ww, mx, px, my, py, mz, pz);
}

// ---- forces -----------------------------------------------------------------

// sin(x) for |x| <= pi (a hair beyond after rounding): the smoothed
// Heaviside's one transcendental.  Host and device libm differ in the last
// bit on about 3% of its arguments, so the pipe mode takes this polynomial
// on both sides instead: reduction to [0, pi/2] by sin(x) = sin(pi - x) with
// a two-part pi (the subtraction from the leading part exact by Sterbenz'
// lemma), then the Taylor series of sin to x^23 (truncation below 1e-20 on
// [0, pi/2]) by Horner's rule, coefficients (-1)^k/(2k+1)! rounded to
// double.  Within 2 ulps of sin; no FMA, so every build gives the same bits.
__device__ __forceinline__ double pipe_sin(double x) {
    const double c[11] = {-0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13, 0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33, -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49, -0x1.2f49b46814157p-57, 0x1.71b8ef6dcf572p-66, -0x1.761b41316381ap-75};
    const double pi_hi = 0x1.921fb54442d18p+1, pi_lo = 0x1.1a62633145c07p-53;
    const bool neg = x < 0;
    double a = neg ? -x : x;
    if (a > pi_hi * 0.5) a = (pi_hi - a) + pi_lo;
    double a2 = a * a, t = c[10];
#pragma unroll
    for (int k = 9; k >= 0; k--) t = t * a2 + c[k];
    double r = a + a * (a2 * t);
    return neg ? -r : r;
}

// EULERIAN_FLUID_SOLVER_3D::HeavisideFunction (:2459-2479), also
// PROJECTION_3D's (:1452-1472)
__device__ __forceinline__ double heav_solver(double phi, double eps) {
    double one_over_epsilon = (double)1 / eps, one_over_pi = (double)1 / FS_PI;
    if (phi < -eps) return 0;
    if (phi > eps) return 1;
    return (double)0.5 + phi * one_over_epsilon * (double)0.5 +
           (double)0.5 * one_over_pi * pipe_sin((double)FS_PI * phi * one_over_epsilon);
}
// VISCOSITY_3D::HeavisideFunction (VISCOSITY_3D.h:1234-1256)
__device__ __forceinline__ double heav_visc(double phi, double eps) {
    double one_over_epsilon = (double)1 / eps, one_over_pi = (double)1 / FS_PI;
    if (phi < -eps) return (double)0;
    if (phi > eps) return (double)1;
    return (double)0.5 * ((double)1 + phi * one_over_epsilon + one_over_pi * pipe_sin(FS_PI * phi * one_over_epsilon));
}

// EULERIAN_FLUID_SOLVER_3D::ApplyGravity, dimensionless; the parallel pipe
// (:2362-2393): inside the pipe, fixed faces excepted, u += R1/V0^2 (dt g_x
// + 1/(rho_half rho_oil) dt f) with rho_half = 1 + (eta - 1) H of the face
// level set (eps 1.5 dx, :2522-2561); v, w += R1/V0^2 dt g
// The vertical pipe (:2327-2360) drives v alone: v += R1/V0^2 (dt g_y +
// 1/(rho_oil rho_half) dt f), rho_half at the y faces (eps 1.5 dy).
struct FsGravPipe {
    double gd[3], drive, eta, ro, R1, V0, eps;
    int vert;
};
__global__ void k_gravity_pipe(double* __restrict__ vel, const double* __restrict__ sfg, const double* __restrict__ b,
                               FsLayout L, FsBox bx, int comp, FsGravPipe G) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= bx.ni || j >= bx.nj) return;
    if (fixed_face(comp, i, j, k, L.nx, L.ny, L.nz, G.vert)) return;
    const long o = L.idx(i, j, k);
    if (!(b[o] < 0)) return;
    if (G.vert) {
        double lc = (double)0.5 * (sfg[o] + sfg[o - L.sx]);
        double dh = (double)1 + (G.eta - (double)1) * heav_solver(lc, G.eps);
        vel[o] += G.R1 / (G.V0 * G.V0) * (G.gd[1] + (double)1 / (G.ro * dh) * G.drive);
    } else if (comp == 1) {
        double lc = (double)0.5 * (sfg[o] + sfg[o - 1]);
        double dh = (double)1 + (G.eta - (double)1) * heav_solver(lc, G.eps);
        vel[o] += G.R1 / (G.V0 * G.V0) * (G.gd[0] + ((double)1 / (dh * G.ro)) * G.drive);
    } else {
        vel[o] += G.R1 / (G.V0 * G.V0) * G.gd[comp - 1];
    }
}

// ---- semi-implicit viscosity: coefficients and explicit terms ----------------

// VISCOSITY_3D::ApplyViscosity, dimensionless (VISCOSITY_3D.h:579-852),
// one component: the level set at the face, edge
// centres and the two cells about it from the shared scalar ghost copy,
// density 1 + (eta - 1) H, viscosities 1 + (m - 1) H (eps 1.5 dx,
// :1258-1347), coefficients 2 dt/(Re rho) mu for the normal pair and
// dt/(Re rho) mu for the tangential ones, and the explicit term from the
// per-axis ghost copies (Y/Z: the two cross-component copies the component
// reads).
struct FsViscPipe {
    double dt, Re, eta, m, eps;
    int vert;  // the vertical pipe's z term takes 1/dz where the parallel one takes 1/dx (:827 vs :835)
};
__global__ void k_visc_coef(double* __restrict__ c0, double* __restrict__ c1, double* __restrict__ c2,
                            double* __restrict__ c3, double* __restrict__ c4, double* __restrict__ c5,
                            double* __restrict__ expl, const double* __restrict__ W, const double* __restrict__ self,
                            const double* __restrict__ A, const double* __restrict__ B, FsLayout L, FsBox bx, int comp,
                            FsGridConst g, FsViscPipe V) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= bx.ni || j >= bx.nj) return;
    const long o = L.idx(i, j, k), sx = L.sx, sy = L.sy;
    const double ox = g.x.inv, oy = g.y.inv, oz = g.z.inv;
    double half, l[6];
    if (comp == 1) {
        half = (double)0.5 * (W[o - 1] + W[o]);
        l[0] = W[o];
        l[1] = W[o - 1];
        l[2] = (double)0.25 * (W[o - 1] + W[o] + W[o - 1 + sx] + W[o + sx]);
        l[3] = (double)0.25 * (W[o - 1] + W[o] + W[o - 1 - sx] + W[o - sx]);
        l[4] = (double)0.25 * (W[o - 1] + W[o] + W[o - 1 + sy] + W[o + sy]);
        l[5] = (double)0.25 * (W[o - 1] + W[o] + W[o - 1 - sy] + W[o - sy]);
    } else if (comp == 2) {
        half = (double)0.5 * (W[o - sx] + W[o]);
        l[0] = (double)0.25 * (W[o - sx] + W[o + 1 - sx] + W[o] + W[o + 1]);
        l[1] = (double)0.25 * (W[o - 1 - sx] + W[o - sx] + W[o - 1] + W[o]);
        l[2] = W[o];
        l[3] = W[o - sx];
        l[4] = (double)0.25 * (W[o - sx + sy] + W[o - sx] + W[o + sy] + W[o]);
        l[5] = (double)0.25 * (W[o - sx - sy] + W[o - sx] + W[o - sy] + W[o]);
    } else {
        half = (double)0.5 * (W[o - sy] + W[o]);
        l[0] = (double)0.25 * (W[o - sy] + W[o + 1 - sy] + W[o] + W[o + 1]);
        l[1] = (double)0.25 * (W[o - sy] + W[o - 1 - sy] + W[o] + W[o - 1]);
        l[2] = (double)0.25 * (W[o - sy] + W[o + sx - sy] + W[o] + W[o + sx]);
        l[3] = (double)0.25 * (W[o - sy] + W[o - sx - sy] + W[o] + W[o - sx]);
        l[4] = W[o];
        l[5] = W[o - sy];
    }
    const double rho_h = (double)1 + (V.eta - (double)1) * heav_visc(half, V.eps);
    double mu[6];
    for (int e = 0; e < 6; e++) mu[e] = (double)1 + (V.m - (double)1) * heav_visc(l[e], V.eps);
    const int n0 = comp == 1 ? 0 : comp == 2 ? 2 : 4;
    double* cf[6] = {c0, c1, c2, c3, c4, c5};
    for (int e = 0; e < 6; e++)
        cf[e][o] = (e == n0 || e == n0 + 1) ? (double)2 * V.dt / (V.Re * rho_h) * mu[e] : V.dt / (V.Re * rho_h) * mu[e];
    const double dt = V.dt, Re = V.Re;
    double ex;
    if (comp == 1) {
        ex = self[o] +
             dt / (Re * rho_h) * (oy * (mu[2] * ox * (A[o + sx] - A[o - 1 + sx]) - mu[3] * ox * (A[o] - A[o - 1]))) +
             dt / (Re * rho_h) * (oz * (mu[4] * ox * (B[o + sy] - B[o - 1 + sy]) - mu[5] * ox * (B[o] - B[o - 1])));
    } else if (comp == 2) {
        ex = self[o] +
             dt / (Re * rho_h) * (ox * (mu[0] * oy * (A[o + 1] - A[o + 1 - sx]) - mu[1] * oy * (A[o] - A[o - sx]))) +
             dt / (Re * rho_h) * (oz * (mu[4] * oy * (B[o + sy] - B[o - sx + sy]) - mu[5] * ox * (B[o] - B[o - sx])));
    } else {
        ex = self[o] +
             dt / (Re * rho_h) * (ox * (oz * (mu[0] * (A[o + 1] - A[o + 1 - sy])) - oz * (mu[1] * (A[o] - A[o - sy])))) +
             dt / (Re * rho_h) * (oy * (oz * (mu[2] * (B[o + sx] - B[o + sx - sy])) - (V.vert ? oz : ox) * (mu[3] * (B[o] - B[o - sy]))));
    }
    expl[o] = ex;
}

// ---- linear systems ------------------------------------------------------------

// ELL matrix with up to 7 entries per row in the reference's CSR order.
#define PSLOTS 7

// POISSON_SOLVER_3D::BuildLinearSystemNodeForSemiImplicitViscosity with
// per-axis ghost arrays (:3641-3828), one row per thread.  bc: the
// component's bc field after row numbering (FULL cells hold their row).
struct FsViscSys {
    double inv_dxdx, inv_dydy, inv_dzdz;
    int comp, ie, je;
};
__global__ void k_visc_assemble(const long* __restrict__ rows, int N, const int* __restrict__ bc,
                                const double* __restrict__ c0, const double* __restrict__ c1,
                                const double* __restrict__ c2, const double* __restrict__ c3,
                                const double* __restrict__ c4, const double* __restrict__ c5,
                                const double* __restrict__ expl, const double* __restrict__ vx,
                                const double* __restrict__ vy, const double* __restrict__ vz,
                                const double* __restrict__ vel, double* __restrict__ val, int* __restrict__ col,
                                int* __restrict__ cnt, double* __restrict__ dg, double* __restrict__ rhs,
                                double* __restrict__ x, FsLayout L, FsViscSys S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const long o = rows[r], sx = L.sx, sy = L.sy;
    const double cf[6] = {c0[o], c1[o], c2[o], c3[o], c4[o], c5[o]};
    // neighbours i-1, i+1, j-1, j+1, k-1, k+1 with coef_2, coef_1, coef_4,
    // coef_3, coef_6, coef_5 and 1/dx^2, 1/dx^2, 1/dy^2, ...
    const long off[6] = {-1, 1, -sx, sx, -sy, sy};
    const int ci[6] = {1, 0, 3, 2, 5, 4};
    const double inv[6] = {S.inv_dxdx, S.inv_dxdx, S.inv_dydy, S.inv_dydy, S.inv_dzdz, S.inv_dzdz};
    int nb[6];
    for (int e = 0; e < 6; e++) nb[e] = bc[o + off[e]];
    int n = 0;
    double coef_ijk = 0;
    for (int e = 0; e < 6; e++)
        if (nb[e] > -1) {
            coef_ijk += cf[ci[e]];
            val[n * N + r] = -cf[ci[e]] * inv[e];
            col[n * N + r] = nb[e];
            n++;
        }
    // periodic neighbours along x, then y (:3721-3770)
    const int i = (int)((o - L.off) % sx) - FS_G, j = (int)(((o - L.off) / sx) % (sy / sx)) - FS_G;
    if (nb[0] == BC_PER) {
        coef_ijk += cf[1];
        int pi = S.comp == 1 ? S.ie + (i - 1 - (0 + 1)) : S.ie + (i - 1 - 0);
        val[n * N + r] = -cf[1] * S.inv_dxdx;
        col[n * N + r] = bc[o + (pi - i)];
        n++;
    }
    if (nb[1] == BC_PER) {
        coef_ijk += cf[0];
        int pi = S.comp == 1 ? 0 + (i + 1 - (S.ie - 1)) : 0 + (i + 1 - S.ie);
        val[n * N + r] = -cf[0] * S.inv_dxdx;
        col[n * N + r] = bc[o + (pi - i)];
        n++;
    }
    if (nb[2] == BC_PER) {
        coef_ijk += cf[3];
        int pj = S.comp == 2 ? S.je + (j - 1 - (0 + 1)) : S.je + (j - 1 - 0);
        val[n * N + r] = -cf[3] * S.inv_dydy;
        col[n * N + r] = bc[o + (long)(pj - j) * sx];
        n++;
    }
    if (nb[3] == BC_PER) {
        coef_ijk += cf[2];
        int pj = S.comp == 2 ? 0 + (j + 1 - (S.je - 1)) : 0 + (j + 1 - S.je);
        val[n * N + r] = -cf[2] * S.inv_dydy;
        col[n * N + r] = bc[o + (long)(pj - j) * sx];
        n++;
    }
    for (int e = 0; e < 6; e++)
        if (nb[e] == BC_DIR) coef_ijk += cf[ci[e]];
    if (coef_ijk == 0) coef_ijk = 1;
    const double d = (double)1 + S.inv_dxdx * (double)coef_ijk;
    val[n * N + r] = d;
    col[n * N + r] = r;
    n++;
    cnt[r] = n;
    dg[r] = d;
    for (int e = n; e < PSLOTS; e++) val[e * N + r] = 0, col[e * N + r] = r;  // padding the PCG loop reads
    double bv = expl[o];
    const double* dir[6] = {vx, vx, vy, vy, vz, vz};
    for (int e = 0; e < 6; e++)
        if (nb[e] == BC_DIR) bv += cf[ci[e]] * inv[e] * dir[e][o + off[e]];
    rhs[r] = bv;
    x[r] = vel[o];
}

// POISSON_SOLVER_3D::BuildLinearSystemNodeJumpConditionVaribleCoefficient,
// oil/water (:1360-1815), with jc_on_derivative = 0: GFM betas from the
// level set L and 1/rho V, the off-diagonals to FULL neighbours, the pin of
// the first full cell (beta_d when its lower neighbour is NEUM), the jump
// terms of the right-hand side, x0 = p.
struct FsPresSys {
    double inv_dxdx, inv_dydy, inv_dzdz, odx, ody, odz, odx2, ody2, odz2;
};
__global__ void k_pres_assemble(const long* __restrict__ rows, int N, const int* __restrict__ bc,
                                const double* __restrict__ Lv, const double* __restrict__ V,
                                const double* __restrict__ JC, const double* __restrict__ div,
                                const double* __restrict__ p, double* __restrict__ val, int* __restrict__ col,
                                int* __restrict__ cnt, double* __restrict__ dg, double* __restrict__ rhs,
                                double* __restrict__ x, FsLayout L, FsPresSys S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const long o = rows[r], sx = L.sx, sy = L.sy;
    const long off[6] = {-1, 1, -sx, sx, -sy, sy};
    const double lc = Lv[o], vc = V[o];
    double ln[6], vn[6], bt[6];
    for (int e = 0; e < 6; e++) ln[e] = Lv[o + off[e]], vn[e] = V[o + off[e]];
    for (int e = 0; e < 6; e++)
        bt[e] = (e % 2 == 0) ? vc * vn[e] * (fabs(ln[e]) + fabs(lc)) / (vc * fabs(ln[e]) + vn[e] * fabs(lc))
                             : vc * vn[e] * (fabs(ln[e]) + fabs(lc)) / (vn[e] * fabs(lc) + vc * fabs(ln[e]));
    const double inv[6] = {S.inv_dxdx, S.inv_dxdx, S.inv_dydy, S.inv_dydy, S.inv_dzdz, S.inv_dzdz};
    int n = 0;
    double coef_ijk = 0;
    for (int e = 0; e < 6; e++) {
        int nb = bc[o + off[e]];
        if (nb > -1) {
            coef_ijk += bt[e];
            val[n * N + r] = -bt[e] * inv[e];
            col[n * N + r] = nb;
            n++;
        }
    }
    if (r == 0 && bc[o - sy] == BC_NEUM) coef_ijk += bt[4];
    if (coef_ijk == 0) coef_ijk = 1;
    const double d = S.inv_dxdx * (double)coef_ijk;
    val[n * N + r] = d;
    col[n * N + r] = r;
    n++;
    cnt[r] = n;
    dg[r] = d;
    for (int e = n; e < PSLOTS; e++) val[e * N + r] = 0, col[e * N + r] = r;  // padding the PCG loop reads
    const double o2[6] = {S.odx2, S.odx2, S.ody2, S.ody2, S.odz2, S.odz2}, o1[6] = {S.odx, S.odx, S.ody, S.ody, S.odz, S.odz};
    const double jcc = JC[o];
    double F[6];
    for (int e = 0; e < 6; e++) {
        const double jn = JC[o + off[e]];
        double subcell = fabs(ln[e]) / (fabs(lc) + fabs(ln[e]));
        double aa = (jcc * fabs(ln[e]) + jn * fabs(lc)) / (fabs(lc) + fabs(ln[e]));
        double bb = (0 * fabs(ln[e]) + 0 * fabs(lc)) / (fabs(lc) + fabs(ln[e]));
        const bool minus = e % 2 == 0;
        if ((lc <= 0 && ln[e] <= 0) || (lc > 0 && ln[e] > 0))
            F[e] = 0;
        else if (lc <= 0 && ln[e] > 0)
            F[e] = minus ? o2[e] * aa * bt[e] - o1[e] * bt[e] * bb * subcell / vn[e]
                         : o2[e] * aa * bt[e] + o1[e] * bt[e] * bb * subcell / vn[e];
        else
            F[e] = minus ? -o2[e] * aa * bt[e] + o1[e] * bt[e] * bb * subcell / vn[e]
                         : -o2[e] * aa * bt[e] - o1[e] * bt[e] * bb * subcell / vn[e];
    }
    const double FX = F[0] + F[1], FY = F[2] + F[3], FZ = F[4] + F[5];
    rhs[r] = -div[o] - FX - FY - FZ;
    x[r] = p[o];
}

// ---- reductions in the fixed order -------------------------------------------

// one block of RBLK rows: the binary tree of the values (rows past N are +0)
__device__ __forceinline__ double block_tree(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int h = RBLK / 2; h >= 1; h /= 2) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    return sh[0];
}

// PCG scalars on the device.  done: 1 converged, 2 max_iteration reached.
struct FsPipePcg {
    double rz, rz_old, pq, alpha, beta, rr, bb, sum;
    double rzs[2];  // r.z of the last two iterations, by parity (the loop kernels' own copy)
    int stop[2];    // the loop's stop decision of iteration k in stop[k & 1] (0 go on, else done's value)
    int iter, done, max_iter, rel;
    double tol2, rtol;
};

// the second level of the fixed order: lane t adds the block sums t, t +
// RBLK, ... in order, then the binary tree over the lanes; every thread of
// the (single, RBLK-thread) workgroup returns the total
__device__ __forceinline__ double sum_parts(const double* __restrict__ part, int nb, double* sh) {
    double v = 0;
    for (int q = threadIdx.x; q < nb; q += RBLK) v += part[q];
    double t = block_tree(v, sh);
    __syncthreads();
    return t;
}

// partial[blk] = tree over the block of a*b (b == NULL: a)
__global__ void __launch_bounds__(RBLK) k_dot_part(const double* __restrict__ a, const double* __restrict__ b, int N,
                                                   double* __restrict__ part) {
    __shared__ double sh[RBLK];
    int r = blockIdx.x * RBLK + threadIdx.x;
    double v = r < N ? (b ? a[r] * b[r] : a[r]) : (double)0;
    double t = block_tree(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
// the total of the block sums into *out (one RBLK-thread workgroup)
__global__ void __launch_bounds__(RBLK) k_sum_parts(const double* __restrict__ part, int nb, double* out) {
    __shared__ double sh[RBLK];
    double t = sum_parts(part, nb, sh);
    if (threadIdx.x == 0) *out = t;
}

// b -= (1/N) sum(b), POISSON_SOLVER_3D.h:221-246
__global__ void k_remove_mean(double* __restrict__ b, int N, const double* __restrict__ sum) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    double one_over_b_dimension = 1 / (double)N;
    b[r] -= one_over_b_dimension * *sum;
}

// y = A x in each row's entry order (CSR_MATRIX::Multiply, CSR_MATRIX.h:
// 236-260)
__global__ void __launch_bounds__(RBLK) k_spmv(const double* __restrict__ val, const int* __restrict__ col,
                                               const int* __restrict__ cnt, int N, const double* __restrict__ x,
                                               double* __restrict__ y) {
    int r = blockIdx.x * RBLK + threadIdx.x;
    if (r >= N) return;
    const int n = cnt[r];
    double v = 0;
    for (int e = 0; e < n; e++) v += val[e * N + r] * x[col[e * N + r]];
    y[r] = v;
}

// r = b - q, z = r / D, p = z; partials of r.z, b.b, r.r
__global__ void __launch_bounds__(RBLK) k_pcg_begin(const double* __restrict__ b, const double* __restrict__ q,
                                                    const double* __restrict__ dg, double* __restrict__ r,
                                                    double* __restrict__ z, double* __restrict__ p, int N,
                                                    double* __restrict__ part, int nb) {
    __shared__ double sh[RBLK];
    int i = blockIdx.x * RBLK + threadIdx.x;
    double rv = 0, zv = 0, bv = 0;
    if (i < N) {
        bv = b[i];
        rv = bv - q[i];
        zv = rv / dg[i];
        r[i] = rv;
        z[i] = zv;
        p[i] = zv;
    }
    double t0 = block_tree(i < N ? rv * zv : (double)0, sh);
    __syncthreads();
    double t1 = block_tree(i < N ? bv * bv : (double)0, sh);
    __syncthreads();
    double t2 = block_tree(i < N ? rv * rv : (double)0, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = t0;
        part[nb + blockIdx.x] = t1;
        part[2 * nb + blockIdx.x] = t2;
    }
}
__global__ void __launch_bounds__(RBLK) k_pcg_begin_scalars(const double* __restrict__ part, int nb, FsPipePcg* S) {
    __shared__ double sh[RBLK];
    const double a = sum_parts(part, nb, sh), b = sum_parts(part + nb, nb, sh), c = sum_parts(part + 2 * nb, nb, sh);
    if (threadIdx.x != 0) return;
    S->rz = a;
    S->rzs[1] = a;
    S->bb = b;
    S->rr = c;
    S->iter = 0;
    S->done = 0;
    double bnorm = sqrt(b);
    if (S->rel && (bnorm == 0 || sqrt(c) <= S->rtol * bnorm)) S->done = 1;
    if (S->max_iter <= 0) S->done = S->done ? S->done : 2;
    S->stop[0] = S->done;  // read by iteration 1's first launch
    S->stop[1] = 0;
}
// ---- the PCG loop in two launches per iteration ------------------------------
//
// The fixed order's second level has RBLK lanes: lane t adds the block sums
// t, t + RBLK, ... in order.  The loop kernels run exactly RBLK workgroups of
// W threads (a multiple of RBLK up to PWG) per system; workgroup w walks the blocks w, w +
// RBLK, ... (W / RBLK of them at a time, added in block order), so its
// running sum IS lane w's, whatever W, and a consumer needs only the binary
// tree of the RBLK workgroup sums -- every workgroup of the next kernel
// re-forms the scalar from them, in the same order, without a fence or atomic.
#define PWG 1024
#define PSUB (PWG / RBLK)

// Up to PMAXSYS systems solved together: launch row y of the loop kernels
// iterates system y with its own rows, sums, scalars and stop flags (the
// semi-implicit viscosity's three independent components share launches)
#define PMAXSYS 3
struct PcgBatch {
    const double* val[PMAXSYS];
    const int* col[PMAXSYS];
    const int* cnt[PMAXSYS];
    const double* dg[PMAXSYS];
    double* x[PMAXSYS];
    double* r[PMAXSYS];
    double* z[PMAXSYS];
    double* p[PMAXSYS][2];  // p_k in p[k & 1]; p_1 = z in p[1]
    double* q[PMAXSYS];
    double* lp[PMAXSYS];  // the loop's workgroup sums: p.q, r.z, r.r (RBLK each)
    FsPipePcg* S[PMAXSYS];
    int N[PMAXSYS], nb[PMAXSYS];
};

// The binary tree of one block's RBLK values (t[q] += t[q + h], h = RBLK/2
// .. 1) from LDS, called by the 64 lanes of one wave: the first two levels
// from the four quarters, the last six by lane shuffles (lane q adds lane
// q + h, the same pairs).  Lane 0 returns the total.
__device__ __forceinline__ double tree_lds(const double* t) {
    const int l = threadIdx.x % 64;
    const double a = t[l] + t[l + 128], b = t[l + 64] + t[l + 192];
    double v = a + b;
    for (int h = 32; h >= 1; h /= 2) v += __shfl_down(v, h, 64);
    return v;
}

// the stop test and beta of the iteration whose r.z (and r.r) workgroup sums
// are `lp` (PCG_METHOD.cpp:81-101); m is that iteration's number.  Returns 0
// to go on (beta set), 1 converged, 2 max_iteration reached.  Every thread
// of the workgroup takes part and gets the same result.
__device__ __forceinline__ int pcg_tail(const double* __restrict__ lp, const FsPipePcg* S, int m, double rz_old,
                                        double* sh, double* res, double& t, double& rr, double& beta) {
    const int rel = S->rel;
    if (threadIdx.x < RBLK) {
        sh[threadIdx.x] = lp[threadIdx.x];
        if (rel) sh[RBLK + threadIdx.x] = lp[RBLK + threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const double a = tree_lds(sh), b = rel ? tree_lds(sh + RBLK) : (double)0;
        if (threadIdx.x == 0) res[0] = a, res[1] = b;
    }
    __syncthreads();
    t = res[0];
    rr = res[1];
    if (rel) {
        if (sqrt(rr) < S->rtol * sqrt(S->bb)) return 1;
    } else if (t < S->tol2) {
        return 1;
    }
    beta = t / rz_old;
    return m >= S->max_iter ? 2 : 0;
}
// the stop rule's constants, read before the launch stores to S
struct PcgRule {
    int rel, max_iter;
    double rtol, bb, tol2;
};
__device__ __forceinline__ PcgRule pcg_rule(const FsPipePcg* S) {
    return PcgRule{S->rel, S->max_iter, S->rtol, S->bb, S->tol2};
}

// pcg_tail with the workgroup sums of r.z and r.r already loaded by threads
// t < RBLK (l0 = lp[t], l1 = lp[RBLK + t] of the sums' array)
__device__ __forceinline__ int pcg_tail_pre(double l0, double l1, const PcgRule& R, int m, double rz_old,
                                            double* sh, double* res, double& t, double& rr, double& beta) {
    const int rel = R.rel;
    if (threadIdx.x < RBLK) {
        sh[threadIdx.x] = l0;
        sh[RBLK + threadIdx.x] = l1;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const double a = tree_lds(sh), b = rel ? tree_lds(sh + RBLK) : (double)0;
        if (threadIdx.x == 0) res[0] = a, res[1] = b;
    }
    __syncthreads();
    t = res[0];
    rr = res[1];
    if (rel) {
        if (sqrt(rr) < R.rtol * sqrt(R.bb)) return 1;
    } else if (t < R.tol2) {
        return 1;
    }
    beta = t / rz_old;
    return m >= R.max_iter ? 2 : 0;
}

// Keeps loads issued above a branch: the compiler may not move a load whose
// value an asm statement reads below that statement
#define PIN2(a, b) asm volatile("" ::"v"(a), "v"(b))

// the host-facing record of a tail decision (one thread)
__device__ __forceinline__ void pcg_record(FsPipePcg* S, int dec, int m, double rz_old, double t, double rr,
                                          double beta) {
    S->rz_old = rz_old;
    S->rz = t;
    if (S->rel) S->rr = rr;
    if (dec == 1) {
        S->iter = m - 1;
        S->done = 1;
        return;
    }
    S->beta = beta;
    S->iter = m;
    if (dec == 2) S->done = 2;
}

// After a pass has put its NS x W row values in sh (step s, sub-block q,
// row l at sh[s * W + q * RBLK + l]): each block's tree, then thread 0
// adds the trees of the blocks that exist (index c0 + RBLK ((W / RBLK) s + q) <
// nb) to acc in block order
template <int NS, int W>
__device__ __forceinline__ void pass_trees(const double* sh, double* res, int c0, int nb, double& acc) {
    __syncthreads();
    if (threadIdx.x % RBLK < 64) {
        const int q = threadIdx.x / RBLK;
        for (int s = 0; s < NS; s++) {
            const double v = tree_lds(sh + s * W + q * RBLK);
            if (threadIdx.x % RBLK == 0) res[s * (W / RBLK) + q] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int s = 0; s < NS; s++)
            for (int q = 0; q < (W / RBLK); q++)
                if (c0 + RBLK * ((W / RBLK) * s + q) < nb) acc += res[s * (W / RBLK) + q];
}

// One row's operands of the fused direction + SpMV step, loaded apart from
// their use so the loads can be issued before the tail that gives beta.
// ELL slots past the row's count hold (column r, value 0), so every load is
// in bounds.
template <bool FIRST>
struct PcgRow {
    int n, c[PSLOTS];
    double a[PSLOTS], pc[PSLOTS], zc[PSLOTS], pr, zr;
    // the row's own entries, count, p and z (no dependent loads)
    __device__ __forceinline__ void load_row(const double* __restrict__ val, const int* __restrict__ col,
                                             const int* __restrict__ cnt, int N, const double* __restrict__ po,
                                             const double* __restrict__ z, int rc) {
        n = cnt[rc];
#pragma unroll
        for (int e = 0; e < PSLOTS; e++) {
            c[e] = col[e * N + rc];
            a[e] = val[e * N + rc];
        }
        pr = po[rc];
        zr = FIRST ? (double)0 : z[rc];
    }
    // the neighbours' p and z, by the row's columns
    __device__ __forceinline__ void gather(const double* __restrict__ po, const double* __restrict__ z) {
#pragma unroll
        for (int e = 0; e < PSLOTS; e++) {
            pc[e] = po[c[e]];
            zc[e] = FIRST ? (double)0 : z[c[e]];
        }
    }
    // the columns are read: the gathers can go out
    __device__ __forceinline__ void pin_columns() const {
        PIN2(c[0], c[1]);
        PIN2(c[2], c[3]);
        PIN2(c[4], c[5]);
        PIN2(c[6], n);
    }
    __device__ __forceinline__ void load(const double* __restrict__ val, const int* __restrict__ col,
                                         const int* __restrict__ cnt, int N, const double* __restrict__ po,
                                         const double* __restrict__ z, int rc) {
        load_row(val, col, cnt, N, po, z, rc);
        gather(po, z);
    }
    // q = sum over the row's entries in order of a_e (p beta + z)_e; the
    // row's own new direction in p_row
    __device__ __forceinline__ double apply(double beta, double& p_row) const {
        double v = 0;
#pragma unroll
        for (int e = 0; e < PSLOTS; e++) {
            double x = pc[e];
            if (!FIRST) {
                x *= beta;
                x += zc[e];
            }
            const double t = v + a[e] * x;
            v = e < n ? t : v;
        }
        p_row = pr;
        if (!FIRST) {
            p_row *= beta;
            p_row += zr;
        }
        return v;
    }
};

// Iteration k, first launch: the tail of iteration k - 1 (k > 1), the new
// direction p_k = p_{k-1} beta + z (PCG_METHOD.cpp:104-108, formed again for
// every neighbour a row reads, the same bits), q = A p_k in each row's entry
// order (CSR_MATRIX.h:236-260), and the workgroup sums of p.q.  p_1 = z is
// in `po` already; for k > 1 p_k goes to `pn`.  S->rzs[k & 1] receives r.z
// of iteration k - 1 for the update.  A pass covers NS steps of W rows,
// their loads in flight together; with NS = 1 (one step covers the system)
// the row's loads go out before the tail.
template <int NS, bool FIRST, int W>
__global__ void __launch_bounds__(W) k_pcg_spmv_dir(const PcgBatch B, int k) {
    const int y = blockIdx.y;
    const double* __restrict__ val = B.val[y];
    const int* __restrict__ col = B.col[y];
    const int* __restrict__ cnt = B.cnt[y];
    const int N = B.N[y], nb = B.nb[y];
    const double* __restrict__ po = B.p[y][k == 1 ? 1 : (k - 1) & 1];
    double* __restrict__ pn = B.p[y][k & 1];
    const double* __restrict__ z = B.z[y];
    double* __restrict__ q = B.q[y];
    double* __restrict__ lp = B.lp[y];
    FsPipePcg* S = B.S[y];
    // the rows' values of a pass, and the tail's two sets of RBLK sums
    __shared__ double sh[NS * W > 2 * RBLK ? NS * W : 2 * RBLK];
    __shared__ double res[NS * (W / RBLK) > 2 ? NS * (W / RBLK) : 2];
    // an earlier stop is carried forward; this launch reads slot k - 1 and
    // writes slot k only, so no flag is read and written in one launch
    const int prev = S->stop[(k - 1) & 1];
    const PcgRule R = pcg_rule(S);
    const double rz_old = FIRST ? (double)0 : S->rzs[(k - 1) & 1];
    const int sub = threadIdx.x / RBLK, l = threadIdx.x % RBLK;
    PcgRow<FIRST> row1;
    const int b1 = blockIdx.x + RBLK * sub, r1 = b1 * RBLK + l;
    const bool in1 = b1 < nb && r1 < N;
    // one-step passes: the flag, the row and the last iteration's sums are
    // independent loads, all in bounds whatever the flag says; they go out
    // together, ahead of the branch on the flag
    double l0 = 0, l1 = 0;
    if (NS == 1) {
        row1.load_row(val, col, cnt, N, po, z, in1 ? r1 : 0);
        if (!FIRST && threadIdx.x < RBLK) {
            l0 = lp[RBLK + threadIdx.x];
            l1 = lp[2 * RBLK + threadIdx.x];
        }
        row1.pin_columns();
        PIN2(l0, l1);
    }
    if (prev) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->stop[k & 1] = prev;
        return;
    }
    if (NS == 1) row1.gather(po, z);
    double beta = 0;
    if (FIRST) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S->stop[k & 1] = 0;
    } else {
        double t, rr;
        const int dec = NS == 1 ? pcg_tail_pre(l0, l1, R, k - 1, rz_old, sh, res, t, rr, beta)
                                : pcg_tail(lp + RBLK, S, k - 1, rz_old, sh, res, t, rr, beta);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            pcg_record(S, dec, k - 1, rz_old, t, rr, beta);
            S->stop[k & 1] = dec;
            if (dec == 0) S->rzs[k & 1] = t;
        }
        // stopped: the reference's last direction is never read
        if (dec) return;
    }
    double acc = 0;
    if (NS == 1) {
        double pr;
        const double v = row1.apply(beta, pr);
        if (in1) {
            q[r1] = v;
            if (!FIRST) pn[r1] = pr;
        }
        sh[threadIdx.x] = in1 ? pr * v : (double)0;
        pass_trees<NS, W>(sh, res, blockIdx.x, nb, acc);
    } else {
        for (int c0 = blockIdx.x; c0 < nb; c0 += RBLK * (W / RBLK) * NS) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int b = c0 + RBLK * ((W / RBLK) * s + sub), r = b * RBLK + l;
                const bool in = b < nb && r < N;
                PcgRow<FIRST> row;
                row.load(val, col, cnt, N, po, z, in ? r : 0);
                double pr;
                const double v = row.apply(beta, pr);
                if (in) {
                    q[r] = v;
                    if (!FIRST) pn[r] = pr;
                }
                sh[s * W + threadIdx.x] = in ? pr * v : (double)0;
            }
            pass_trees<NS, W>(sh, res, c0, nb, acc);
        }
    }
    if (threadIdx.x == 0) lp[blockIdx.x] = acc;
}

// Iteration k, second launch: alpha = r.z / p.q, x += alpha p, r -= alpha q,
// z = r / D, and the workgroup sums of r.z (and r.r) for the next tail
template <int NS, int W>
__global__ void __launch_bounds__(W) k_pcg_update2(const PcgBatch B, int k) {
    const int y = blockIdx.y;
    double* __restrict__ x = B.x[y];
    double* __restrict__ r = B.r[y];
    double* __restrict__ z = B.z[y];
    const double* __restrict__ p = B.p[y][k & 1];
    const double* __restrict__ q = B.q[y];
    const double* __restrict__ dg = B.dg[y];
    const int N = B.N[y], nb = B.nb[y];
    double* __restrict__ lp = B.lp[y];
    FsPipePcg* S = B.S[y];
    __shared__ double sh[2 * NS * W];
    __shared__ double res[NS * (W / RBLK)];
    const int stop = S->stop[k & 1];
    const int rel = S->rel;
    const double rz = S->rzs[k & 1];
    const int sub = threadIdx.x / RBLK, l = threadIdx.x % RBLK;
    // one-step passes: the row's loads and the sums of p.q go out together,
    // ahead of the branch on the stop flag (all in bounds whatever it says)
    const int b1 = blockIdx.x + RBLK * sub, i1 = b1 * RBLK + l;
    const bool in1 = b1 < nb && i1 < N;
    const int ic1 = in1 ? i1 : 0;
    double x1 = 0, p1 = 0, r1 = 0, q1 = 0, d1 = 1, l0 = 0;
    if (NS == 1) {
        x1 = x[ic1], p1 = p[ic1], r1 = r[ic1], q1 = q[ic1], d1 = dg[ic1];
        if (threadIdx.x < RBLK) l0 = lp[threadIdx.x];
        PIN2(p1, q1);
        PIN2(l0, r1);
        PIN2(rz, rel);
    }
    if (stop) return;
    if (threadIdx.x < RBLK) sh[threadIdx.x] = NS == 1 ? l0 : lp[threadIdx.x];
    __syncthreads();
    if (threadIdx.x < 64) {
        const double t = tree_lds(sh);
        if (threadIdx.x == 0) res[0] = t;
    }
    __syncthreads();
    const double pq = res[0];
    const double alpha = rz / pq;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->pq = pq;
        S->alpha = alpha;
    }
    double acc0 = 0, acc1 = 0;
    if (NS == 1) {
        const double xv = x1 + alpha * p1;
        const double rv = r1 - alpha * q1;
        const double zv = rv / d1;
        if (in1) {
            x[i1] = xv;
            r[i1] = rv;
            z[i1] = zv;
        }
        sh[threadIdx.x] = in1 ? rv * zv : (double)0;
        sh[W + threadIdx.x] = in1 ? rv * rv : (double)0;
        pass_trees<NS, W>(sh, res, blockIdx.x, nb, acc0);
        if (rel) pass_trees<NS, W>(sh + W, res, blockIdx.x, nb, acc1);
    } else {
        for (int c0 = blockIdx.x; c0 < nb; c0 += RBLK * (W / RBLK) * NS) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int b = c0 + RBLK * ((W / RBLK) * s + sub), i = b * RBLK + l;
                const bool in = b < nb && i < N;
                const int ic = in ? i : 0;
                const double xv = x[ic] + alpha * p[ic];
                const double rv = r[ic] - alpha * q[ic];
                const double zv = rv / dg[ic];
                if (in) {
                    x[i] = xv;
                    r[i] = rv;
                    z[i] = zv;
                }
                sh[s * W + threadIdx.x] = in ? rv * zv : (double)0;
                sh[(NS + s) * W + threadIdx.x] = in ? rv * rv : (double)0;
            }
            pass_trees<NS, W>(sh, res, c0, nb, acc0);
            if (rel) pass_trees<NS, W>(sh + NS * W, res, c0, nb, acc1);
        }
    }
    if (threadIdx.x == 0) {
        lp[RBLK + blockIdx.x] = acc0;
        lp[2 * RBLK + blockIdx.x] = acc1;
    }
}

// the tail of iteration m at the end of a launch batch, so the host sees the
// stop without another batch (one RBLK-thread workgroup per system; the next
// batch's first launch repeats the same decision when it goes on)
__global__ void __launch_bounds__(RBLK) k_pcg_tail_end(const PcgBatch B, int m) {
    const double* __restrict__ lp = B.lp[blockIdx.y];
    FsPipePcg* S = B.S[blockIdx.y];
    __shared__ double sh[2 * RBLK];
    __shared__ double res[2];
    if (m < 1 || S->stop[m & 1]) return;
    const double rz_old = S->rzs[m & 1];
    double t, rr, beta = 0;
    const int dec = pcg_tail(lp + RBLK, S, m, rz_old, sh, res, t, rr, beta);
    if (threadIdx.x == 0) {
        pcg_record(S, dec, m, rz_old, t, rr, beta);
        if (dec) S->stop[m & 1] = dec;  // the next batch's first launch reads it
    }
}

// r = b - q (the true residual after the solve)
__global__ void k_residual(const double* __restrict__ b, const double* __restrict__ q, double* __restrict__ r, int N) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) r[i] = b[i] - q[i];
}

// VectorToGrid (POISSON_SOLVER_3D.h:4060-4078): FULL cells take x, fixed
// faces excepted
__global__ void k_scatter(double* __restrict__ dst, const double* __restrict__ x, const long* __restrict__ rows, int N,
                          FsLayout L, int comp, int vert) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const long o = rows[r];
    if (comp) {
        long q = o - L.off;
        int i = (int)(q % L.sx) - FS_G, j = (int)((q / L.sx) % (L.sy / L.sx)) - FS_G, k = (int)(q / L.sy) - FS_G;
        if (fixed_face(comp, i, j, k, L.nx, L.ny, L.nz, vert)) return;
    }
    dst[o] = x[r];
}

// after the solve of the face component normal to the pipe
// (POISSON_SOLVER_3D.h:354-376): u(ie, j, k) = u(1, j, k) for the parallel
// pipe, v(i, je, k) = v(i, 1, k) for the vertical one
__global__ void k_fix_last(double* __restrict__ vel, FsLayout L, FsBox bx, int vert) {
    int q = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (vert) {
        if (q >= bx.ni) return;
        vel[L.idx(q, bx.nj - 1, k)] = vel[L.idx(q, 0 + 1, k)];
    } else {
        if (q >= bx.nj) return;
        vel[L.idx(bx.ni - 1, q, k)] = vel[L.idx(0 + 1, q, k)];
    }
}

// ---- projection -----------------------------------------------------------------

// DetermineProjectionDensity, dimensionless (PROJECTION_3D.h:642-668)
__global__ void k_pdensity(const double* __restrict__ phi, double* __restrict__ rho, double* __restrict__ irho,
                           FsLayout L, double rr) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const long o = L.idx(i, j, k);
    double d = phi[o] <= 0 ? 1 : rr;
    rho[o] = d;
    irho[o] = 1 / d;
}

// DetermineDivergence, oil/water (PROJECTION_3D.h:697-725): 0 outside the
// pipe, else the difference of the ghost copies; times 1/dt
__global__ void k_pdiv(const double* __restrict__ gu, const double* __restrict__ gv, const double* __restrict__ gw,
                       const double* __restrict__ bls, double* __restrict__ div, FsLayout L, FsGridConst g,
                       double one_over_dt) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const long o = L.idx(i, j, k);
    double dv;
    if (bls[o] > 0)
        dv = (double)0;
    else {
        dv = g.x.inv * (gu[o + 1] - gu[o]);
        dv += g.y.inv * (gv[o + L.sx] - gv[o]);
        dv += g.z.inv * (gw[o + L.sy] - gw[o]);
    }
    div[o] = dv * one_over_dt;
}

// UpdateVelocityByPressureGradientVariableDensity, oil/water
// (PROJECTION_3D.h:1117-1427), one component; vmax: bits of max |u| over the
// faces updated
struct FsPresUpd {
    double dt, od, eps, rr;
};
__global__ void k_pupdate(double* __restrict__ vel, const double* __restrict__ psfg, const double* __restrict__ phi,
                          const double* __restrict__ bls, const double* __restrict__ jc,
                          const double* __restrict__ rho, const double* __restrict__ p, const int* __restrict__ bc,
                          FsLayout L, FsBox bx, int comp, FsPresUpd U, unsigned long long* vmax, int vert) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    double mine = 0;
    if (i < bx.ni && j < bx.nj && !fixed_face(comp, i, j, k, L.nx, L.ny, L.nz, vert)) {
        const long o = L.idx(i, j, k);
        const long n = o - (comp == 1 ? 1 : comp == 2 ? L.sx : L.sy);
        if (bc[o] >= 0) {
            const double lcf = (double)0.5 * (psfg[o] + psfg[n]);
            const double dh = 1 + (U.rr - (double)1) * heav_solver(lcf, U.eps);
            double vp = vel[o];
            const double one_over_density_ijk = (double)1 / rho[o];
            const double lv = phi[o], ln = phi[n];
            const double gamma = (jc[o] * fabs(ln) + jc[n] * fabs(lv)) / (fabs(ln) + fabs(lv));
            const double one_over_half = 1 / dh;
            const double bcv = (double)bc[o], bcn = (double)bc[n];
            const double dp = p[o] - p[n];
            if (vert && comp != 3) {
                // the vertical pipe's x and y faces (:1188-1205, 1297-1318);
                // the y faces skip the first and the last face
                if (bcv * bcn >= 0 && (comp == 1 || (j != 0 && j <= bx.nj - 1 - 1))) {
                    if (lv * ln > 0)
                        vp -= U.dt * dp * one_over_density_ijk * U.od;
                    else if (lv <= 0)
                        vp -= U.dt * (dp + gamma) * one_over_half * U.od;
                    else if (lv > 0)
                        vp -= U.dt * (dp - gamma) * one_over_half * U.od;
                }
            } else if (comp == 1) {
                if (bls[o] * bls[n] >= 0) {
                    if (bcv * bcn > 0)
                        vp -= U.dt * dp * one_over_density_ijk * U.od;
                    else if (lv <= 0)
                        vp -= U.dt * (dp + gamma) * one_over_half * U.od;
                    else if (lv > 0)
                        vp -= U.dt * (dp - gamma) * one_over_half * U.od;
                }
            } else if (comp == 2) {
                if (bcv * bcn >= 0) {
                    if (lv * ln > 0)
                        vp -= U.dt * dp * one_over_density_ijk * U.od;
                    else if (lv < 0)
                        vp -= U.dt * (dp + gamma) * one_over_half * U.od;
                    else if (lv > 0)
                        vp -= U.dt * (dp - gamma) * one_over_half * U.od;
                }
            } else {
                if (bcv * bcn >= 0) {
                    if (lv * ln > 0)
                        vp -= U.dt * dp * one_over_density_ijk * U.od;
                    else if (lv <= 0)
                        vp -= U.dt * (dp + gamma) * one_over_half * U.od;
                    else if (lv > 0)
                        vp -= U.dt * (dp - gamma) * one_over_half * U.od;
                }
            }
            vel[o] = vp;
            mine = fabs(vp);
        }
    }
    // block max, then one atomic (values >= 0 order as their bit patterns)
    __shared__ double sh[PBX * PBY];
    const int t = threadIdx.y * PBX + threadIdx.x;
    sh[t] = mine;
    __syncthreads();
    for (int h = PBX * PBY / 2; h >= 1; h /= 2) {
        if (t < h) sh[t] = fmax(sh[t], sh[t + h]);
        __syncthreads();
    }
    if (t == 0) atomicMax(vmax, (unsigned long long)__double_as_longlong(sh[0]));
}

// the periodic copies of the end faces after the update (PROJECTION_3D.h:
// 1238-1256, 1345-1363, 1429-1444): parallel pipe, x faces u(ie) = u(1),
// u(0) = u(ie-1), y and z faces v(ie) = v(0); vertical pipe, the same along
// y with the y faces in the role of the x faces
__global__ void k_pfix(double* __restrict__ vel, FsLayout L, FsBox bx, int comp, int vert) {
    if (vert) {
        int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
        if (i >= bx.ni) return;
        const int je = bx.nj - 1;
        if (comp == 2) {
            double a = vel[L.idx(i, 0 + 1, k)], b = vel[L.idx(i, je - 1, k)];
            vel[L.idx(i, je, k)] = a;
            vel[L.idx(i, 0, k)] = b;
        } else
            vel[L.idx(i, je, k)] = vel[L.idx(i, 0, k)];
        return;
    }
    int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (j >= bx.nj) return;
    const int ie = bx.ni - 1;
    if (comp == 1) {
        double a = vel[L.idx(0 + 1, j, k)], b = vel[L.idx(ie - 1, j, k)];
        vel[L.idx(ie, j, k)] = a;
        vel[L.idx(0, j, k)] = b;
    } else
        vel[L.idx(ie, j, k)] = vel[L.idx(0, j, k)];
}

// max |u| of a face box (bits of the FP64 value)
__global__ void k_absmax(const double* __restrict__ a, FsLayout L, FsBox bx, unsigned long long* out) {
    int i = blockIdx.x * PBX + threadIdx.x, j = blockIdx.y * PBY + threadIdx.y, k = blockIdx.z;
    double mine = 0;
    if (i < bx.ni && j < bx.nj) mine = fabs(a[L.idx(i, j, k)]);
    __shared__ double sh[PBX * PBY];
    const int t = threadIdx.y * PBX + threadIdx.x;
    sh[t] = mine;
    __syncthreads();
    for (int h = PBX * PBY / 2; h >= 1; h /= 2) {
        if (t < h) sh[t] = fmax(sh[t], sh[t + h]);
        __syncthreads();
    }
    if (t == 0) atomicMax(out, (unsigned long long)__double_as_longlong(sh[0]));
}

// ---- host data of one linear system ---------------------------------------------

struct PipeSys {
    int N = 0;
    int* bc = nullptr;     // device: the bc field after numbering (padded layout)
    long* rows = nullptr;  // device: row -> padded index
};

}  // namespace

// =====================================================================================
// the solver
// =====================================================================================

struct FsPipe {
    fs_params P;
    fs_pipe_params Q;
    fs_grid G;
    FsLayout L;
    FsGridConst cell, mac[4];
    double x0f[4], y0f[4], z0f[4];
    int vert = 0;  // 1: the vertical pipe (along y), 0: the parallel pipe (along x)
    // crossing tables of the masked y fill's extrapolation, per face component:
    // [0] x rows' first right crossing, [1] last left crossing (nk x nj),
    // [2] z columns' first top crossing, [3] last bottom crossing (nj x ni)
    int* xz[4][4] = {};
    double R1, f, A, V0, J, Re, We, A0;
    int device = 0;
    hipStream_t st = nullptr;
    bool async = false;
    // fields
    double *phi = nullptr, *bls = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *p = nullptr;
    double *sfg = nullptr, *g[4] = {nullptr, nullptr, nullptr, nullptr};
    double* ag[4][2] = {};
    double* vg[4][2] = {};
    double* zero = nullptr;  // the parallel pipe reads VISCOSITY_3D's y_x and z_x copies, never written there
    double* bp[4] = {};
    double *psfg = nullptr, *rho = nullptr, *irho = nullptr, *div = nullptr, *jc = nullptr, *kappa = nullptr;
    double *sgn = nullptr, *tmp = nullptr, *scr = nullptr, *rk[4] = {};
    double* coef[6] = {};
    double* expl = nullptr;
    // systems: 0 pressure, 1..3 viscosity x, y, z
    PipeSys sys[4];
    int Nmax = 0;
    // one PCG system's buffers (ELL rows of at most Nmax, its vectors and
    // sums); set 0 takes the pressure, sets 0-2 the viscosity components,
    // which are solved together
    struct PcgSet {
        double *val = nullptr, *dg = nullptr, *rhs = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr,
               *p2 = nullptr, *q = nullptr, *part = nullptr, *lp = nullptr;
        int *col = nullptr, *cnt = nullptr;
        int N = 0;
    };
    PcgSet ps[PMAXSYS];
    FsPipePcg* pcg = nullptr;  // PMAXSYS scalar records, one per set
    // iterations of the last pressure (0) and viscosity (1) solve: the next
    // solve of the kind checks for its stop first after about that many
    int pcg_hint[2] = {0, 0};
    double* dscal = nullptr;             // device scalars: [0] a sum of the fixed-order reductions
    unsigned long long* bits = nullptr;  // [0] kappa max, [1] curvature failures, [2..4] max |u|, |v|, |w|
    FsPipePcg* host_pcg = nullptr;
    double* host_bits = nullptr;
    std::vector<double*> owned;
    double kmax = 0;
    int last_iters = 0, visc_iters = 0;
    double last_res_abs = 0, last_res_rel = 0;
    bool not_converged = false;
    // in-loop samples of the PCG loop kernels (0 spmv_dir, 1 update2)
    bool sampling = false;
    hipEvent_t kev[3] = {};
    double ks_ms[2] = {}, ks_bytes[2] = {};
    long long ks_n[2] = {}, ks_launch[2] = {};

    FsBox box(int c) const { return fs_box(L, c); }
    double* field(int c) { return c == 1 ? u : c == 2 ? v : w; }

    double* alloc_field() {
        double* d = nullptr;
        PCHECK(hipMalloc(&d, sizeof(double) * L.total));
        owned.push_back(d);
        PCHECK(hipMemsetAsync(d, 0, sizeof(double) * L.total, st));
        return d;
    }

    ~FsPipe() {
        if (st) (void)hipStreamSynchronize(st);
        for (double* d : owned) (void)hipFree(d);
        for (int s = 0; s < 4; s++) {
            (void)hipFree(sys[s].bc);
            (void)hipFree(sys[s].rows);
            for (int t = 0; t < 4; t++) (void)hipFree(xz[s][t]);
        }
        for (PcgSet& q : ps)
            for (void* d : {(void*)q.val, (void*)q.dg, (void*)q.rhs, (void*)q.x, (void*)q.r, (void*)q.z, (void*)q.p,
                            (void*)q.p2, (void*)q.q, (void*)q.part, (void*)q.lp, (void*)q.col, (void*)q.cnt})
                (void)hipFree(d);
        for (void* d : {(void*)pcg, (void*)bits, (void*)dscal}) (void)hipFree(d);
        if (host_pcg) (void)hipHostFree(host_pcg);
        if (host_bits) (void)hipHostFree(host_bits);
        for (hipEvent_t e : kev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }

    void sync() { PCHECK(hipStreamSynchronize(st)); }
    void check_launch() { PCHECK(hipGetLastError()); }

    // ---- set-up --------------------------------------------------------------
    void init(const fs_params& P_, const fs_pipe_params& Q_, const fs_grid& G_, int dev) {
        P = P_, Q = Q_, G = G_;
        if ((Q.is_parallel != 0 && Q.is_parallel != 1) || !Q.dimensionless_form || !Q.cylindrical_boundary ||
            Q.boundary_order_velocity != 1)
            throw FsPipeError{FS_ERR_UNSUPPORTED,
                              "pipe mode: only the parallel or vertical pipe with a cylindrical boundary, the "
                              "dimensionless form and first-order wall extrapolation are implemented"};
        vert = !Q.is_parallel;
        if (P.precision != FS_FP64 || P.schedule != FS_SCHEDULE_COMPAT)
            throw FsPipeError{FS_ERR_UNSUPPORTED, "pipe mode: FS_FP64 and FS_SCHEDULE_COMPAT only"};
        if (P.csf_model) throw FsPipeError{FS_ERR_UNSUPPORTED, "pipe mode: CSF_model is not implemented"};
        if (P.curvature_method != FS_CURV_LEVELSET)
            throw FsPipeError{FS_ERR_UNSUPPORTED, "pipe mode: curvature by level set only"};
        if (P.fast_weno) throw FsPipeError{FS_ERR_UNSUPPORTED, "pipe mode: fast_weno is not implemented"};
        if (G.nx < 8 || G.ny < 8 || G.nz < 8) throw FsPipeError{FS_ERR_BAD_ARG, "pipe mode: every extent must be >= 8"};
        if (P.order_for_time_advancing < 1 || P.order_for_time_advancing > 3)
            throw FsPipeError{FS_ERR_BAD_ARG, "order_for_time_advancing must be 1, 2 or 3"};
        if (!(Q.a > 0) || !(Q.R2 > 0) || !(Q.oil_viscosity > 0) || !(Q.oil_density > 0) ||
            !(P.water_viscosity > 0) || !(P.water_density > 0))
            throw FsPipeError{FS_ERR_BAD_ARG, "pipe mode: a, R2, densities and viscosities must be positive"};
        if ((long)G.nx * G.ny * G.nz > (1L << 28)) throw FsPipeError{FS_ERR_OOM, "grid too large"};
        if (dev >= 0) PCHECK(hipSetDevice(dev));
        PCHECK(hipGetDevice(&device));
        PCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        L = fs_make_layout(G.nx, G.ny, G.nz);
        cell.x = fs_make_axis(G.x_min, G.x_max, G.nx);
        cell.y = fs_make_axis(G.y_min, G.y_max, G.ny);
        cell.z = fs_make_axis(G.z_min, G.z_max, G.nz);
        const double dx = cell.x.d, dy = cell.y.d, dz = cell.z.d;
        for (int c = 0; c < 4; c++) mac[c] = cell, x0f[c] = G.x_min, y0f[c] = G.y_min, z0f[c] = G.z_min;
        // MAC grids, EULERIAN_FLUID_SOLVER_3D.h:407,415,423
        mac[1].x = fs_make_axis(G.x_min - (double)0.5 * dx, G.x_max + (double)0.5 * dx, G.nx + 1);
        mac[2].y = fs_make_axis(G.y_min - (double)0.5 * dy, G.y_max + (double)0.5 * dy, G.ny + 1);
        mac[3].z = fs_make_axis(G.z_min - (double)0.5 * dz, G.z_max + (double)0.5 * dz, G.nz + 1);
        x0f[1] = G.x_min - (double)0.5 * dx;
        y0f[2] = G.y_min - (double)0.5 * dy;
        z0f[3] = G.z_min - (double)0.5 * dz;
        // EULERIAN_FLUID_SOLVER_3D.h:496-520
        const double a = Q.a, m = Q.m, K = Q.K, gg = Q.g, rw = P.water_density, ro = Q.oil_density;
        R1 = Q.R2 / a;
        f = (double)1 / (K - (double)1) * gg * (ro - K * rw);
        A = (a * a - (double)1) + (m * K + (double)2 * (K - 1) * std::log(a));
        V0 = (f + rw * gg) * (R1 * R1) / ((double)4 * P.water_viscosity) * A;
        J = P.surface_tension * R1 * ro / (Q.oil_viscosity * Q.oil_viscosity);
        Re = (ro * R1 * std::fabs(V0)) / Q.oil_viscosity;
        We = ((double)1 / J) * (Re * Re);
        A0 = Q.A_0 / (Q.R2 / a);  // SIMULATION_WORLD.h:126-127
        if (!std::isfinite(Re) || !std::isfinite(We) || !(V0 != 0))
            throw FsPipeError{FS_ERR_BAD_ARG, "pipe mode: the derived constants V0, Re, We are not finite"};
        // fields
        phi = alloc_field(); bls = alloc_field(); u = alloc_field(); v = alloc_field(); w = alloc_field();
        p = alloc_field(); sfg = alloc_field();
        for (int c = 1; c < 4; c++) {
            g[c] = alloc_field();
            ag[c][0] = alloc_field();
            ag[c][1] = alloc_field();
            vg[c][0] = alloc_field();
            vg[c][1] = alloc_field();
            bp[c] = alloc_field();
            rk[c] = alloc_field();
        }
        zero = alloc_field();
        psfg = alloc_field(); rho = alloc_field(); irho = alloc_field(); div = alloc_field(); jc = alloc_field();
        kappa = alloc_field(); sgn = alloc_field(); tmp = alloc_field(); scr = alloc_field();
        for (int e = 0; e < 6; e++) coef[e] = alloc_field();
        expl = alloc_field();
        PCHECK(hipMalloc(&pcg, PMAXSYS * sizeof(FsPipePcg)));
        PCHECK(hipMalloc(&bits, 8 * sizeof(unsigned long long)));
        PCHECK(hipMalloc(&dscal, 8 * sizeof(double)));
        PCHECK(hipHostMalloc(&host_pcg, PMAXSYS * sizeof(FsPipePcg)));
        PCHECK(hipHostMalloc(&host_bits, 8 * sizeof(double)));
        static_geometry();
        sync();
    }

    // boundary_phi_<c> and the pipe's level set (computed on the host, as the
    // reference computes them on each module's interior), and the static bc
    // fields and row numberings of the four systems
    void static_geometry() {
        const long n = L.total;
        std::vector<double> h(n);
        // boundary_phi_<c> on the interior of each face grid: the distance from
        // the pipe's axis less a, with r^2 = y^2 + z^2 for the parallel pipe
        // and x^2 + z^2 for the vertical one (ADVECTION_3D.h:380-432 and its
        // copies in the other modules)
        std::vector<std::vector<double>> hb(4);
        for (int c = 1; c < 4; c++) {
            std::fill(h.begin(), h.end(), 0.0);
            FsBox b = box(c);
            for (int k = 0; k < b.nk; k++)
                for (int j = 0; j < b.nj; j++)
                    for (int i = 0; i < b.ni; i++) {
                        double w1 = vert ? x0f[c] + i * mac[c].x.d : y0f[c] + j * mac[c].y.d, z = z0f[c] + k * mac[c].z.d;
                        h[L.idx(i, j, k)] = std::sqrt(w1 * w1 + z * z) - Q.a;
                    }
            PCHECK(hipMemcpy(bp[c], h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
            hb[c] = h;
            if (vert) crossing_tables(c, h);
        }
        // the pipe inside the box: the first and last two cells and faces of
        // every line across it outside the wall
        for (int c = 0; c < 4; c++) {
            FsBox b = box(c);
            for (int k = 0; k < b.nk; k++)
                for (int j = 0; j < b.nj; j++)
                    for (int i = 0; i < b.ni; i++) {
                        const int q = vert ? i : j, nq = vert ? b.ni : b.nj;
                        if (!(q < 2 || q >= nq - 2 || k < 2 || k >= b.nk - 2)) continue;
                        double w1 = vert ? x0f[c] + i * mac[c].x.d : y0f[c] + j * mac[c].y.d, z = z0f[c] + k * mac[c].z.d;
                        if (!(std::sqrt(w1 * w1 + z * z) - Q.a > 0))
                            throw FsPipeError{FS_ERR_BAD_ARG, "pipe mode: the pipe wall must lie inside the box, two "
                                                              "cells from every face across the pipe"};
                    }
        }
        // the pipe's level set, SIMULATION_WORLD.cpp:305-319: r - a on the
        // cells, ghosts by linear extrapolation
        std::fill(h.begin(), h.end(), 0.0);
        for (int k = 0; k < G.nz; k++)
            for (int j = 0; j < G.ny; j++)
                for (int i = 0; i < G.nx; i++) {
                    double w1 = vert ? G.x_min + i * cell.x.d : G.y_min + j * cell.y.d, z = G.z_min + k * cell.z.d;
                    h[L.idx(i, j, k)] = std::sqrt(w1 * w1 + z * z) - Q.a;
                }
        std::vector<double> hbls = linear_ghosts(h);
        PCHECK(hipMemcpy(bls, hbls.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        // pressure bc, PROJECTION_3D::SetupBoundaryConditions, Neumann,
        // cylindrical (:1017-1065)
        std::vector<int> bc(n, 0);
        auto B = [&](int i, int j, int k) { return hbls[L.idx(i, j, k)]; };
        for (int k = -1; k <= G.nz; k++)
            for (int j = -1; j <= G.ny; j++)
                for (int i = -1; i <= G.nx; i++) {
                    int val;
                    if (vert)  // is_vertical (:1025-1043): outside and the y ghosts NEUM
                        val = (B(i, j, k) > 0 || j < 0 || j > G.ny - 1) ? BC_NEUM : BC_FULL;
                    else if (B(i, j, k) > 0)
                        val = (B(i, j - 1, k) < 0 || B(i, j + 1, k) < 0 || B(i, j, k - 1) < 0 || B(i, j, k + 1) < 0)
                                  ? BC_NEUM : BC_OTHER;
                    else if (i == -1 || i == G.nx)
                        val = BC_NEUM;
                    else
                        val = BC_FULL;
                    bc[L.idx(i, j, k)] = val;
                }
        upload_system(0, bc, box(0));
        // viscosity bc, VISCOSITY_3D::SetupBoundaryConditionsForVelocity,
        // cylindrical (:1422-1539): PER beyond the ends along the pipe
        for (int c = 1; c < 4; c++) {
            std::fill(bc.begin(), bc.end(), 0);
            FsBox b = box(c);
            const std::vector<double>& hbp = hb[c];
            for (int k = -1; k <= b.nk; k++)
                for (int j = -1; j <= b.nj; j++)
                    for (int i = -1; i <= b.ni; i++) {
                        int val;
                        if (hbp[L.idx(i, j, k)] > 0)
                            val = BC_DIR;
                        else if (vert ? j < 0 : i < 0)
                            val = BC_PER;
                        else if (vert ? (c == 2 ? j >= b.nj - 1 : j > b.nj - 1) : (c == 1 ? i >= b.ni - 1 : i > b.ni - 1))
                            val = BC_PER;
                        else
                            val = BC_FULL;
                        bc[L.idx(i, j, k)] = val;
                    }
            upload_system(c, bc, b);
        }
        // ELL storage for the largest system
        const size_t N = (size_t)std::max(Nmax, 1);
        const int nb = (int)((N + RBLK - 1) / RBLK);
        for (PcgSet& q : ps) {
            PCHECK(hipMalloc(&q.val, sizeof(double) * N * PSLOTS));
            PCHECK(hipMalloc(&q.col, sizeof(int) * N * PSLOTS));
            PCHECK(hipMalloc(&q.cnt, sizeof(int) * N));
            for (double** d : {&q.dg, &q.rhs, &q.x, &q.r, &q.z, &q.p, &q.p2, &q.q})
                PCHECK(hipMalloc(d, sizeof(double) * N));
            PCHECK(hipMalloc(&q.part, sizeof(double) * 3 * (nb + 1)));
            PCHECK(hipMalloc(&q.lp, sizeof(double) * 3 * RBLK));  // the loop's workgroup sums: p.q, r.z, r.r
        }
    }

    // the masked y fill's crossings (k_periodic_y_ext) of face component c
    // from its boundary_phi hb (zero ghosts, as on the device)
    void crossing_tables(int c, const std::vector<double>& hb) {
        FsBox b = box(c);
        const int big = 1 << 30;
        std::vector<int> xr((size_t)b.nk * b.nj, big), xl((size_t)b.nk * b.nj, -big), zt((size_t)b.nj * b.ni, big),
            zb((size_t)b.nj * b.ni, -big);
        auto M = [&](int i, int j, int k) { return hb[L.idx(i, j, k)]; };
        for (int k = 0; k < b.nk; k++)
            for (int j = 0; j < b.nj; j++)
                for (int i = 0; i < b.ni; i++) {
                    if (!(M(i, j, k) > 0)) continue;
                    if (M(i - 1, j, k) < 0) xr[k * b.nj + j] = std::min(xr[k * b.nj + j], i);
                    if (M(i + 1, j, k) < 0) xl[k * b.nj + j] = std::max(xl[k * b.nj + j], i);
                    if (M(i, j, k - 1) < 0) zt[j * b.ni + i] = std::min(zt[j * b.ni + i], k);
                    if (M(i, j, k + 1) < 0) zb[j * b.ni + i] = std::max(zb[j * b.ni + i], k);
                }
        const std::vector<int>* tabs[4] = {&xr, &xl, &zt, &zb};
        for (int t = 0; t < 4; t++) {
            PCHECK(hipMalloc(&xz[c][t], sizeof(int) * tabs[t]->size()));
            PCHECK(hipMemcpy(xz[c][t], tabs[t]->data(), sizeof(int) * tabs[t]->size(), hipMemcpyHostToDevice));
        }
    }

    // FillGhostCellsContinuousDerivativesFrom of a cell field onto itself,
    // depth 2, faces then the three edge blocks (FIELD_STRUCTURE_3D.h:
    // 1776-2199), on the host
    std::vector<double> linear_ghosts(const std::vector<double>& in) {
        std::vector<double> a = in;
        const int ni = G.nx, nj = G.ny, nk = G.nz, d = 2;
        auto A_ = [&](int i, int j, int k) -> double& { return a[L.idx(i, j, k)]; };
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++) {
                for (int i = -1; i >= -d; i--) A_(i, j, k) = (double)2 * A_(i + 1, j, k) - A_(i + 2, j, k);
                for (int i = ni; i <= ni - 1 + d; i++) A_(i, j, k) = (double)2 * A_(i - 1, j, k) - A_(i - 2, j, k);
            }
        for (int k = 0; k < nk; k++)
            for (int i = 0; i < ni; i++) {
                for (int j = -1; j >= -d; j--) A_(i, j, k) = (double)2 * A_(i, j + 1, k) - A_(i, j + 2, k);
                for (int j = nj; j <= nj - 1 + d; j++) A_(i, j, k) = (double)2 * A_(i, j - 1, k) - A_(i, j - 2, k);
            }
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                for (int k = -1; k >= -d; k--) A_(i, j, k) = (double)2 * A_(i, j, k + 1) - A_(i, j, k + 2);
                for (int k = nk; k <= nk - 1 + d; k++) A_(i, j, k) = (double)2 * A_(i, j, k - 1) - A_(i, j, k - 2);
            }
        for (int k = 0; k < nk; k++) {
            for (int j = -1; j >= -d; j--)
                for (int i = -1; i >= -d; i--)
                    A_(i, j, k) = (double)0.5 * (((double)2 * A_(i + 1, j, k) - A_(i + 2, j, k)) +
                                                 ((double)2 * A_(i, j + 1, k) - A_(i, j + 2, k)));
            for (int j = nj; j <= nj - 1 + d; j++)
                for (int i = ni; i <= ni - 1 + d; i++)
                    A_(i, j, k) = (double)0.5 * (((double)2 * A_(i - 1, j, k) - A_(i - 2, j, k)) +
                                                 ((double)2 * A_(i, j - 1, k) - A_(i, j - 2, k)));
            for (int j = nj; j <= nj - 1 + d; j++)
                for (int i = -1; i >= -d; i--)
                    A_(i, j, k) = (double)0.5 * (((double)2 * A_(i + 1, j, k) - A_(i + 2, j, k)) +
                                                 ((double)2 * A_(i, j - 1, k) - A_(i, j - 2, k)));
        }
        return a;
    }

    // AssignSequentialindicesToFullCells (POISSON_SOLVER_3D.h:527-545): FULL
    // cells of the interior renumbered in k, j, i order
    void upload_system(int s, std::vector<int>& bc, FsBox b) {
        std::vector<long> rows;
        for (int k = 0; k < b.nk; k++)
            for (int j = 0; j < b.nj; j++)
                for (int i = 0; i < b.ni; i++) {
                    long o = L.idx(i, j, k);
                    if (bc[o] > -1) {
                        bc[o] = (int)rows.size();
                        rows.push_back(o);
                    }
                }
        sys[s].N = (int)rows.size();
        if (sys[s].N == 0) throw FsPipeError{FS_ERR_BAD_ARG, "pipe mode: no cell inside the pipe"};
        Nmax = std::max(Nmax, sys[s].N);
        PCHECK(hipMalloc(&sys[s].bc, sizeof(int) * L.total));
        PCHECK(hipMemcpy(sys[s].bc, bc.data(), sizeof(int) * L.total, hipMemcpyHostToDevice));
        PCHECK(hipMalloc(&sys[s].rows, sizeof(long) * rows.size()));
        PCHECK(hipMemcpy(sys[s].rows, rows.data(), sizeof(long) * rows.size(), hipMemcpyHostToDevice));
    }

    // ---- the scene's initial state --------------------------------------------
    void initial() {
        const long n = L.total;
        std::vector<double> hu(n, 0.0), hphi(n, 0.0);
        const double a = Q.a, K = Q.K, m = Q.m;
        // EULERIAN_FLUID_SOLVER_3D.h:566-611: the vertical pipe's y faces,
        // interior only, with the sign of V0
        if (vert)
            for (int k = 0; k <= G.nz - 1; k++)
                for (int j = 0; j <= G.ny; j++)
                    for (int i = 0; i <= G.nx - 1; i++) {
                        double x = G.x_min + i * mac[2].x.d, z = G.z_min + k * mac[2].z.d;
                        double magnitude = std::sqrt(x * x + z * z), val = 0;
                        if ((magnitude >= 1) && (magnitude <= a)) {
                            if (V0 > 0)
                                val = (double)1 / A * (a * a - (x * x + z * z) - (double)2 * (K - 1) * std::log(std::sqrt(x * x + z * z) / a));
                            else
                                val = -(double)1 / A * (a * a - (x * x + z * z) - (double)2 * (K - 1) * std::log(std::sqrt(x * x + z * z) / a));
                        } else if (magnitude < 1) {
                            if (V0 > 0)
                                val = (double)1 - (double)1 / A * (m * (x * x + z * z) * K);
                            else
                                val = -((double)1 - (double)1 / A * (m * (x * x + z * z) * K));
                        }
                        hu[L.idx(i, j, k)] = val;
                    }
        // EULERIAN_FLUID_SOLVER_3D.h:536-563: the parallel pipe's x faces,
        // ghost layers of width 2
        if (!vert)
        for (int k = -2; k <= G.nz - 1 + 2; k++)
            for (int j = -2; j <= G.ny - 1 + 2; j++)
                for (int i = -2; i <= G.nx + 2; i++) {
                    double y = G.y_min + j * mac[1].y.d, z = G.z_min + k * mac[1].z.d;
                    double magnitude = std::sqrt(y * y + z * z), val;
                    if ((magnitude >= 1) && (magnitude <= a))
                        val = (double)1 / A * (a * a - (y * y + z * z) - (double)2 * (K - 1) * std::log(std::sqrt(y * y + z * z) / a));
                    else if (magnitude < 1)
                        val = (double)1 - (double)1 / A * (m * (y * y + z * z) * K);
                    else
                        val = 0;
                    hu[L.idx(i, j, k)] = val;
                }
        // SIMULATION_WORLD.cpp:290-303: the interface; its ghosts from the pipe
        // level set's values at that moment, all 3 dx (:755-757)
        const double three_dx = cell.x.d * (double)3;
        std::fill(hphi.begin(), hphi.end(), three_dx);
        for (int k = 0; k < G.nz; k++)
            for (int j = 0; j < G.ny; j++)
                for (int i = 0; i < G.nx; i++) {
                    double x = G.x_min + i * cell.x.d, y = G.y_min + j * cell.y.d, z = G.z_min + k * cell.z.d;
                    hphi[L.idx(i, j, k)] = vert ? std::sqrt(x * x + z * z) - (A0 * std::cos(Q.alpha * y) + (double)1)
                                                : std::sqrt(y * y + z * z) - (A0 * std::cos(Q.alpha * x) + (double)1);
                }
        PCHECK(hipMemsetAsync(u, 0, sizeof(double) * n, st));
        PCHECK(hipMemsetAsync(v, 0, sizeof(double) * n, st));
        PCHECK(hipMemcpyAsync(vert ? v : u, hu.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
        PCHECK(hipMemcpyAsync(phi, hphi.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
        PCHECK(hipMemsetAsync(w, 0, sizeof(double) * n, st));
        PCHECK(hipMemsetAsync(p, 0, sizeof(double) * n, st));
        sync();
    }

    // interior-only host arrays in and out
    void put(double* dst, const double* src, int c) {
        FsBox b = box(c);
        std::vector<double> h(L.total, 0.0);
        PCHECK(hipMemcpy(h.data(), dst, sizeof(double) * L.total, hipMemcpyDeviceToHost));
        long q = 0;
        for (int k = 0; k < b.nk; k++)
            for (int j = 0; j < b.nj; j++)
                for (int i = 0; i < b.ni; i++) h[L.idx(i, j, k)] = src[q++];
        PCHECK(hipMemcpy(dst, h.data(), sizeof(double) * L.total, hipMemcpyHostToDevice));
    }
    void take(double* dst, const double* src, int c) {
        FsBox b = box(c);
        std::vector<double> h(L.total);
        PCHECK(hipMemcpy(h.data(), src, sizeof(double) * L.total, hipMemcpyDeviceToHost));
        long q = 0;
        for (int k = 0; k < b.nk; k++)
            for (int j = 0; j < b.nj; j++)
                for (int i = 0; i < b.ni; i++) dst[q++] = h[L.idx(i, j, k)];
    }

    // ---- launch helpers ----------------------------------------------------------
    // the periodic fill along the pipe: x (parallel) or y (vertical, the
    // masked fill then extrapolates beyond the wall)
    void periodic(double* dst, const double* src, int c, int gw, int copy, const double* mask) {
        FsBox b = box(c);
        if (!vert) {
            hipLaunchKernelGGL(k_periodic_x, box_grid(b, gw), dim3(PBX, PBY), 0, st, dst, src, mask, L, b, c, gw,
                               copy);
            return;
        }
        hipLaunchKernelGGL(k_periodic_y, dim3((b.ni + PBX - 1) / PBX, (b.nj + 2 * gw + PBY - 1) / PBY, b.nk),
                           dim3(PBX, PBY), 0, st, dst, src, mask, L, b, c, gw, copy);
        if (mask)
            hipLaunchKernelGGL(k_periodic_y_ext, dim3((b.ni + 2 * gw + PBX - 1) / PBX, (b.nj + PBY - 1) / PBY,
                                                      b.nk + 2 * gw),
                               dim3(PBX, PBY), 0, st, dst, src, xz[c][0], xz[c][1], xz[c][2], xz[c][3], L, b, gw);
    }
    void masked_copies(double* const* dsts3) {  // dsts3[1..3] <- u, v, w masked by their boundary_phi
        for (int c = 1; c < 4; c++) periodic(dsts3[c], field(c), c, 3, 1, bp[c]);
    }

    // the stride of the wall's first axis: y (parallel pipe) or x (vertical)
    long s1() const { return vert ? 1L : (long)L.sx; }

    // ---- phases ------------------------------------------------------------------
    // ADVECTION_3D::Solve_Levelset, oil/water (ADVECTION_3D.h:204-242)
    void LevelsetAdvection(double dt) {
        masked_copies(g);
        if (vert) {
            hipLaunchKernelGGL(k_lsbc_x, box_grid(box(1)), dim3(PBX, PBY), 0, st, u, g[1], bp[1], L, box(1));
        } else {
            PCHECK(hipMemcpyAsync(scr, g[2], sizeof(double) * L.total, hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(k_lsbc_y, box_grid(box(2)), dim3(PBX, PBY), 0, st, v, g[2], scr, bp[2], L, box(2));
        }
        hipLaunchKernelGGL(k_lsbc_z, box_grid(box(3)), dim3(PBX, PBY), 0, st, w, g[3], bp[3], L, box(3), s1());
        periodic(sfg, phi, 0, 3, 1, nullptr);
        hipLaunchKernelGGL(k_weno5_pipe, box_grid(box(0)), dim3(PBX, PBY), 0, st, phi, sfg, g[1], g[2], g[3], bls, L,
                           cell.x.inv, cell.y.inv, cell.z.inv, P.levelset_epsilon, dt);
    }

    // ADVECTION_3D::Solve_Velocity, ENO3 with the pipe (ADVECTION_3D.h:244-357)
    void VelocityAdvection(double dt) {
        masked_copies(g);
        for (int c = 1; c < 4; c++) {
            periodic(ag[c][0], field(c), c, 3, 1, bp[c]);
            periodic(ag[c][1], field(c), c, 3, 1, bp[c]);
        }
        const int order[3] = {1, 3, 2};
        for (int oc = 0; oc < 3; oc++) {
            int c = order[oc];
            hipLaunchKernelGGL(k_velbc_cross, box_grid(box(c)), dim3(PBX, PBY), 0, st, ag[c][0], ag[c][1], g[c], bp[c],
                               L, box(c), s1());
        }
        // derivatives along the pipe from the shared copy, across it from the
        // per-axis copies (ADVECTION_3D.h:316-330)
        for (int c = 1; c < 4; c++)
            hipLaunchKernelGGL(k_eno3_pipe, box_grid(box(c)), dim3(PBX, PBY), 0, st, field(c), vert ? ag[c][0] : g[c],
                               vert ? g[c] : ag[c][0], ag[c][1], g[1], g[2], g[3], bp[c], L, box(c), c, mac[c], dt,
                               vert);
    }

    // EULERIAN_FLUID_SOLVER_3D::ApplyGravity, dimensionless
    void ApplyGravity(double dt) {
        FsGravPipe Gp;
        Gp.gd[0] = dt * P.gravity[0];
        Gp.gd[1] = dt * P.gravity[1];
        Gp.gd[2] = dt * P.gravity[2];
        Gp.drive = dt * f;
        Gp.eta = Q.eta, Gp.ro = Q.oil_density, Gp.R1 = R1, Gp.V0 = V0;
        Gp.eps = (double)1.5 * (vert ? cell.y.d : cell.x.d);
        Gp.vert = vert;
        periodic(sfg, phi, 0, 3, 1, nullptr);
        for (int c = vert ? 2 : 1; c <= (vert ? 2 : 3); c++)
            hipLaunchKernelGGL(k_gravity_pipe, box_grid(box(c)), dim3(PBX, PBY), 0, st, field(c), sfg, bp[c], L, box(c),
                               c, Gp);
    }

    // Adds the timed launches of iteration k to the samples, counting the rows
    // of the systems for which they ran in full: iteration k's launches
    // return at once for a system whose tail of iteration k - 1 stopped it
    void sample(int k, int s0, int ns) {
        double rows = 0;
        for (int i = 0; i < ns; i++) {
            const FsPipePcg& h = host_pcg[i];
            if (h.done == 0 || (h.done == 1 ? k - 1 <= h.iter : k <= h.iter)) rows += ps[s0 + i].N;
        }
        if (rows == 0) return;
        float a = 0, b = 0;
        PCHECK(hipEventElapsedTime(&a, kev[0], kev[1]));
        PCHECK(hipEventElapsedTime(&b, kev[1], kev[2]));
        ks_ms[0] += a, ks_ms[1] += b;
        ks_bytes[0] += 120.0 * rows, ks_bytes[1] += 64.0 * rows;
        ks_n[0]++, ks_n[1]++;
    }

    // The Jacobi-PCG of PCG_METHOD.cpp:4-160 on the assembled ELL systems of
    // sets s0 .. s0 + ns - 1 (each x holds its x0), solved together: every
    // launch of the loop iterates each system still running, each with its
    // own sums and stop test, so each takes the iterates of a solve alone.
    // iters[i] receives set s0 + i's iterations; kind 0 pressure (whose true
    // residual the step reports), 1 viscosity (no report: the next
    // projection's replaces it)
    // (the host checks for the stop between launch batches: the first one
    // runs to a little short of the last solve of the kind, at most 1024
    // iterations, the next ones 8, 16, 32, 64).
    void PcgSolve(int s0, int ns, int* iters, int kind) {
        PcgBatch B{};
        int nbmax = 0;
        for (int i = 0; i < ns; i++) {
            PcgSet& q = ps[s0 + i];
            const int N = q.N, nb = (N + RBLK - 1) / RBLK;
            nbmax = std::max(nbmax, nb);
            FsPipePcg& hp = host_pcg[i];
            hp = FsPipePcg{};
            hp.max_iter = P.max_iteration;
            hp.rel = P.stop_rule == FS_STOP_REL_RESIDUAL;
            hp.tol2 = P.tolerance * P.tolerance;
            hp.rtol = P.rel_tolerance;
            B.val[i] = q.val, B.col[i] = q.col, B.cnt[i] = q.cnt, B.dg[i] = q.dg;
            B.x[i] = q.x, B.r[i] = q.r, B.z[i] = q.z, B.q[i] = q.q, B.lp[i] = q.lp;
            // p_1 = z, which k_pcg_begin writes to q.p
            B.p[i][0] = q.p2, B.p[i][1] = q.p;
            B.S[i] = pcg + i;
            B.N[i] = N, B.nb[i] = nb;
        }
        PCHECK(hipMemcpyAsync(pcg, host_pcg, ns * sizeof(FsPipePcg), hipMemcpyHostToDevice, st));
        for (int i = 0; i < ns; i++) {
            PcgSet& q = ps[s0 + i];
            const int N = q.N, nb = B.nb[i];
            hipLaunchKernelGGL(k_spmv, dim3(nb), dim3(RBLK), 0, st, q.val, q.col, q.cnt, N, q.x, q.q);
            hipLaunchKernelGGL(k_pcg_begin, dim3(nb), dim3(RBLK), 0, st, q.rhs, q.q, q.dg, q.r, q.z, q.p, N, q.part, nb);
            hipLaunchKernelGGL(k_pcg_begin_scalars, dim3(1), dim3(RBLK), 0, st, q.part, nb, pcg + i);
        }
        // one step of a workgroup's rows covers every system (one-step
        // passes) in the narrowest workgroup that does -- RBLK threads per
        // block of a step -- so that three systems' workgroups fit the GPU at
        // once; larger systems take steps of PWG rows
        const int steps = nbmax <= RBLK ? 1 : nbmax <= 2 * RBLK ? 2 : nbmax <= 3 * RBLK ? 3 : 4;
        const bool one = nbmax <= RBLK * PSUB;
        const dim3 grid(RBLK, ns), wg(one ? steps * RBLK : PWG);
        using KA = void (*)(const PcgBatch, int);
        const KA spmv1[4] = {k_pcg_spmv_dir<1, true, RBLK>, k_pcg_spmv_dir<1, true, 2 * RBLK>,
                             k_pcg_spmv_dir<1, true, 3 * RBLK>, k_pcg_spmv_dir<1, true, 4 * RBLK>};
        const KA spmvk[4] = {k_pcg_spmv_dir<1, false, RBLK>, k_pcg_spmv_dir<1, false, 2 * RBLK>,
                             k_pcg_spmv_dir<1, false, 3 * RBLK>, k_pcg_spmv_dir<1, false, 4 * RBLK>};
        const KA upd[4] = {k_pcg_update2<1, RBLK>, k_pcg_update2<1, 2 * RBLK>, k_pcg_update2<1, 3 * RBLK>,
                           k_pcg_update2<1, 4 * RBLK>};
        int batch = std::min(1024, std::max(8, pcg_hint[kind] - 8)), next = 8, k = 1;
        while (true) {
            int ks = 0;  // the iteration whose two launches are timed (0: none)
            for (int j = 0; j < batch; j++, k++) {
                const bool timed = sampling && j == 1;
                if (timed) {
                    ks = k;
                    PCHECK(hipEventRecord(kev[0], st));
                }
                const KA ka = one ? (k == 1 ? spmv1 : spmvk)[steps - 1]
                                  : (k == 1 ? k_pcg_spmv_dir<4, true, PWG> : k_pcg_spmv_dir<4, false, PWG>);
                const KA kb = one ? upd[steps - 1] : k_pcg_update2<4, PWG>;
                hipLaunchKernelGGL(ka, grid, wg, 0, st, B, k);
                if (timed) PCHECK(hipEventRecord(kev[1], st));
                hipLaunchKernelGGL(kb, grid, wg, 0, st, B, k);
                if (timed) PCHECK(hipEventRecord(kev[2], st));
            }
            if (sampling) ks_launch[0] += batch, ks_launch[1] += batch;
            hipLaunchKernelGGL(k_pcg_tail_end, dim3(1, ns), dim3(RBLK), 0, st, B, k - 1);
            PCHECK(hipMemcpyAsync(host_pcg, pcg, ns * sizeof(FsPipePcg), hipMemcpyDeviceToHost, st));
            sync();
            if (ks) sample(ks, s0, ns);
            bool done = true;
            for (int i = 0; i < ns; i++) done = done && host_pcg[i].done;
            if (done) break;
            if (k > P.max_iteration + 3) throw FsPipeError{FS_ERR_HIP, "pipe PCG: iteration count out of step"};
            batch = next;
            next = std::min(64, next * 2);
        }
        check_launch();
        pcg_hint[kind] = 0;
        for (int i = 0; i < ns; i++) {
            const FsPipePcg& hp = host_pcg[i];
            iters[i] = hp.iter;
            pcg_hint[kind] = std::max(pcg_hint[kind], hp.iter);
            if (hp.done == 2) not_converged = true;
        }
        if (kind != 0) return;
        // the true residuals for the report
        for (int i = 0; i < ns; i++) {
            PcgSet& q = ps[s0 + i];
            const int N = q.N, nb = B.nb[i];
            hipLaunchKernelGGL(k_spmv, dim3(nb), dim3(RBLK), 0, st, q.val, q.col, q.cnt, N, q.x, q.q);
            hipLaunchKernelGGL(k_residual, dim3(nb), dim3(RBLK), 0, st, q.rhs, q.q, q.r, N);
            hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(RBLK), 0, st, q.r, q.r, N, q.part);
            hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(RBLK), 0, st, q.part, nb, dscal + i);
        }
        PCHECK(hipMemcpyAsync(host_bits, dscal, ns * sizeof(double), hipMemcpyDeviceToHost, st));
        sync();
        // the report keeps the last system's residuals, as one solve after
        // another would
        for (int i = 0; i < ns; i++) {
            const FsPipePcg& hp = host_pcg[i];
            last_res_abs = std::sqrt(hp.rz);
            const double bnorm = std::sqrt(hp.bb);
            last_res_rel = bnorm > 0 ? std::sqrt(host_bits[i]) / bnorm : 0;
        }
    }

    // VISCOSITY_3D::ApplyViscosity, semi-implicit, dimensionless
    // (VISCOSITY_3D.h:565-852): the boundary conditions (VISCOSITY_3D.h:
    // 2048-2106, 2235-2356) once, then each component's coefficients,
    // explicit term, system and solve
    void Viscosity(double dt) {
        periodic(sfg, phi, 0, 3, 1, nullptr);
        masked_copies(g);
        for (int c = 1; c < 4; c++) {
            periodic(vg[c][0], field(c), c, 3, 1, bp[c]);
            periodic(vg[c][1], field(c), c, 3, 1, bp[c]);
        }
        const int order[3] = {1, 3, 2};
        for (int oc = 0; oc < 3; oc++) {
            int c = order[oc];
            hipLaunchKernelGGL(k_viscbc, box_grid(box(c)), dim3(PBX, PBY), 0, st, vg[c][0], vg[c][1],
                               vert && c == 2 ? v : g[c], bp[c], L, box(c), s1());
        }
        FsViscPipe Vp{dt, Re, Q.eta, Q.m, (double)1.5 * cell.x.d, vert};
        for (int c = 1; c < 4; c++) {
            // cross-component copies each explicit term reads (VISCOSITY_3D.h:
            // 615-631, 754-767, 821-834): parallel pipe, x reads y_x and z_x
            // (never written: zero), y reads x_y and z_y, z reads x_z and y_z;
            // vertical pipe, x reads y_x and z_x, y reads x_x and z_z, z reads
            // x_z and y_z (index 0: the wall's first axis, y or x)
            const double* A_ = c == 1 ? (vert ? vg[2][0] : zero) : c == 2 ? vg[1][0] : vg[1][1];
            const double* B_ = c == 1 ? (vert ? vg[3][0] : zero) : c == 2 ? (vert ? vg[3][1] : vg[3][0]) : vg[2][1];
            hipLaunchKernelGGL(k_visc_coef, box_grid(box(c)), dim3(PBX, PBY), 0, st, coef[0], coef[1], coef[2], coef[3],
                               coef[4], coef[5], expl, sfg, g[c], A_, B_, L, box(c), c, mac[c], Vp);
            FsViscSys S;
            S.inv_dxdx = (double)1 / (mac[c].x.d * mac[c].x.d);
            S.inv_dydy = (double)1 / (mac[c].y.d * mac[c].y.d);
            S.inv_dzdz = (double)1 / (mac[c].z.d * mac[c].z.d);
            S.comp = c;
            S.ie = box(c).ni - 1;
            S.je = box(c).nj - 1;
            const int N = sys[c].N, nb = (N + RBLK - 1) / RBLK;
            // the Dirichlet values along x, y, z: the shared copy along the
            // pipe, the per-axis copies across it (:712-716, 778-782, 844-848)
            const double* dx_ = vert ? vg[c][0] : g[c];
            const double* dy_ = vert ? g[c] : vg[c][0];
            PcgSet& q = ps[c - 1];
            q.N = N;
            hipLaunchKernelGGL(k_visc_assemble, dim3(nb), dim3(RBLK), 0, st, sys[c].rows, N, sys[c].bc, coef[0],
                               coef[1], coef[2], coef[3], coef[4], coef[5], expl, dx_, dy_, vg[c][1], field(c), q.val,
                               q.col, q.cnt, q.dg, q.rhs, q.x, L, S);
        }
        // the components' systems read only the copies made above, and each
        // solve writes its own component: the three are solved together
        int it[3];
        PcgSolve(0, 3, it, 1);
        visc_iters += it[0] + it[1] + it[2];
        for (int c = 1; c < 4; c++) {
            const int N = sys[c].N, nb = (N + RBLK - 1) / RBLK;
            hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(RBLK), 0, st, field(c), ps[c - 1].x, sys[c].rows, N, L, c,
                               vert);
            if (c == (vert ? 2 : 1)) {
                FsBox b = box(c);
                hipLaunchKernelGGL(k_fix_last, dim3(((vert ? b.ni : b.nj) + 63) / 64, b.nk), dim3(64), 0, st, field(c), L,
                                   b, vert);
            }
        }
    }

    // LEVELSET_3D::ComputeCurvatures by level set (LEVELSET_3D.h:963-1082):
    // phi's ghosts by linear extrapolation, kappa and its maximum
    void Curvatures(double* ph) {
        fs_fill_linear(st, ph, ph, L, box(0), 2, 1);
        PCHECK(hipMemsetAsync(bits, 0, 2 * sizeof(unsigned long long), st));
        fs_curvature(st, ph, kappa, jc, (double)1 / We, L, cell, 0, bits);
        PCHECK(hipMemcpyAsync(host_bits, bits, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        sync();
        kmax = host_bits[0];
    }

    // PROJECTION_3D::Solve, oil/water (PROJECTION_3D.h:414-432)
    void Projection(double dt, fs_step_stats* stt) {
        // phi's own x ghosts and first cell from the periodic image (:424-427)
        periodic(phi, phi, 0, 2, 0, nullptr);
        const double rr = P.water_density / Q.oil_density;
        hipLaunchKernelGGL(k_pdensity, box_grid(box(0)), dim3(PBX, PBY), 0, st, phi, rho, irho, L, rr);
        // DetermineDivergence (:697-725)
        masked_copies(g);
        if (vert) {  // x faces, then z (:1566-1627)
            hipLaunchKernelGGL(k_divbc, box_grid(box(1)), dim3(PBX, PBY), 0, st, g[1], bp[1], L, box(1), 1L, 1L);
            hipLaunchKernelGGL(k_divbc, box_grid(box(3)), dim3(PBX, PBY), 0, st, g[3], bp[3], L, box(3), (long)L.sy,
                               1L);
        } else {  // z faces, then y (:1629-1685)
            hipLaunchKernelGGL(k_divbc, box_grid(box(3)), dim3(PBX, PBY), 0, st, g[3], bp[3], L, box(3), (long)L.sy,
                               (long)L.sx);
            hipLaunchKernelGGL(k_divbc, box_grid(box(2)), dim3(PBX, PBY), 0, st, g[2], bp[2], L, box(2), (long)L.sx,
                               (long)L.sx);
        }
        hipLaunchKernelGGL(k_pdiv, box_grid(box(0)), dim3(PBX, PBY), 0, st, g[1], g[2], g[3], bls, div, L, cell, 1 / dt);
        // DetermineJumpConditionField (:746-792): jc = kappa/We, periodic along the pipe
        Curvatures(phi);
        periodic(jc, jc, 0, 1, 1, nullptr);
        // DeterminePressure (:794-834)
        fs_fill_const(st, phi, phi, L, box(0), 2);  // water_levelset.ComputeNormals, LEVELSET_3D.h:931
        periodic(irho, irho, 0, 2, 0, nullptr);
        fs_fill_const(st, bls, bls, L, box(0), 2);  // boundary_levelset.ComputeNormals
        FsPresSys S;
        S.inv_dxdx = (double)1 / (cell.x.d * cell.x.d);
        S.inv_dydy = (double)1 / (cell.y.d * cell.y.d);
        S.inv_dzdz = (double)1 / (cell.z.d * cell.z.d);
        S.odx = cell.x.inv, S.ody = cell.y.inv, S.odz = cell.z.inv;
        S.odx2 = cell.x.inv_d2, S.ody2 = cell.y.inv_d2, S.odz2 = cell.z.inv_d2;
        const int N = sys[0].N, nb = (N + RBLK - 1) / RBLK;
        PcgSet& q = ps[0];
        q.N = N;
        hipLaunchKernelGGL(k_pres_assemble, dim3(nb), dim3(RBLK), 0, st, sys[0].rows, N, sys[0].bc, phi, irho, jc, div,
                           p, q.val, q.col, q.cnt, q.dg, q.rhs, q.x, L, S);
        // Neumann compatibility, POISSON_SOLVER_3D.h:221-246
        hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(RBLK), 0, st, q.rhs, nullptr, N, q.part);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(RBLK), 0, st, q.part, nb, dscal);
        hipLaunchKernelGGL(k_remove_mean, dim3(nb), dim3(RBLK), 0, st, q.rhs, N, dscal);
        PcgSolve(0, 1, &last_iters, 0);
        hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(RBLK), 0, st, p, q.x, sys[0].rows, N, L, 0, vert);
        // UpdateVelocity (:1101-1449)
        periodic(psfg, phi, 0, 2, 1, nullptr);
        PCHECK(hipMemsetAsync(bits + 2, 0, 3 * sizeof(unsigned long long), st));
        for (int c = 1; c < 4; c++) {
            FsPresUpd U;
            U.dt = dt;
            U.od = c == 1 ? mac[1].x.inv : c == 2 ? mac[2].y.inv : mac[3].z.inv;
            U.eps = (double)1.5 * (c == 1 ? cell.x.d : c == 2 ? cell.y.d : cell.z.d);
            U.rr = rr;
            hipLaunchKernelGGL(k_pupdate, box_grid(box(c)), dim3(PBX, PBY), 0, st, field(c), psfg, phi, bls, jc, rho, p,
                               sys[0].bc, L, box(c), c, U, bits + 1 + c, vert);
            hipLaunchKernelGGL(k_pfix, dim3(((vert ? box(c).ni : box(c).nj) + 63) / 64, box(c).nk), dim3(64), 0, st,
                               field(c), L, box(c), c, vert);
        }
        PCHECK(hipMemcpyAsync(host_bits + 2, bits + 2, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        sync();
        if (stt) {
            stt->pcg_iterations += last_iters;
            stt->residual_precond = last_res_abs;
            stt->residual_rel = last_res_rel;
            for (int c = 0; c < 3; c++) stt->max_velocity[c] = host_bits[2 + c];
        }
    }

    // EULERIAN_FLUID_SOLVER_3D::Reinitialization with ReinitializationBySussman,
    // oil/water (ADVECTION_3D.h:360-378): each sweep copies phi into the
    // shared scalar ghost copy, periodic along the pipe, and takes one WENO5 Sussman step
    void Reinitialization() {
        if (!P.sussmanstyle_reinitialization) return;
        const double fictious_dt = cell.x.d / P.scaling_number_for_reinitialization;
        fs_sign(st, phi, sgn, L, cell.x.d * cell.x.d);
        for (int it = 0; it < P.iteration_number_for_reinitialization; it++) {
            periodic(sfg, phi, 0, 3, 1, nullptr);
            fs_reinit_sweep(st, phi, sfg, sgn, L, cell, P.levelset_epsilon, fictious_dt, 0, false);
        }
    }

    // SolveForRK, EULERIAN_FLUID_SOLVER_3D.h:1453-1476
    void SolveForRK(double dt, fs_step_stats* stt) {
        if (P.is_velocity_advection_active) VelocityAdvection(dt);
        if (P.is_external_force_active) ApplyGravity(dt);
        if (P.is_sourcing_active) {
            visc_iters = 0;
            Viscosity(dt);
            if (stt) stt->viscosity_iterations += visc_iters;
        }
        if (P.is_projection_active) Projection(dt, stt);
    }

    // AdvanceOneTimeStepThread, orders 1-3 (EULERIAN_FLUID_SOLVER_3D.h:803-1111)
    void Step(double dt, fs_step_stats* stt) {
        const int order = P.order_for_time_advancing;
        if (order == 1 || order == 2) {
            if (P.is_levelset_advection_active) {
                fs_copy_box(st, tmp, phi, L, box(0));
                LevelsetAdvection(dt);
                LevelsetAdvection(dt);
                fs_average(st, phi, tmp, L);
            }
            if (order == 1)
                SolveForRK(dt, stt);
            else {
                for (int c = 1; c < 4; c++) fs_copy_box(st, rk[c], field(c), L, box(c));
                SolveForRK(dt, stt);
                SolveForRK(dt, stt);
                for (int c = 1; c < 4; c++) fs_combine(st, field(c), rk[c], (double)0.5, (double)0.5, L, box(c));
            }
            fs_copy_box(st, tmp, phi, L, box(0));
            Reinitialization();
            Reinitialization();
            fs_average(st, phi, tmp, L);
        } else {
            if (P.is_levelset_advection_active) LevelsetAdvection(dt);
            for (int c = 1; c < 4; c++) fs_copy_box(st, rk[c], field(c), L, box(c));
            SolveForRK(dt, stt);
            SolveForRK(dt, stt);
            for (int c = 1; c < 4; c++) fs_combine(st, field(c), rk[c], (double)0.75, (double)0.25, L, box(c));
            SolveForRK(dt, stt);
            for (int c = 1; c < 4; c++)
                fs_combine(st, field(c), rk[c], (double)1 / (double)3, (double)2 / (double)3, L, box(c));
            fs_copy_box(st, tmp, phi, L, box(0));
            Reinitialization();
            Reinitialization();
            fs_combine(st, phi, tmp, (double)0.75, (double)0.25, L, box(0));
            Reinitialization();
            fs_combine(st, phi, tmp, (double)1 / (double)3, (double)2 / (double)3, L, box(0));
        }
    }

    // CFLOneTimeStep, oil/water, dimensionless, semi-implicit
    // (EULERIAN_FLUID_SOLVER_3D.h:2673-2785), the curvature maximum of the
    // current level set (the frame's ComputeCurvatures, SIMULATION_WORLD.h:324)
    double Cfl(double cfl) {
        PCHECK(hipMemsetAsync(bits + 2, 0, 3 * sizeof(unsigned long long), st));
        for (int c = 1; c < 4; c++)
            hipLaunchKernelGGL(k_absmax, box_grid(box(c)), dim3(PBX, PBY), 0, st, field(c), L, box(c), bits + 1 + c);
        PCHECK(hipMemcpyAsync(scr, phi, sizeof(double) * L.total, hipMemcpyDeviceToDevice, st));
        // the curvature pass writes kappa and jc: keep the projection's copies
        PCHECK(hipMemcpyAsync(tmp, jc, sizeof(double) * L.total, hipMemcpyDeviceToDevice, st));
        Curvatures(scr);
        PCHECK(hipMemcpyAsync(jc, tmp, sizeof(double) * L.total, hipMemcpyDeviceToDevice, st));
        PCHECK(hipMemcpyAsync(host_bits + 2, bits + 2, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        sync();
        const double c_f = host_bits[2] * mac[1].x.inv + host_bits[3] * mac[2].y.inv + host_bits[4] * mac[3].z.inv;
        double g_f = 0;
        if (P.is_external_force_active) {
            double mn = std::min(Q.oil_density, P.water_density);
            g_f = std::sqrt(R1 / (V0 * V0) * (std::fabs(P.gravity[1]) + std::fabs((double)1 / mn * f)) * mac[2].y.inv);
        }
        const double rr = P.water_density / Q.oil_density;
        const double mind = std::min(std::min(mac[1].x.d, mac[2].y.d), mac[3].z.d);
        const double s_f = std::sqrt(((double)1 / We * kmax) / ((rr < 1 ? rr : 1) * (mind * mind)));
        const double v_f = 0;
        const double rate =
            (double)0.5 * (c_f + v_f + std::sqrt((c_f + v_f) * (c_f + v_f) + (double)4 * (g_f * g_f) + (double)4 * (s_f * s_f)));
        return cfl * ((double)1 / rate);
    }
};

// ---- entry points used by the C-ABI ------------------------------------------------

FsPipe* fs_pipe_new(const fs_params& P, const fs_pipe_params& Q, const fs_grid& g, int device) {
    FsPipe* s = new FsPipe();
    try {
        s->init(P, Q, g, device);
    } catch (...) {
        delete s;
        throw;
    }
    return s;
}

void fs_pipe_delete(FsPipe* s) { delete s; }

void fs_pipe_initial(FsPipe* s) { s->initial(); }

void fs_pipe_consts(FsPipe* s, double* out) {
    out[0] = s->R1, out[1] = s->f, out[2] = s->A, out[3] = s->V0, out[4] = s->J, out[5] = s->Re, out[6] = s->We;
}

void fs_pipe_set(FsPipe* s, const double* phi, const double* u, const double* v, const double* w, const double* p) {
    s->sync();
    if (phi) s->put(s->phi, phi, 0);
    if (u) s->put(s->u, u, 1);
    if (v) s->put(s->v, v, 2);
    if (w) s->put(s->w, w, 3);
    if (p) s->put(s->p, p, 0);
}

void fs_pipe_get(FsPipe* s, double* phi, double* u, double* v, double* w, double* p) {
    s->sync();
    if (phi) s->take(phi, s->phi, 0);
    if (u) s->take(u, s->u, 1);
    if (v) s->take(v, s->v, 2);
    if (w) s->take(w, s->w, 3);
    if (p) s->take(p, s->p, 0);
}

int fs_pipe_advance(FsPipe* s, double dt, fs_step_stats* stt, bool frame) {
    if (!(dt > 0) || !std::isfinite(dt)) throw FsPipeError{FS_ERR_BAD_ARG, "dt must be positive and finite"};
    fs_step_stats local{};
    hipEvent_t e0, e1;
    PCHECK(hipEventCreate(&e0));
    PCHECK(hipEventCreate(&e1));
    PCHECK(hipEventRecord(e0, s->st));
    s->not_converged = false;
    if (frame) s->Curvatures(s->phi);
    s->Step(dt, &local);
    PCHECK(hipEventRecord(e1, s->st));
    s->check_launch();
    float ms = 0;
    if (!s->async) {
        s->sync();
        PCHECK(hipEventElapsedTime(&ms, e0, e1));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    local.step_ms = ms;
    if (stt) *stt = local;
    return s->not_converged ? FS_ERR_NOT_CONVERGED : FS_OK;
}

int fs_pipe_run_phase(FsPipe* s, int phase, double dt, fs_step_stats* stt) {
    fs_step_stats local{};
    s->not_converged = false;
    switch (phase) {
        case FS_PIPE_LEVELSET_ADVECTION: s->LevelsetAdvection(dt); break;
        case FS_PIPE_VELOCITY_ADVECTION: s->VelocityAdvection(dt); break;
        case FS_PIPE_GRAVITY: s->ApplyGravity(dt); break;
        case FS_PIPE_SOURCING:
            s->visc_iters = 0;
            s->Viscosity(dt);
            local.viscosity_iterations = s->visc_iters;
            break;
        case FS_PIPE_PROJECTION: s->Projection(dt, &local); break;
        case FS_PIPE_REINITIALIZATION: s->Reinitialization(); break;
        default: throw FsPipeError{FS_ERR_BAD_ARG, "unknown phase"};
    }
    s->check_launch();
    s->sync();
    if (stt) *stt = local;
    return s->not_converged ? FS_ERR_NOT_CONVERGED : FS_OK;
}

double fs_pipe_cfl(FsPipe* s, double cfl) { return s->Cfl(cfl); }

void fs_pipe_array(FsPipe* s, int which, double* dst) {
    double* src = nullptr;
    int c = 0;
    switch (which) {
        case FS_PIPE_ARR_PHI: src = s->phi; break;
        case FS_PIPE_ARR_SCALAR_GHOST: src = s->sfg; break;
        case FS_PIPE_ARR_GHOST_U: src = s->g[1], c = 1; break;
        case FS_PIPE_ARR_GHOST_V: src = s->g[2], c = 2; break;
        case FS_PIPE_ARR_GHOST_W: src = s->g[3], c = 3; break;
        case FS_PIPE_ARR_ADV_U_Y: src = s->ag[1][0], c = 1; break;
        case FS_PIPE_ARR_ADV_U_Z: src = s->ag[1][1], c = 1; break;
        case FS_PIPE_ARR_ADV_V_Y: src = s->ag[2][0], c = 2; break;
        case FS_PIPE_ARR_ADV_V_Z: src = s->ag[2][1], c = 2; break;
        case FS_PIPE_ARR_ADV_W_Y: src = s->ag[3][0], c = 3; break;
        case FS_PIPE_ARR_ADV_W_Z: src = s->ag[3][1], c = 3; break;
        case FS_PIPE_ARR_VIS_U_Y: src = s->vg[1][0], c = 1; break;
        case FS_PIPE_ARR_VIS_U_Z: src = s->vg[1][1], c = 1; break;
        case FS_PIPE_ARR_VIS_V_Y: src = s->vg[2][0], c = 2; break;
        case FS_PIPE_ARR_VIS_V_Z: src = s->vg[2][1], c = 2; break;
        case FS_PIPE_ARR_VIS_W_Y: src = s->vg[3][0], c = 3; break;
        case FS_PIPE_ARR_VIS_W_Z: src = s->vg[3][1], c = 3; break;
        case FS_PIPE_ARR_BOUNDARY: src = s->bls; break;
        case FS_PIPE_ARR_DIVERGENCE: src = s->div; break;
        case FS_PIPE_ARR_JUMP: src = s->jc; break;
        case FS_PIPE_ARR_PRESSURE: src = s->p; break;
        default: throw FsPipeError{FS_ERR_BAD_ARG, "unknown array"};
    }
    s->sync();
    std::vector<double> h(s->L.total);
    PCHECK(hipMemcpy(h.data(), src, sizeof(double) * s->L.total, hipMemcpyDeviceToHost));
    FsBox b = s->box(c);
    long q = 0;
    for (int k = -FS_G; k < b.nk + FS_G; k++)
        for (int j = -FS_G; j < b.nj + FS_G; j++)
            for (int i = -FS_G; i < b.ni + FS_G; i++) dst[q++] = h[s->L.idx(i, j, k)];
}

void fs_pipe_async(FsPipe* s, bool async) { s->async = async; }
void fs_pipe_sync(FsPipe* s) {
    s->sync();
    s->check_launch();
}
hipStream_t fs_pipe_stream(FsPipe* s) { return s->st; }
void fs_pipe_sampling(FsPipe* s, bool on) {
    s->sync();
    for (hipEvent_t& e : s->kev)
        if (!e) PCHECK(hipEventCreate(&e));
    s->sampling = on;
    for (int i = 0; i < 2; i++) s->ks_ms[i] = s->ks_bytes[i] = 0, s->ks_n[i] = s->ks_launch[i] = 0;
}
void fs_pipe_samples(FsPipe* s, int kernel, double* avg_ms, double* avg_bytes, long long* samples,
                     long long* launches) {
    if (kernel < 0 || kernel > 1) throw FsPipeError{FS_ERR_BAD_ARG, "pipe kernel samples: kernel must be 0 or 1"};
    const long long n = s->ks_n[kernel];
    *avg_ms = n ? s->ks_ms[kernel] / n : 0;
    *avg_bytes = n ? s->ks_bytes[kernel] / n : 0;
    *samples = n;
    *launches = s->ks_launch[kernel];
}
