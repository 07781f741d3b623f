// fs_layout.h -- device memory layout shared by every kernel.
//
// Every field (cell-centred or MAC face) lives in one padded box of ghost
// width G = 3 around the largest extent (nx+1, ny+1, nz+1), x fastest.  A
// cell (i,j,k) and the faces (i,j,k) of u, v, w share one linear offset, so
// a kernel computes one index and steps +-1, +-sx, +-sy to its neighbours.
// The row pitch sx is a multiple of 16 elements and the base offset puts
// (0, j, k) on a 128-byte boundary for FP64 (64 bytes for FP32).
#pragma once
#include <cstdint>

#define FS_G 3

struct FsLayout {
    int nx, ny, nz;   // interior cells
    long sx;          // x pitch (elements)
    long sy;          // plane pitch (elements) = sx * (ny + 1 + 2G)
    long off;         // offset of (-G, -G, -G)
    long total;       // elements per allocated field
    __host__ __device__ inline long idx(int i, int j, int k) const {
        return off + (long)(k + FS_G) * sy + (long)(j + FS_G) * sx + (i + FS_G);
    }
};

inline FsLayout fs_make_layout(int nx, int ny, int nz) {
    FsLayout L;
    L.nx = nx;
    L.ny = ny;
    L.nz = nz;
    long w = nx + 1 + 2 * FS_G;
    L.sx = (w + 15) / 16 * 16;
    L.sy = L.sx * (ny + 1 + 2 * FS_G);
    L.off = 16 - FS_G;  // (0,0,0) lands on a multiple of 16 elements
    L.total = L.off + L.sy * (long)(nz + 1 + 2 * FS_G) + 16;
    return L;
}

// Grid constants of one axis, computed as GRID_STRUCTURE_3D::Initialize
// does (GRID_STRUCTURE_3D.h:114-133).
struct FsAxis {
    double d, d2, inv, inv2d, inv_d2;
};
struct FsGridConst {
    FsAxis x, y, z;
};

inline FsAxis fs_make_axis(double mn, double mx, int res) {
    FsAxis a;
    a.d = (mx - mn) / (res - 1);
    a.d2 = a.d * a.d;
    a.inv = (double)1 / a.d;
    a.inv2d = (double)0.5 / a.d;
    a.inv_d2 = (double)1 / a.d2;
    return a;
}
