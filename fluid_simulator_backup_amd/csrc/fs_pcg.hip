// fs_pcg.hip -- matrix-free Jacobi-PCG kernels (FP64 and FP32 vectors).
//
// Layout: the padded grid of fs_layout.h; every vector keeps zero ghosts so
// the 7-point stencil needs no bounds tests.  Work unit: a tile of (64*V) x 4
// cells in x-y marching zc planes in z (Tiling::zc).  Each thread owns V
// consecutive x-cells (V = 2 in FP64, 4 in FP32: 16-byte loads and stores),
// keeps p(k-1), p(k), p(k+1) and beta_z(k), beta_z(k+1) in registers, and
// reads its y- and x-neighbours from the plane's tile rows staged in LDS (on
// large grids; small ones take them from L1/L2 with their own loads).
//
// One iteration is two kernels: K1 runs the loop tail of the previous
// iteration (stop test, beta), forms its direction p = p_old*beta + s on the
// fly for every cell the stencil touches, stores it, and computes q = A p
// with the partial p.q; K2 updates x, r, s = r/D with the partials of r.s and
// r.r.  The two directions alternate between two buffers.
// A fixed grid of NB workgroups walks the tiles; workgroup b lives on XCD
// b % 8 and takes tiles from that XCD's contiguous share, so y-neighbour rows
// are reused in one L2.
//
// Operator row (CSR_MATRIX::Multiply, CSR_MATRIX.h:236-260, on the rows of
// POISSON_SOLVER_3D.h:1088-1213): entries in the order l, r, b, t, d, u,
// diagonal; off-diagonals -beta*inv_d2; A(c,c) = inv_dx2 * sum of the six
// betas of full neighbours (+ the Neumann pin on cell (0,0,0)).  Faces to
// non-full neighbours carry beta = 0, which adds an exact -0 term.  The
// per-cell operation order does not depend on V.
//
// Reductions use no device-scope fence or atomic: each workgroup stores its
// partial sum (fixed wave/LDS tree), and every workgroup of the consuming
// kernel re-reduces the NB partials in the same fixed order, so all of them
// see the same scalar and results do not depend on scheduling.  Dot products
// accumulate in FP64 for both precisions.
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "fs_device.h"
#include "fs_pcg.h"

#define TBX 64
#define TBY 4
#define NT (TBX * TBY)
// vector width of the fused SpMV (at most the update kernels' width)
#ifndef FS_SPMV_V32
#define FS_SPMV_V32 4
#endif
#ifndef FS_SPMV_V64
#define FS_SPMV_V64 2
#endif

// partial-sum slots inside FsPcgWork::partials
#define SLOT_PQ 0
#define SLOT_RS 1
#define SLOT_RR 2
#define SLOT_BB 3
#define SLOT_SUM 4

// ---- vectors ----------------------------------------------------------------
template <typename T, int V>
struct VecT;
template <>
struct VecT<double, 1> {
    typedef double type;
};
template <>
struct VecT<double, 2> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct VecT<float, 1> {
    typedef float type;
};
template <>
struct VecT<float, 2> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct VecT<float, 4> {
    typedef float type __attribute__((ext_vector_type(4)));
};

template <typename T, int V>
__device__ __forceinline__ void vld(const T* __restrict__ p, T (&o)[V]) {
    typedef typename VecT<T, V>::type vt;
    vt t = *(const vt*)p;
    memcpy(o, &t, sizeof(o));
}
template <typename T, int V>
__device__ __forceinline__ void vst(T* __restrict__ p, const T (&o)[V]) {
    typedef typename VecT<T, V>::type vt;
    vt t;
    memcpy(&t, o, sizeof(o));
    *(vt*)p = t;
}
// Streaming (nontemporal) forms, or the plain ones when NTF is 0.
template <int NTF, typename T, int V>
__device__ __forceinline__ void vld_s(const T* __restrict__ p, T (&o)[V]) {
    if (NTF) {
        typedef typename VecT<T, V>::type vt;
        vt t = __builtin_nontemporal_load((const vt*)p);
        memcpy(o, &t, sizeof(o));
    } else {
        vld<T, V>(p, o);
    }
}
template <int NTF, typename T, int V>
__device__ __forceinline__ void vst_s(T* __restrict__ p, const T (&o)[V]) {
    if (NTF) {
        typedef typename VecT<T, V>::type vt;
        vt t;
        memcpy(&t, o, sizeof(o));
        __builtin_nontemporal_store(t, (vt*)p);
    } else {
        vst<T, V>(p, o);
    }
}
// Which PCG vector accesses stream past the caches (bit set: nontemporal):
// K2 loads x 0, p 1, r 2, q 3, D 4; K2 stores x 5, r 6, s 7; K1 stores q 8,
// p 9.  The default streams every vector that the next kernel does not
// stencil over and keeps D and s (and the face coefficients, read with plain
// loads) cacheable: at 256^3 in FP32 the three coefficient fields then stay
// in the 256 MB last-level cache across iterations (measured: K1 136 -> 111
// us, K2 107 -> 91 us; loads of the direction in K1 streamed were slower,
// its halo rows are re-read by neighbouring tiles).
#ifndef FS_NT_MASK
#define FS_NT_MASK 0x36F
#endif
#define NTB(b) ((FS_NT_MASK >> (b)) & 1)

// ---- tiling -------------------------------------------------------------------
#ifndef FS_MIN_TILES
#define FS_MIN_TILES 512
#endif
// The fused SpMV stages its operand's tile rows in LDS on grids of at least
// FS_LDS_MIN_CELLS cells (smaller grids are latency-bound and pay for the
// barriers), with tiles of its own (at least FS_K1_MIN_TILES of them; its
// partial count then differs from the update's, and each consumer is told
// its producer's count).
// (Measured: at 256^3 K1 110 -> 96 us in FP32 and 226 -> 210 us in FP64 with
// LDS and 1024 tiles; neutral at 128^3, slower at 64^3.)
#ifndef FS_LDS_MIN_CELLS
#define FS_LDS_MIN_CELLS (128L * 128 * 128)
#endif
#ifndef FS_K1_MIN_TILES
#define FS_K1_MIN_TILES 1024
#endif
struct Tiling {
    int zc;  // z-planes per tile
    int ntx, nty, ntz, ntiles, nb, v;
};

// zc > 0 forces the z-planes per tile and nbmax > 0 caps the workgroups (the
// tuned tiling); otherwise the fixed rule below.
static Tiling make_tiling(const FsLayout& L, int v, int min_tiles = FS_MIN_TILES, int zc = 0, int nbmax = 0) {
    Tiling t;
    t.v = v;
    t.ntx = (L.nx + TBX * v - 1) / (TBX * v);
    t.nty = (L.ny + TBY - 1) / TBY;
    if (zc > 0) {
        t.zc = zc;
    } else {
        // the deepest z-chunk that still leaves at least min_tiles tiles: a
        // longer z-march re-reads fewer halo planes (measured at 256^3: 32
        // planes beat 16 and 8 for the plain kernels), enough tiles keep
        // every CU busy
        for (t.zc = 32; t.zc > 4; t.zc >>= 1)
            if ((long)t.ntx * t.nty * ((L.nz + t.zc - 1) / t.zc) >= min_tiles) break;
    }
    t.ntz = (L.nz + t.zc - 1) / t.zc;
    t.ntiles = t.ntx * t.nty * t.ntz;
    const int cap = nbmax > 0 && nbmax < FS_PCG_MAX_BLOCKS ? nbmax : FS_PCG_MAX_BLOCKS;
    int nb = t.ntiles < cap ? t.ntiles : cap;
    if (nb >= 8) nb = nb / 8 * 8;
    t.nb = nb;
    return t;
}

// The tiling of iteration kernel k (0 the fused SpMV, 1 the update and the
// kernels that write partials the SpMV reads) of workspace w.
static Tiling work_tiling(const FsLayout& L, int v, const FsPcgWork& w, int k, int min_tiles = FS_MIN_TILES) {
    return make_tiling(L, v, min_tiles, w.zc[k], w.nbmax[k]);
}

// Vector width for a precision: the widest that divides nx.
static int vec_width(const FsLayout& L, int is_float) {
    if (is_float) return L.nx % 4 == 0 ? 4 : (L.nx % 2 == 0 ? 2 : 1);
    return L.nx % 2 == 0 ? 2 : 1;
}

// Tiles of workgroup b: XCD x = b % 8 owns tiles [x*T/8, (x+1)*T/8); its
// workgroups stride through that range.
struct TileIter {
    int t, hi, step;
    __device__ TileIter(const Tiling& tl) {
        int b = blockIdx.x;
        if (tl.nb >= 8) {
            int x = b & 7, slot = b >> 3, nslot = tl.nb >> 3;
            int lo = (int)((long)x * tl.ntiles / 8);
            hi = (int)((long)(x + 1) * tl.ntiles / 8);
            t = lo + slot;
            step = nslot;
        } else {
            t = b;
            hi = tl.ntiles;
            step = tl.nb;
        }
    }
    __device__ bool valid() const { return t < hi; }
    __device__ void next() { t += step; }
};

// First cell (i0, j, k0) of this thread in tile t.
template <int V>
__device__ __forceinline__ void tile_origin(const Tiling& tl, int t, int& i0, int& j, int& k0) {
    int tx = t % tl.ntx, r = t / tl.ntx;
    int ty = r % tl.nty, tz = r / tl.nty;
    i0 = (tx * TBX + threadIdx.x) * V;
    j = ty * TBY + threadIdx.y;
    k0 = tz * tl.zc;
}

// ---- reductions ---------------------------------------------------------------
// Sum over the workgroup, returned to every thread; fixed order.
__device__ __forceinline__ double wg_sum_all(double v, double* lds) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    int tid = threadIdx.y * TBX + threadIdx.x;
    __syncthreads();
    if ((tid & 63) == 0) lds[tid >> 6] = v;
    __syncthreads();
    double t = ((lds[0] + lds[1]) + lds[2]) + lds[3];
    __syncthreads();
    return t;
}

// Store this workgroup's partial of a slot (plain store; consumers run in a
// later kernel).
__device__ __forceinline__ void put_partial(double* part, int slot, double v, double* lds) {
    double t = wg_sum_all(v, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[slot * FS_PCG_MAX_BLOCKS + blockIdx.x] = t;
}

// Total of a slot over the nb workgroups of the producing kernel.
__device__ __forceinline__ double get_total(const double* part, int slot, int nb, double* lds) {
    int tid = threadIdx.y * TBX + threadIdx.x;
    double a = 0;
    for (int b = tid; b < nb; b += NT) a += part[slot * FS_PCG_MAX_BLOCKS + b];
    return wg_sum_all(a, lds);
}

// ---- operator rows ----------------------------------------------------------
template <typename T>
struct OpT {
    const T* __restrict__ bx;
    const T* __restrict__ by;
    const T* __restrict__ bz;
    T idx2, idy2, idz2;
    bool cubic;
    bool axis;  // per-axis diagonal on non-cubic cells
};

template <typename T>
__device__ __forceinline__ OpT<T> make_op(const FsPcgOp& o) {
    OpT<T> r;
    r.bx = (const T*)o.bx;
    r.by = (const T*)o.by;
    r.bz = (const T*)o.bz;
    r.idx2 = (T)o.idx2;
    r.idy2 = (T)o.idy2;
    r.idz2 = (T)o.idz2;
    r.cubic = o.idx2 == o.idy2 && o.idx2 == o.idz2;
    r.axis = o.axis_diag != 0 && !r.cubic;
    return r;
}

// The neighbourhood of one cell.
template <typename T>
struct Nb {
    T xc, xl, xr, xb, xt, xd, xu;
    T bl, br, bb, bt, bd, bu;
};

// Diagonal with each face's beta over its own spacing squared: the betas of
// x-, y- and z-faces summed per axis (the pin's extras on their own axes,
// y for beta_b and z for beta_d), then idx2*cx + idy2*cy + idz2*cz; 1/dx^2
// for a row without faces, as the reference's coef == 0 -> 1.
__host__ __device__ __forceinline__ double axis_diag(double idx2, double idy2, double idz2, double bl, double br,
                                                     double bb, double bt, double bd, double bu, bool pin,
                                                     double pin_b, double pin_d) {
    double cx = 0, cy = 0, cz = 0;
    cx += bl;
    cx += br;
    cy += bb;
    cy += bt;
    cz += bd;
    cz += bu;
    if (pin) {
        cy += pin_b;
        cz += pin_d;
    }
    if (cx + cy + cz == 0) return idx2;
    return idx2 * cx + idy2 * cy + idz2 * cz;
}

// FP64 row in the CSR order of the reference.
__device__ __forceinline__ double row_val(const OpT<double>& op, const Nb<double>& n, bool pin, double pin_b,
                                          double pin_d) {
    double coef = 0;
    coef += n.bl;
    coef += n.br;
    coef += n.bb;
    coef += n.bt;
    coef += n.bd;
    coef += n.bu;
    if (pin) {
        coef += pin_b;
        coef += pin_d;
    }
    if (coef == 0) coef = 1;
    double v = 0;
    v += -n.bl * op.idx2 * n.xl;
    v += -n.br * op.idx2 * n.xr;
    v += -n.bb * op.idy2 * n.xb;
    v += -n.bt * op.idy2 * n.xt;
    v += -n.bd * op.idz2 * n.xd;
    v += -n.bu * op.idz2 * n.xu;
    if (op.axis) v += axis_diag(op.idx2, op.idy2, op.idz2, n.bl, n.br, n.bb, n.bt, n.bd, n.bu, pin, pin_b, pin_d) * n.xc;
    else v += op.idx2 * coef * n.xc;
    return v;
}

// FP32 inner operator in flux form: sum over faces of beta*(x_c - x_n)/d^2
// (+ the pin on cell (0,0,0)).  Constants are then an exact null vector of
// every non-pinned row in any precision, as they are for the FP64 matrix up
// to its own rounding; the diagonal-sum form rounded to FP32 would shift the
// near-null eigenvalue of this singly-pinned Neumann system by O(eps32) per
// row and stall the refinement.  Non-cubic cells keep the FP64 row's
// scaling (1/dx^2 on the diagonal for every face).
__device__ __forceinline__ float row_val(const OpT<float>& op, const Nb<float>& n, bool pin, float pin_b,
                                         float pin_d) {
    float v;
    if (op.cubic) {
        float s = op.idx2;
        v = n.bl * (s * (n.xc - n.xl)) + n.br * (s * (n.xc - n.xr)) + n.bb * (s * (n.xc - n.xb)) +
            n.bt * (s * (n.xc - n.xt)) + n.bd * (s * (n.xc - n.xd)) + n.bu * (s * (n.xc - n.xu));
    } else if (op.axis) {
        v = n.bl * (op.idx2 * (n.xc - n.xl)) + n.br * (op.idx2 * (n.xc - n.xr)) + n.bb * (op.idy2 * (n.xc - n.xb)) +
            n.bt * (op.idy2 * (n.xc - n.xt)) + n.bd * (op.idz2 * (n.xc - n.xd)) + n.bu * (op.idz2 * (n.xc - n.xu));
        if (pin) v += pin_b * (op.idy2 * n.xc) + pin_d * (op.idz2 * n.xc);
        return v;
    } else {
        float cx = op.idx2 * n.xc;
        v = n.bl * (cx - op.idx2 * n.xl) + n.br * (cx - op.idx2 * n.xr) + n.bb * (cx - op.idy2 * n.xb) +
            n.bt * (cx - op.idy2 * n.xt) + n.bd * (cx - op.idz2 * n.xd) + n.bu * (cx - op.idz2 * n.xu);
    }
    if (pin) v += (pin_b + pin_d) * (op.idx2 * n.xc);
    return v;
}

// Operand loaders for march_apply: the vector x itself, or the direction
// p = p_old*beta + s formed on the fly (the p-update of PCG_METHOD.cpp,
// p *= beta; p += s, with the same two roundings).
template <typename T, int V>
struct LdPlain {
    const T* __restrict__ x;
    __device__ __forceinline__ void vec(long c, T (&out)[V]) const { vld<T, V>(x + c, out); }
    __device__ __forceinline__ T at(long c) const { return x[c]; }
};
template <typename T, int V>
struct LdDir {
    const T* __restrict__ po;
    const T* __restrict__ s;
    T beta;
    __device__ __forceinline__ void vec(long c, T (&out)[V]) const {
        T sv[V];
        vld<T, V>(po + c, out);
        vld<T, V>(s + c, sv);
#pragma unroll
        for (int v = 0; v < V; v++) {
            out[v] *= beta;
            out[v] += sv[v];
        }
    }
    __device__ __forceinline__ T at(long c) const {
        T r = po[c];
        r *= beta;
        r += s[c];
        return r;
    }
};

// Marches one thread's V cells through planes k0..k1-1, calling f(c, v, k,
// x_c, Ax) for each cell with the operator applied to the loader's vector.
// (Issuing the next plane's loads ahead of the current plane's arithmetic
// measured no faster at 256^3 and cost FP64 occupancy.)
template <typename T, int V, typename LD, typename F>
__device__ __forceinline__ void march_apply(const FsLayout& L, const OpT<T>& op, const LD& x, int i0, int j, int k0,
                                            int zc, T pin_b, T pin_d, F&& f) {
    int k1 = min(k0 + zc, L.nz);
    long c = L.idx(i0, j, k0);
    T xm[V], xc[V], bzd[V];
    x.vec(c - L.sy, xm);
    x.vec(c, xc);
    vld<T, V>(op.bz + c, bzd);
    for (int k = k0; k < k1; k++, c += L.sy) {
        T xp[V], bzu[V], xb[V], xt[V], bl[V], bb[V], bt[V];
        x.vec(c + L.sy, xp);
        vld<T, V>(op.bz + c + L.sy, bzu);
        x.vec(c - L.sx, xb);
        x.vec(c + L.sx, xt);
        vld<T, V>(op.bx + c, bl);
        vld<T, V>(op.by + c, bb);
        vld<T, V>(op.by + c + L.sx, bt);
        T xleft = x.at(c - 1), xright = x.at(c + V), bxr = op.bx[c + V];
#pragma unroll
        for (int v = 0; v < V; v++) {
            Nb<T> n;
            n.xc = xc[v];
            n.xl = v == 0 ? xleft : xc[v - 1];
            n.xr = v == V - 1 ? xright : xc[v + 1];
            n.xb = xb[v];
            n.xt = xt[v];
            n.xd = xm[v];
            n.xu = xp[v];
            n.bl = bl[v];
            n.br = v == V - 1 ? bxr : bl[v + 1];
            n.bb = bb[v];
            n.bt = bt[v];
            n.bd = bzd[v];
            n.bu = bzu[v];
            bool pin = (v == 0) && ((i0 | j | k) == 0);
            f(c, v, k, xc[v], row_val(op, n, pin, pin_b, pin_d));
        }
#pragma unroll
        for (int v = 0; v < V; v++) {
            xm[v] = xc[v];
            xc[v] = xp[v];
            bzd[v] = bzu[v];
        }
    }
}

// march_apply for a whole workgroup with the operand staged in LDS: each
// plane's tile rows j0-1 .. j0+TBY (and one cell either side in x) are formed
// once by the loader and written to LDS, where the y- and x-neighbours are
// read; the z-neighbours stay in registers.  Two plane buffers alternate, so
// one barrier per plane separates a plane's writes from its reads (a
// workgroup cannot run two planes ahead of its slowest thread).  Every
// thread of the workgroup must call it with the same tile (barriers); cells
// outside the grid are staged but not computed.  Operand values and the
// operation order are those of march_apply.
#define LDS_ROW(V) (8 + TBX * (V))
template <typename T, int V>
struct StageLds {
    T a[2][TBY + 2][LDS_ROW(V)];
};
template <typename T, int V, typename LD, typename F>
__device__ __forceinline__ void march_apply_lds(const FsLayout& L, const OpT<T>& op, const LD& x, int tile_x0, int j0,
                                                int k0, int zc, T pin_b, T pin_d, StageLds<T, V>& sh, F&& f) {
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = tile_x0 + tx * V, j = j0 + ty;
    const bool inside = i0 < L.nx && j < L.ny;
    const int k1 = min(k0 + zc, L.nz);
    // halo duties: rows j0-1 and j0+TBY (threads of rows 0 and 1), and the
    // x-halo cell either side of the 6 rows (12 threads of row 2)
    const int hrow = ty == 0 ? j0 - 1 : j0 + TBY;
    const bool hload = ty < 2 && i0 < L.nx;
    // the right x-halo cell lies past the tile; when the tile overhangs the
    // grid its right neighbour column is a centre slot (staged as the zero
    // ghost) and is not loaded
    const int xr = tile_x0 + TBX * V;
    const int xrow = tx % (TBY + 2), xside = tx / (TBY + 2);
    const bool xload = ty == 2 && tx < 2 * (TBY + 2) && (xside == 0 || xr <= L.nx);
    const int xcol = xside ? 4 + TBX * V : 3;
    const long xoff = L.idx(xside ? xr : tile_x0 - 1, j0 - 1 + xrow, 0) - L.idx(0, 0, 0);
    long c = L.idx(i0, j, k0);
    const long hc = L.idx(i0, hrow, k0);
    T xm[V], xc[V], bzd[V];
    auto stage = [&](int b, const T (&ctr)[V], const T (&hal)[V], T xh) {
#pragma unroll
        for (int v = 0; v < V; v++) sh.a[b][ty + 1][4 + tx * V + v] = inside ? ctr[v] : (T)0;
        if (hload) {
#pragma unroll
            for (int v = 0; v < V; v++) sh.a[b][ty == 0 ? 0 : TBY + 1][4 + tx * V + v] = hal[v];
        }
        if (xload) sh.a[b][xrow][xcol] = xh;
    };
    T hal[V] = {}, xh = 0;
    // the previous tile's last plane may still be read from buffer 0
    __syncthreads();
    if (inside) {
        x.vec(c - L.sy, xm);
        x.vec(c, xc);
        vld<T, V>(op.bz + c, bzd);
    } else {
#pragma unroll
        for (int v = 0; v < V; v++) xm[v] = xc[v] = bzd[v] = 0;
    }
    if (hload) x.vec(hc, hal);
    if (xload) xh = x.at(xoff + L.idx(0, 0, k0));
    stage(0, xc, hal, xh);
    for (int k = k0; k < k1; k++, c += L.sy) {
        const int b = (k - k0) & 1;
        const long kk = (long)(k - k0 + 1) * L.sy;
        T xp[V], bzu[V], bl[V], bb[V], bt[V], hn[V], xhn = 0;
#pragma unroll
        for (int v = 0; v < V; v++) xp[v] = bzu[v] = bl[v] = bb[v] = bt[v] = hn[v] = 0;
        const bool more = k + 1 < k1;
        if (inside) {
            x.vec(c + L.sy, xp);
            vld<T, V>(op.bz + c + L.sy, bzu);
            vld<T, V>(op.bx + c, bl);
            vld<T, V>(op.by + c, bb);
            vld<T, V>(op.by + c + L.sx, bt);
        }
        if (more && hload) x.vec(hc + kk, hn);
        if (more && xload) xhn = x.at(xoff + L.idx(0, 0, k + 1));
        __syncthreads();
        if (inside) {
            T xb[V], xt[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                xb[v] = sh.a[b][ty][4 + tx * V + v];
                xt[v] = sh.a[b][ty + 2][4 + tx * V + v];
            }
            T xleft = sh.a[b][ty + 1][3 + tx * V], xright = sh.a[b][ty + 1][4 + tx * V + V];
            T bxr = op.bx[c + V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                Nb<T> n;
                n.xc = xc[v];
                n.xl = v == 0 ? xleft : xc[v - 1];
                n.xr = v == V - 1 ? xright : xc[v + 1];
                n.xb = xb[v];
                n.xt = xt[v];
                n.xd = xm[v];
                n.xu = xp[v];
                n.bl = bl[v];
                n.br = v == V - 1 ? bxr : bl[v + 1];
                n.bb = bb[v];
                n.bt = bt[v];
                n.bd = bzd[v];
                n.bu = bzu[v];
                bool pin = (v == 0) && ((i0 | j | k) == 0);
                f(c, v, k, xc[v], row_val(op, n, pin, pin_b, pin_d));
            }
        }
        if (more) stage(b ^ 1, xp, hn, xhn);
        if (inside) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                xm[v] = xc[v];
                xc[v] = xp[v];
                bzd[v] = bzu[v];
            }
        }
    }
}

// The GFM face coefficient of fs_assemble (gfm_beta, k_beta) with v = 1/rho
// from the sign of phi (k_density).  FP64: the same formula, so the same
// bits.  FP32 (the inner operator of the refinement, which only has to
// approximate A): a face within one phase takes v itself, which the formula
// equals up to rounding, and an interface face divides by a hardware
// reciprocal (about 1 ulp).  Both cells of a face form it from the same
// operands in the same order, so the operator stays symmetric.
__device__ __forceinline__ double gfm_t(double vc, double vn, double fn, double fc) { return gfm_beta(vc, vn, fn, fc); }
__device__ __forceinline__ float gfm_t(float vc, float vn, float fn, float fc) {
    if (vc == vn) return vc;
    return vc * vn * (fabsf(fn) + fabsf(fc)) * __builtin_amdgcn_rcpf(vc * fabsf(fn) + vn * fabsf(fc));
}

// march_apply_lds with the face coefficients formed from the level set:
// the operand and phi are staged together (same rows, same halo duties),
// each cell forms its left, bottom, top and up faces (the right face of
// all but the last of its V cells is the next cell's left face, the down
// face is the previous plane's up face), wall faces are 0.  Three words per
// cell fewer than reading bx, by, bz.
template <typename T, int V>
struct StageLdsPhi {
    T a[2][TBY + 2][LDS_ROW(V)];
    T f[2][TBY + 2][LDS_ROW(V)];
};
template <typename T, int V, typename LD, typename F>
__device__ __forceinline__ void march_apply_lds_phi(const FsLayout& L, const OpT<T>& op, const T* __restrict__ phi,
                                                    T vneg, T vpos, const LD& x, int tile_x0, int j0, int k0, int zc,
                                                    T pin_b, T pin_d, StageLdsPhi<T, V>& sh, F&& f) {
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = tile_x0 + tx * V, j = j0 + ty;
    const bool inside = i0 < L.nx && j < L.ny;
    const int k1 = min(k0 + zc, L.nz);
    const int hrow = ty == 0 ? j0 - 1 : j0 + TBY;
    const bool hload = ty < 2 && i0 < L.nx;
    const int xr = tile_x0 + TBX * V;
    const int xrow = tx % (TBY + 2), xside = tx / (TBY + 2);
    const bool xload = ty == 2 && tx < 2 * (TBY + 2) && (xside == 0 || xr <= L.nx);
    const int xcol = xside ? 4 + TBX * V : 3;
    const long xoff = L.idx(xside ? xr : tile_x0 - 1, j0 - 1 + xrow, 0) - L.idx(0, 0, 0);
    long c = L.idx(i0, j, k0);
    const long hc = L.idx(i0, hrow, k0);
    auto vof = [&](T fv) { return fv <= (T)0 ? vneg : vpos; };
    T xm[V], xc[V], fc[V], bzd[V];
    auto stage = [&](int b, const T (&ctr)[V], const T (&hal)[V], T xh, const T (&fctr)[V], const T (&fhal)[V],
                     T fxh) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            sh.a[b][ty + 1][4 + tx * V + v] = inside ? ctr[v] : (T)0;
            sh.f[b][ty + 1][4 + tx * V + v] = inside ? fctr[v] : (T)0;
        }
        if (hload) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                sh.a[b][ty == 0 ? 0 : TBY + 1][4 + tx * V + v] = hal[v];
                sh.f[b][ty == 0 ? 0 : TBY + 1][4 + tx * V + v] = fhal[v];
            }
        }
        if (xload) {
            sh.a[b][xrow][xcol] = xh;
            sh.f[b][xrow][xcol] = fxh;
        }
    };
    T hal[V] = {}, fhal[V] = {}, xh = 0, fxh = 0;
    __syncthreads();
    if (inside) {
        T fm[V];
        x.vec(c - L.sy, xm);
        x.vec(c, xc);
        vld<T, V>(phi + c, fc);
        vld<T, V>(phi + c - L.sy, fm);
#pragma unroll
        for (int v = 0; v < V; v++) bzd[v] = k0 > 0 ? gfm_t(vof(fc[v]), vof(fm[v]), fm[v], fc[v]) : (T)0;
    } else {
#pragma unroll
        for (int v = 0; v < V; v++) xm[v] = xc[v] = fc[v] = bzd[v] = 0;
    }
    if (hload) {
        x.vec(hc, hal);
        vld<T, V>(phi + hc, fhal);
    }
    if (xload) {
        xh = x.at(xoff + L.idx(0, 0, k0));
        fxh = phi[xoff + L.idx(0, 0, k0)];
    }
    stage(0, xc, hal, xh, fc, fhal, fxh);
    for (int k = k0; k < k1; k++, c += L.sy) {
        const int b = (k - k0) & 1;
        const long kk = (long)(k - k0 + 1) * L.sy;
        T xp[V], fp[V], hn[V], fhn[V], xhn = 0, fxhn = 0;
#pragma unroll
        for (int v = 0; v < V; v++) xp[v] = fp[v] = hn[v] = fhn[v] = 0;
        const bool more = k + 1 < k1;
        if (inside) {
            x.vec(c + L.sy, xp);
            vld<T, V>(phi + c + L.sy, fp);
        }
        if (more && hload) {
            x.vec(hc + kk, hn);
            vld<T, V>(phi + hc + kk, fhn);
        }
        if (more && xload) {
            xhn = x.at(xoff + L.idx(0, 0, k + 1));
            fxhn = phi[xoff + L.idx(0, 0, k + 1)];
        }
        __syncthreads();
        if (inside) {
            T xb[V], xt[V], fb[V], ft[V], bl[V], bzu[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                xb[v] = sh.a[b][ty][4 + tx * V + v];
                xt[v] = sh.a[b][ty + 2][4 + tx * V + v];
                fb[v] = sh.f[b][ty][4 + tx * V + v];
                ft[v] = sh.f[b][ty + 2][4 + tx * V + v];
            }
            const T xleft = sh.a[b][ty + 1][3 + tx * V], xright = sh.a[b][ty + 1][4 + tx * V + V];
            const T fleft = sh.f[b][ty + 1][3 + tx * V], fright = sh.f[b][ty + 1][4 + tx * V + V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                const T fl = v == 0 ? fleft : fc[v - 1];
                bl[v] = i0 + v > 0 ? gfm_t(vof(fc[v]), vof(fl), fl, fc[v]) : (T)0;
                bzu[v] = k + 1 < L.nz ? gfm_t(vof(fp[v]), vof(fc[v]), fc[v], fp[v]) : (T)0;
            }
            const T bxr = i0 + V < L.nx ? gfm_t(vof(fright), vof(fc[V - 1]), fc[V - 1], fright) : (T)0;
#pragma unroll
            for (int v = 0; v < V; v++) {
                Nb<T> n;
                n.xc = xc[v];
                n.xl = v == 0 ? xleft : xc[v - 1];
                n.xr = v == V - 1 ? xright : xc[v + 1];
                n.xb = xb[v];
                n.xt = xt[v];
                n.xd = xm[v];
                n.xu = xp[v];
                n.bl = bl[v];
                n.br = v == V - 1 ? bxr : bl[v + 1];
                n.bb = j > 0 ? gfm_t(vof(fc[v]), vof(fb[v]), fb[v], fc[v]) : (T)0;
                n.bt = j + 1 < L.ny ? gfm_t(vof(ft[v]), vof(fc[v]), fc[v], ft[v]) : (T)0;
                n.bd = bzd[v];
                n.bu = bzu[v];
                bool pin = (v == 0) && ((i0 | j | k) == 0);
                f(c, v, k, xc[v], row_val(op, n, pin, pin_b, pin_d));
            }
#pragma unroll
            for (int v = 0; v < V; v++) bzd[v] = bzu[v];
        }
        if (more) stage(b ^ 1, xp, hn, xhn, fp, fhn, fxhn);
        if (inside) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                xm[v] = xc[v];
                xc[v] = xp[v];
                fc[v] = fp[v];
            }
        }
    }
}

// ---- begin: r = b - A x, s = r/D, p = s; partials of r.s, r.r, b.b -------
template <typename T, int V>
__global__ void __launch_bounds__(NT) k_begin(FsLayout L, Tiling tl, FsPcgOp o, const T* __restrict__ b,
                                              const T* __restrict__ D, const T* __restrict__ x, T* __restrict__ r,
                                              T* __restrict__ s, T* __restrict__ p, double* part, FsPcgState* S) {
    __shared__ double lds[4];
    OpT<T> op = make_op<T>(o);
    double a_rs = 0, a_rr = 0, a_bb = 0;
    T pin_b = (T)S->pin_b, pin_d = (T)S->pin_d;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        if (i0 >= L.nx || j >= L.ny) continue;
        march_apply<T, V>(L, op, LdPlain<T, V>{x}, i0, j, k0, tl.zc, pin_b, pin_d, [&](long c, int v, int, T, T ax) {
            T bv = b[c + v];
            T rv = bv - ax;
            T sv = rv / D[c + v];
            r[c + v] = rv;
            s[c + v] = sv;
            p[c + v] = sv;
            a_rs += (double)rv * (double)sv;
            a_rr += (double)rv * (double)rv;
            a_bb += (double)bv * (double)bv;
        });
    }
    put_partial(part, SLOT_RS, a_rs, lds);
    put_partial(part, SLOT_RR, a_rr, lds);
    put_partial(part, SLOT_BB, a_bb, lds);
}

// One workgroup: the begin scalars into the state.
__global__ void __launch_bounds__(NT) k_begin_final(const double* part, int nb, FsPcgState* S) {
    __shared__ double lds[4];
    double rs = get_total(part, SLOT_RS, nb, lds);
    double rr = get_total(part, SLOT_RR, nb, lds);
    double bb = get_total(part, SLOT_BB, nb, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        S->rz[0] = rs;
        S->rs_last = rs;
        S->rr = rr;
        S->bb = bb;
        S->iter = 0;
        S->done = 0;
        if (S->stop_rule == 1 && (bb == 0 || rr <= S->rtol2 * bb)) S->done = 1;
        if (S->max_iter <= 0) S->done = 2;
    }
}

// ---- state step shared by the fused K1 and the chunk-end test ------------
// The reference's loop tail for iteration m = n-1 (PCG_METHOD.cpp:81-104):
// with r.s and r.r of K2(m), stop or take r.s as the next rz.  Workgroup 0
// records the outcome.  Returns true if iteration n is not to run.
__device__ __forceinline__ bool loop_tail(const double* part, int nb, FsPcgState* S, int n, double& rs,
                                          double* lds) {
    rs = get_total(part, SLOT_RS, nb, lds);
    double rr = get_total(part, SLOT_RR, nb, lds);
    bool stop = S->stop_rule == 1 ? (rr < S->rtol2 * S->bb) : (rs < S->tol2);
    bool cap = !stop && n >= S->max_iter;
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        S->rs_last = rs;
        S->rr = rr;
        if (stop) {
            S->iter = n - 1;
            S->done = 1;
        } else {
            S->rz[n & 1] = rs;
            S->iter = n;
            if (cap) S->done = 2;
        }
    }
    return stop || cap;
}

// ---- K1 (fused with the p-update of the previous iteration) ----------------
// n = 0: q = A p with p from begin.  n > 0: loop tail of iteration n-1, then
// p(n) = p(n-1)*beta + s formed on the fly for every cell the stencil
// touches (p(n-1) from pold, written to pnew: the two directions live in
// separate buffers so no workgroup reads a value another has replaced),
// q = A p(n) and the partial p.q.
template <typename T, int V, bool LDS, bool PHI = false>
__global__ void __launch_bounds__(NT) k_pspmv_dot(FsLayout L, Tiling tl, FsPcgOp o, const T* __restrict__ s,
                                                  const T* __restrict__ pold, T* __restrict__ pnew, T* __restrict__ q,
                                                  double* part, FsPcgState* S, int n, int nb_rs) {
    if (S->done) return;
    __shared__ double lds[4];
    T beta = 0;
    if (n > 0) {
        double rz_old = S->rz[(n - 1) & 1];
        double rs;
        if (loop_tail(part, nb_rs, S, n, rs, lds)) return;
        beta = (T)(rs / rz_old);
    }
    OpT<T> op = make_op<T>(o);
    double acc = 0;
    T pin_b = (T)S->pin_b, pin_d = (T)S->pin_d;
    if constexpr (PHI) {
    __shared__ StageLdsPhi<T, V> sh;
    const T* phi = (const T*)o.phi;
    const T vneg = (T)o.v_neg, vpos = (T)o.v_pos;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        const int tile_x0 = i0 - threadIdx.x * V, j0 = j - threadIdx.y;
        T qv[V], pv[V];
        auto body = [&](long c, int v, int, T pc, T ap) {
            qv[v] = ap;
            pv[v] = pc;
            acc += (double)pc * (double)ap;
            if (v == V - 1) {
                vst_s<NTB(8), T, V>(q + c, qv);
                if (n > 0) vst_s<NTB(9), T, V>(pnew + c, pv);
            }
        };
        if (n > 0)
            march_apply_lds_phi<T, V>(L, op, phi, vneg, vpos, LdDir<T, V>{pold, s, beta}, tile_x0, j0, k0, tl.zc,
                                      pin_b, pin_d, sh, body);
        else
            march_apply_lds_phi<T, V>(L, op, phi, vneg, vpos, LdPlain<T, V>{pold}, tile_x0, j0, k0, tl.zc, pin_b,
                                      pin_d, sh, body);
    }
    } else if constexpr (LDS) {
    __shared__ StageLds<T, V> sh;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        const int tile_x0 = i0 - threadIdx.x * V, j0 = j - threadIdx.y;
        T qv[V], pv[V];
        auto body = [&](long c, int v, int, T pc, T ap) {
            qv[v] = ap;
            pv[v] = pc;
            acc += (double)pc * (double)ap;
            if (v == V - 1) {
                vst_s<NTB(8), T, V>(q + c, qv);
                if (n > 0) vst_s<NTB(9), T, V>(pnew + c, pv);
            }
        };
        if (n > 0)
            march_apply_lds<T, V>(L, op, LdDir<T, V>{pold, s, beta}, tile_x0, j0, k0, tl.zc, pin_b, pin_d, sh, body);
        else
            march_apply_lds<T, V>(L, op, LdPlain<T, V>{pold}, tile_x0, j0, k0, tl.zc, pin_b, pin_d, sh, body);
    }
    } else {
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        if (i0 >= L.nx || j >= L.ny) continue;
        T qv[V], pv[V];
        auto body = [&](long c, int v, int, T pc, T ap) {
            qv[v] = ap;
            pv[v] = pc;
            acc += (double)pc * (double)ap;
            if (v == V - 1) {
                vst_s<NTB(8), T, V>(q + c, qv);
                if (n > 0) vst_s<NTB(9), T, V>(pnew + c, pv);
            }
        };
        if (n > 0)
            march_apply<T, V>(L, op, LdDir<T, V>{pold, s, beta}, i0, j, k0, tl.zc, pin_b, pin_d, body);
        else
            march_apply<T, V>(L, op, LdPlain<T, V>{pold}, i0, j, k0, tl.zc, pin_b, pin_d, body);
    }
    }
    put_partial(part, SLOT_PQ, acc, lds);
}

// The loop tail of iteration n alone (one workgroup), after the last K2 of a
// launch batch: the host sees the stop and r.r without another batch.
__global__ void __launch_bounds__(NT) k_tail(const double* part, int nb, FsPcgState* S, int n) {
    if (S->done) return;
    __shared__ double lds[4];
    double rs = get_total(part, SLOT_RS, nb, lds);
    double rr = get_total(part, SLOT_RR, nb, lds);
    bool stop = S->stop_rule == 1 ? (rr < S->rtol2 * S->bb) : (rs < S->tol2);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        S->rr = rr;
        if (stop) {
            S->rs_last = rs;
            S->iter = n;
            S->done = 1;
        }
    }
}

// ---- K2: alpha = rz/pq; x += a p, r -= a q, s = r/D; partials r.s, r.r ----
template <typename T, int V>
__global__ void __launch_bounds__(NT) k_update(FsLayout L, Tiling tl, const T* __restrict__ D, T* __restrict__ x,
                                               T* __restrict__ r, T* __restrict__ s, const T* __restrict__ p,
                                               const T* __restrict__ q, double* part, const FsPcgState* S, int n,
                                               int nb_pq) {
    if (S->done) return;
    __shared__ double lds[4];
    double pq = get_total(part, SLOT_PQ, nb_pq, lds);
    T alpha = (T)(S->rz[n & 1] / pq);
    double a_rs = 0, a_rr = 0;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        if (i0 >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + tl.zc, L.nz);
        long c = L.idx(i0, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy) {
            T xv[V], pv[V], rv[V], qv[V], dv[V], sv[V];
            vld_s<NTB(0), T, V>(x + c, xv);
            vld_s<NTB(1), T, V>(p + c, pv);
            vld_s<NTB(2), T, V>(r + c, rv);
            vld_s<NTB(3), T, V>(q + c, qv);
            vld_s<NTB(4), T, V>(D + c, dv);
#pragma unroll
            for (int v = 0; v < V; v++) {
                xv[v] += alpha * pv[v];
                rv[v] -= alpha * qv[v];
                sv[v] = rv[v] / dv[v];
                a_rs += (double)rv[v] * (double)sv[v];
                a_rr += (double)rv[v] * (double)rv[v];
            }
            vst_s<NTB(5), T, V>(x + c, xv);
            vst_s<NTB(6), T, V>(r + c, rv);
            vst_s<NTB(7), T, V>(s + c, sv);
        }
    }
    put_partial(part, SLOT_RS, a_rs, lds);
    put_partial(part, SLOT_RR, a_rr, lds);
}

// ---- FS_OP_ROWS operator (FP64): explicit entries per row -----------------
template <typename LD>
__device__ __forceinline__ double rows_val(const FsPcgOp& o, const LD& x, long c, long sx, long sy) {
    double v = 0;
    v += o.a[0][c] * x.at(c - 1);
    v += o.a[1][c] * x.at(c + 1);
    v += o.a[2][c] * x.at(c - sx);
    v += o.a[3][c] * x.at(c + sx);
    v += o.a[4][c] * x.at(c - sy);
    v += o.a[5][c] * x.at(c + sy);
    v += o.dg[c] * x.at(c);
    return v;
}

// begin on FS_OP_ROWS: r = b - A x, s = r/D, p = s and the partials.  The
// tiling (hence the partial count) is the one the update kernels use.
template <int V>
__global__ void __launch_bounds__(NT) k_rows_begin(FsLayout L, Tiling tl, FsPcgOp o, const double* __restrict__ b,
                                                   const double* __restrict__ D, const double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ s,
                                                   double* __restrict__ p, double* part) {
    __shared__ double lds[4];
    double a_rs = 0, a_rr = 0, a_bb = 0;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        if (i0 >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + tl.zc, L.nz);
        long c = L.idx(i0, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy)
            for (int v = 0; v < V; v++) {
                double bv = b[c + v];
                double rv = bv - rows_val(o, LdPlain<double, 1>{x}, c + v, L.sx, L.sy);
                double sv = rv / D[c + v];
                r[c + v] = rv;
                s[c + v] = sv;
                p[c + v] = sv;
                a_rs += rv * sv;
                a_rr += rv * rv;
                a_bb += bv * bv;
            }
    }
    put_partial(part, SLOT_RS, a_rs, lds);
    put_partial(part, SLOT_RR, a_rr, lds);
    put_partial(part, SLOT_BB, a_bb, lds);
}

// K1 on FS_OP_ROWS, fused like k_pspmv_dot: loop tail of n-1, p(n) on the
// fly, q = A p(n), partial p.q.
template <int V>
__global__ void __launch_bounds__(NT) k_rows_pspmv_dot(FsLayout L, Tiling tl, FsPcgOp o, const double* __restrict__ s,
                                                       const double* __restrict__ pold, double* __restrict__ pnew,
                                                       double* __restrict__ q, double* part, FsPcgState* S, int n) {
    if (S->done) return;
    __shared__ double lds[4];
    double beta = 0;
    if (n > 0) {
        double rz_old = S->rz[(n - 1) & 1];
        double rs;
        if (loop_tail(part, tl.nb, S, n, rs, lds)) return;
        beta = rs / rz_old;
    }
    LdDir<double, 1> dir{pold, s, beta};
    LdPlain<double, 1> plain{pold};
    double acc = 0;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        if (i0 >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + tl.zc, L.nz);
        long c = L.idx(i0, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy)
            for (int v = 0; v < V; v++) {
                double pc, ap;
                if (n > 0) {
                    pc = dir.at(c + v);
                    ap = rows_val(o, dir, c + v, L.sx, L.sy);
                    pnew[c + v] = pc;
                } else {
                    pc = plain.at(c + v);
                    ap = rows_val(o, plain, c + v, L.sx, L.sy);
                }
                q[c + v] = ap;
                acc += pc * ap;
            }
    }
    put_partial(part, SLOT_PQ, acc, lds);
}

// ---- launch helpers ---------------------------------------------------------
template <typename T, int V>
static void begin_v(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* rhs, const T* diag, T* x,
                    FsPcgWork& w, FsPcgState* S) {
    Tiling tl = work_tiling(L, V, w, 1);
    if (op.kind == FS_OP_ROWS)
        hipLaunchKernelGGL((k_rows_begin<V>), dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, (const double*)rhs,
                           (const double*)diag, (const double*)x, (double*)w.r, (double*)w.s, (double*)w.p,
                           w.partials);
    else
        hipLaunchKernelGGL((k_begin<T, V>), dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, rhs, diag, x, (T*)w.r,
                           (T*)w.s, (T*)w.p, w.partials, S);
    hipLaunchKernelGGL(k_begin_final, dim3(1), dim3(TBX, TBY), 0, st, w.partials, tl.nb, S);
}

// Direction of iteration n (begin writes p(0) into w.p).
static inline void* dir_buf(FsPcgWork& w, int n) { return (n & 1) ? w.p2 : w.p; }

// Kernel `which` of iteration n: 0 the fused K1, 1 the update K2, 2 the
// one-workgroup loop tail of n.
template <typename T, int V>
// e0/e1 (may be null) are recorded by the dispatch itself around the
// kernel (hipExtLaunchKernelGGL), so a probed kernel's time excludes the gaps
// that separate event packets would add.
static void launch_one(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* diag, T* x, FsPcgWork& w,
                       FsPcgState* S, int which, int n, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    Tiling tl = work_tiling(L, V, w, 1);
    T* pn = (T*)dir_buf(w, n);
    const T* po = (const T*)dir_buf(w, n > 0 ? n - 1 : 0);
    // the fused SpMV keeps three planes and six coefficient vectors live:
    // it may run narrower than the update (fewer registers, more waves)
    constexpr int VW = sizeof(T) == 4 ? FS_SPMV_V32 : FS_SPMV_V64;
    constexpr int VK = VW < V ? VW : V;
    const bool lds = (long)L.nx * L.ny * L.nz >= FS_LDS_MIN_CELLS;
    const Tiling tk = work_tiling(L, VK, w, 0, lds ? FS_K1_MIN_TILES : FS_MIN_TILES);
    const int nb_pq = op.kind == FS_OP_ROWS ? tl.nb : tk.nb;  // workgroups of this iteration's K1
    if (which == 0 && op.kind == FS_OP_ROWS)
        hipExtLaunchKernelGGL((k_rows_pspmv_dot<V>), dim3(tl.nb), dim3(TBX, TBY), 0, st, e0, e1, 0, L, tl, op,
                              (const double*)w.s, (const double*)po, (double*)pn, (double*)w.q, w.partials, S, n);
    else if (which == 0 && lds && op.phi)
        hipExtLaunchKernelGGL((k_pspmv_dot<T, VK, true, true>), dim3(tk.nb), dim3(TBX, TBY), 0, st, e0, e1, 0, L, tk,
                              op, (const T*)w.s, po, pn, (T*)w.q, w.partials, S, n, tl.nb);
    else if (which == 0 && lds)
        hipExtLaunchKernelGGL((k_pspmv_dot<T, VK, true>), dim3(tk.nb), dim3(TBX, TBY), 0, st, e0, e1, 0, L, tk, op,
                              (const T*)w.s, po, pn, (T*)w.q, w.partials, S, n, tl.nb);
    else if (which == 0)
        hipExtLaunchKernelGGL((k_pspmv_dot<T, VK, false>), dim3(tk.nb), dim3(TBX, TBY), 0, st, e0, e1, 0, L, tk, op,
                              (const T*)w.s, po, pn, (T*)w.q, w.partials, S, n, tl.nb);
    else if (which == 1)
        hipExtLaunchKernelGGL((k_update<T, V>), dim3(tl.nb), dim3(TBX, TBY), 0, st, e0, e1, 0, L, tl, diag, x,
                              (T*)w.r, (T*)w.s, (const T*)pn, (const T*)w.q, w.partials, S, n, nb_pq);
    else
        hipExtLaunchKernelGGL(k_tail, dim3(1), dim3(TBX, TBY), 0, st, e0, e1, 0, (const double*)w.partials, tl.nb, S,
                              n);
}

template <typename T>
static void begin_t(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* rhs, const T* diag, T* x,
                    FsPcgWork& w, FsPcgState* S) {
    int v = vec_width(L, sizeof(T) == 4);
    if (v == 4) begin_v<T, (sizeof(T) == 4 ? 4 : 2)>(st, L, op, rhs, diag, x, w, S);
    else if (v == 2) begin_v<T, 2>(st, L, op, rhs, diag, x, w, S);
    else begin_v<T, 1>(st, L, op, rhs, diag, x, w, S);
}

template <typename T>
static void one_t(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* diag, T* x, FsPcgWork& w,
                  FsPcgState* S, int which, int n, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    int v = vec_width(L, sizeof(T) == 4);
    if (v == 4) launch_one<T, (sizeof(T) == 4 ? 4 : 2)>(st, L, op, diag, x, w, S, which, n, e0, e1);
    else if (v == 2) launch_one<T, 2>(st, L, op, diag, x, w, S, which, n, e0, e1);
    else launch_one<T, 1>(st, L, op, diag, x, w, S, which, n, e0, e1);
}

// Kernel `which` of iteration n, bracketed by the probe's events when n is
// the probed iteration.
template <typename T>
static void one_probed(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* diag, T* x, FsPcgWork& w,
                       FsPcgState* S, int which, int n, FsKernelProbe* probe) {
    bool rec = probe && probe->iter == n;
    if (rec)
        one_t<T>(st, L, op, diag, x, w, S, which, n, probe->ev[2 * which], probe->ev[2 * which + 1]);
    else
        one_t<T>(st, L, op, diag, x, w, S, which, n);
    if (rec && which == FS_PCG_KERNELS - 1) probe->recorded = true;
}

// Iterations n0 .. n0+count-1 (K1, K2 each) and the loop tail of the last.
template <typename T>
static void iterations_t(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* diag, T* x, FsPcgWork& w,
                         FsPcgState* S, int n0, int count, FsKernelProbe* probe) {
    if (probe) {
        probe->iter = n0;
        probe->recorded = false;
    }
    for (int n = n0; n < n0 + count; n++)
        for (int which = 0; which < FS_PCG_KERNELS; which++) one_probed<T>(st, L, op, diag, x, w, S, which, n, probe);
    one_t<T>(st, L, op, diag, x, w, S, 2, n0 + count - 1);
}

void fs_pcg64_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* rhs, const double* diag,
                    double* x, FsPcgWork& w, FsPcgState* S) {
    begin_t<double>(st, L, op, rhs, diag, x, w, S);
}
void fs_pcg64_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* diag, double* x,
                         FsPcgWork& w, FsPcgState* S, int n0, int count, FsKernelProbe* probe) {
    iterations_t<double>(st, L, op, diag, x, w, S, n0, count, probe);
}
void fs_pcg32_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* rhs, const float* diag,
                    float* x, FsPcgWork& w, FsPcgState* S) {
    begin_t<float>(st, L, op, rhs, diag, x, w, S);
}
void fs_pcg32_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* diag, float* x,
                         FsPcgWork& w, FsPcgState* S, int n0, int count) {
    iterations_t<float>(st, L, op, diag, x, w, S, n0, count, nullptr);
}

void fs_pcg_time_kernels(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const void* diag, void* x,
                         FsPcgWork& w, FsPcgState* S, int precision, int which, int n) {
    if (which < 0 || which >= FS_PCG_KERNELS) return;
    if (precision == 0)
        one_t<double>(st, L, op, (const double*)diag, (double*)x, w, S, which, n);
    else
        one_t<float>(st, L, op, (const float*)diag, (float*)x, w, S, which, n);
}

// ---- tiling autotuner ----------------------------------------------------
// The state the timed launches run with: never stops (relative rule against
// a zero target, a cap out of reach), r.z of one.
__global__ void k_tune_state(FsPcgState* S) {
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        S->rz[0] = 1;
        S->rz[1] = 1;
        S->done = 0;
        S->max_iter = 1 << 30;
        S->stop_rule = 1;
        S->rtol2 = 0;
        S->bb = 1;
    }
}

__global__ void k_tune_zero(double* a, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = 0;
}

namespace {
struct TuneKey {
    int dev, nx, ny, nz, prec;
    bool operator<(const TuneKey& o) const {
        const int a[5] = {dev, nx, ny, nz, prec}, b[5] = {o.dev, o.nx, o.ny, o.nz, o.prec};
        return std::lexicographical_compare(a, a + 5, b, b + 5);
    }
};
struct TuneVal {
    int zc[2], nbmax[2];
};
std::mutex tune_mu;
std::map<TuneKey, TuneVal> tune_cache;
}  // namespace

void fs_pcg_tune(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const void* diag, FsPcgWork& w,
                 FsPcgState* S, int precision) {
    if (op.kind != FS_OP_GFM) return;
    const char* env = std::getenv("FS_PCG_TUNE");
    if (env && std::strcmp(env, "0") == 0) {
        w.zc[0] = w.zc[1] = w.nbmax[0] = w.nbmax[1] = 0;
        return;
    }
    // FS_PCG_TILING="zc0,nb0,zc1,nb1" sets the tiling instead (measuring the
    // spread that the partial sums' order leaves in the results)
    if (const char* fix = std::getenv("FS_PCG_TILING")) {
        int a[4] = {0, 0, 0, 0};
        if (std::sscanf(fix, "%d,%d,%d,%d", &a[0], &a[1], &a[2], &a[3]) == 4 && a[0] > 0 && a[2] > 0 && a[1] > 0 &&
            a[3] > 0) {
            w.zc[0] = a[0];
            w.nbmax[0] = a[1];
            w.zc[1] = a[2];
            w.nbmax[1] = a[3];
            return;
        }
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    const TuneKey key{dev, L.nx, L.ny, L.nz, precision};
    // one tuning per shape at a time; other handles of the shape wait for it
    std::lock_guard<std::mutex> lk(tune_mu);
    auto hit = tune_cache.find(key);
    if (hit != tune_cache.end()) {
        for (int k = 0; k < 2; k++) {
            w.zc[k] = hit->second.zc[k];
            w.nbmax[k] = hit->second.nbmax[k];
        }
        return;
    }
    const size_t bytes = (size_t)L.total * (precision == 0 ? sizeof(double) : sizeof(float));
    void* x = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto fail = [&](hipError_t e) {
        if (x) (void)hipFree(x);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipGetLastError();
        w.zc[0] = w.zc[1] = w.nbmax[0] = w.nbmax[1] = 0;
        (void)e;
    };
    hipError_t e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess || (e = hipMalloc(&x, bytes)) != hipSuccess ||
        (e = hipMemsetAsync(x, 0, bytes, st)) != hipSuccess || (e = hipEventCreate(&e0)) != hipSuccess ||
        (e = hipEventCreate(&e1)) != hipSuccess)
        return fail(e);
    hipLaunchKernelGGL(k_tune_zero, dim3(64), dim3(256), 0, st, w.partials, (long)FS_PCG_SLOTS * FS_PCG_MAX_BLOCKS);
    const int v = vec_width(L, precision != 0);
    const bool debug = std::getenv("FS_DEBUG") != nullptr;
    TuneVal best{{0, 0}, {0, 0}};
    for (int k = 0; k < FS_PCG_KERNELS; k++) {
        float best_ms = 0;
        for (int zc : {8, 16, 32, 64}) {
            if (zc > 8 && (zc >> 1) >= L.nz) break;
            for (int nbmax : {256, 512, 1024, 2048}) {
                const Tiling t = make_tiling(L, v, FS_MIN_TILES, zc, nbmax);
                if (nbmax > 256 && t.ntiles <= (nbmax >> 1)) break;  // same grid as the smaller cap
                w.zc[k] = zc;
                w.nbmax[k] = nbmax;
                float ms_min = 0;
                for (int rep = 0; rep < 4; rep++) {
                    hipLaunchKernelGGL(k_tune_state, dim3(1), dim3(64), 0, st, S);
                    if ((e = hipEventRecord(e0, st)) != hipSuccess) return fail(e);
                    for (int q = 0; q < 3; q++) fs_pcg_time_kernels(st, L, op, diag, x, w, S, precision, k, 1);
                    if ((e = hipEventRecord(e1, st)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess)
                        return fail(e);
                    float ms = 0;
                    if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return fail(e);
                    if (rep == 1 || (rep > 1 && ms < ms_min)) ms_min = ms;  // rep 0 warms up
                }
                if (debug)
                    fprintf(stderr, "[fs] pcg tune %dx%dx%d prec %d kernel %d: zc %d nb %d (%d tiles) %.1f us\n",
                            L.nx, L.ny, L.nz, precision, k, zc, t.nb, t.ntiles, ms_min * 1e3f / 3);
                if (best.zc[k] == 0 || ms_min < best_ms) {
                    best_ms = ms_min;
                    best.zc[k] = zc;
                    best.nbmax[k] = nbmax;
                }
            }
        }
        w.zc[k] = best.zc[k];
        w.nbmax[k] = best.nbmax[k];
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e);
    (void)hipFree(x);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tune_cache[key] = best;
    if (debug)
        fprintf(stderr, "[fs] pcg tiling %dx%dx%d prec %d: spmv zc %d nb<=%d, update zc %d nb<=%d\n", L.nx, L.ny, L.nz,
                precision, best.zc[0], best.nbmax[0], best.zc[1], best.nbmax[1]);
}

// ---- plain operator application (FP64) ----------------------------------
template <int V>
__global__ void __launch_bounds__(NT) k_apply64(FsLayout L, Tiling tl, FsPcgOp o, const double* __restrict__ x,
                                                double* __restrict__ q, const FsPcgState* S) {
    OpT<double> op = make_op<double>(o);
    double pin_b = S->pin_b, pin_d = S->pin_d;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        if (i0 >= L.nx || j >= L.ny) continue;
        march_apply<double, V>(L, op, LdPlain<double, V>{x}, i0, j, k0, tl.zc, pin_b, pin_d,
                               [&](long c, int v, int, double, double ax) { q[c + v] = ax; });
    }
}

void fs_apply_op64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* x, double* q,
                   const FsPcgState* S) {
    int v = vec_width(L, 0);
    Tiling tl = make_tiling(L, v);
    if (v == 2)
        hipLaunchKernelGGL(k_apply64<2>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, x, q, S);
    else
        hipLaunchKernelGGL(k_apply64<1>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, x, q, S);
}

// ---- refinement pieces ----------------------------------------------------
template <int V>
__global__ void __launch_bounds__(NT) k_residual64(FsLayout L, Tiling tl, FsPcgOp o, const double* __restrict__ b,
                                                   const double* __restrict__ x, double* __restrict__ r,
                                                   float* __restrict__ r32, double* part, const FsPcgState* S) {
    __shared__ double lds[4];
    OpT<double> op = make_op<double>(o);
    double a_rr = 0, a_bb = 0, a_sum = 0;
    double pin_b = S->pin_b, pin_d = S->pin_d;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        if (i0 >= L.nx || j >= L.ny) continue;
        march_apply<double, V>(L, op, LdPlain<double, V>{x}, i0, j, k0, tl.zc, pin_b, pin_d, [&](long c, int v, int, double, double ax) {
            double bv = b[c + v];
            double rv = bv - ax;
            r[c + v] = rv;
            r32[c + v] = (float)rv;
            a_rr += rv * rv;
            a_bb += bv * bv;
            a_sum += rv;
        });
    }
    put_partial(part, SLOT_RR, a_rr, lds);
    put_partial(part, SLOT_BB, a_bb, lds);
    put_partial(part, SLOT_SUM, a_sum, lds);
}

__global__ void __launch_bounds__(NT) k_residual_final(const double* part, int nb, FsPcgState* S) {
    __shared__ double lds[4];
    double rr = get_total(part, SLOT_RR, nb, lds);
    double bb = get_total(part, SLOT_BB, nb, lds);
    double sum = get_total(part, SLOT_SUM, nb, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        S->rr = rr;
        S->bb = bb;
        S->rsum = sum;
    }
}

// x += rsum/pin on every interior cell; the pinned cell's residual drops by
// rsum (its row is the only one a constant reaches), r.r follows.
__global__ void __launch_bounds__(NT) k_shift(FsLayout L, double* __restrict__ x, double* __restrict__ r,
                                              float* __restrict__ r32, FsPcgState* S, double pin) {
    double sum = S->rsum;
    double cst = sum / pin;
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    long c = L.idx(i, j, k);
    x[c] += cst;
    if ((i | j | k) == 0) {
        double r0 = r[c], r1 = r0 - sum;
        r[c] = r1;
        r32[c] = (float)r1;
        // rsum itself stays: every workgroup of this kernel reads it
        S->rr = S->rr - r0 * r0 + r1 * r1;
    }
}

void fs_deflate_constant(hipStream_t st, const FsLayout& L, double* x, double* r, float* r32, FsPcgState* S,
                         double pin) {
    dim3 g((L.nx + TBX - 1) / TBX, (L.ny + TBY - 1) / TBY, L.nz);
    hipLaunchKernelGGL(k_shift, g, dim3(TBX, TBY), 0, st, L, x, r, r32, S, pin);
}

void fs_residual64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* b, const double* x, double* r,
                   float* r32, FsPcgWork& w, FsPcgState* S) {
    int v = vec_width(L, 0);
    Tiling tl = make_tiling(L, v);
    if (v == 2)
        hipLaunchKernelGGL(k_residual64<2>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, b, x, r, r32, w.partials,
                           S);
    else
        hipLaunchKernelGGL(k_residual64<1>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, b, x, r, r32, w.partials,
                           S);
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(TBX, TBY), 0, st, w.partials, tl.nb, S);
}

// Reliable update of the FP32 recursion, between the update (K2) of an
// iteration and the next fused K1: the recursion's residual is replaced by
// the rounded FP64 residual, s = r/D and the partials of r.s and r.r are
// rewritten, so that K1 takes its stop test and beta from the refreshed
// residual.
template <int V>
__global__ void __launch_bounds__(NT) k_reprecond(FsLayout L, Tiling tl, const float* __restrict__ r,
                                                  const float* __restrict__ D, float* __restrict__ s,
                                                  float* __restrict__ rrec, double* part) {
    __shared__ double lds[4];
    double a_rs = 0, a_rr = 0;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<V>(tl, it.t, i0, j, k0);
        if (i0 >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + tl.zc, L.nz);
        long c = L.idx(i0, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy) {
            float rv[V], dv[V], sv[V];
            vld<float, V>(r + c, rv);
            vld<float, V>(D + c, dv);
#pragma unroll
            for (int v = 0; v < V; v++) {
                sv[v] = rv[v] / dv[v];
                a_rs += (double)rv[v] * (double)sv[v];
                a_rr += (double)rv[v] * (double)rv[v];
            }
            vst<float, V>(s + c, sv);
            vst<float, V>(rrec + c, rv);  // the recursion continues from the FP64 residual
        }
    }
    put_partial(part, SLOT_RS, a_rs, lds);
    put_partial(part, SLOT_RR, a_rr, lds);
}

// The r.r of the refreshed residual into the state; clears `done` when
// asked (a recursion that met its target resumes).
__global__ void __launch_bounds__(NT) k_peek(const double* part, int nb, FsPcgState* S, int clear_done) {
    __shared__ double lds[4];
    double rr = get_total(part, SLOT_RR, nb, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        S->rr = rr;
        if (clear_done) S->done = 0;
    }
}

void fs_pcg32_reliable_update(hipStream_t st, const FsLayout& L, const float* r32, const float* diag, FsPcgWork& w,
                              FsPcgState* S) {
    int v = vec_width(L, 1);
    Tiling tl = work_tiling(L, v, w, 1);
    if (v == 4)
        hipLaunchKernelGGL(k_reprecond<4>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, r32, diag, (float*)w.s,
                           (float*)w.r, w.partials);
    else if (v == 2)
        hipLaunchKernelGGL(k_reprecond<2>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, r32, diag, (float*)w.s,
                           (float*)w.r, w.partials);
    else
        hipLaunchKernelGGL(k_reprecond<1>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, r32, diag, (float*)w.s,
                           (float*)w.r, w.partials);
    hipLaunchKernelGGL(k_peek, dim3(1), dim3(TBX, TBY), 0, st, w.partials, tl.nb, S, 1);
}

// FP32 iterations n0 .. n0+count-1 and the loop tail of the last, which
// leaves its r.r in the state for the host's reliable-update decision.
void fs_pcg32_chunk(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* diag, float* x, FsPcgWork& w,
                    FsPcgState* S, int n0, int count, FsKernelProbe* probe) {
    iterations_t<float>(st, L, op, diag, x, w, S, n0, count, probe);
}

__global__ void __launch_bounds__(NT) k_add_corr(FsLayout L, Tiling tl, double* __restrict__ x,
                                                 const float* __restrict__ e) {
    for (TileIter it(tl); it.valid(); it.next()) {
        int i0, j, k0;
        tile_origin<1>(tl, it.t, i0, j, k0);
        if (i0 >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + tl.zc, L.nz);
        long c = L.idx(i0, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy) x[c] += (double)e[c];
    }
}

void fs_add_correction(hipStream_t st, const FsLayout& L, double* x, const float* e) {
    Tiling tl = make_tiling(L, 1);
    hipLaunchKernelGGL(k_add_corr, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, x, e);
}

__global__ void k_make_op32(FsLayout L, const double* bx, const double* by, const double* bz, const double* diag,
                            float* bx32, float* by32, float* bz32, float* d32) {
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k = blockIdx.z;
    if (i > L.nx || j > L.ny) return;
    long c = L.idx(i, j, k);
    bx32[c] = (float)bx[c];
    by32[c] = (float)by[c];
    bz32[c] = (float)bz[c];
    d32[c] = (float)diag[c];
}

__global__ void k_to_float(const double* __restrict__ src, float* __restrict__ dst, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

void fs_to_float(hipStream_t st, const double* src, float* dst, long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
}

void fs_make_op32(hipStream_t st, const FsLayout& L, const double* bx, const double* by, const double* bz,
                  const double* diag, float* bx32, float* by32, float* bz32, float* diag32) {
    dim3 g((L.nx + 1 + TBX - 1) / TBX, (L.ny + 1 + TBY - 1) / TBY, L.nz + 1);
    hipLaunchKernelGGL(k_make_op32, g, dim3(TBX, TBY), 0, st, L, bx, by, bz, diag, bx32, by32, bz32, diag32);
}
