// fs_pcg.hip -- matrix-free Jacobi-PCG kernels (FP64 and FP32 vectors).
//
// Layout: the padded grid of fs_layout.h; every vector keeps zero ghosts so
// the 7-point stencil needs no bounds tests.  Work unit: a tile of 64 x 4
// cells in x-y marching ZC planes in z (p(k-1), p(k), p(k+1) and
// beta_z(k), beta_z(k+1) stay in registers).  A fixed grid of NB workgroups
// walks the tiles; workgroup b lives on XCD b % 8 and takes tiles from that
// XCD's contiguous share, so y-neighbour rows are reused in one L2.
//
// Operator row (CSR_MATRIX::Multiply, CSR_MATRIX.h:236-260, on the rows of
// POISSON_SOLVER_3D.h:1088-1213): entries in the order l, r, b, t, d, u,
// diagonal; off-diagonals -beta*inv_d2; A(c,c) = inv_dx2 * sum of the six
// betas of full neighbours (+ the Neumann pin on cell (0,0,0)).  Faces to
// non-full neighbours carry beta = 0, which adds an exact -0 term.
//
// Reductions use no device-scope fence or atomic: each workgroup stores its
// partial sum (fixed wave/LDS tree), and every workgroup of the consuming
// kernel re-reduces the NB partials in the same fixed order, so all of them
// see the same scalar and results do not depend on scheduling.  Dot products
// accumulate in FP64 for both precisions.
#include "fs_pcg.h"

#define TBX 64
#define TBY 4
#define ZC 8
#define NT (TBX * TBY)

// partial-sum slots inside FsPcgWork::partials
#define SLOT_PQ 0
#define SLOT_RS 1
#define SLOT_RR 2
#define SLOT_BB 3

struct Tiling {
    int ntx, nty, ntz, ntiles, nb;
};

static Tiling make_tiling(const FsLayout& L) {
    Tiling t;
    t.ntx = (L.nx + TBX - 1) / TBX;
    t.nty = (L.ny + TBY - 1) / TBY;
    t.ntz = (L.nz + ZC - 1) / ZC;
    t.ntiles = t.ntx * t.nty * t.ntz;
    int nb = t.ntiles < FS_PCG_MAX_BLOCKS ? t.ntiles : FS_PCG_MAX_BLOCKS;
    if (nb >= 8) nb = nb / 8 * 8;
    t.nb = nb;
    return t;
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

__device__ __forceinline__ void tile_origin(const Tiling& tl, int t, int& i, int& j, int& k0) {
    int tx = t % tl.ntx, r = t / tl.ntx;
    int ty = r % tl.nty, tz = r / tl.nty;
    i = tx * TBX + threadIdx.x;
    j = ty * TBY + threadIdx.y;
    k0 = tz * ZC;
}

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

// Store this workgroup's partial of slot s (plain store; consumers run in a
// later kernel).
__device__ __forceinline__ void put_partial(double* part, int slot, double v, double* lds) {
    double t = wg_sum_all(v, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[slot * FS_PCG_MAX_BLOCKS + blockIdx.x] = t;
}

// Total of slot s over the nb workgroups of the producing kernel.
__device__ __forceinline__ double get_total(const double* part, int slot, int nb, double* lds) {
    int tid = threadIdx.y * TBX + threadIdx.x;
    double a = 0;
    for (int b = tid; b < nb; b += NT) a += part[slot * FS_PCG_MAX_BLOCKS + b];
    return wg_sum_all(a, lds);
}

template <typename T>
struct OpT {
    const T* __restrict__ bx;
    const T* __restrict__ by;
    const T* __restrict__ bz;
    T idx2, idy2, idz2;
};

// q = A x at cell c; xm/xc/xp = x at k-1, k, k+1; bzd/bzu = beta_z below/above.
template <typename T>
__device__ __forceinline__ T apply_row(const OpT<T>& op, const T* __restrict__ x, long c, long sx, T xm, T xc, T xp,
                                       T bzd, T bzu, bool pin, T pin_b, T pin_d) {
    T bl = op.bx[c], br = op.bx[c + 1], bb = op.by[c], bt = op.by[c + sx];
    T coef = 0;
    coef += bl;
    coef += br;
    coef += bb;
    coef += bt;
    coef += bzd;
    coef += bzu;
    if (pin) {
        coef += pin_b;
        coef += pin_d;
    }
    if (coef == 0) coef = 1;
    T v = 0;
    v += -bl * op.idx2 * x[c - 1];
    v += -br * op.idx2 * x[c + 1];
    v += -bb * op.idy2 * x[c - sx];
    v += -bt * op.idy2 * x[c + sx];
    v += -bzd * op.idz2 * xm;
    v += -bzu * op.idz2 * xp;
    v += op.idx2 * coef * xc;
    return v;
}

// FP32 inner operator in flux form: sum over faces of beta*(x_c - x_n)/d^2
// (+ the pin on cell (0,0,0)).  Constants are then an exact null vector of
// every non-pinned row in any precision, as they are for the FP64 matrix up
// to its own rounding; the diagonal-sum form rounded to FP32 would shift the
// near-null eigenvalue of this singly-pinned Neumann system by O(eps32) per
// row and stall the refinement.  Non-cubic cells keep the FP64 row's
// scaling (1/dx^2 on the diagonal for every face).
__device__ __forceinline__ float apply_row_flux(const OpT<float>& op, const float* __restrict__ x, long c, long sx,
                                                float xm, float xc, float xp, float bzd, float bzu, bool pin,
                                                float pin_b, float pin_d, bool cubic) {
    float bl = op.bx[c], br = op.bx[c + 1], bb = op.by[c], bt = op.by[c + sx];
    float v;
    if (cubic) {
        float s = op.idx2;
        v = bl * (s * (xc - x[c - 1])) + br * (s * (xc - x[c + 1])) + bb * (s * (xc - x[c - sx])) +
            bt * (s * (xc - x[c + sx])) + bzd * (s * (xc - xm)) + bzu * (s * (xc - xp));
    } else {
        float cx = op.idx2 * xc;
        v = bl * (cx - op.idx2 * x[c - 1]) + br * (cx - op.idx2 * x[c + 1]) + bb * (cx - op.idy2 * x[c - sx]) +
            bt * (cx - op.idy2 * x[c + sx]) + bzd * (cx - op.idz2 * xm) + bzu * (cx - op.idz2 * xp);
    }
    if (pin) v += (pin_b + pin_d) * (op.idx2 * xc);
    return v;
}

template <typename T>
__device__ __forceinline__ T row(const OpT<T>& op, const T* __restrict__ x, long c, long sx, T xm, T xc, T xp, T bzd,
                                 T bzu, bool pin, T pin_b, T pin_d) {
    return apply_row<T>(op, x, c, sx, xm, xc, xp, bzd, bzu, pin, pin_b, pin_d);
}
template <>
__device__ __forceinline__ float row<float>(const OpT<float>& op, const float* __restrict__ x, long c, long sx,
                                            float xm, float xc, float xp, float bzd, float bzu, bool pin, float pin_b,
                                            float pin_d) {
    bool cubic = op.idx2 == op.idy2 && op.idx2 == op.idz2;
    return apply_row_flux(op, x, c, sx, xm, xc, xp, bzd, bzu, pin, pin_b, pin_d, cubic);
}

template <typename T>
__device__ __forceinline__ OpT<T> make_op(const FsPcgOp& o) {
    OpT<T> r;
    r.bx = (const T*)o.bx;
    r.by = (const T*)o.by;
    r.bz = (const T*)o.bz;
    r.idx2 = (T)o.idx2;
    r.idy2 = (T)o.idy2;
    r.idz2 = (T)o.idz2;
    return r;
}

// ---- begin: r = b - A x, s = r/D, p = s; partials of r.s, r.r, b.b -------
template <typename T>
__global__ void __launch_bounds__(NT) k_begin(FsLayout L, Tiling tl, FsPcgOp o, const T* __restrict__ b,
                                              const T* __restrict__ D, const T* __restrict__ x, T* __restrict__ r,
                                              T* __restrict__ s, T* __restrict__ p, double* part, FsPcgState* S) {
    __shared__ double lds[4];
    OpT<T> op = make_op<T>(o);
    double a_rs = 0, a_rr = 0, a_bb = 0;
    T pin_b = (T)S->pin_b, pin_d = (T)S->pin_d;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i, j, k0;
        tile_origin(tl, it.t, i, j, k0);
        if (i >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        T xm = x[c - L.sy], xc = x[c], bzd = op.bz[c];
        for (int k = k0; k < k1; k++, c += L.sy) {
            T xp = x[c + L.sy], bzu = op.bz[c + L.sy];
            T ax = row<T>(op, x, c, L.sx, xm, xc, xp, bzd, bzu, (i | j | k) == 0, pin_b, pin_d);
            T rv = b[c] - ax;
            T sv = rv / D[c];
            r[c] = rv;
            s[c] = sv;
            p[c] = sv;
            a_rs += (double)rv * (double)sv;
            a_rr += (double)rv * (double)rv;
            a_bb += (double)b[c] * (double)b[c];
            xm = xc;
            xc = xp;
            bzd = bzu;
        }
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

// ---- K1: q = A p, partial p.q ---------------------------------------------
template <typename T>
__global__ void __launch_bounds__(NT) k_spmv_dot(FsLayout L, Tiling tl, FsPcgOp o, const T* __restrict__ p,
                                                 T* __restrict__ q, double* part, const FsPcgState* S) {
    if (S->done) return;
    __shared__ double lds[4];
    OpT<T> op = make_op<T>(o);
    double acc = 0;
    T pin_b = (T)S->pin_b, pin_d = (T)S->pin_d;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i, j, k0;
        tile_origin(tl, it.t, i, j, k0);
        if (i >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        T xm = p[c - L.sy], xc = p[c], bzd = op.bz[c];
        for (int k = k0; k < k1; k++, c += L.sy) {
            T xp = p[c + L.sy], bzu = op.bz[c + L.sy];
            T v = row<T>(op, p, c, L.sx, xm, xc, xp, bzd, bzu, (i | j | k) == 0, pin_b, pin_d);
            q[c] = v;
            acc += (double)xc * (double)v;
            xm = xc;
            xc = xp;
            bzd = bzu;
        }
    }
    put_partial(part, SLOT_PQ, acc, lds);
}

// ---- K2: alpha = rz/pq; x += a p, r -= a q, s = r/D; partials r.s, r.r ----
template <typename T>
__global__ void __launch_bounds__(NT) k_update(FsLayout L, Tiling tl, const T* __restrict__ D, T* __restrict__ x,
                                               T* __restrict__ r, T* __restrict__ s, const T* __restrict__ p,
                                               const T* __restrict__ q, double* part, const FsPcgState* S, int n) {
    if (S->done) return;
    __shared__ double lds[4];
    double pq = get_total(part, SLOT_PQ, tl.nb, lds);
    T alpha = (T)(S->rz[n & 1] / pq);
    double a_rs = 0, a_rr = 0;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i, j, k0;
        tile_origin(tl, it.t, i, j, k0);
        if (i >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy) {
            x[c] += alpha * p[c];
            T rv = r[c];
            rv -= alpha * q[c];
            r[c] = rv;
            T sv = rv / D[c];
            s[c] = sv;
            a_rs += (double)rv * (double)sv;
            a_rr += (double)rv * (double)rv;
        }
    }
    put_partial(part, SLOT_RS, a_rs, lds);
    put_partial(part, SLOT_RR, a_rr, lds);
}

// ---- K3: stop test, beta = rz_new/rz, p = p*beta + s ----------------------
// Workgroup 0 records the outcome for the host and the next iteration.
template <typename T>
__global__ void __launch_bounds__(NT) k_pupdate(FsLayout L, Tiling tl, T* __restrict__ p, const T* __restrict__ s,
                                                const double* part, FsPcgState* S, int n) {
    if (S->done) return;
    __shared__ double lds[4];
    double rs = get_total(part, SLOT_RS, tl.nb, lds);
    double rr = get_total(part, SLOT_RR, tl.nb, lds);
    double rz_old = S->rz[n & 1];
    bool stop = S->stop_rule == 1 ? (rr < S->rtol2 * S->bb) : (rs < S->tol2);
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        S->rs_last = rs;
        S->rr = rr;
        if (stop) {
            S->iter = n;
            S->done = 1;
        } else {
            S->rz[(n + 1) & 1] = rs;
            S->iter = n + 1;
            if (n + 1 >= S->max_iter) S->done = 2;
        }
    }
    if (stop) return;
    T beta = (T)(rs / rz_old);
    for (TileIter it(tl); it.valid(); it.next()) {
        int i, j, k0;
        tile_origin(tl, it.t, i, j, k0);
        if (i >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy) {
            T pv = p[c];
            pv *= beta;
            pv += s[c];
            p[c] = pv;
        }
    }
}

template <typename T>
static void begin_t(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* rhs, const T* diag, T* x,
                    FsPcgWork& w, FsPcgState* S) {
    Tiling tl = make_tiling(L);
    hipLaunchKernelGGL(k_begin<T>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, rhs, diag, x, (T*)w.r, (T*)w.s,
                       (T*)w.p, w.partials, S);
    hipLaunchKernelGGL(k_begin_final, dim3(1), dim3(TBX, TBY), 0, st, w.partials, tl.nb, S);
}

template <typename T>
static void launch_one(hipStream_t st, const FsLayout& L, const Tiling& tl, const FsPcgOp& op, const T* diag, T* x,
                       FsPcgWork& w, FsPcgState* S, int which, int n) {
    if (which == 0)
        hipLaunchKernelGGL(k_spmv_dot<T>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, (const T*)w.p, (T*)w.q,
                           w.partials, S);
    else if (which == 1)
        hipLaunchKernelGGL(k_update<T>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, diag, x, (T*)w.r, (T*)w.s,
                           (const T*)w.p, (const T*)w.q, w.partials, S, n);
    else
        hipLaunchKernelGGL(k_pupdate<T>, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, (T*)w.p, (const T*)w.s,
                           w.partials, S, n);
}

template <typename T>
static void iterations_t(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* diag, T* x, FsPcgWork& w,
                         FsPcgState* S, int n0, int count) {
    Tiling tl = make_tiling(L);
    for (int n = n0; n < n0 + count; n++)
        for (int which = 0; which < 3; which++) launch_one<T>(st, L, tl, op, diag, x, w, S, which, n);
}

void fs_pcg64_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* rhs, const double* diag,
                    double* x, FsPcgWork& w, FsPcgState* S) {
    begin_t<double>(st, L, op, rhs, diag, x, w, S);
}
void fs_pcg64_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* diag, double* x,
                         FsPcgWork& w, FsPcgState* S, int n0, int count) {
    iterations_t<double>(st, L, op, diag, x, w, S, n0, count);
}
void fs_pcg32_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* rhs, const float* diag,
                    float* x, FsPcgWork& w, FsPcgState* S) {
    begin_t<float>(st, L, op, rhs, diag, x, w, S);
}
void fs_pcg32_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* diag, float* x,
                         FsPcgWork& w, FsPcgState* S, int n0, int count) {
    iterations_t<float>(st, L, op, diag, x, w, S, n0, count);
}

void fs_pcg_time_kernels(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const void* diag, void* x,
                         FsPcgWork& w, FsPcgState* S, int precision, int which, int n) {
    Tiling tl = make_tiling(L);
    if (precision == 0)
        launch_one<double>(st, L, tl, op, (const double*)diag, (double*)x, w, S, which, n);
    else
        launch_one<float>(st, L, tl, op, (const float*)diag, (float*)x, w, S, which, n);
}

// ---- plain operator application (FP64) ----------------------------------
__global__ void __launch_bounds__(NT) k_apply64(FsLayout L, Tiling tl, FsPcgOp o, const double* __restrict__ x,
                                                double* __restrict__ q, const FsPcgState* S) {
    OpT<double> op = make_op<double>(o);
    for (TileIter it(tl); it.valid(); it.next()) {
        int i, j, k0;
        tile_origin(tl, it.t, i, j, k0);
        if (i >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        double xm = x[c - L.sy], xc = x[c], bzd = op.bz[c];
        for (int k = k0; k < k1; k++, c += L.sy) {
            double xp = x[c + L.sy], bzu = op.bz[c + L.sy];
            q[c] = apply_row<double>(op, x, c, L.sx, xm, xc, xp, bzd, bzu, (i | j | k) == 0, S->pin_b, S->pin_d);
            xm = xc;
            xc = xp;
            bzd = bzu;
        }
    }
}

void fs_apply_op64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* x, double* q,
                   const FsPcgState* S) {
    Tiling tl = make_tiling(L);
    hipLaunchKernelGGL(k_apply64, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, x, q, S);
}

// ---- refinement pieces ----------------------------------------------------
__global__ void __launch_bounds__(NT) k_residual64(FsLayout L, Tiling tl, FsPcgOp o, const double* __restrict__ b,
                                                   const double* __restrict__ x, double* __restrict__ r,
                                                   float* __restrict__ r32, double* part, const FsPcgState* S) {
    __shared__ double lds[4];
    OpT<double> op = make_op<double>(o);
    double a_rr = 0, a_bb = 0;
    for (TileIter it(tl); it.valid(); it.next()) {
        int i, j, k0;
        tile_origin(tl, it.t, i, j, k0);
        if (i >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        double xm = x[c - L.sy], xc = x[c], bzd = op.bz[c];
        for (int k = k0; k < k1; k++, c += L.sy) {
            double xp = x[c + L.sy], bzu = op.bz[c + L.sy];
            double ax = apply_row<double>(op, x, c, L.sx, xm, xc, xp, bzd, bzu, (i | j | k) == 0, S->pin_b, S->pin_d);
            double rv = b[c] - ax;
            r[c] = rv;
            r32[c] = (float)rv;
            a_rr += rv * rv;
            a_bb += b[c] * b[c];
            xm = xc;
            xc = xp;
            bzd = bzu;
        }
    }
    put_partial(part, SLOT_RR, a_rr, lds);
    put_partial(part, SLOT_BB, a_bb, lds);
}

__global__ void __launch_bounds__(NT) k_residual_final(const double* part, int nb, FsPcgState* S) {
    __shared__ double lds[4];
    double rr = get_total(part, SLOT_RR, nb, lds);
    double bb = get_total(part, SLOT_BB, nb, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        S->rr = rr;
        S->bb = bb;
    }
}

void fs_residual64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* b, const double* x, double* r,
                   float* r32, FsPcgWork& w, FsPcgState* S) {
    Tiling tl = make_tiling(L);
    hipLaunchKernelGGL(k_residual64, dim3(tl.nb), dim3(TBX, TBY), 0, st, L, tl, op, b, x, r, r32, w.partials, S);
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(TBX, TBY), 0, st, w.partials, tl.nb, S);
}

__global__ void __launch_bounds__(NT) k_add_corr(FsLayout L, Tiling tl, double* __restrict__ x,
                                                 const float* __restrict__ e) {
    for (TileIter it(tl); it.valid(); it.next()) {
        int i, j, k0;
        tile_origin(tl, it.t, i, j, k0);
        if (i >= L.nx || j >= L.ny) continue;
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy) x[c] += (double)e[c];
    }
}

void fs_add_correction(hipStream_t st, const FsLayout& L, double* x, const float* e) {
    Tiling tl = make_tiling(L);
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

void fs_make_op32(hipStream_t st, const FsLayout& L, const double* bx, const double* by, const double* bz,
                  const double* diag, float* bx32, float* by32, float* bz32, float* diag32) {
    dim3 g((L.nx + 1 + TBX - 1) / TBX, (L.ny + 1 + TBY - 1) / TBY, L.nz + 1);
    hipLaunchKernelGGL(k_make_op32, g, dim3(TBX, TBY), 0, st, L, bx, by, bz, diag, bx32, by32, bz32, diag32);
}
