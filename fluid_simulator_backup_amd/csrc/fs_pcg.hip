// fs_pcg.hip -- matrix-free Jacobi-PCG kernels (FP64 and FP32 vectors).
//
// Layout: the padded grid of fs_layout.h; every vector keeps zero ghosts so
// the 7-point stencil needs no bounds tests.  A workgroup is 64 x 4 threads
// over an x-y tile and marches ZC planes in z, keeping p(k-1), p(k), p(k+1)
// and beta_z(k), beta_z(k+1) in registers.
//
// Operator row (CSR_MATRIX::Multiply, CSR_MATRIX.h:236-260, on the rows of
// POISSON_SOLVER_3D.h:1088-1213): entries in the order l, r, b, t, d, u,
// diagonal; off-diagonals -beta*inv_d2; A(c,c) = inv_dx2 * sum of the six
// betas of full neighbours (+ the Neumann pin on cell (0,0,0)).  Faces to
// non-full neighbours carry beta = 0, which adds an exact -0 term.
//
// Reductions are deterministic: fixed wave/LDS tree per workgroup, block
// partials stored, and the last workgroup (atomic ticket) sums them in
// block order.  Dot products accumulate in FP64 for both precisions.
#include "fs_pcg.h"

#define TBX 64
#define TBY 4
#define ZC 8
#define NT (TBX * TBY)

int fs_pcg_blocks(const FsLayout& L) {
    return ((L.nx + TBX - 1) / TBX) * ((L.ny + TBY - 1) / TBY) * ((L.nz + ZC - 1) / ZC);
}
static dim3 pcg_grid(const FsLayout& L) {
    return dim3((L.nx + TBX - 1) / TBX, (L.ny + TBY - 1) / TBY, (L.nz + ZC - 1) / ZC);
}

__device__ __forceinline__ void store_partial(double* part, int b, double v) {
    __hip_atomic_store(&part[b], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_partial(const double* part, int b) {
    return __hip_atomic_load(&part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of v over the workgroup, valid in thread 0; fixed order.
__device__ __forceinline__ double wg_sum(double v, double* lds) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    int tid = threadIdx.y * TBX + threadIdx.x;
    __syncthreads();
    if ((tid & 63) == 0) lds[tid >> 6] = v;
    __syncthreads();
    double t = 0;
    if (tid == 0) t = ((lds[0] + lds[1]) + lds[2]) + lds[3];
    return t;
}

// NV values per thread reduced across the grid.  Returns true in thread 0 of
// the last workgroup, with the grid totals in out[].
template <int NV>
__device__ bool grid_sums(const double (&v)[NV], double* part, int nblocks, unsigned int* counter, double* out) {
    __shared__ double lds[4];
    __shared__ int last;
    int tid = threadIdx.y * TBX + threadIdx.x;
    int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    double t[NV];
    for (int n = 0; n < NV; n++) t[n] = wg_sum(v[n], lds);
    if (tid == 0) {
        for (int n = 0; n < NV; n++) store_partial(part, n * nblocks + bid, t[n]);
        __threadfence();
        unsigned int prev = atomicAdd(counter, 1u);
        last = (prev == (unsigned int)nblocks - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    for (int n = 0; n < NV; n++) {
        double a = 0;
        for (int b = tid; b < nblocks; b += NT) a += load_partial(part, n * nblocks + b);
        out[n] = wg_sum(a, lds);
    }
    if (tid == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return tid == 0;
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

// ---- begin: r = b - A x, s = r/D, p = s, rz = r.s, rr = r.r, bb = b.b ------
template <typename T>
__global__ void __launch_bounds__(NT) k_begin(FsLayout L, FsPcgOp o, const T* __restrict__ b, const T* __restrict__ D,
                                              const T* __restrict__ x, T* __restrict__ r, T* __restrict__ s,
                                              T* __restrict__ p, double* part, int nblocks, FsPcgState* S) {
    OpT<T> op = make_op<T>(o);
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k0 = blockIdx.z * ZC;
    double acc[3] = {0, 0, 0};
    T pin_b = (T)S->pin_b, pin_d = (T)S->pin_d;
    if (i < L.nx && j < L.ny) {
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
            acc[0] += (double)rv * (double)sv;
            acc[1] += (double)rv * (double)rv;
            acc[2] += (double)b[c] * (double)b[c];
            xm = xc;
            xc = xp;
            bzd = bzu;
        }
    }
    double out[3];
    if (grid_sums<3>(acc, part, nblocks, &S->counter[0], out)) {
        S->rz = out[0];
        S->rr = out[1];
        S->bb = out[2];
        S->iter = 0;
        S->done = 0;
        if (S->stop_rule == 1 && (out[2] == 0 || out[1] <= S->rtol2 * out[2])) S->done = 1;
        if (S->max_iter <= 0) S->done = 2;
    }
}

// ---- K1: q = A p, pq = p.q, alpha = rz/pq ---------------------------------
template <typename T>
__global__ void __launch_bounds__(NT) k_spmv_dot(FsLayout L, FsPcgOp o, const T* __restrict__ p, T* __restrict__ q,
                                                 double* part, int nblocks, FsPcgState* S) {
    if (S->done) return;
    OpT<T> op = make_op<T>(o);
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k0 = blockIdx.z * ZC;
    double acc[1] = {0};
    T pin_b = (T)S->pin_b, pin_d = (T)S->pin_d;
    if (i < L.nx && j < L.ny) {
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        T xm = p[c - L.sy], xc = p[c], bzd = op.bz[c];
        for (int k = k0; k < k1; k++, c += L.sy) {
            T xp = p[c + L.sy], bzu = op.bz[c + L.sy];
            T v = row<T>(op, p, c, L.sx, xm, xc, xp, bzd, bzu, (i | j | k) == 0, pin_b, pin_d);
            q[c] = v;
            acc[0] += (double)xc * (double)v;
            xm = xc;
            xc = xp;
            bzd = bzu;
        }
    }
    double out[1];
    if (grid_sums<1>(acc, part, nblocks, &S->counter[1], out)) {
        S->pq = out[0];
        S->alpha = S->rz / out[0];
    }
}

// ---- K2: x += a p, r -= a q, s = r/D, rs, rr; stop test; beta ------------
template <typename T>
__global__ void __launch_bounds__(NT) k_update(FsLayout L, const T* __restrict__ D, T* __restrict__ x,
                                               T* __restrict__ r, T* __restrict__ s, const T* __restrict__ p,
                                               const T* __restrict__ q, double* part, int nblocks, FsPcgState* S) {
    if (S->done) return;
    T alpha = (T)S->alpha;
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k0 = blockIdx.z * ZC;
    double acc[2] = {0, 0};
    if (i < L.nx && j < L.ny) {
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        for (int k = k0; k < k1; k++, c += L.sy) {
            x[c] += alpha * p[c];
            T rv = r[c];
            rv -= alpha * q[c];
            r[c] = rv;
            T sv = rv / D[c];
            s[c] = sv;
            acc[0] += (double)rv * (double)sv;
            acc[1] += (double)rv * (double)rv;
        }
    }
    double out[2];
    if (grid_sums<2>(acc, part, nblocks, &S->counter[2], out)) {
        S->rz_old = S->rz;
        S->rz = out[0];
        S->rr = out[1];
        bool stop = S->stop_rule == 1 ? (out[1] < S->rtol2 * S->bb) : (out[0] < S->tol2);
        if (stop) {
            S->done = 1;
        } else {
            S->beta = S->rz / S->rz_old;
            S->iter += 1;
            if (S->iter >= S->max_iter) S->done = 2;
        }
    }
}

// ---- K3: p = p*beta + s ---------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(NT) k_pupdate(FsLayout L, T* __restrict__ p, const T* __restrict__ s,
                                                FsPcgState* S) {
    if (S->done) return;
    T beta = (T)S->beta;
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k0 = blockIdx.z * ZC;
    if (i >= L.nx || j >= L.ny) return;
    int k1 = min(k0 + ZC, L.nz);
    long c = L.idx(i, j, k0);
    for (int k = k0; k < k1; k++, c += L.sy) {
        T pv = p[c];
        pv *= beta;
        pv += s[c];
        p[c] = pv;
    }
}

template <typename T>
static void begin_t(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* rhs, const T* diag, T* x,
                    FsPcgWork& w, FsPcgState* S) {
    hipLaunchKernelGGL(k_begin<T>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, op, rhs, diag, x, (T*)w.r, (T*)w.s,
                       (T*)w.p, w.partials, w.nblocks, S);
}

template <typename T>
static void iterations_t(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const T* diag, T* x, FsPcgWork& w,
                         FsPcgState* S, int count) {
    for (int n = 0; n < count; n++) {
        hipLaunchKernelGGL(k_spmv_dot<T>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, op, (const T*)w.p, (T*)w.q,
                           w.partials, w.nblocks, S);
        hipLaunchKernelGGL(k_update<T>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, diag, x, (T*)w.r, (T*)w.s,
                           (const T*)w.p, (const T*)w.q, w.partials, w.nblocks, S);
        hipLaunchKernelGGL(k_pupdate<T>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, (T*)w.p, (const T*)w.s, S);
    }
}

void fs_pcg64_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* rhs, const double* diag,
                    double* x, FsPcgWork& w, FsPcgState* S) {
    begin_t<double>(st, L, op, rhs, diag, x, w, S);
}
void fs_pcg64_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* diag, double* x,
                         FsPcgWork& w, FsPcgState* S, int count) {
    iterations_t<double>(st, L, op, diag, x, w, S, count);
}
void fs_pcg32_begin(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* rhs, const float* diag,
                    float* x, FsPcgWork& w, FsPcgState* S) {
    begin_t<float>(st, L, op, rhs, diag, x, w, S);
}
void fs_pcg32_iterations(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const float* diag, float* x,
                         FsPcgWork& w, FsPcgState* S, int count) {
    iterations_t<float>(st, L, op, diag, x, w, S, count);
}

// ---- plain operator application (FP64) ----------------------------------
__global__ void __launch_bounds__(NT) k_apply64(FsLayout L, FsPcgOp o, const double* __restrict__ x,
                                                double* __restrict__ q, const FsPcgState* S) {
    OpT<double> op = make_op<double>(o);
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k0 = blockIdx.z * ZC;
    if (i >= L.nx || j >= L.ny) return;
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

void fs_apply_op64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* x, double* q,
                   const FsPcgState* S) {
    hipLaunchKernelGGL(k_apply64, pcg_grid(L), dim3(TBX, TBY), 0, st, L, op, x, q, S);
}

// ---- refinement pieces ----------------------------------------------------
__global__ void __launch_bounds__(NT) k_residual64(FsLayout L, FsPcgOp o, const double* __restrict__ b,
                                                   const double* __restrict__ x, double* __restrict__ r,
                                                   float* __restrict__ r32, double* part, int nblocks,
                                                   FsPcgState* S) {
    OpT<double> op = make_op<double>(o);
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k0 = blockIdx.z * ZC;
    double acc[2] = {0, 0};
    if (i < L.nx && j < L.ny) {
        int k1 = min(k0 + ZC, L.nz);
        long c = L.idx(i, j, k0);
        double xm = x[c - L.sy], xc = x[c], bzd = op.bz[c];
        for (int k = k0; k < k1; k++, c += L.sy) {
            double xp = x[c + L.sy], bzu = op.bz[c + L.sy];
            double ax = apply_row<double>(op, x, c, L.sx, xm, xc, xp, bzd, bzu, (i | j | k) == 0, S->pin_b, S->pin_d);
            double rv = b[c] - ax;
            r[c] = rv;
            r32[c] = (float)rv;
            acc[0] += rv * rv;
            acc[1] += b[c] * b[c];
            xm = xc;
            xc = xp;
            bzd = bzu;
        }
    }
    double out[2];
    if (grid_sums<2>(acc, part, nblocks, &S->counter[3], out)) {
        S->rr = out[0];
        S->bb = out[1];
    }
}

void fs_residual64(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const double* b, const double* x, double* r,
                   float* r32, FsPcgWork& w, FsPcgState* S) {
    hipLaunchKernelGGL(k_residual64, pcg_grid(L), dim3(TBX, TBY), 0, st, L, op, b, x, r, r32, w.partials, w.nblocks,
                       S);
}

__global__ void k_add_corr(FsLayout L, double* __restrict__ x, const float* __restrict__ e) {
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k0 = blockIdx.z * ZC;
    if (i >= L.nx || j >= L.ny) return;
    int k1 = min(k0 + ZC, L.nz);
    long c = L.idx(i, j, k0);
    for (int k = k0; k < k1; k++, c += L.sy) x[c] += (double)e[c];
}

void fs_add_correction(hipStream_t st, const FsLayout& L, double* x, const float* e) {
    hipLaunchKernelGGL(k_add_corr, pcg_grid(L), dim3(TBX, TBY), 0, st, L, x, e);
}

__global__ void k_make_op32(FsLayout L, const double* bx, const double* by, const double* bz, const double* diag,
                            float* bx32, float* by32, float* bz32, float* d32) {
    int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y, k0 = blockIdx.z * ZC;
    if (i > L.nx || j > L.ny) return;
    int k1 = min(k0 + ZC, L.nz + 1);
    long c = L.idx(i, j, k0);
    for (int k = k0; k < k1; k++, c += L.sy) {
        bx32[c] = (float)bx[c];
        by32[c] = (float)by[c];
        bz32[c] = (float)bz[c];
        d32[c] = (float)diag[c];
    }
}

void fs_make_op32(hipStream_t st, const FsLayout& L, const double* bx, const double* by, const double* bz,
                  const double* diag, float* bx32, float* by32, float* bz32, float* diag32) {
    dim3 g((L.nx + 1 + TBX - 1) / TBX, (L.ny + 1 + TBY - 1) / TBY, (L.nz + 1 + ZC - 1) / ZC);
    hipLaunchKernelGGL(k_make_op32, g, dim3(TBX, TBY), 0, st, L, bx, by, bz, diag, bx32, by32, bz32, diag32);
}

// Timing hook: launches kernel `which` (0 SpMV+dot, 1 update, 2 p update)
// once.  The caller has set state->done = 0 and max_iter large.
void fs_pcg_time_kernels(hipStream_t st, const FsLayout& L, const FsPcgOp& op, const void* diag, void* x,
                         FsPcgWork& w, FsPcgState* S, int precision, int which) {
    if (precision == 0) {
        if (which == 0)
            hipLaunchKernelGGL(k_spmv_dot<double>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, op, (const double*)w.p,
                               (double*)w.q, w.partials, w.nblocks, S);
        else if (which == 1)
            hipLaunchKernelGGL(k_update<double>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, (const double*)diag,
                               (double*)x, (double*)w.r, (double*)w.s, (const double*)w.p, (const double*)w.q,
                               w.partials, w.nblocks, S);
        else
            hipLaunchKernelGGL(k_pupdate<double>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, (double*)w.p,
                               (const double*)w.s, S);
    } else {
        if (which == 0)
            hipLaunchKernelGGL(k_spmv_dot<float>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, op, (const float*)w.p,
                               (float*)w.q, w.partials, w.nblocks, S);
        else if (which == 1)
            hipLaunchKernelGGL(k_update<float>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, (const float*)diag,
                               (float*)x, (float*)w.r, (float*)w.s, (const float*)w.p, (const float*)w.q,
                               w.partials, w.nblocks, S);
        else
            hipLaunchKernelGGL(k_pupdate<float>, pcg_grid(L), dim3(TBX, TBY), 0, st, L, (float*)w.p,
                               (const float*)w.s, S);
    }
}
