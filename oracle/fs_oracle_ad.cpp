// fs_oracle_ad.cpp -- test infrastructure only: the reverse-mode derivative
// of the CPU checker (fs_oracle.c) by operator overloading.
//
// The checker's own source is compiled a second time with every `double`
// replaced by `adouble`, a value with a slot on a tape.  Each arithmetic
// operation and libm call computes its value exactly as the plain build does
// (same IEEE operations, same order, -ffp-contract=off) and records its local
// partial derivatives; one backward sweep over the tape then gives the
// gradient of a scalar with respect to every input (Griewank and Walther,
// "Evaluating Derivatives", 2nd ed., SIAM 2008, ch. 3 and 6).  Branches
// (upwind sides, Godunov cases, the clamp, cell classification, the PCG's
// stop test) follow the recorded values, so the result is the derivative of
// the branch taken at the point -- the same convention as the GPU adjoint.
// The pressure solve is differentiated through its PCG iterations; with the
// solve converged tightly this is the derivative of the solution.
//
// The FP64 reference gradient of a checkpointed rollout: fsa_rollout_gradient
// runs `steps` fixed-dt steps of fso_step from (phi, u, v, w) and the loss
// sum_cells (H(phi) - H(target))^2 with the projection's smoothed Heaviside
// at 1.5 dx (the GPU's fs_levelset_loss), and returns d loss / d (phi, u, v,
// w).  `visc_inplace` selects the explicit viscosity's order as
// fso_set_viscosity_inplace does.  Single-threaded (built without OpenMP:
// one tape).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fs_abi.h"

namespace fsa {
struct Node {
    int a, b;       // argument slots (0: a constant)
    double da, db;  // partial derivatives with respect to them
};
static std::vector<Node> tape(1, Node{0, 0, 0, 0});
static int push(int a, double da, int b, double db) {
    if (a == 0 && b == 0) return 0;
    tape.push_back(Node{a, b, da, db});
    return (int)tape.size() - 1;
}
}  // namespace fsa

struct adouble {
    double v;
    int i;  // tape slot, 0 for a constant
    adouble() = default;
    adouble(double x) : v(x), i(0) {}
    adouble(double x, int slot) : v(x), i(slot) {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
};

static inline adouble operator+(adouble a, adouble b) { return adouble(a.v + b.v, fsa::push(a.i, 1, b.i, 1)); }
static inline adouble operator-(adouble a, adouble b) { return adouble(a.v - b.v, fsa::push(a.i, 1, b.i, -1)); }
static inline adouble operator*(adouble a, adouble b) { return adouble(a.v * b.v, fsa::push(a.i, b.v, b.i, a.v)); }
static inline adouble operator/(adouble a, adouble b) {
    const double q = a.v / b.v;
    return adouble(q, fsa::push(a.i, 1 / b.v, b.i, -q / b.v));
}
static inline adouble operator-(adouble a) { return adouble(-a.v, fsa::push(a.i, -1, 0, 0)); }
static inline adouble operator+(adouble a) { return a; }
static inline adouble& operator+=(adouble& a, adouble b) { return a = a + b; }
static inline adouble& operator-=(adouble& a, adouble b) { return a = a - b; }
static inline adouble& operator*=(adouble& a, adouble b) { return a = a * b; }
static inline adouble& operator/=(adouble& a, adouble b) { return a = a / b; }
static inline bool operator<(adouble a, adouble b) { return a.v < b.v; }
static inline bool operator>(adouble a, adouble b) { return a.v > b.v; }
static inline bool operator<=(adouble a, adouble b) { return a.v <= b.v; }
static inline bool operator>=(adouble a, adouble b) { return a.v >= b.v; }
static inline bool operator==(adouble a, adouble b) { return a.v == b.v; }
static inline bool operator!=(adouble a, adouble b) { return a.v != b.v; }
// sqrt at 0 takes the zero subgradient (a vanishing Godunov norm or GFM
// distance stays finite)
static inline adouble sqrt(adouble x) {
    const double r = std::sqrt(x.v);
    return adouble(r, fsa::push(x.i, r > 0 ? 0.5 / r : 0, 0, 0));
}
static inline adouble sin(adouble x) { return adouble(std::sin(x.v), fsa::push(x.i, std::cos(x.v), 0, 0)); }
static inline adouble cos(adouble x) { return adouble(std::cos(x.v), fsa::push(x.i, -std::sin(x.v), 0, 0)); }
static inline adouble fabs(adouble x) { return adouble(std::fabs(x.v), fsa::push(x.i, x.v < 0 ? -1 : 1, 0, 0)); }
static inline adouble floor(adouble x) { return adouble(std::floor(x.v)); }
static inline adouble fmin(adouble a, adouble b) { return b.v < a.v || std::isnan(a.v) ? b : a; }
static inline adouble fmax(adouble a, adouble b) { return b.v > a.v || std::isnan(a.v) ? b : a; }
static inline bool isfinite(adouble x) { return std::isfinite(x.v); }

#define FSO_VALUE(x) ((x).v)
#define double adouble
#include "fs_oracle.c"
#undef double

namespace {
// a fresh input slot for each value
void inputs(std::vector<adouble>& out, const double* src, long n) {
    out.resize((size_t)n);
    for (long q = 0; q < n; q++) {
        fsa::tape.push_back(fsa::Node{0, 0, 0, 0});
        out[(size_t)q] = adouble(src[q], (int)fsa::tape.size() - 1);
    }
}
void gather(const std::vector<double>& adj, const std::vector<adouble>& in, double* g) {
    for (size_t q = 0; q < in.size(); q++) g[q] = adj[(size_t)in[q].i];
}
}  // namespace

extern "C" __attribute__((visibility("default"))) int fsa_rollout_gradient(
    const fs_params* P, const fs_grid* g, const double* phi, const double* u, const double* v, const double* w,
    const double* p, int visc_inplace, int steps, double dt, const double* phi_target, double* loss, double* gphi,
    double* gu, double* gv, double* gw, long* tape_nodes) {
    if (!P || !g || !phi || !u || !v || !w || !phi_target || !loss || steps < 0 || !(dt > 0)) return FS_ERR_BAD_ARG;
    fsa::tape.assign(1, fsa::Node{0, 0, 0, 0});
    fso_state* s = nullptr;
    int rc = fso_create(P, g, &s);
    if (rc != FS_OK) return rc;
    fso_set_viscosity_inplace(s, visc_inplace);
    const long nx = g->nx, ny = g->ny, nz = g->nz, nc = nx * ny * nz;
    std::vector<adouble> aphi, au, av, aw, ap((size_t)nc, adouble(0.0));
    inputs(aphi, phi, nc);
    inputs(au, u, (nx + 1) * ny * nz);
    inputs(av, v, nx * (ny + 1) * nz);
    inputs(aw, w, nx * ny * (nz + 1));
    if (p)
        for (long q = 0; q < nc; q++) ap[(size_t)q] = adouble(p[q]);  // the warm start carries no derivative
    rc = fso_set_fields(s, aphi.data(), au.data(), av.data(), aw.data(), ap.data());
    int step_rc = FS_OK;
    for (int k = 0; k < steps && rc == FS_OK; k++) {
        int r = fso_step(s, adouble(dt), nullptr);
        if (r != FS_OK && step_rc == FS_OK) step_rc = r;
    }
    if (rc == FS_OK) {
        const adouble eps = (adouble)1.5 * s->cell.x.d;
        adouble total(0.0);
        for (int k = 0; k < s->nz; k++)
            for (int j = 0; j < s->ny; j++)
                for (int i = 0; i < s->nx; i++) {
                    adouble d = heaviside_proj(A3(s->phi, i, j, k), eps) -
                                heaviside_proj(adouble(phi_target[((long)k * ny + j) * nx + i]), eps);
                    total = total + d * d;
                }
        *loss = total.v;
        std::vector<double> adj(fsa::tape.size(), 0.0);
        if (total.i) adj[(size_t)total.i] = 1;
        for (size_t n = fsa::tape.size() - 1; n > 0; n--) {
            const double a = adj[n];
            if (a == 0) continue;
            const fsa::Node& e = fsa::tape[n];
            if (e.a) adj[(size_t)e.a] += e.da * a;
            if (e.b) adj[(size_t)e.b] += e.db * a;
        }
        if (gphi) gather(adj, aphi, gphi);
        if (gu) gather(adj, au, gu);
        if (gv) gather(adj, av, gv);
        if (gw) gather(adj, aw, gw);
    }
    if (tape_nodes) *tape_nodes = (long)fsa::tape.size();
    fso_destroy(s);
    fsa::tape.assign(1, fsa::Node{0, 0, 0, 0});
    fsa::tape.shrink_to_fit();
    return rc != FS_OK ? rc : step_rc;
}

// The forward alone through the active type (values only), for checking
// that it reproduces the plain build bit for bit.
extern "C" __attribute__((visibility("default"))) int fsa_rollout_values(const fs_params* P, const fs_grid* g,
                                                                          double* phi, double* u, double* v,
                                                                          double* w, double* p, int visc_inplace,
                                                                          int steps, double dt) {
    if (!P || !g || !phi || !u || !v || !w || !p || steps < 0 || !(dt > 0)) return FS_ERR_BAD_ARG;
    fsa::tape.assign(1, fsa::Node{0, 0, 0, 0});
    fso_state* s = nullptr;
    int rc = fso_create(P, g, &s);
    if (rc != FS_OK) return rc;
    fso_set_viscosity_inplace(s, visc_inplace);
    const long nx = g->nx, ny = g->ny, nz = g->nz;
    const long n[5] = {nx * ny * nz, (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1), nx * ny * nz};
    double* f[5] = {phi, u, v, w, p};
    std::vector<adouble> a[5];
    for (int c = 0; c < 5; c++) inputs(a[c], f[c], n[c]);
    rc = fso_set_fields(s, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data());
    int step_rc = FS_OK;
    for (int k = 0; k < steps && rc == FS_OK; k++) {
        int r = fso_step(s, adouble(dt), nullptr);
        if (r != FS_OK && step_rc == FS_OK) step_rc = r;
    }
    if (rc == FS_OK) {
        rc = fso_get_fields(s, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data());
        for (int c = 0; c < 5; c++)
            for (long q = 0; q < n[c]; q++) f[c][q] = a[c][(size_t)q].v;
    }
    fso_destroy(s);
    fsa::tape.assign(1, fsa::Node{0, 0, 0, 0});
    fsa::tape.shrink_to_fit();
    return rc != FS_OK ? rc : step_rc;
}
