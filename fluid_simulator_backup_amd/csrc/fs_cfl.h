// fs_cfl.h -- the CFL rate of EULERIAN_FLUID_SOLVER_3D::CFLOneTimeStep
// (EULERIAN_FLUID_SOLVER_3D.h:2673-2785), written once for the host
// (fs_cfl_dt) and the device (fs_cfl_dt_noalloc), so both round alike.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

// The terms that depend on the parameters and the grid only.
struct FsCflTerms {
    double inv_mx, inv_my, inv_mz;  // 1/dx of the x-, y-, z-face grids along their normal
    double g_f, s_f, v_f;           // gravity, surface tension and viscosity terms
};

// 1 / dt of CFL number 1: 0.5 (c + v + sqrt((c + v)^2 + 4 g^2 + 4 s^2)) with
// c = sum_d max|u_d| / dx_d.
__host__ __device__ inline double fs_cfl_rate(const FsCflTerms& t, double mu, double mv, double mw) {
    double c_f = mu * t.inv_mx + mv * t.inv_my + mw * t.inv_mz;
    return (double)0.5 * (c_f + t.v_f + sqrt((c_f + t.v_f) * (c_f + t.v_f) + (double)4 * (t.g_f * t.g_f) +
                                             (double)4 * (t.s_f * t.s_f)));
}

#include "../../include/fs_abi.h"
#include "fs_layout.h"

// The parameter and grid terms of CFLOneTimeStep with zero surface tension
// (the caller adds s_f): c's spacings are those of the MAC grids
// (EULERIAN_FLUID_SOLVER_3D.h:407-423), g = |g_y| as the reference's rule
// reads it (:2717) or |g| with gravity_norm, v from the larger kinematic
// viscosity (:2766), none with the semi-implicit viscosity (:2770-2773).
inline FsCflTerms fs_cfl_terms(const fs_params& P, const fs_grid& G, int gravity_norm) {
    const FsAxis cx = fs_make_axis(G.x_min, G.x_max, G.nx), cy = fs_make_axis(G.y_min, G.y_max, G.ny),
                 cz = fs_make_axis(G.z_min, G.z_max, G.nz);
    const FsAxis mx = fs_make_axis(G.x_min - (double)0.5 * cx.d, G.x_max + (double)0.5 * cx.d, G.nx + 1);
    const FsAxis my = fs_make_axis(G.y_min - (double)0.5 * cy.d, G.y_max + (double)0.5 * cy.d, G.ny + 1);
    const FsAxis mz = fs_make_axis(G.z_min - (double)0.5 * cz.d, G.z_max + (double)0.5 * cz.d, G.nz + 1);
    FsCflTerms t{};
    t.inv_mx = mx.inv;
    t.inv_my = my.inv;
    t.inv_mz = mz.inv;
    if (P.is_external_force_active) {
        double g = gravity_norm ? std::sqrt(P.gravity[0] * P.gravity[0] + P.gravity[1] * P.gravity[1] +
                                            P.gravity[2] * P.gravity[2])
                                : std::fabs(P.gravity[1]);
        t.g_f = std::sqrt(g * my.inv);
    }
    double a = P.water_viscosity / P.water_density, b = P.air_viscosity / P.air_density;
    t.v_f = (a > b ? a : b) * ((double)2 * cx.inv_d2 + (double)2 * cy.inv_d2 + (double)2 * cz.inv_d2);
    if (P.semi_implicit_viscosity) t.v_f = 0;
    return t;
}
