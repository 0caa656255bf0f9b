// fsgh_user_strength.h -- the user strength model, the counterpart of the
// reference's user_strength_model (K/user_mat.cpp:47-90).  Materials whose
// strength model is FSGH_STRENGTH_HYPO (fsgh_set_material_models) call it
// from the corner-force kernels after the forces are formed, as
// get_force_sgh does (K/force_sgh.cpp:369-391, 2D :774-801); materials with
// FSGH_STRENGTH_HYPER call it from the state update before the EOS
// (K/properties.cpp:46-95).  A maintainer replaces the body and rebuilds.
//
// Inputs: material id, the material's parameters (fsgh_set_strength_params,
// up to 8), density, specific internal energy at t_n (the reference passes
// elem_sie(elem_gid), bin 0 of the [2][E] array), the velocity gradient
// L_ij = du_i/dx_j (row-major 3 x 3; in 2D RZ with L_22 = v_r / r), the
// volume, dt, the RK weight and the stress at t_n (bin 0; the reference's
// model sees both stress bins).  In/out: the element's current Cauchy stress
// (row-major 3 x 3, symmetric), pressure and sound speed.
//
// The shipped body is the reference's: empty.  Built with
// FSGH_TEST_STRENGTH_MODEL (a test-only variant of the library, never the
// product build) the body is a hypoelastic update
//     sigma = sigma_n + rk_alpha dt 2 G dev(D),  D = (L + L^T)/2, G = params[0],
// the model the strength tests run through every call site.
#pragma once

#include <cstdint>

#ifndef FSGH_HD
#define FSGH_HD __host__ __device__
#endif

namespace fsgh {

FSGH_HD inline void user_strength(int32_t mat_id, const double* params, double den, double sie,
                                  const double vel_grad[9], double vol, double dt,
                                  double rk_alpha, const double stress_n[9], double stress[9],
                                  double& pres, double& sspd)
{
    (void)mat_id;
    (void)den;
    (void)sie;
    (void)vol;
    (void)pres;
    (void)sspd;
#ifdef FSGH_TEST_STRENGTH_MODEL
    const double G = params[0];
    const double tr = vel_grad[0] + vel_grad[4] + vel_grad[8];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double d = 0.5 * (vel_grad[3 * i + j] + vel_grad[3 * j + i]) -
                             (i == j ? tr / 3.0 : 0.0);
            stress[3 * i + j] = stress_n[3 * i + j] + rk_alpha * dt * 2.0 * G * d;
        }
#else
    (void)params;
    (void)vel_grad;
    (void)dt;
    (void)rk_alpha;
    (void)stress_n;
    (void)stress;
#endif
}

}  // namespace fsgh
