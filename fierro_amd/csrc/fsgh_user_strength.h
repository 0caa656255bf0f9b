// fsgh_user_strength.h -- the user strength model, the counterpart of the
// reference's user_strength_model (K/user_mat.cpp:47-90).  Materials whose
// strength model is FSGH_STRENGTH_HYPO (fsgh_set_material_models) call it
// from the corner-force kernels after the forces are formed, as
// get_force_sgh does (K/force_sgh.cpp:369-391, 2D :774-801); materials with
// FSGH_STRENGTH_HYPER call it from the state update before the EOS
// (K/properties.cpp:46-95).  A maintainer replaces the body and rebuilds.
//
// Inputs: material id, the material's parameters (fsgh_set_strength_params,
// up to 8), density, specific internal energy at t_n, the velocity gradient
// L_ij = du_i/dx_j (row-major 3 x 3; in 2D RZ with L_22 = v_r / r), the
// volume, dt and the RK weight.  In/out: the element's Cauchy stress
// (row-major 3 x 3, symmetric), pressure and sound speed.
//
// The shipped body is the reference's: empty.
#pragma once

#include <cstdint>

#ifndef FSGH_HD
#define FSGH_HD __host__ __device__
#endif

namespace fsgh {

FSGH_HD inline void user_strength(int32_t mat_id, const double* params, double den, double sie,
                                  const double vel_grad[9], double vol, double dt,
                                  double rk_alpha, double stress[9], double& pres, double& sspd)
{
    (void)mat_id;
    (void)params;
    (void)den;
    (void)sie;
    (void)vel_grad;
    (void)vol;
    (void)dt;
    (void)rk_alpha;
    (void)stress;
    (void)pres;
    (void)sspd;
}

}  // namespace fsgh
