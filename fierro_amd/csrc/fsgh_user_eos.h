// fsgh_user_eos.h -- the user equation of state, the counterpart of the
// reference's user_eos_model (K/user_mat.cpp:14-40).  Elements whose material
// selects FSGH_EOS_USER (fsgh_set_material_models) get their pressure and
// sound speed from user_eos in every kernel that evaluates an EOS: set-up,
// update_state, the fused energy/EOS kernel, and -- since the function is a
// template over the scalar type -- the tangent-linear model and the adjoint,
// which differentiate it with dual numbers.  A maintainer replaces the body
// and rebuilds the library.
//
// Inputs: gamma = statev(e, 0), csmin = statev(e, 1), density, specific
// internal energy, and the element's Cauchy stress (row-major 3 x 3), or
// nullptr where no stress field exists (the stress is then zero).  Outputs:
// pressure and sound speed.
//
// The shipped body is the reference's: pressure = -tr(stress)/3 and sound
// speed 1e-8.  (The reference reads elem_stress(i,i) with two indices of its
// four-index stress view; here the model sees the element's own stress,
// which gives the same result whenever the stress is zero -- always with the
// reference's empty strength model.)
#pragma once

#ifndef FSGH_HD
#define FSGH_HD __host__ __device__
#endif

namespace fsgh {

template <typename T>
FSGH_HD inline void user_eos(T gamma, double csmin, T den, T sie, const double* stress, T& pres,
                             T& sspd)
{
    (void)gamma;
    (void)csmin;
    (void)den;
    (void)sie;
    T p = 0.0;
    if (stress != nullptr)
        for (int i = 0; i < 3; ++i) p = p - stress[4 * i];
    p = p * (1.0 / 3.0);
    pres = p;
    sspd = 1.0e-8;
}

}  // namespace fsgh
