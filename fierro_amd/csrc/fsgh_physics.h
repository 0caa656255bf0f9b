// fsgh_physics.h -- the per-element / per-node physics of one SGH RK stage,
// templated on the scalar type (double for the forward kernels, Dual for the
// tangent-linear model).  Every expression keeps the reference's FP64
// operation order (file:line at each function); with -ffp-contract=off the
// double instantiation is bit-identical to the reference arithmetic.
#pragma once

#include <cmath>
#include <cstdint>

#ifndef FSGH_HD
#define FSGH_HD __host__ __device__
#endif
#include "fsgh_dual.h"
#include "fsgh_user_eos.h"
#include "fsgh_user_strength.h"
#include "hex_geometry.inc"

namespace fsgh {

// Taylor-Green vortex initial condition (velocity fill kind 5,
// K/setup.cpp:326-350), with pi as the reference defines it (K/mesh.h:8).
// Node velocity from the node's coordinates; the element's pressure from its
// centroid and the energy that pressure implies with the element's state
// variable 4 taken as gamma (the reference reads elem_statev(e, 4), which is
// zero unless the material sets at least five state variables).
constexpr double kTgPi = 3.141592653589793;
FSGH_HD inline void tg_vortex_velocity(const double* X, bool three_d, double v[3])
{
    v[0] = sin(kTgPi * X[0]) * cos(kTgPi * X[1]);
    v[1] = -1.0 * cos(kTgPi * X[0]) * sin(kTgPi * X[1]);
    if (three_d) v[2] = 0.0;
}
FSGH_HD inline void tg_vortex_state(const double* centroid, double den, double statev4,
                                    double& pres, double& sie)
{
    pres = 0.25 * (cos(2.0 * kTgPi * centroid[0]) + cos(2.0 * kTgPi * centroid[1])) + 1.0;
    sie = pres / (den * (statev4 - 1.0));
}

// velocity divergence from B, v and V (K/momentum.cpp:265-286)
template <typename T>
FSGH_HD inline T divergence(const T* B, const T* u, const T* v, const T* w, T vol)
{
    const T inverse_vol = 1.0 / vol;
    T sx = u[0] * B[0], sy = v[0] * B[1], sz = w[0] * B[2];
#pragma unroll
    for (int a = 1; a < 8; ++a) {
        sx = sx + u[a] * B[3 * a + 0];
        sy = sy + v[a] * B[3 * a + 1];
        sz = sz + w[a] * B[3 * a + 2];
    }
    T div = 0.0;
    div += sx * inverse_vol;
    div += sy * inverse_vol;
    div += sz * inverse_vol;
    return div;
}

// ideal gas (K/eos.cpp:30-43)
template <typename T>
FSGH_HD inline void ideal_gas(T gamma, double csmin, T den, T sie, T& pres, T& sspd)
{
    pres = (gamma - 1.0) * sie * den;
    T c = sqrt(gamma * (gamma - 1.0) * sie);
    if (c < csmin) c = csmin;
    sspd = c;
}


// the material's equation of state (the eos_model pointer of K/state.h:140-147)
constexpr int32_t kEosIdealGas = 0;
constexpr int32_t kEosUser = 1;
// stress: the element's Cauchy stress (row-major 3 x 3) or nullptr (zero)
template <typename T>
FSGH_HD inline void eos(int32_t kind, T gamma, double csmin, T den, T sie, T& pres, T& sspd,
                        const double* stress = nullptr)
{
    if (kind == kEosUser)
        user_eos(gamma, csmin, den, sie, stress, pres, sspd);
    else
        ideal_gas(gamma, csmin, den, sie, pres, sspd);
}

// strength models of a material (the strength_type of K/state.h:166-167)
constexpr int32_t kStrengthNone = 0;
constexpr int32_t kStrengthHypo = 1;
constexpr int32_t kStrengthHyper = 2;

// velocity gradient L_ij = sum_p u_p^i B_pj / V with the outward B
// (K/momentum.cpp:60-148)
template <typename T>
FSGH_HD inline void velgrad(const T* B, const T* u, const T* v, const T* w, T vol, T* L)
{
    const T inverse_vol = 1.0 / vol;
    const T* c[3] = {u, v, w};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T acc = c[i][0] * B[j];
#pragma unroll
            for (int p = 1; p < 8; ++p) acc = acc + c[i][p] * B[3 * p + j];
            L[3 * i + j] = acc * inverse_vol;
        }
}

// Element-only terms of the corner force (K/force_sgh.cpp:217-224, :318-321):
// the impedance coefficients chosen by compression or expansion (qas = qa c,
// qb) and the shock limiter alpha = min(1, 20 V^(1/3) |div| / (c + fuzz)).
// They depend on no node value, so a kernel can form them while its node
// values are still in flight.
template <typename T>
struct ForceCoef {
    T den, pres, qas, sspd, vol, div;
    double qb;
};

template <typename T>
FSGH_HD inline ForceCoef<T> force_coef(T den, T pres, T sspd, T vol, T div, double q1,
                                       double q1ex, double q2, double q2ex, double fuzz)
{
    ForceCoef<T> k;
    const double qa = div < 0.0 ? q1 : q1ex;
    k.qb = div < 0.0 ? q2 : q2ex;
    k.qas = qa * sspd;
    k.sspd = sspd;
    k.vol = vol;
    k.den = den;
    k.pres = pres;
    k.div = div;
    return k;
}

// Corner forces of one element, 3D, strength-free (K/force_sgh.cpp:35-359),
// from the element terms of force_coef.  `ext` is the extreme neighbour
// divergence chosen by the shock detector: min over neighbours when
// div + small >= 0, max otherwise (the kernels find it; r_face is monotone
// in the neighbour divergence).
template <typename T>
FSGH_HD inline void force_elem_coef(const T* x, const T* y, const T* z, const T* u, const T* v,
                                    const T* w, const ForceCoef<T>& k, T ext, double fuzz,
                                    double small, T* f)
{
    T B[24];
    hex_bmatrix(x, y, z, B);
    // MARS: cell-average velocity estimate (:147-164)
    T vsx = 0.0, vsy = 0.0, vsz = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        vsx += 0.125 * u[a];
        vsy += 0.125 * v[a];
        vsz += 0.125 * w[a];
    }
    // impedance x |corner area| per node (:177-255): den (qa c + qb |v - v*|)
    T sum0 = 0.0, sum1 = 0.0, sum2 = 0.0, sum3 = 0.0;
    T mu[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const T dx = u[a] - vsx, dy = v[a] - vsy, dz = w[a] - vsz;
        const T mag_vel = sqrt(dx * dx + dy * dy + dz * dz);
        const T muc = k.den * (k.qas + k.qb * mag_vel);
        const T bx = B[3 * a], by = B[3 * a + 1], bz = B[3 * a + 2];
        const T mu_term = muc * sqrt(bx * bx + by * by + bz * bz);
        sum0 += mu_term * u[a];
        sum1 += mu_term * v[a];
        sum2 += mu_term * w[a];
        sum3 += mu_term;
        mu[a] = mu_term;
    }
    // Riemann velocity (:261-270)
    if (sum3 > fuzz) {
        vsx = sum0 / sum3;
        vsy = sum1 / sum3;
        vsz = sum2 / sum3;
    } else {
        vsx = 0.0;
        vsy = 0.0;
        vsz = 0.0;
    }
    // shock detector (:287-328); pow(phi, 1.0) == phi
    const T b = k.div + small;
    const T r_face = 0.9 * (ext + small) / b;
    const T r_min = fmin(r_face, 1.0);
    T phi = 1.0 - fmax(0.0, r_min);
    const T c_length = pow(k.vol, 1.0 / 3.0);
    const T alpha = fmin(1.0, 20.0 * (c_length * fabs(k.div)) / (k.sspd + fuzz));
    phi = alpha * phi;
    // corner forces (:338-359): a = -B, tau = -p I, so a_i . tau_d = B_id p
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const T pm = phi * mu[a];
        f[3 * a + 0] = B[3 * a + 0] * k.pres + pm * (vsx - u[a]);
        f[3 * a + 1] = B[3 * a + 1] * k.pres + pm * (vsy - v[a]);
        f[3 * a + 2] = B[3 * a + 2] * k.pres + pm * (vsz - w[a]);
    }
}

// Corner forces of one element, 3D, strength-free (K/force_sgh.cpp:35-359).
template <typename T>
FSGH_HD inline void force_elem(const T* x, const T* y, const T* z, const T* u, const T* v,
                               const T* w, T den, T pres, T sspd, T vol, T div, T ext,
                               double q1, double q1ex, double q2, double q2ex, double fuzz,
                               double small, T* f)
{
    const ForceCoef<T> k = force_coef(den, pres, sspd, vol, div, q1, q1ex, q2, q2ex, fuzz);
    force_elem_coef(x, y, z, u, v, w, k, ext, fuzz, small, f);
}

// Corner forces of one element, 3D, with a stress field: tau = stress - p I
// (K/force_sgh.cpp:127-137) and F_i,d = a_i . tau_:,d + phi mu_i (v* - v_i)
// with the inward normals a = -B (:103-107, :338-359), term by term as the
// reference.  B (outward) is returned for the velocity gradient.
FSGH_HD inline void force_elem_stressed(const double* x, const double* y, const double* z,
                                        const double* u, const double* v, const double* w,
                                        double den, double pres, double sspd, double vol,
                                        double div, double ext, double q1, double q1ex,
                                        double q2, double q2ex, double fuzz, double small,
                                        const double* stress, double* B, double* f)
{
    hex_bmatrix(x, y, z, B);
    double vsx = 0.0, vsy = 0.0, vsz = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        vsx += 0.125 * u[a];
        vsy += 0.125 * v[a];
        vsz += 0.125 * w[a];
    }
    const double qa = div < 0.0 ? q1 : q1ex;
    const double qb = div < 0.0 ? q2 : q2ex;
    double sum0 = 0.0, sum1 = 0.0, sum2 = 0.0, sum3 = 0.0;
    double mu[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const double dx = u[a] - vsx, dy = v[a] - vsy, dz = w[a] - vsz;
        const double mag_vel = sqrt(dx * dx + dy * dy + dz * dz);
        const double muc = den * (qa * sspd + qb * mag_vel);
        const double bx = B[3 * a], by = B[3 * a + 1], bz = B[3 * a + 2];
        const double mu_term = muc * sqrt(bx * bx + by * by + bz * bz);
        sum0 += mu_term * u[a];
        sum1 += mu_term * v[a];
        sum2 += mu_term * w[a];
        sum3 += mu_term;
        mu[a] = mu_term;
    }
    if (sum3 > fuzz) {
        vsx = sum0 / sum3;
        vsy = sum1 / sum3;
        vsz = sum2 / sum3;
    } else {
        vsx = 0.0;
        vsy = 0.0;
        vsz = 0.0;
    }
    const double r_face = 0.9 * (ext + small) / (div + small);
    const double r_min = fmin(r_face, 1.0);
    double phi = 1.0 - fmax(0.0, r_min);
    const double c_length = pow(vol, 1.0 / 3.0);
    const double alpha = fmin(1.0, 20.0 * (c_length * fabs(div)) / (sspd + fuzz));
    phi = alpha * phi;
    double tau[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) tau[k] = stress[k];
    tau[0] -= pres;
    tau[4] -= pres;
    tau[8] -= pres;
    const double vs[3] = {vsx, vsy, vsz};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const double ax = (-1.0) * B[3 * a], ay = (-1.0) * B[3 * a + 1], az = (-1.0) * B[3 * a + 2];
        const double vel[3] = {u[a], v[a], w[a]};
#pragma unroll
        for (int d = 0; d < 3; ++d)
            f[3 * a + d] = ax * tau[d] + ay * tau[3 + d] + az * tau[6 + d] +
                           phi * mu[a] * (vs[d] - vel[d]);
    }
}

// Node velocity + boundary + position (K/momentum.cpp:44-47, K/boundary.cpp:16-55,
// K/geometry.cpp:19-22); f = sum of the node's corner forces in CSR order.
template <typename T>
FSGH_HD inline void node_update(const T* f, T mass, const T* v0, const T* x0, uint8_t mask,
                                double alpha, double dt, T* v1, T* x1)
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T vn = v0[d] + alpha * dt * f[d] / mass;
        if (mask & (1u << d)) vn = 0.0;
        v1[d] = vn;
        x1[d] = x0[d] + alpha * dt * ((vn + v0[d]) * 0.5);
    }
}

// the node's mean velocity over the stage, (v1 + v0) / 2 per component, as
// the energy update evaluates it (K/energy_sgh.cpp:34-43)
template <typename T>
FSGH_HD inline T half_sum(T a1, T a0)
{
    return (a1 + a0) * 0.5;
}

// Element energy update (K/energy_sgh.cpp:18-50) from the nodes' mean
// velocities hu, hv, hw = half_sum(v1, v0): F . (v1+v0)/2, corner by corner,
// then e1 = e0 - alpha dt / M * power.
template <typename T>
FSGH_HD inline T energy_update_h(const T* f, const T* hu, const T* hv, const T* hw, T e0,
                                 T mass, double alpha, double dt)
{
    T power = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        power += f[3 * a + 0] * hu[a];
        power += f[3 * a + 1] * hv[a];
        power += f[3 * a + 2] * hw[a];
    }
    return e0 - alpha * dt / mass * power;
}

// Element energy update (K/energy_sgh.cpp:18-50) from both velocity bins.
template <typename T>
FSGH_HD inline T energy_update(const T* f, const T* u1, const T* v1, const T* w1, const T* u0,
                               const T* v0, const T* w0, T e0, T mass, double alpha, double dt)
{
    T hu[8], hv[8], hw[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        hu[a] = half_sum(u1[a], u0[a]);
        hv[a] = half_sum(v1[a], v0[a]);
        hw[a] = half_sum(w1[a], w0[a]);
    }
    return energy_update_h(f, hu, hv, hw, e0, mass, alpha, dt);
}

// Element state from coordinates, velocities, mass and energy: volume
// (K/geometry.cpp:298-306), density and the material's EOS
// (K/properties.cpp:39,99), divergence (K/momentum.cpp:229-287).
template <typename T>
FSGH_HD inline void derived_state(const T* x, const T* y, const T* z, const T* u, const T* v,
                                  const T* w, T mass, T sie, T gamma, double csmin,
                                  int32_t eos_kind, T& vol, T& den, T& pres, T& sspd, T& div)
{
    T B[24];
    vol = hex_bmatrix_volume(x, y, z, B);
    den = mass / vol;
    div = divergence(B, u, v, w, vol);
    // the EOS last: the B matrix and velocities are dead by then
    eos(eos_kind, gamma, csmin, den, sie, pres, sspd);
}

}  // namespace fsgh
