/*
 * rz_oracle.c -- CPU restatement of the reference's 2D axisymmetric (RZ) SGH
 * path.  TEST INFRASTRUCTURE ONLY (see rz_oracle.h).  Each function cites the
 * reference routine it restates; K/ = Explicit-Lagrange-Kokkos/SGH_solver/.
 */
#define _POSIX_C_SOURCE 200809L
#include "rz_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* node_coords(bin, n, d) and node_vel(bin, n, d) with 2 components */
#define XC(s, N, bin, n, d) ((s)->node_coords[((bin) * (N) + (n)) * 2 + (d)])
#define VC(s, N, bin, n, d) ((s)->node_vel[((bin) * (N) + (n)) * 2 + (d)])

/* ------------------------------------------------------------------------ */
/* geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* get_bmatrix2D (K/geometry.cpp:315-378): outward corner normals, B[i*2+d] */
void orz_get_bmatrix2D(double B[8], const double x[4], const double y[4])
{
    B[0 * 2 + 0] = -0.5 * (y[3] - y[1]);
    B[1 * 2 + 0] = -0.5 * (y[0] - y[2]);
    B[2 * 2 + 0] = -0.5 * (y[1] - y[3]);
    B[3 * 2 + 0] = -0.5 * (y[2] - y[0]);
    B[0 * 2 + 1] = -0.5 * (x[1] - x[3]);
    B[1 * 2 + 1] = -0.5 * (x[2] - x[0]);
    B[2 * 2 + 1] = -0.5 * (x[3] - x[1]);
    B[3 * 2 + 1] = -0.5 * (x[0] - x[2]);
}

/* get_area_quad (K/geometry.cpp:435-468) */
double orz_get_area_quad(const double x[4], const double y[4])
{
    return 0.5 * ((x[0] - x[2]) * (y[1] - y[3]) + (x[3] - x[1]) * (y[0] - y[2]));
}

/* get_vol_quad (K/geometry.cpp:381-431): the RZ volume per radian */
double orz_get_vol_quad(const double x[4], const double y[4])
{
    return ((y[2] + y[3] + y[0]) * ((y[2] - y[3]) * (x[0] - x[3]) - (y[0] - y[3]) * (x[2] - x[3])) +
            (y[0] + y[1] + y[2]) * ((y[0] - y[1]) * (x[2] - x[1]) - (y[2] - y[1]) * (x[0] - x[1]))) /
           6.0;
}

/* heron (K/geometry.cpp:472-498) */
double orz_heron(double x1, double y1, double x2, double y2, double x3, double y3)
{
    double S = 0.0, a, b, c;
    a = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    S += a;
    b = sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
    S += b;
    c = sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1));
    S += c;
    S *= 0.5;
    return sqrt(S * (S - a) * (S - b) * (S - c));
}

/* get_area_weights2D (K/geometry.cpp:503-543): Barlow's corner areas */
void orz_get_area_weights2D(double A[4], const double x[4], const double y[4])
{
    double rc = 0.0, zc = 0.0;
    for (int a = 0; a < 4; ++a) {
        rc += 0.25 * y[a];
        zc += 0.25 * x[a];
    }
    double A12 = orz_heron(x[0], y[0], zc, rc, x[1], y[1]);
    double A23 = orz_heron(x[1], y[1], zc, rc, x[2], y[2]);
    double A34 = orz_heron(x[2], y[2], zc, rc, x[3], y[3]);
    double A41 = orz_heron(x[3], y[3], zc, rc, x[0], y[0]);
    A[0] = (5. * A41 + 5. * A12 + A23 + A34) / 12.;
    A[1] = (A41 + 5. * A12 + 5. * A23 + A34) / 12.;
    A[2] = (A41 + A12 + 5. * A23 + 5. * A34) / 12.;
    A[3] = (5. * A41 + A12 + A23 + 5. * A34) / 12.;
}

/* get_velgrad2D (K/momentum.cpp:151-213) with the outward B */
void orz_get_velgrad2D(double L[9], const double u[4], const double v[4], const double B[8],
                       double elem_vol, double elem_area)
{
    for (int i = 0; i < 9; ++i) L[i] = 0.0;
    double mean_radius = elem_vol / elem_area;
    double elem_vel_r = 0.25 * (v[0] + v[1] + v[2] + v[3]);
    double inverse_area = 1.0 / elem_area;
    L[0] = (u[0] * B[0] + u[1] * B[2] + u[2] * B[4] + u[3] * B[6]) * inverse_area;
    L[1] = (u[0] * B[1] + u[1] * B[3] + u[2] * B[5] + u[3] * B[7]) * inverse_area;
    L[3] = (v[0] * B[0] + v[1] * B[2] + v[2] * B[4] + v[3] * B[6]) * inverse_area;
    L[4] = (v[0] * B[1] + v[1] * B[3] + v[2] * B[5] + v[3] * B[7]) * inverse_area;
    L[8] = elem_vel_r / mean_radius;
}

static void elem_xy(const orz_mesh *m, const orc_state *s, int64_t e, double *x, double *y)
{
    const int64_t N = m->num_nodes;
    for (int a = 0; a < 4; ++a) {
        int64_t n = m->nodes_in_elem[e * 4 + a];
        x[a] = XC(s, N, 1, n, 0);
        y[a] = XC(s, N, 1, n, 1);
    }
}

/* get_vol, 2D (K/geometry.cpp:244-255) */
void orz_get_vol(const orz_mesh *m, orc_state *s)
{
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->num_elems; ++e) {
        double x[4], y[4];
        elem_xy(m, s, e, x, y);
        s->elem_vol[e] = orz_get_vol_quad(x, y);
    }
}

/* get_divergence2D (K/momentum.cpp:298-378) */
void orz_get_divergence(const orz_mesh *m, orc_state *s)
{
    const int64_t N = m->num_nodes;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->num_elems; ++e) {
        double x[4], y[4], B[8], u[4], v[4];
        elem_xy(m, s, e, x, y);
        orz_get_bmatrix2D(B, x, y);
        double elem_area = orz_get_area_quad(x, y);
        for (int a = 0; a < 4; ++a) {
            int64_t n = m->nodes_in_elem[e * 4 + a];
            u[a] = VC(s, N, 1, n, 0);
            v[a] = VC(s, N, 1, n, 1);
        }
        double inverse_area = 1.0 / elem_area;
        double mean_radius = s->elem_vol[e] / elem_area;
        double elem_vel_r = 0.25 * (v[0] + v[1] + v[2] + v[3]);
        double d = 0.0;
        d += (u[0] * B[0] + u[1] * B[2] + u[2] * B[4] + u[3] * B[6]) * inverse_area;
        d += (v[0] * B[1] + v[1] * B[3] + v[2] * B[5] + v[3] * B[7]) * inverse_area +
             elem_vel_r / mean_radius;
        s->elem_div[e] = d;
    }
}

/* get_force_sgh2D (K/force_sgh.cpp:407-808), strength-free: the stress is
 * zero, tau = -p I; the velocity gradient (curl) and the shock direction do
 * not reach the corner force and are omitted. */
void orz_get_force(const orz_mesh *m, orc_state *s, const orc_material *mat, double fuzz,
                   double small, double dt, double rk_alpha)
{
    const int64_t N = m->num_nodes;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->num_elems; ++e) {
        double x[4], y[4], B[8], area[4][2], vel[4][2], muc[4], corner_areas[4];
        elem_xy(m, s, e, x, y);
        orz_get_bmatrix2D(B, x, y);
        double elem_area = orz_get_area_quad(x, y);
        orz_get_area_weights2D(corner_areas, x, y);
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 2; ++d) area[a][d] = (-1.0) * B[a * 2 + d];
        for (int a = 0; a < 4; ++a) {
            int64_t n = m->nodes_in_elem[e * 4 + a];
            vel[a][0] = VC(s, N, 1, n, 0);
            vel[a][1] = VC(s, N, 1, n, 1);
        }
        const double div = s->elem_div[e];
        const double den = s->elem_den[e];
        const double sspd = s->elem_sspd[e];
        const double pres = s->elem_pres[e];
        double *stress = s->elem_stress ? s->elem_stress + (m->num_elems + e) * 9 : NULL;
        double tau[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) tau[i][j] = stress ? stress[i * 3 + j] : 0.0;
        for (int i = 0; i < 3; ++i) tau[i][i] -= pres;

        double vel_star[2] = {0.0, 0.0};
        for (int a = 0; a < 4; ++a) {
            vel_star[0] += 0.25 * vel[a][0];
            vel_star[1] += 0.25 * vel[a][1];
        }
        double sum[4] = {0.0, 0.0, 0.0, 0.0};
        const orc_material *mm = &mat[s->elem_mat_id[e]];
        for (int a = 0; a < 4; ++a) {
            double mag_vel = sqrt((vel[a][0] - vel_star[0]) * (vel[a][0] - vel_star[0]) +
                                  (vel[a][1] - vel_star[1]) * (vel[a][1] - vel_star[1]));
            double mu;
            if (div < 0) mu = den * (mm->q1 * sspd + mm->q2 * mag_vel);
            else         mu = den * (mm->q1ex * sspd + mm->q2ex * mag_vel);
            double mu_term = mu * sqrt(area[a][0] * area[a][0] + area[a][1] * area[a][1]);
            sum[0] += mu_term * vel[a][0];
            sum[1] += mu_term * vel[a][1];
            sum[3] += mu_term;
            muc[a] = mu_term;
        }
        if (sum[3] > fuzz) {
            for (int d = 0; d < 2; ++d) vel_star[d] = sum[d] / sum[3];
        } else {
            for (int d = 0; d < 2; ++d) vel_star[d] = 0.0;
        }
        double r_min = 1.0;
        const double r_coef = 0.9;
        for (int64_t q = m->elems_in_elem_start[e]; q < m->elems_in_elem_start[e + 1]; ++q) {
            double div_neighbor = s->elem_div[m->elems_in_elem[q]];
            double r_face = r_coef * (div_neighbor + small) / (div + small);
            r_min = fmin(r_face, r_min);
        }
        double phi = 1.0 - fmax(0.0, r_min);
        phi = pow(phi, 1.0);
        const double omega = 20.0;
        double c_length = sqrt(elem_area);
        double alpha = fmin(1.0, omega * (c_length * fabs(div)) / (sspd + fuzz));
        phi = alpha * phi;
        for (int a = 0; a < 4; ++a) {
            int64_t n = m->nodes_in_elem[e * 4 + a];
            double *F = &s->corner_force[(e * 4 + a) * 2];
            for (int d = 0; d < 2; ++d)
                F[d] = area[a][0] * tau[0][d] + area[a][1] * tau[1][d] +
                       phi * muc[a] * (vel_star[d] - VC(s, N, 1, n, d));
            /* hoop stress terms (:755-769) */
            double node_radius = XC(s, N, 1, n, 1);
            if (node_radius > 1e-14) {
                F[0] += tau[1][0] * corner_areas[a] / node_radius;
                F[1] += (tau[1][1] - tau[2][2]) * corner_areas[a] / node_radius;
            }
        }
        /* hypo-elastic strength (:774-801) with the velocity gradient of the
         * outward B (:488-501) */
        if (stress && s->mat_strength && s->mat_strength[s->elem_mat_id[e]] == 1) {
            double u[4], v[4], L[9];
            for (int a = 0; a < 4; ++a) {
                u[a] = vel[a][0];
                v[a] = vel[a][1];
            }
            orz_get_velgrad2D(L, u, v, B, s->elem_vol[e], elem_area);
            orc_user_strength(s->elem_mat_id[e], orc_strength_params(s, e), den, s->elem_sie[e],
                              L, s->elem_vol[e], dt, rk_alpha, s->elem_stress + e * 9, stress,
                              &s->elem_pres[e], &s->elem_sspd[e]);
        }
    }
}

/* update_velocity_sgh, num_dims = 2 (K/momentum.cpp:11-53) */
void orz_update_velocity(const orz_mesh *m, orc_state *s, double rk_alpha, double dt)
{
    const int64_t N = m->num_nodes;
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
        double f[2] = {0.0, 0.0};
        for (int64_t q = m->corners_in_node_start[n]; q < m->corners_in_node_start[n + 1]; ++q) {
            int64_t c = m->corners_in_node[q];
            for (int d = 0; d < 2; ++d) f[d] += s->corner_force[c * 2 + d];
        }
        for (int d = 0; d < 2; ++d)
            VC(s, N, 1, n, d) = VC(s, N, 0, n, d) + rk_alpha * dt * f[d] / s->node_mass[n];
    }
}

/* boundary_velocity, num_dims = 2 (K/boundary.cpp:10-55) */
void orz_boundary_velocity(const orz_mesh *m, orc_state *s, const orc_boundary *bc)
{
    const int64_t N = m->num_nodes;
    for (int64_t set = 0; set < m->num_bdy_sets; ++set) {
        for (int64_t q = m->bdy_nodes_in_set_start[set]; q < m->bdy_nodes_in_set_start[set + 1];
             ++q) {
            int64_t n = m->bdy_nodes_in_set[q];
            if (bc[set].hydro_bc == 1) {
                VC(s, N, 1, n, bc[set].surface) = 0.0;
            } else if (bc[set].hydro_bc == 0) {
                for (int d = 0; d < 2; ++d) VC(s, N, 1, n, d) = 0.0;
            }
        }
    }
}

/* update_energy_sgh, num_dims = 2 (K/energy_sgh.cpp:6-55): power weighted by
 * the node radius */
void orz_update_energy(const orz_mesh *m, orc_state *s, double rk_alpha, double dt)
{
    const int64_t N = m->num_nodes, E = m->num_elems;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
        double elem_power = 0.0;
        for (int a = 0; a < 4; ++a) {
            int64_t n = m->nodes_in_elem[e * 4 + a];
            double node_radius = XC(s, N, 1, n, 1);
            for (int d = 0; d < 2; ++d) {
                double half_vel = (VC(s, N, 1, n, d) + VC(s, N, 0, n, d)) * 0.5;
                elem_power += s->corner_force[(e * 4 + a) * 2 + d] * node_radius * half_vel;
            }
        }
        s->elem_sie[E + e] = s->elem_sie[e] - rk_alpha * dt / s->elem_mass[e] * elem_power;
    }
}

/* update_position_sgh, num_dims = 2 (K/geometry.cpp:9-26) */
void orz_update_position(const orz_mesh *m, orc_state *s, double rk_alpha, double dt)
{
    const int64_t N = m->num_nodes;
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n)
        for (int d = 0; d < 2; ++d) {
            double half_vel = (VC(s, N, 1, n, d) + VC(s, N, 0, n, d)) * 0.5;
            XC(s, N, 1, n, d) = XC(s, N, 0, n, d) + rk_alpha * dt * half_vel;
        }
}

/* the element's EOS (ideal_gas, K/eos.cpp:12-46, or the user model of
 * K/user_mat.cpp:14-40 with zero stress) */
static void eos(orc_state *s, int64_t E, int64_t e, double den, double sie)
{
    if (s->mat_eos && s->mat_eos[s->elem_mat_id[e]] == 1) {
        double p = 0.0;
        if (s->elem_stress)
            for (int i = 0; i < 3; ++i) p -= s->elem_stress[(E + e) * 9 + i * 3 + i];
        p *= 1.0 / 3.0;
        s->elem_pres[e] = p;
        s->elem_sspd[e] = 1.0e-8;
        return;
    }
    double gamma = s->elem_statev[e * s->num_state_vars + 0];
    double csmin = s->elem_statev[e * s->num_state_vars + 1];
    s->elem_pres[e] = (gamma - 1.0) * sie * den;
    s->elem_sspd[e] = sqrt(gamma * (gamma - 1.0) * sie);
    if (s->elem_sspd[e] < csmin) s->elem_sspd[e] = csmin;
}

/* update_state2D, no strength (K/properties.cpp:118-234) */
void orz_update_state(const orz_mesh *m, orc_state *s)
{
    const int64_t E = m->num_elems;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
        s->elem_den[e] = s->elem_mass[e] / s->elem_vol[e];
        eos(s, E, e, s->elem_den[e], s->elem_sie[E + e]);
    }
}

/* rk_init (K/time_integration.cpp:9-44), num_dims = 2 */
void orz_rk_init(const orz_mesh *m, orc_state *s)
{
    const int64_t N = m->num_nodes, E = m->num_elems;
    memcpy(s->elem_sie, s->elem_sie + E, (size_t)E * sizeof(double));
    if (s->elem_stress) memcpy(s->elem_stress, s->elem_stress + 9 * E, (size_t)(9 * E) * sizeof(double));
    memcpy(s->node_coords, s->node_coords + 2 * N, (size_t)(2 * N) * sizeof(double));
    memcpy(s->node_vel, s->node_vel + 2 * N, (size_t)(2 * N) * sizeof(double));
}

/* get_timestep2D (K/time_integration.cpp:161-246): 6 node pairs; pow(x, 2.0)
 * is x*x */
void orz_get_timestep(const orz_mesh *m, orc_state *s, double time_value, double graphics_time,
                      double time_final, double dt_max, double dt_min, double dt_cfl, double *dt,
                      double fuzz)
{
    *dt = *dt * 1.1;
    double min_dt_calc = __DBL_MAX__;
#pragma omp parallel for schedule(static) reduction(min : min_dt_calc)
    for (int64_t e = 0; e < m->num_elems; ++e) {
        double x[4], y[4], dist[6];
        elem_xy(m, s, e, x, y);
        int count = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = i + 1; j <= 3; ++j) {
                double dx = x[i] - x[j], dy = y[i] - y[j];
                dist[count++] = fabs(sqrt(dx * dx + dy * dy));
            }
        double dist_min = dist[0];
        for (int i = 0; i < 6; ++i) dist_min = fmin(dist[i], dist_min);
        double dt_lcl = dt_cfl * dist_min / (s->elem_sspd[e] + fuzz);
        dt_lcl = fmin(dt_lcl, dt_max);
        dt_lcl = fmax(dt_lcl, dt_min);
        if (dt_lcl < min_dt_calc) min_dt_calc = dt_lcl;
    }
    if (min_dt_calc < *dt) *dt = min_dt_calc;
    *dt = fmin(*dt, (graphics_time - time_value) + fuzz);
    *dt = fmin(*dt, time_final - time_value);
}

void orz_extensive_mass(const orz_mesh *m, const orc_state *s, double *ext)
{
    const int64_t N = m->num_nodes;
    for (int64_t n = 0; n < N; ++n) ext[n] = s->node_mass[n] * XC(s, N, 1, n, 1);
}

void orz_update_node_mass(const orz_mesh *m, orc_state *s, const double *ext, double tiny)
{
    const int64_t N = m->num_nodes;
    for (int64_t n = 0; n < N; ++n) {
        s->node_mass[n] = 0.0;
        if (XC(s, N, 1, n, 1) > tiny) s->node_mass[n] = ext[n] / XC(s, N, 1, n, 1);
    }
    /* on-axis boundary nodes take half the largest off-axis neighbour mass */
    for (int64_t b = 0; b < m->num_bdy_nodes; ++b) {
        int64_t n = m->bdy_nodes[b];
        if (XC(s, N, 1, n, 1) < tiny) {
            for (int64_t q = m->nodes_in_node_start[n]; q < m->nodes_in_node_start[n + 1]; ++q) {
                int64_t nb = m->nodes_in_node[q];
                if (XC(s, N, 1, nb, 1) > tiny)
                    s->node_mass[n] = fmax(s->node_mass[n], s->node_mass[nb] / 2.0);
            }
        }
    }
}

void orz_energy_tallies(const orz_mesh *m, const orc_state *s, double *ie, double *ke)
{
    const int64_t N = m->num_nodes, E = m->num_elems;
    double ie_sum = 0.0, ke_sum = 0.0;
    for (int64_t e = 0; e < E; ++e) ie_sum += s->elem_mass[e] * s->elem_sie[E + e];
    for (int64_t n = 0; n < N; ++n) {
        double k = 0;
        for (int d = 0; d < 2; ++d) k += VC(s, N, 1, n, d) * VC(s, N, 1, n, d);
        ke_sum += s->node_mass[n] * XC(s, N, 1, n, 1) * k;
    }
    *ie = ie_sum;
    *ke = 0.5 * ke_sum;
}

static double now_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* sgh_solve (K/sgh_solve.cpp:8-692), 2D branches, outputs disabled */
void orz_sgh_solve(const orz_mesh *m, orc_state *s, const orc_material *mat,
                   const orc_boundary *bc, const orc_run_params *p, double tiny,
                   orc_run_result *r)
{
    double time_value = p->time_value;
    double graphics_time = p->graphics_time;
    double dt = p->dt_start;
    double *ext = (double *)malloc((size_t)m->num_nodes * sizeof(double));
    orz_energy_tallies(m, s, &r->ie0, &r->ke0);
    r->te0 = r->ie0 + r->ke0;
    orz_extensive_mass(m, s, ext);
    int64_t cycle = 0, done = 0;
    double t0 = now_seconds();
    for (cycle = 0; cycle < p->cycle_stop; ++cycle) {
        if (p->dt_replay) dt = p->dt_replay[cycle];
        else orz_get_timestep(m, s, time_value, graphics_time, p->time_final, p->dt_max,
                              p->dt_min, p->dt_cfl, &dt, p->fuzz);
        if (p->dt_record) p->dt_record[cycle] = dt;
        orz_rk_init(m, s);
        for (int64_t stage = 0; stage < p->rk_num_stages; ++stage) {
            double rk_alpha = orc_rk_alpha(p->rk_num_stages, stage);
            orz_get_divergence(m, s);
            orz_get_force(m, s, mat, p->fuzz, p->small, dt, rk_alpha);
            orz_update_velocity(m, s, rk_alpha, dt);
            orz_boundary_velocity(m, s, bc);
            orz_update_energy(m, s, rk_alpha, dt);
            orz_update_position(m, s, rk_alpha, dt);
            orz_get_vol(m, s);
            orz_update_state(m, s);
            orz_update_node_mass(m, s, ext, tiny);
        }
        time_value += dt;
        int write = 0;
        if ((cycle + 1) % p->graphics_cyc_ival == 0 && cycle > 0) write = 1;
        else if (cycle == p->cycle_stop) write = 1;
        else if (time_value >= p->time_final) write = 1;
        else if (time_value >= graphics_time) write = 1;
        if (write) graphics_time = time_value + p->graphics_dt_ival;
        if (p->write_record) p->write_record[cycle] = write;
        if (time_value >= p->time_final) { done = 1; break; }
    }
    r->seconds = now_seconds() - t0;
    r->cycles = done ? cycle + 1 : cycle;
    r->time_value = time_value;
    r->dt = dt;
    r->graphics_time = graphics_time;
    orz_energy_tallies(m, s, &r->ie1, &r->ke1);
    r->te1 = r->ie1 + r->ke1;
    free(ext);
}

/* ------------------------------------------------------------------------ */
/* mesh input and connectivity                                               */
/* ------------------------------------------------------------------------ */
void orz_gen_rect(int64_t nx, int64_t ny, double lx_num, double ly_num, double l_den,
                  double *coords, int64_t *nodes_in_elem)
{
    for (int64_t j = 0; j <= ny; ++j)
        for (int64_t i = 0; i <= nx; ++i) {
            int64_t n = i + (nx + 1) * j;
            coords[2 * n + 0] = ((double)i * lx_num) / ((double)nx * l_den);
            coords[2 * n + 1] = ((double)j * ly_num) / ((double)ny * l_den);
        }
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t e = i + nx * j;
            int64_t n0 = i + (nx + 1) * j;
            nodes_in_elem[e * 4 + 0] = n0;
            nodes_in_elem[e * 4 + 1] = n0 + 1;
            nodes_in_elem[e * 4 + 2] = n0 + 1 + (nx + 1);
            nodes_in_elem[e * 4 + 3] = n0 + (nx + 1);
        }
}

static int skip_line(FILE *f)
{
    int ch;
    while ((ch = fgetc(f)) != '\n')
        if (ch == EOF) return -1;
    return 0;
}

int orz_read_ensight(const char *path, int64_t *num_nodes, int64_t *num_elems, double *coords,
                     int64_t *nodes_in_elem)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    for (int j = 0; j < 8; ++j)
        if (skip_line(f)) { fclose(f); return -2; }
    long long nn = 0, ne = 0;
    if (fscanf(f, "%lld", &nn) != 1) { fclose(f); return -3; }
    for (int d = 0; d < 3; ++d)
        for (long long n = 0; n < nn; ++n) {
            double v;
            if (fscanf(f, "%le", &v) != 1) { fclose(f); return -4; }
            if (coords && d < 2) coords[2 * n + d] = v;
        }
    fgetc(f);
    if (skip_line(f)) { fclose(f); return -5; }
    if (fscanf(f, "%lld", &ne) != 1) { fclose(f); return -6; }
    if (nodes_in_elem)
        for (long long e = 0; e < ne; ++e)
            for (int a = 0; a < 4; ++a) {
                long long id;
                if (fscanf(f, "%lld", &id) != 1) { fclose(f); return -7; }
                nodes_in_elem[e * 4 + a] = id - 1;
            }
    fclose(f);
    *num_nodes = nn;
    *num_elems = ne;
    return 0;
}

/* build_corner_connectivity (K/mesh.h:180-245), 4 corners per element */
void orz_build_corner_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                   int64_t *start, int64_t *corners_in_node,
                                   int64_t *elems_in_node)
{
    int64_t *count = (int64_t *)calloc((size_t)N, sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e)
        for (int a = 0; a < 4; ++a) count[nodes_in_elem[e * 4 + a]]++;
    start[0] = 0;
    for (int64_t n = 0; n < N; ++n) start[n + 1] = start[n] + count[n];
    memset(count, 0, (size_t)N * sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e)
        for (int a = 0; a < 4; ++a) {
            int64_t n = nodes_in_elem[e * 4 + a];
            int64_t j = count[n]++;
            corners_in_node[start[n] + j] = a + e * 4;
            elems_in_node[start[n] + j] = e;
        }
    free(count);
}

/* build_elem_elem_connectivity (K/mesh.h:249-338) with 4 nodes per element */
int64_t orz_build_elem_elem_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                         const int64_t *start, const int64_t *elems_in_node,
                                         int64_t *ee_start, int64_t *ee)
{
    (void)N;
    int64_t total = 0;
    ee_start[0] = 0;
    for (int64_t e = 0; e < E; ++e) {
        int64_t saved[64];
        int64_t ns = 0;
        for (int a = 0; a < 4; ++a) {
            int64_t n = nodes_in_elem[e * 4 + a];
            for (int64_t q = start[n]; q < start[n + 1]; ++q) {
                int64_t nb = elems_in_node[q];
                int save = nb != e;
                for (int64_t i = 0; i < ns && save; ++i)
                    if (saved[i] == nb) save = 0;
                if (save && ns < 64) saved[ns++] = nb;
            }
        }
        if (ee)
            for (int64_t i = 0; i < ns; ++i) ee[total + i] = saved[i];
        total += ns;
        ee_start[e + 1] = total;
    }
    return total;
}

/* the 4 edges of a quad as node lids (K/mesh.h:366-377) */
static const int EDGE_LIDS[4][2] = {{0, 3}, {1, 2}, {0, 1}, {3, 2}};

/* build_patch_connectivity, num_dims = 2 (K/mesh.h:342-649): key = s0 + N s1
 * of the sorted node pair, matched against the neighbours in elems_in_elem
 * order, interior patches numbered as found, boundary patches after each
 * element's interior ones. */
int64_t orz_build_patch_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                     const int64_t *ee_start, const int64_t *ee,
                                     int64_t *nodes_in_patch, int64_t *bdy_patches,
                                     int64_t *bdy_nodes, int64_t *counts)
{
    long long *hash = (long long *)malloc((size_t)E * 4 * sizeof(long long));
    int64_t *pie = (int64_t *)calloc((size_t)E * 4, sizeof(int64_t));
    int64_t *tbdy = (int64_t *)malloc((size_t)E * 4 * sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e)
        for (int pl = 0; pl < 4; ++pl) {
            int64_t a = nodes_in_elem[e * 4 + EDGE_LIDS[pl][0]];
            int64_t b = nodes_in_elem[e * 4 + EDGE_LIDS[pl][1]];
            int64_t s0 = a < b ? a : b, s1 = a < b ? b : a;
            hash[e * 4 + pl] = -(long long)(s0 + N * s1);
        }
    int64_t patch_gid = 0, bdy_patch_gid = 0;
    for (int64_t e = 0; e < E; ++e) {
        for (int pl = 0; pl < 4; ++pl) {
            if (hash[e * 4 + pl] < 0) {
                int found = 0;
                for (int64_t q = ee_start[e]; q < ee_start[e + 1] && !found; ++q) {
                    int64_t nb = ee[q];
                    for (int np = 0; np < 4; ++np)
                        if (hash[nb * 4 + np] == hash[e * 4 + pl]) {
                            hash[e * 4 + pl] = nb;
                            hash[nb * 4 + np] = e;
                            pie[e * 4 + pl] = patch_gid;
                            pie[nb * 4 + np] = patch_gid;
                            patch_gid++;
                            found = 1;
                            break;
                        }
                }
            }
        }
        for (int pl = 0; pl < 4; ++pl)
            if (hash[e * 4 + pl] < 0) {
                hash[e * 4 + pl] = e;
                pie[e * 4 + pl] = patch_gid;
                tbdy[bdy_patch_gid] = patch_gid;
                patch_gid++;
                bdy_patch_gid++;
            }
    }
    int64_t P = patch_gid;
    counts[0] = P;
    counts[1] = bdy_patch_gid;
    int64_t *nip = (int64_t *)calloc((size_t)P * 2, sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e)
        for (int pl = 0; pl < 4; ++pl) {
            int64_t pg = pie[e * 4 + pl];
            for (int q = 0; q < 2; ++q) nip[pg * 2 + q] = nodes_in_elem[e * 4 + EDGE_LIDS[pl][q]];
        }
    if (nodes_in_patch) memcpy(nodes_in_patch, nip, (size_t)P * 2 * sizeof(int64_t));
    if (bdy_patches) memcpy(bdy_patches, tbdy, (size_t)bdy_patch_gid * sizeof(int64_t));
    char *seen = (char *)calloc((size_t)N, 1);
    int64_t nb_nodes = 0;
    for (int64_t b = 0; b < bdy_patch_gid; ++b) {
        int64_t pg = tbdy[b];
        for (int q = 0; q < 2; ++q) {
            int64_t n = nip[pg * 2 + q];
            if (!seen[n]) {
                seen[n] = 1;
                if (bdy_nodes) bdy_nodes[nb_nodes] = n;
                nb_nodes++;
            }
        }
    }
    counts[2] = nb_nodes;
    free(seen);
    free(nip);
    free(hash);
    free(pie);
    free(tbdy);
    return P;
}

/* build_node_node_connectivity, 2D branch (K/mesh.h:732-773): both ends of
 * every patch, in patch order */
void orz_build_node_node_connectivity(int64_t N, int64_t P, const int64_t *nodes_in_patch,
                                      int64_t *start, int64_t *out)
{
    int64_t *cnt = (int64_t *)calloc((size_t)N, sizeof(int64_t));
    for (int64_t p = 0; p < P; ++p) {
        cnt[nodes_in_patch[2 * p]]++;
        cnt[nodes_in_patch[2 * p + 1]]++;
    }
    start[0] = 0;
    for (int64_t n = 0; n < N; ++n) start[n + 1] = start[n] + cnt[n];
    memset(cnt, 0, (size_t)N * sizeof(int64_t));
    for (int64_t p = 0; p < P; ++p) {
        int64_t a = nodes_in_patch[2 * p], b = nodes_in_patch[2 * p + 1];
        out[start[a] + cnt[a]++] = b;
        out[start[b] + cnt[b]++] = a;
    }
    free(cnt);
}

static int check_bdy2(const double *xy, const int64_t *nodes_in_patch, int64_t pg, int tag,
                      double val)
{
    int on = 0;
    for (int q = 0; q < 2; ++q) {
        const double *c = xy + nodes_in_patch[pg * 2 + q] * 2;
        if (tag == 0) { if (fabs(c[0] - val) <= 1.0e-7) on += 1; }
        else if (tag == 1) { if (fabs(c[1] - val) <= 1.0e-7) on += 1; }
        else if (tag == 3) {
            double R = sqrt(c[0] * c[0] + c[1] * c[1]);
            if (fabs(R - val) <= 1.0e-7) on += 1;
        }
    }
    return on == 2;
}

/* tag_bdys + build_boundry_node_sets (K/setup.cpp:431-665) over edges */
int64_t orz_build_bdy_sets(const double *xy, int64_t num_bdy_patches, const int64_t *bdy_patches,
                           const int64_t *nodes_in_patch, int64_t nsets, const orc_boundary *bc,
                           int64_t N, int64_t *out_start, int64_t *out_nodes)
{
    char *seen = (char *)malloc((size_t)N);
    int64_t total = 0;
    out_start[0] = 0;
    for (int64_t set = 0; set < nsets; ++set) {
        memset(seen, 0, (size_t)N);
        for (int64_t b = 0; b < num_bdy_patches; ++b) {
            int64_t pg = bdy_patches[b];
            if (!check_bdy2(xy, nodes_in_patch, pg, (int)bc[set].surface, bc[set].value)) continue;
            for (int q = 0; q < 2; ++q) {
                int64_t n = nodes_in_patch[pg * 2 + q];
                if (!seen[n]) {
                    seen[n] = 1;
                    out_nodes[total++] = n;
                }
            }
        }
        out_start[set + 1] = total;
    }
    free(seen);
    return total;
}

/* setup fills, 2D (K/setup.cpp:106-357): centroid of the 4 nodes (third
 * coordinate 0), velocities on 2 components; velocity kinds 0 (cartesian), 1
 * (radial), 3 and 4 (no-ops), 5 (Taylor-Green vortex, :326-350: the two
 * in-plane components, and the element pressure and energy with gamma read
 * from state variable 4) */
void orz_setup_fills(const orz_mesh *m, orc_state *s, int64_t num_fills, const orc_fill *fills,
                     const double *state_vars, const int64_t *mat_num_state_vars)
{
    const double PI = 3.14159265358979323846;   /* K/mesh.h:8 */
    const int64_t N = m->num_nodes, E = m->num_elems;
    for (int64_t f = 0; f < num_fills; ++f) {
        const orc_fill *F = &fills[f];
        for (int64_t e = 0; e < E; ++e) {
            double c[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < 4; ++a) {
                int64_t n = m->nodes_in_elem[e * 4 + a];
                c[0] += XC(s, N, 1, n, 0);
                c[1] += XC(s, N, 1, n, 1);
                c[2] = 0.0;
            }
            c[0] = c[0] / 4;
            c[1] = c[1] / 4;
            c[2] = c[2] / 4;
            double radius = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            double radius_cyl = sqrt(c[0] * c[0] + c[1] * c[1]);
            int fill = 0;
            switch (F->volume) {
            case 0: fill = 1; break;
            case 1:
                fill = c[0] >= F->x1 && c[0] <= F->x2 && c[1] >= F->y1 && c[1] <= F->y2 &&
                       c[2] >= F->z1 && c[2] <= F->z2;
                break;
            case 2: fill = radius_cyl >= F->radius1 && radius_cyl <= F->radius2; break;
            case 3: fill = radius >= F->radius1 && radius <= F->radius2; break;
            }
            if (!fill) continue;
            s->elem_den[e] = F->den;
            s->elem_mass[e] = s->elem_den[e] * s->elem_vol[e];
            s->elem_sie[E + e] = F->sie;
            s->elem_mat_id[e] = F->mat_id;
            int64_t mid = F->mat_id;
            for (int64_t v = 0; v < mat_num_state_vars[mid]; ++v)
                s->elem_statev[e * s->num_state_vars + v] = state_vars[mid * s->num_state_vars + v];
            eos(s, E, e, s->elem_den[e], s->elem_sie[E + e]);
            for (int a = 0; a < 4; ++a) {
                int64_t n = m->nodes_in_elem[e * 4 + a];
                if (F->velocity == 0) {
                    VC(s, N, 1, n, 0) = F->u;
                    VC(s, N, 1, n, 1) = F->v;
                } else if (F->velocity == 1) {
                    double dir[2] = {0.0, 0.0}, rv = 0.0;
                    for (int d = 0; d < 2; ++d) {
                        dir[d] = XC(s, N, 1, n, d);
                        rv += XC(s, N, 1, n, d) * XC(s, N, 1, n, d);
                    }
                    rv = sqrt(rv);
                    for (int d = 0; d < 2; ++d) dir[d] = rv > 1.0e-14 ? dir[d] / rv : 0.0;
                    VC(s, N, 1, n, 0) = F->speed * dir[0];
                    VC(s, N, 1, n, 1) = F->speed * dir[1];
                } else if (F->velocity == 5) {
                    VC(s, N, 1, n, 0) = sin(PI * XC(s, N, 1, n, 0)) * cos(PI * XC(s, N, 1, n, 1));
                    VC(s, N, 1, n, 1) =
                        -1.0 * cos(PI * XC(s, N, 1, n, 0)) * sin(PI * XC(s, N, 1, n, 1));
                }
            }
            if (F->velocity == 5) {
                s->elem_pres[e] = 0.25 * (cos(2.0 * PI * c[0]) + cos(2.0 * PI * c[1])) + 1.0;
                double gamma = s->elem_statev[e * s->num_state_vars + 4];
                s->elem_sie[E + e] = s->elem_pres[e] / (F->den * (gamma - 1.0));
            }
        }
    }
}

/* corner masses and areal node mass, 2D (K/setup.cpp:365-420) */
void orz_node_mass(const orz_mesh *m, orc_state *s)
{
    const int64_t N = m->num_nodes, E = m->num_elems;
    double *cm = (double *)malloc((size_t)(4 * E) * sizeof(double));
    for (int64_t e = 0; e < E; ++e) {
        double x[4], y[4], A[4];
        elem_xy(m, s, e, x, y);
        orz_get_area_weights2D(A, x, y);
        for (int a = 0; a < 4; ++a) cm[e * 4 + a] = A[a] * s->elem_den[e];
    }
    for (int64_t n = 0; n < N; ++n) {
        double acc = 0.0;
        for (int64_t q = m->corners_in_node_start[n]; q < m->corners_in_node_start[n + 1]; ++q)
            acc += cm[m->corners_in_node[q]];
        s->node_mass[n] = acc;
    }
    free(cm);
}
