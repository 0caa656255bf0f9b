/*
 * sgh_oracle.c -- CPU restatement of Fierro's 3D SGH cycle (test infrastructure).
 *
 * See sgh_oracle.h.  This file is the checker for the HIP path; it is never
 * linked into the product library.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off -fopenmp).  Every loop that the reference runs
 * as a Kokkos FOR_ALL is an OpenMP loop here; nothing depends on the thread
 * count except the order of the two REDUCE_SUM energy tallies, which this file
 * evaluates serially in ascending index order (the reference leaves that order
 * to Kokkos: parity for the tallies is therefore to round-off, not bitwise).
 */
#include "sgh_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define X1(s, n, d) ((s)->node_coords[((int64_t)1 * N + (n)) * 3 + (d)])

/* --------------------------------------------------------------------------
 * Hex corner normals, restating get_bmatrix (K/geometry.cpp:40-234).
 *
 * B(p,0) = (1/12) * sum_j y_j * (sum_k +-z_k), B(p,1) is the same pattern
 * with (z,x), B(p,2) with (x,y).  Each row below lists, for node p, the
 * terms in the reference's order: j, count, signed k codes (+-(k+1)).
 * -------------------------------------------------------------------------- */
static const int BTAB[8][6][6] = {
    /* p=0  K/geometry.cpp:66-71 */
    {{1, 4, -3, -4, +5, +6}, {2, 2, +2, -4, 0, 0}, {3, 4, +2, +3, -5, -8},
     {4, 4, -2, +4, -6, +8}, {5, 2, -2, +5, 0, 0}, {7, 2, +4, -5, 0, 0}},
    /* p=1  :73-78 */
    {{0, 4, +3, +4, -5, -6}, {2, 4, -1, -4, +6, +7}, {3, 2, -1, +3, 0, 0},
     {4, 2, +1, -6, 0, 0}, {5, 4, +1, -3, +5, -7}, {6, 2, -3, +6, 0, 0}},
    /* p=2  :80-85 */
    {{0, 2, -2, +4, 0, 0}, {1, 4, +1, +4, -6, -7}, {3, 4, -1, -2, +7, +8},
     {5, 2, +2, -7, 0, 0}, {6, 4, +2, -4, +6, -8}, {7, 2, -4, +7, 0, 0}},
    /* p=3  :87-92 */
    {{0, 4, -2, -3, +5, +8}, {1, 2, +1, -3, 0, 0}, {2, 4, +1, +2, -7, -8},
     {4, 2, -1, +8, 0, 0}, {6, 2, +3, -8, 0, 0}, {7, 4, -1, +3, -5, +7}},
    /* p=4  :94-99 */
    {{0, 4, +2, -4, +6, -8}, {1, 2, -1, +6, 0, 0}, {3, 2, +1, -8, 0, 0},
     {5, 4, -1, -2, +7, +8}, {6, 2, -6, +8, 0, 0}, {7, 4, +1, +4, -6, -7}},
    /* p=5  :101-106 */
    {{0, 2, +2, -5, 0, 0}, {1, 4, -1, +3, -5, +7}, {2, 2, -2, +7, 0, 0},
     {4, 4, +1, +2, -7, -8}, {6, 4, -2, -3, +5, +8}, {7, 2, +5, -7, 0, 0}},
    /* p=6  :108-113 */
    {{1, 2, +3, -6, 0, 0}, {2, 4, -2, +4, -6, +8}, {3, 2, -3, +8, 0, 0},
     {4, 2, +6, -8, 0, 0}, {5, 4, +2, +3, -5, -8}, {7, 4, -3, -4, +5, +6}},
    /* p=7  :115-120 */
    {{0, 2, -4, +5, 0, 0}, {2, 2, +4, -7, 0, 0}, {3, 4, +1, -3, +5, -7},
     {4, 4, -1, -4, +6, +7}, {5, 2, -5, +7, 0, 0}, {6, 4, +3, +4, -5, -6}},
};

/* get_vol_hex (K/geometry.cpp:298-306) sums x_i * (sum_j y_j * inner(i,j)) with
 * inner(i,j) identical to the B-matrix inner sums, in this outer/inner order. */
static const int VOL_OUTER[8] = {1, 7, 3, 5, 6, 0, 2, 4};
static const int VOL_JORDER[8][6] = {
    {3, 4, 0, 6, 5, 2}, /* x1 :299 */
    {0, 6, 2, 3, 5, 4}, /* x7 :300 */
    {1, 7, 6, 2, 4, 0}, /* x3 :301 */
    {0, 7, 2, 1, 4, 6}, /* x5 :302 */
    {1, 7, 5, 4, 3, 2}, /* x6 :303 */
    {2, 7, 5, 1, 3, 4}, /* x0 :304 */
    {0, 5, 1, 7, 6, 3}, /* x2 :305 */
    {1, 7, 3, 0, 6, 5}, /* x4 :306 */
};

static double inner_sum(const int *t, const double *c)
{
    /* t = {j, n, k codes...}; evaluates (+-c[k0]) +- c[k1] ... left to right */
    int n = t[1];
    int code = t[2];
    double acc = code > 0 ? c[code - 1] : -c[-code - 1];
    for (int q = 1; q < n; ++q) {
        code = t[2 + q];
        if (code > 0) acc = acc + c[code - 1];
        else          acc = acc - c[-code - 1];
    }
    return acc;
}

static double brow(int p, const double *a, const double *b)
{
    /* sum_j a_j * inner_j(b), outer terms left to right */
    double acc = a[BTAB[p][0][0]] * inner_sum(BTAB[p][0], b);
    for (int t = 1; t < 6; ++t) acc = acc + a[BTAB[p][t][0]] * inner_sum(BTAB[p][t], b);
    return acc;
}

void orc_get_bmatrix(double B[24], const double x[8], const double y[8], const double z[8])
{
    const double twelth = 1. / 12.;
    for (int p = 0; p < 8; ++p) {
        B[p * 3 + 0] = brow(p, y, z) * twelth;
        B[p * 3 + 1] = brow(p, z, x) * twelth;
        B[p * 3 + 2] = brow(p, x, y) * twelth;
    }
}

static const int *btab_term(int p, int j)
{
    for (int t = 0; t < 6; ++t)
        if (BTAB[p][t][0] == j) return BTAB[p][t];
    return NULL;
}

double orc_get_vol_hex(const double x[8], const double y[8], const double z[8])
{
    const double twelth = 1. / 12.;
    double outer = 0.0;
    for (int o = 0; o < 8; ++o) {
        int i = VOL_OUTER[o];
        double s = 0.0;
        for (int q = 0; q < 6; ++q) {
            int j = VOL_JORDER[o][q];
            double term = y[j] * inner_sum(btab_term(i, j), z);
            s = (q == 0) ? term : s + term;
        }
        double term = x[i] * s;
        outer = (o == 0) ? term : outer + term;
    }
    return outer * twelth;
}

static void elem_xyz(const orc_mesh *m, const orc_state *s, int64_t e, double *x, double *y,
                     double *z)
{
    const int64_t N = m->num_nodes;
    for (int a = 0; a < 8; ++a) {
        int64_t n = m->nodes_in_elem[e * 8 + a];
        x[a] = X1(s, n, 0);
        y[a] = X1(s, n, 1);
        z[a] = X1(s, n, 2);
    }
}

/* get_vol (K/geometry.cpp:238-267) */
void orc_get_vol(const orc_mesh *m, orc_state *s)
{
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->num_elems; ++e) {
        double x[8], y[8], z[8];
        elem_xyz(m, s, e, x, y, z);
        s->elem_vol[e] = orc_get_vol_hex(x, y, z);
    }
}

/* get_divergence (K/momentum.cpp:221-292) */
void orc_get_divergence(const orc_mesh *m, orc_state *s)
{
    const int64_t N = m->num_nodes;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->num_elems; ++e) {
        double x[8], y[8], z[8], B[24], u[8], v[8], w[8];
        elem_xyz(m, s, e, x, y, z);
        orc_get_bmatrix(B, x, y, z);
        for (int a = 0; a < 8; ++a) {
            int64_t n = m->nodes_in_elem[e * 8 + a];
            u[a] = s->node_vel[(N + n) * 3 + 0];
            v[a] = s->node_vel[(N + n) * 3 + 1];
            w[a] = s->node_vel[(N + n) * 3 + 2];
        }
        double inverse_vol = 1.0 / s->elem_vol[e];
        double comp[3];
        const double *f[3] = {u, v, w};
        for (int d = 0; d < 3; ++d) {
            double acc = f[d][0] * B[0 * 3 + d];
            for (int a = 1; a < 8; ++a) acc = acc + f[d][a] * B[a * 3 + d];
            comp[d] = acc;
        }
        double div = 0.0;
        div += comp[0] * inverse_vol;
        div += comp[1] * inverse_vol;
        div += comp[2] * inverse_vol;
        s->elem_div[e] = div;
    }
}

void orc_user_strength(int64_t mat_id, const double *params, double den, double sie,
                       const double vel_grad[9], double vol, double dt, double rk_alpha,
                       const double stress_n[9], double stress[9], double *pres, double *sspd)
{
    (void)mat_id; (void)den; (void)sie; (void)vol; (void)pres; (void)sspd;
#ifdef ORC_TEST_STRENGTH_MODEL
    const double G = params[0];
    const double tr = vel_grad[0] + vel_grad[4] + vel_grad[8];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double d = 0.5 * (vel_grad[3 * i + j] + vel_grad[3 * j + i]) -
                             (i == j ? tr / 3.0 : 0.0);
            stress[3 * i + j] = stress_n[3 * i + j] + rk_alpha * dt * 2.0 * G * d;
        }
#else
    (void)params; (void)vel_grad; (void)dt; (void)rk_alpha; (void)stress_n; (void)stress;
#endif
}

void orc_get_velgrad(double L[9], const double u[8], const double v[8], const double w[8],
                     const double B[24], double vol)
{
    const double *c[3] = {u, v, w};
    double inverse_vol = 1.0 / vol;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = c[i][0] * B[0 * 3 + j];
            for (int p = 1; p < 8; ++p) acc = acc + c[i][p] * B[p * 3 + j];
            L[i * 3 + j] = acc * inverse_vol;
        }
}

static int strength_of(const orc_state *s, int64_t e)
{
    return s->mat_strength ? (int)s->mat_strength[s->elem_mat_id[e]] : 0;
}

/* get_force_sgh, 3D (K/force_sgh.cpp:12-399): tau = stress - p I
 * (K/force_sgh.cpp:127-137), the stress being zero unless a stress field is
 * present; a hypo-elastic material's strength model then updates the
 * element's stress (:369-391).  The curl (:111-123) and the shock direction
 * (:185-212) do not reach the corner force and are omitted. */
void orc_get_force_sgh(const orc_mesh *m, orc_state *s, const orc_material *mat, double fuzz,
                       double small, double dt, double rk_alpha)
{
    const int64_t N = m->num_nodes;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->num_elems; ++e) {
        double x[8], y[8], z[8], B[24], vel[8][3], area[8][3], muc[8];
        elem_xyz(m, s, e, x, y, z);
        orc_get_bmatrix(B, x, y, z);
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) area[a][d] = (-1.0) * B[a * 3 + d];
        for (int a = 0; a < 8; ++a) {
            int64_t n = m->nodes_in_elem[e * 8 + a];
            for (int d = 0; d < 3; ++d) vel[a][d] = s->node_vel[(N + n) * 3 + d];
        }
        const double vol = s->elem_vol[e];
        const double div = s->elem_div[e];
        const double den = s->elem_den[e];
        const double sspd = s->elem_sspd[e];
        const double pres = s->elem_pres[e];

        double *stress = s->elem_stress ? s->elem_stress + (m->num_elems + e) * 9 : NULL;
        double tau[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) tau[i][j] = stress ? stress[i * 3 + j] : 0.0;
        for (int i = 0; i < 3; ++i) tau[i][i] -= pres;

        /* MARS: cell-average velocity (:147-164) */
        double vel_star[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) vel_star[d] += 0.125 * vel[a][d];

        double sum[4] = {0.0, 0.0, 0.0, 0.0};
        const orc_material *mm = &mat[s->elem_mat_id[e]];
        for (int a = 0; a < 8; ++a) {
            double dx = vel[a][0] - vel_star[0];
            double dy = vel[a][1] - vel_star[1];
            double dz = vel[a][2] - vel_star[2];
            double mag_vel = sqrt(dx * dx + dy * dy + dz * dz);
            double mu;
            if (div < 0) mu = den * (mm->q1 * sspd + mm->q2 * mag_vel);        /* :217-220 */
            else         mu = den * (mm->q1ex * sspd + mm->q2ex * mag_vel);    /* :221-224 */
            double mu_term = mu * sqrt(area[a][0] * area[a][0] + area[a][1] * area[a][1] +
                                       area[a][2] * area[a][2]);               /* :242-245 */
            sum[0] += mu_term * vel[a][0];
            sum[1] += mu_term * vel[a][1];
            sum[2] += mu_term * vel[a][2];
            sum[3] += mu_term;
            muc[a] = mu_term;
        }
        if (sum[3] > fuzz) {                                                   /* :261-270 */
            for (int d = 0; d < 3; ++d) vel_star[d] = sum[d] / sum[3];
        } else {
            for (int d = 0; d < 3; ++d) vel_star[d] = 0.0;
        }

        /* shock detector (:287-328) */
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
        const double third = 1.0 / 3.0;
        double c_length = pow(vol, third);
        double alpha = fmin(1.0, omega * (c_length * fabs(div)) / (sspd + fuzz));
        phi = alpha * phi;

        /* corner forces (:338-359) */
        for (int a = 0; a < 8; ++a) {
            int64_t n = m->nodes_in_elem[e * 8 + a];
            for (int d = 0; d < 3; ++d) {
                s->corner_force[(e * 8 + a) * 3 + d] =
                    area[a][0] * tau[0][d] + area[a][1] * tau[1][d] + area[a][2] * tau[2][d] +
                    phi * muc[a] * (vel_star[d] - s->node_vel[(N + n) * 3 + d]);
            }
        }
        /* hypo-elastic strength: the model sees the velocity gradient of the
         * outward B (:92-99) and updates the stress in place */
        if (stress && strength_of(s, e) == 1) {
            double u[8], v[8], w[8], L[9];
            for (int a = 0; a < 8; ++a) {
                u[a] = vel[a][0];
                v[a] = vel[a][1];
                w[a] = vel[a][2];
            }
            orc_get_velgrad(L, u, v, w, B, vol);
            orc_user_strength(s->elem_mat_id[e], orc_strength_params(s, e), den, s->elem_sie[e], L,
                              vol, dt, rk_alpha, s->elem_stress + e * 9, stress, &s->elem_pres[e],
                              &s->elem_sspd[e]);
        }
    }
}

/* update_velocity_sgh (K/momentum.cpp:11-53) */
void orc_update_velocity_sgh(const orc_mesh *m, orc_state *s, double rk_alpha, double dt)
{
    const int64_t N = m->num_nodes;
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
        double f[3] = {0.0, 0.0, 0.0};
        for (int64_t q = m->corners_in_node_start[n]; q < m->corners_in_node_start[n + 1]; ++q) {
            int64_t c = m->corners_in_node[q];
            for (int d = 0; d < 3; ++d) f[d] += s->corner_force[c * 3 + d];
        }
        for (int d = 0; d < 3; ++d)
            s->node_vel[(N + n) * 3 + d] =
                s->node_vel[n * 3 + d] + rk_alpha * dt * f[d] / s->node_mass[n];
    }
}

/* boundary_velocity (K/boundary.cpp:10-55) */
void orc_boundary_velocity(const orc_mesh *m, orc_state *s, const orc_boundary *bc)
{
    const int64_t N = m->num_nodes;
    for (int64_t set = 0; set < m->num_bdy_sets; ++set) {
        int64_t b = m->bdy_nodes_in_set_start[set], e = m->bdy_nodes_in_set_start[set + 1];
        for (int64_t q = b; q < e; ++q) {
            int64_t n = m->bdy_nodes_in_set[q];
            if (bc[set].hydro_bc == 1) {
                s->node_vel[(N + n) * 3 + bc[set].surface] = 0.0;
            } else if (bc[set].hydro_bc == 0) {
                for (int d = 0; d < 3; ++d) s->node_vel[(N + n) * 3 + d] = 0.0;
            }
        }
    }
}

/* update_energy_sgh (K/energy_sgh.cpp:6-55) */
void orc_update_energy_sgh(const orc_mesh *m, orc_state *s, double rk_alpha, double dt)
{
    const int64_t N = m->num_nodes;
    const int64_t E = m->num_elems;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
        double elem_power = 0.0;
        for (int a = 0; a < 8; ++a) {
            int64_t n = m->nodes_in_elem[e * 8 + a];
            double node_radius = 1;
            for (int d = 0; d < 3; ++d) {
                double half_vel = (s->node_vel[(N + n) * 3 + d] + s->node_vel[n * 3 + d]) * 0.5;
                elem_power += s->corner_force[(e * 8 + a) * 3 + d] * node_radius * half_vel;
            }
        }
        s->elem_sie[E + e] = s->elem_sie[e] - rk_alpha * dt / s->elem_mass[e] * elem_power;
    }
}

/* update_position_sgh (K/geometry.cpp:9-26) */
void orc_update_position_sgh(const orc_mesh *m, orc_state *s, double rk_alpha, double dt)
{
    const int64_t N = m->num_nodes;
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
        for (int d = 0; d < 3; ++d) {
            double half_vel = (s->node_vel[(N + n) * 3 + d] + s->node_vel[n * 3 + d]) * 0.5;
            s->node_coords[(N + n) * 3 + d] = s->node_coords[n * 3 + d] + rk_alpha * dt * half_vel;
        }
    }
}

/* The element's EOS, selected by its material (the eos_model pointer,
 * K/state.h:140-147): ideal_gas (K/eos.cpp:12-46) or the user model
 * user_eos_model (K/user_mat.cpp:14-40), which with the strength-free stress
 * gives p = -tr(stress)/3 = 0 and c = 1e-8. */
static void ideal_gas(orc_state *s, int64_t E, int64_t e, double den, double sie)
{
    if (s->mat_eos && s->mat_eos[s->elem_mat_id[e]] == 1) {
        /* p = -tr(stress)/3 of the element's current stress (zero without a
         * stress field) */
        double p = 0.0;
        if (s->elem_stress) {
            const double *st = s->elem_stress + (E + e) * 9;
            for (int i = 0; i < 3; ++i) p -= st[i * 3 + i];
        }
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

/* update_state, 3D (K/properties.cpp:9-114): density, then for a
 * hyper-elastic material the strength model with the velocity gradient of
 * the current state (:46-95), then the EOS */
void orc_update_state(const orc_mesh *m, orc_state *s, double dt, double rk_alpha)
{
    const int64_t E = m->num_elems, N = m->num_nodes;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
        s->elem_den[e] = s->elem_mass[e] / s->elem_vol[e];
        if (s->elem_stress && strength_of(s, e) == 2) {
            double x[8], y[8], z[8], B[24], u[8], v[8], w[8], L[9];
            elem_xyz(m, s, e, x, y, z);
            orc_get_bmatrix(B, x, y, z);
            for (int a = 0; a < 8; ++a) {
                int64_t n = m->nodes_in_elem[e * 8 + a];
                u[a] = s->node_vel[(N + n) * 3 + 0];
                v[a] = s->node_vel[(N + n) * 3 + 1];
                w[a] = s->node_vel[(N + n) * 3 + 2];
            }
            orc_get_velgrad(L, u, v, w, B, s->elem_vol[e]);
            orc_user_strength(s->elem_mat_id[e], orc_strength_params(s, e), s->elem_den[e],
                              s->elem_sie[e], L, s->elem_vol[e], dt, rk_alpha,
                              s->elem_stress + e * 9, s->elem_stress + (E + e) * 9,
                              &s->elem_pres[e], &s->elem_sspd[e]);
        }
        ideal_gas(s, E, e, s->elem_den[e], s->elem_sie[E + e]);
    }
}

/* rk_init (K/time_integration.cpp:9-44) */
void orc_rk_init(const orc_mesh *m, orc_state *s)
{
    const int64_t N = m->num_nodes, E = m->num_elems;
    memcpy(s->elem_sie, s->elem_sie + E, (size_t)E * sizeof(double));
    if (s->elem_stress) memcpy(s->elem_stress, s->elem_stress + 9 * E, (size_t)(9 * E) * sizeof(double));
    memcpy(s->node_coords, s->node_coords + 3 * N, (size_t)(3 * N) * sizeof(double));
    memcpy(s->node_vel, s->node_vel + 3 * N, (size_t)(3 * N) * sizeof(double));
}

/* get_timestep, 3D (K/time_integration.cpp:54-154).  pow(x, 2.0) is x*x. */
void orc_get_timestep(const orc_mesh *m, orc_state *s, double time_value, double graphics_time,
                      double time_final, double dt_max, double dt_min, double dt_cfl, double *dt,
                      double fuzz)
{
    *dt = *dt * 1.1;
    double min_dt_calc = __DBL_MAX__;
#pragma omp parallel for schedule(static) reduction(min : min_dt_calc)
    for (int64_t e = 0; e < m->num_elems; ++e) {
        double x[8], y[8], z[8];
        elem_xyz(m, s, e, x, y, z);
        double dist_min = 0.0;
        int first = 1;
        /* all 28 node pairs a<b (the reference's :101-121 enumeration) */
        for (int a = 0; a < 8; ++a)
            for (int b = a + 1; b < 8; ++b) {
                double dx = x[b] - x[a], dy = y[b] - y[a], dz = z[b] - z[a];
                double dist = fabs(sqrt(dx * dx + dy * dy + dz * dz));
                if (first) { dist_min = dist; first = 0; }
                else dist_min = fmin(dist, dist_min);
            }
        double dt_lcl = dt_cfl * dist_min / (s->elem_sspd[e] + fuzz);
        dt_lcl = fmin(dt_lcl, dt_max);
        dt_lcl = fmax(dt_lcl, dt_min);
        if (dt_lcl < min_dt_calc) min_dt_calc = dt_lcl;
    }
    if (min_dt_calc < *dt) *dt = min_dt_calc;
    *dt = fmin(*dt, (graphics_time - time_value) + fuzz);
    *dt = fmin(*dt, time_final - time_value);
}

/* IE = sum M e, KE = 1/2 sum m |v|^2 (K/sgh_solve.cpp:66-104), ascending order */
void orc_energy_tallies(const orc_mesh *m, const orc_state *s, double *ie, double *ke)
{
    const int64_t N = m->num_nodes, E = m->num_elems;
    double ie_sum = 0.0, ke_sum = 0.0;
    for (int64_t e = 0; e < E; ++e) ie_sum += s->elem_mass[e] * s->elem_sie[E + e];
    for (int64_t n = 0; n < N; ++n) {
        double k = 0;
        for (int d = 0; d < 3; ++d) k += s->node_vel[(N + n) * 3 + d] * s->node_vel[(N + n) * 3 + d];
        ke_sum += s->node_mass[n] * k;
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

/* Threads of the OpenMP loops (for timing the CPU baseline at several
 * counts in one process); returns the count now in effect. */
int orc_set_num_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

/* RK coefficient (K/sgh_solve.cpp:353).  ORC_PLANTED_MISTAKE selects a
 * deliberately wrong coefficient; only the mutant builds made by
 * tests/test_oracle_time_integration.py define it, to show that the temporal
 * convergence tests reject such an integrator. */
double orc_rk_alpha(int64_t num_stages, int64_t stage)
{
#if defined(ORC_PLANTED_MISTAKE) && ORC_PLANTED_MISTAKE == 1
    (void)num_stages;
    (void)stage;
    return 1.0;                                          /* every stage a full step */
#elif defined(ORC_PLANTED_MISTAKE) && ORC_PLANTED_MISTAKE == 2
    return 1.0 / ((double)num_stages + (double)stage);   /* weights 1/2, 1/3 */
#elif defined(ORC_PLANTED_MISTAKE) && ORC_PLANTED_MISTAKE == 3
    return 0.5 / ((double)num_stages - (double)stage);   /* weights 1/4, 1/2: second
                                                             order, but for y' = f/2 */
#else
    return 1.0 / ((double)num_stages - (double)stage);
#endif
}

/* sgh_solve (K/sgh_solve.cpp:8-692), 3D branch, outputs disabled. */
void orc_sgh_solve(const orc_mesh *m, orc_state *s, const orc_material *mat,
                   const orc_boundary *bc, const orc_run_params *p, orc_run_result *r)
{
    double time_value = p->time_value;
    double graphics_time = p->graphics_time;
    double dt = p->dt_start;
    orc_energy_tallies(m, s, &r->ie0, &r->ke0);
    r->te0 = r->ie0 + r->ke0;
    int64_t cycle = 0, done = 0;
    double t0 = now_seconds();
    for (cycle = 0; cycle < p->cycle_stop; ++cycle) {
        if (p->dt_replay) dt = p->dt_replay[cycle];
        else orc_get_timestep(m, s, time_value, graphics_time, p->time_final, p->dt_max, p->dt_min,
                              p->dt_cfl, &dt, p->fuzz);
        if (p->dt_record) p->dt_record[cycle] = dt;
        orc_rk_init(m, s);
        for (int64_t stage = 0; stage < p->rk_num_stages; ++stage) {
            double rk_alpha = orc_rk_alpha(p->rk_num_stages, stage);
            orc_get_divergence(m, s);
            orc_get_force_sgh(m, s, mat, p->fuzz, p->small, dt, rk_alpha);
            orc_update_velocity_sgh(m, s, rk_alpha, dt);
            orc_boundary_velocity(m, s, bc);
            orc_update_energy_sgh(m, s, rk_alpha, dt);
            orc_update_position_sgh(m, s, rk_alpha, dt);
            orc_get_vol(m, s);
            orc_update_state(m, s, dt, rk_alpha);
        }
        time_value += dt;
        int write = 0;                                     /* :594-606 */
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
    orc_energy_tallies(m, s, &r->ie1, &r->ke1);
    r->te1 = r->ie1 + r->ke1;
}

/* ------------------------------------------------------------------------ */
/* mesh input                                                                */
/* ------------------------------------------------------------------------ */
void orc_gen_box(int64_t nx, int64_t ny, int64_t nz, double lx_num, double ly_num,
                 double lz_num, double l_den, double *coords, int64_t *nodes_in_elem)
{
    const int64_t nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
    for (int64_t k = 0; k < nz1; ++k)
        for (int64_t j = 0; j < ny1; ++j)
            for (int64_t i = 0; i < nx1; ++i) {
                int64_t n = i + nx1 * (j + ny1 * k);
                coords[n * 3 + 0] = ((double)i * lx_num) / ((double)nx * l_den);
                coords[n * 3 + 1] = ((double)j * ly_num) / ((double)ny * l_den);
                coords[n * 3 + 2] = ((double)k * lz_num) / ((double)nz * l_den);
            }
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t e = i + nx * (j + ny * k);
                int64_t n0 = i + nx1 * (j + ny1 * k);
                int64_t *c = nodes_in_elem + e * 8;
                c[0] = n0;
                c[1] = n0 + 1;
                c[2] = n0 + 1 + nx1;
                c[3] = n0 + nx1;
                c[4] = c[0] + nx1 * ny1;
                c[5] = c[1] + nx1 * ny1;
                c[6] = c[2] + nx1 * ny1;
                c[7] = c[3] + nx1 * ny1;
            }
}

static int skip_line(FILE *f)
{
    int ch;
    while ((ch = fgetc(f)) != '\n')
        if (ch == EOF) return -1;
    return 0;
}

/* read_mesh_ensight (K/read_mesh.cpp:12-147): 8 header lines, N, x/y/z blocks,
 * one line ("hexa8"), E, 8 one-based ids per element. */
int orc_read_ensight(const char *path, int64_t *num_nodes, int64_t *num_elems, double *coords,
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
            if (coords) coords[n * 3 + d] = v;
        }
    fgetc(f);
    if (skip_line(f)) { fclose(f); return -5; }
    if (fscanf(f, "%lld", &ne) != 1) { fclose(f); return -6; }
    if (nodes_in_elem) {
        for (long long e = 0; e < ne; ++e)
            for (int a = 0; a < 8; ++a) {
                long long id;
                if (fscanf(f, "%lld", &id) != 1) { fclose(f); return -7; }
                nodes_in_elem[e * 8 + a] = id - 1;
            }
    }
    fclose(f);
    *num_nodes = nn;
    *num_elems = ne;
    return 0;
}

/* build_corner_connectivity (K/mesh.h:180-245): ascending element order per node */
void orc_build_corner_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                   int64_t *start, int64_t *corners_in_node,
                                   int64_t *elems_in_node, int64_t *corners_in_elem)
{
    int64_t *count = (int64_t *)calloc((size_t)N, sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e)
        for (int a = 0; a < 8; ++a) count[nodes_in_elem[e * 8 + a]]++;
    start[0] = 0;
    for (int64_t n = 0; n < N; ++n) start[n + 1] = start[n] + count[n];
    memset(count, 0, (size_t)N * sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e)
        for (int a = 0; a < 8; ++a) {
            int64_t n = nodes_in_elem[e * 8 + a];
            int64_t j = count[n]++;
            int64_t c = a + e * 8;
            corners_in_node[start[n] + j] = c;
            elems_in_node[start[n] + j] = e;
            corners_in_elem[e * 8 + a] = c;
        }
    free(count);
}

/* build_elem_elem_connectivity (K/mesh.h:249-338): per element, node_lid 0..7,
 * then that node's elements in stored order, skipping self and repeats. */
int64_t orc_build_elem_elem_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                         const int64_t *start, const int64_t *elems_in_node,
                                         int64_t *ee_start, int64_t *ee)
{
    (void)N;
    int64_t total = 0;
    ee_start[0] = 0;
    for (int64_t e = 0; e < E; ++e) {
        int64_t saved[64];
        int64_t ns = 0;
        for (int a = 0; a < 8; ++a) {
            int64_t n = nodes_in_elem[e * 8 + a];
            for (int64_t q = start[n]; q < start[n + 1]; ++q) {
                int64_t nb = elems_in_node[q];
                int save = nb != e;
                for (int64_t i = 0; i < ns && save; ++i)
                    if (saved[i] == nb) save = 0;
                if (save) saved[ns++] = nb;
            }
        }
        if (ee)
            for (int64_t i = 0; i < ns; ++i) ee[total + i] = saved[i];
        total += ns;
        ee_start[e + 1] = total;
    }
    return total;
}

/* node lids of the 6 patches of a hex (K/mesh.h:353-358) */
static const int PATCH_LIDS[6][4] = {{0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4},
                                     {2, 3, 7, 6}, {0, 3, 2, 1}, {4, 5, 6, 7}};

static void sort4(uint64_t *v)
{
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 3 - i; ++j)
            if (v[j] > v[j + 1]) { uint64_t t = v[j]; v[j] = v[j + 1]; v[j + 1] = t; }
}

/* build_patch_connectivity (K/mesh.h:342-649).  The hash key is formed exactly
 * as the reference does, in 64-bit unsigned arithmetic (which wraps for
 * N >= 2^21, a reference defect that only matters on meshes larger than the
 * ones this oracle is run on). */
int64_t orc_build_patch_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                     const int64_t *ee_start, const int64_t *ee,
                                     int64_t *patches_in_elem, int64_t *nodes_in_patch,
                                     int64_t *bdy_patches, int64_t *bdy_nodes, int64_t *counts)
{
    long long *hash = (long long *)malloc((size_t)E * 6 * sizeof(long long));
    int64_t *pie = (int64_t *)calloc((size_t)E * 6, sizeof(int64_t));
    int64_t *tbdy = (int64_t *)malloc((size_t)E * 6 * sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e)
        for (int pl = 0; pl < 6; ++pl) {
            uint64_t v[4];
            for (int q = 0; q < 4; ++q) v[q] = (uint64_t)nodes_in_elem[e * 8 + PATCH_LIDS[pl][q]];
            sort4(v);
            uint64_t un = (uint64_t)N;
            uint64_t key = v[1] + un * v[2] + un * un * v[3];
            hash[e * 6 + pl] = -(long long)key;
        }
    int64_t patch_gid = 0, bdy_patch_gid = 0;
    for (int64_t e = 0; e < E; ++e) {
        for (int pl = 0; pl < 6; ++pl) {
            long long key = hash[e * 6 + pl];
            if (key < 0) {
                int found = 0;
                for (int64_t q = ee_start[e]; q < ee_start[e + 1] && !found; ++q) {
                    int64_t nb = ee[q];
                    for (int np = 0; np < 6; ++np) {
                        if (hash[nb * 6 + np] == hash[e * 6 + pl]) {
                            hash[e * 6 + pl] = nb;
                            hash[nb * 6 + np] = e;
                            pie[e * 6 + pl] = patch_gid;
                            pie[nb * 6 + np] = patch_gid;
                            patch_gid++;
                            found = 1;
                            break;
                        }
                    }
                }
            }
        }
        for (int pl = 0; pl < 6; ++pl) {
            if (hash[e * 6 + pl] < 0) {
                hash[e * 6 + pl] = e;
                pie[e * 6 + pl] = patch_gid;
                tbdy[bdy_patch_gid] = patch_gid;
                patch_gid++;
                bdy_patch_gid++;
            }
        }
    }
    int64_t P = patch_gid;
    counts[0] = P;
    counts[1] = bdy_patch_gid;
    if (patches_in_elem) memcpy(patches_in_elem, pie, (size_t)E * 6 * sizeof(int64_t));
    /* nodes_in_patch: the last element that visits a patch writes its nodes
     * (K/mesh.h:556-581 visits elements in ascending order) */
    int64_t *nip = (int64_t *)calloc((size_t)P * 4, sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e)
        for (int pl = 0; pl < 6; ++pl) {
            int64_t pg = pie[e * 6 + pl];
            for (int q = 0; q < 4; ++q) nip[pg * 4 + q] = nodes_in_elem[e * 8 + PATCH_LIDS[pl][q]];
        }
    if (nodes_in_patch) memcpy(nodes_in_patch, nip, (size_t)P * 4 * sizeof(int64_t));
    if (bdy_patches) memcpy(bdy_patches, tbdy, (size_t)bdy_patch_gid * sizeof(int64_t));
    /* boundary nodes in first-seen order (K/mesh.h:594-645) */
    char *seen = (char *)calloc((size_t)N, 1);
    int64_t nb_nodes = 0;
    for (int64_t b = 0; b < bdy_patch_gid; ++b) {
        int64_t pg = tbdy[b];
        for (int q = 0; q < 4; ++q) {
            int64_t n = nip[pg * 4 + q];
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

/* check_bdy (K/setup.cpp:487-571) */
static int check_bdy(const double *xyz, const int64_t *nodes_in_patch, int64_t pg, int tag,
                     double val)
{
    int on = 0;
    for (int q = 0; q < 4; ++q) {
        const double *c = xyz + nodes_in_patch[pg * 4 + q] * 3;
        if (tag == 0) { if (fabs(c[0] - val) <= 1.0e-7) on += 1; }
        else if (tag == 1) { if (fabs(c[1] - val) <= 1.0e-7) on += 1; }
        else if (tag == 2) { if (fabs(c[2] - val) <= 1.0e-7) on += 1; }
        else if (tag == 3) {
            double R = sqrt(c[0] * c[0] + c[1] * c[1]);
            if (fabs(R - val) <= 1.0e-7) on += 1;
        } else if (tag == 4) {
            double R = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            if (fabs(R - val) <= 1.0e-7) on += 1;
        }
    }
    return on == 4;
}

/* tag_bdys + build_boundry_node_sets (K/setup.cpp:431-665) */
int64_t orc_build_bdy_sets(const double *xyz, int64_t num_bdy_patches, const int64_t *bdy_patches,
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
            if (!check_bdy(xyz, nodes_in_patch, pg, (int)bc[set].surface, bc[set].value)) continue;
            for (int q = 0; q < 4; ++q) {
                int64_t n = nodes_in_patch[pg * 4 + q];
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

/* setup fills (K/setup.cpp:106-357); PI as in the reference's tg_vortex */
void orc_setup_fills(const orc_mesh *m, orc_state *s, int64_t num_fills, const orc_fill *fills,
                     const double *state_vars, const int64_t *mat_num_state_vars)
{
    const int64_t N = m->num_nodes, E = m->num_elems;
    const double PI = 3.14159265358979323846;
    for (int64_t f = 0; f < num_fills; ++f) {
        const orc_fill *F = &fills[f];
        /* serial: several elements write the same node velocity (a benign race
         * in the reference, K/setup.cpp:236-337); ascending order here */
        for (int64_t e = 0; e < E; ++e) {
            double c[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < 8; ++a) {
                int64_t n = m->nodes_in_elem[e * 8 + a];
                c[0] += X1(s, n, 0);
                c[1] += X1(s, n, 1);
                c[2] += X1(s, n, 2);
            }
            c[0] = c[0] / 8;
            c[1] = c[1] / 8;
            c[2] = c[2] / 8;
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
            ideal_gas(s, E, e, s->elem_den[e], s->elem_sie[E + e]);
            for (int a = 0; a < 8; ++a) {
                int64_t n = m->nodes_in_elem[e * 8 + a];
                double *v = &s->node_vel[(N + n) * 3];
                if (F->velocity == 0) {
                    v[0] = F->u;
                    v[1] = F->v;
                    v[2] = F->w;
                } else if (F->velocity == 1 || F->velocity == 2) {
                    int nd = F->velocity == 1 ? 2 : 3;
                    double dir[3] = {0.0, 0.0, 0.0}, rv = 0.0;
                    for (int d = 0; d < nd; ++d) {
                        dir[d] = X1(s, n, d);
                        rv += X1(s, n, d) * X1(s, n, d);
                    }
                    rv = sqrt(rv);
                    for (int d = 0; d < nd; ++d) dir[d] = rv > 1.0e-14 ? dir[d] / rv : 0.0;
                    v[0] = F->speed * dir[0];
                    v[1] = F->speed * dir[1];
                    v[2] = nd == 3 ? F->speed * dir[2] : 0.0;
                } else if (F->velocity == 5) {
                    v[0] = sin(PI * X1(s, n, 0)) * cos(PI * X1(s, n, 1));
                    v[1] = -1.0 * cos(PI * X1(s, n, 0)) * sin(PI * X1(s, n, 1));
                    v[2] = 0.0;
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

/* node mass, 3D (K/setup.cpp:393-420): m_n = sum over elems_in_node of M_e/8 */
void orc_node_mass(const orc_mesh *m, orc_state *s)
{
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < m->num_nodes; ++n) {
        double acc = 0.0;
        for (int64_t q = m->corners_in_node_start[n]; q < m->corners_in_node_start[n + 1]; ++q)
            acc += 1.0 / 8.0 * s->elem_mass[m->elems_in_node[q]];
        s->node_mass[n] = acc;
    }
}
