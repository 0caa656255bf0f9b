/*
 * sgh_oracle.h -- CPU restatement of Fierro's 3D explicit Lagrangian SGH cycle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in fierro_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls it.
 *
 * Every function restates one reference routine (paths relative to the
 * reference root, K/ = Explicit-Lagrange-Kokkos/SGH_solver/) in the same
 * floating-point operation order, compiled with -ffp-contract=off, so that the
 * HIP kernels (built the same way) can be compared bit-for-bit or within the
 * stated tolerance.
 *
 * Data layout mirrors the reference's MATAR arrays (row-major):
 *   node_coords, node_vel : [2 bins][N][3]       (K/driver.cpp:172-180)
 *   elem_sie              : [2 bins][E]          (K/driver.cpp:202-204)
 *   corner_force          : [8E][3], corner = 8e+i (K/mesh.h:228)
 *   nodes_in_elem         : [E][8], 0-based      (K/read_mesh.cpp:109-117)
 * Indices are int64_t (the reference uses size_t).
 */
#ifndef FIERRO_SGH_ORACLE_H
#define FIERRO_SGH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mesh connectivity, ragged arrays stored as CSR (start has n+1 entries). */
typedef struct orc_mesh {
    int64_t num_nodes;
    int64_t num_elems;
    const int64_t *nodes_in_elem;          /* E x 8 */
    const int64_t *corners_in_node_start;  /* N+1 */
    const int64_t *corners_in_node;        /* 8E  */
    const int64_t *elems_in_node;          /* 8E, same strides as corners_in_node */
    const int64_t *elems_in_elem_start;    /* E+1 */
    const int64_t *elems_in_elem;          /* NB  */
    int64_t num_bdy_sets;
    const int64_t *bdy_nodes_in_set_start; /* num_bdy_sets+1 */
    const int64_t *bdy_nodes_in_set;
} orc_mesh;

typedef struct orc_state {
    double *node_coords;   /* 2 x N x 3 */
    double *node_vel;      /* 2 x N x 3 */
    double *node_mass;     /* N */
    double *elem_den;      /* E */
    double *elem_pres;     /* E */
    double *elem_sspd;     /* E */
    double *elem_sie;      /* 2 x E */
    double *elem_vol;      /* E */
    double *elem_div;      /* E */
    double *elem_mass;     /* E */
    int64_t *elem_mat_id;  /* E */
    double *elem_statev;   /* E x num_state_vars; col 0 = gamma, col 1 = min sound speed */
    int64_t num_state_vars;
    double *corner_force;  /* 8E x 3 */
    const int64_t *mat_eos; /* per material: 0 ideal gas, 1 user model; NULL = all ideal gas */
    double *elem_stress;   /* 2 x E x 3 x 3 Cauchy stress (K/state.h:64), or NULL = zero */
    const int64_t *mat_strength; /* per material: 0 none, 1 hypo, 2 hyper; NULL = none */
    const double *mat_strength_params; /* per material ORC_STRENGTH_PARAMS values, or NULL */
} orc_state;

#define ORC_STRENGTH_PARAMS 8

/* the strength parameters of element e's material (zeros when none are set) */
static inline const double *orc_strength_params(const orc_state *s, int64_t e)
{
    static const double zeros[ORC_STRENGTH_PARAMS] = {0.0};
    return s->mat_strength_params
               ? s->mat_strength_params + s->elem_mat_id[e] * ORC_STRENGTH_PARAMS
               : zeros;
}

/* material_t (K/state.h:130-178), ideal-gas EOS, no strength model */
typedef struct orc_material {
    double q1, q1ex, q2, q2ex;
} orc_material;

/* boundary_t (K/mesh.h:924-935); surface: 0 x-plane, 1 y-plane, 2 z-plane,
 * 3 cylinder, 4 sphere; hydro_bc: 0 fixed, 1 reflected (others: no-op) */
typedef struct orc_boundary {
    int64_t surface;
    double value;
    int64_t hydro_bc;
} orc_boundary;

/* mat_fill_t (K/mesh.h:858-890); volume: 0 global, 1 box, 2 cylinder, 3 sphere;
 * velocity: 0 cartesian, 1 radial, 2 spherical, 5 Taylor-Green */
typedef struct orc_fill {
    int64_t volume;
    int64_t mat_id;
    double x1, x2, y1, y2, z1, z2;
    double radius1, radius2;
    int64_t velocity;
    double u, v, w;
    double speed;
    double sie;
    double den;
} orc_fill;

/* Controls of sgh_solve (K/sgh_solve.cpp:8-41, K/driver.cpp:35-62). */
typedef struct orc_run_params {
    double time_value;      /* start time */
    double time_final;
    double dt_start;
    double dt_max, dt_min, dt_cfl;
    double graphics_time;   /* next graphics time */
    double graphics_dt_ival;
    int64_t graphics_cyc_ival;
    int64_t cycle_stop;
    int64_t rk_num_stages;
    double fuzz, small;
    const double *dt_replay;  /* non-NULL: use dt_replay[cycle] instead of get_timestep
                                 (frozen time steps, for derivative checks) */
    double *dt_record;        /* non-NULL: store the dt of every cycle */
    int64_t *write_record;    /* non-NULL: 1 for each cycle after which graphics output
                                 is written (K/sgh_solve.cpp:594-628), else 0 */
} orc_run_params;

typedef struct orc_run_result {
    int64_t cycles;         /* cycles executed */
    double time_value;      /* final time */
    double dt;              /* last dt */
    double graphics_time;
    double ie0, ke0, te0;   /* tallies at t0 (K/sgh_solve.cpp:66-104) */
    double ie1, ke1, te1;   /* tallies at the end (K/sgh_solve.cpp:645-684) */
    double seconds;         /* wall time of the cycle loop */
} orc_run_result;

/* ---- strength model hook (K/user_mat.cpp:47-90) ---- */
/* user_strength_model restated: the reference ships an empty body.  sie is
 * the energy at t_n (the reference passes elem_sie(elem_gid), bin 0 of the
 * [2][E] array), stress_n the t_n stress (bin 0), stress the current stress,
 * updated in place.  Built with ORC_TEST_STRENGTH_MODEL (the oracle's test
 * variant) the body is the hypoelastic update
 * stress = stress_n + rk_alpha dt 2 G dev((L + L^T)/2), G = params[0]. */
void orc_user_strength(int64_t mat_id, const double *params, double den, double sie,
                       const double vel_grad[9], double vol, double dt, double rk_alpha,
                       const double stress_n[9], double stress[9], double *pres, double *sspd);
/* get_velgrad (K/momentum.cpp:59-144): L_ij = sum_p u_p^i B_pj / V with the
 * outward B */
void orc_get_velgrad(double L[9], const double u[8], const double v[8], const double w[8],
                     const double B[24], double vol);

/* ---- geometry (K/geometry.cpp) ---- */
void   orc_get_bmatrix(double B[24], const double x[8], const double y[8], const double z[8]);
double orc_get_vol_hex(const double x[8], const double y[8], const double z[8]);

/* ---- per-stage operators, same semantics as the reference ---- */
void orc_get_vol(const orc_mesh *m, orc_state *s);
void orc_get_divergence(const orc_mesh *m, orc_state *s);
void orc_get_force_sgh(const orc_mesh *m, orc_state *s, const orc_material *mat,
                       double fuzz, double small, double dt, double rk_alpha);
void orc_update_velocity_sgh(const orc_mesh *m, orc_state *s, double rk_alpha, double dt);
void orc_boundary_velocity(const orc_mesh *m, orc_state *s, const orc_boundary *bc);
void orc_update_energy_sgh(const orc_mesh *m, orc_state *s, double rk_alpha, double dt);
void orc_update_position_sgh(const orc_mesh *m, orc_state *s, double rk_alpha, double dt);
void orc_update_state(const orc_mesh *m, orc_state *s, double dt, double rk_alpha);
void orc_rk_init(const orc_mesh *m, orc_state *s);
void orc_get_timestep(const orc_mesh *m, orc_state *s, double time_value, double graphics_time,
                      double time_final, double dt_max, double dt_min, double dt_cfl,
                      double *dt, double fuzz);
void orc_energy_tallies(const orc_mesh *m, const orc_state *s, double *ie, double *ke);

/* Threads of the OpenMP loops; returns the count in effect (n <= 0: query). */
int orc_set_num_threads(int n);

/* RK coefficient of stage s of an n-stage cycle: 1/(n - s) (K/sgh_solve.cpp:353). */
double orc_rk_alpha(int64_t num_stages, int64_t stage);

/* ---- cycle driver (K/sgh_solve.cpp:63-688) ---- */
void orc_sgh_solve(const orc_mesh *m, orc_state *s, const orc_material *mat,
                   const orc_boundary *bc, const orc_run_params *p, orc_run_result *r);

/* ---- mesh input and connectivity (K/read_mesh.cpp, K/mesh.h) ---- */
/* Structured box of nx*ny*nz hexes on [0,Lx]x[0,Ly]x[0,Lz] numbered like the
 * reference's EnSight meshes: node (i,j,k) = i + (nx+1)*(j + (ny+1)*k),
 * element nodes (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k), then the k+1 layer.
 * Coordinates are k*L/n with L = Lnum/Lden computed as (k*Lnum)/(n*Lden). */
void orc_gen_box(int64_t nx, int64_t ny, int64_t nz,
                 double lx_num, double ly_num, double lz_num, double l_den,
                 double *coords /* N x 3 */, int64_t *nodes_in_elem /* E x 8 */);
/* EnSight Gold ASCII reader; call with coords/nodes NULL to get sizes. */
int  orc_read_ensight(const char *path, int64_t *num_nodes, int64_t *num_elems,
                      double *coords /* N x 3 */, int64_t *nodes_in_elem /* E x 8 */);

void orc_build_corner_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                   int64_t *corners_in_node_start, int64_t *corners_in_node,
                                   int64_t *elems_in_node, int64_t *corners_in_elem);
/* returns the total number of elems_in_elem entries; fills start (E+1) and,
 * when out is non-NULL, the entries */
int64_t orc_build_elem_elem_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                         const int64_t *corners_in_node_start,
                                         const int64_t *elems_in_node,
                                         int64_t *elems_in_elem_start, int64_t *elems_in_elem);
/* Patch numbering (K/mesh.h:342-649).  patches_in_elem: E x 6.  Returns the number of
 * patches; counts[0] = patches, counts[1] = bdy patches, counts[2] = bdy nodes.
 * nodes_in_patch (P x 4), bdy_patches, bdy_nodes may be NULL on a sizing call. */
int64_t orc_build_patch_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                     const int64_t *elems_in_elem_start,
                                     const int64_t *elems_in_elem,
                                     int64_t *patches_in_elem, int64_t *nodes_in_patch,
                                     int64_t *bdy_patches, int64_t *bdy_nodes,
                                     int64_t *counts);
/* Boundary sets (K/setup.cpp:431-665): patches tagged by check_bdy then nodes in
 * first-seen order.  out_start has nsets+1 entries; out_nodes capacity must be
 * >= nsets*num_bdy_patches*4.  Returns the total number of set entries. */
int64_t orc_build_bdy_sets(const double *node_coords_bin1 /* N x 3 */,
                           int64_t num_bdy_patches, const int64_t *bdy_patches,
                           const int64_t *nodes_in_patch, int64_t nsets,
                           const orc_boundary *bc, int64_t N,
                           int64_t *out_start, int64_t *out_nodes);

/* ---- initial conditions (K/setup.cpp:106-424) ---- */
void orc_setup_fills(const orc_mesh *m, orc_state *s, int64_t num_fills, const orc_fill *fills,
                     const double *state_vars /* nmat x num_state_vars */,
                     const int64_t *mat_num_state_vars);
void orc_node_mass(const orc_mesh *m, orc_state *s);

#ifdef __cplusplus
}
#endif
#endif
