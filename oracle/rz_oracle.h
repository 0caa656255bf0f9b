/*
 * rz_oracle.h -- CPU restatement of Fierro's 2D axisymmetric (RZ) SGH path.
 *
 * TEST INFRASTRUCTURE ONLY, like sgh_oracle.h: only tests/, the smoke check
 * and the bench's CPU baseline load it; the product path never calls it.
 *
 * Every function restates one reference routine (K/ = Explicit-Lagrange-
 * Kokkos/SGH_solver/) in the same floating-point operation order, compiled
 * with -ffp-contract=off.  Coordinates are (x, y) = (axial z, radius R): the
 * reference takes the radius from node_coords(1, node, 1).
 *
 * Data layout (reference MATAR arrays with num_dims = 2, row-major):
 *   node_coords, node_vel : [2 bins][N][2]
 *   elem_sie              : [2 bins][E]
 *   corner_force          : [4E][2], corner = 4e+i
 *   nodes_in_elem         : [E][4], counter-clockwise 0-1-2-3
 * The orc_state structure of sgh_oracle.h is reused with these shapes.
 */
#ifndef FIERRO_RZ_ORACLE_H
#define FIERRO_RZ_ORACLE_H

#include <stdint.h>

#include "sgh_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orz_mesh {
    int64_t num_nodes;
    int64_t num_elems;
    const int64_t *nodes_in_elem;          /* E x 4 */
    const int64_t *corners_in_node_start;  /* N+1 */
    const int64_t *corners_in_node;        /* 4E */
    const int64_t *elems_in_elem_start;    /* E+1 */
    const int64_t *elems_in_elem;
    int64_t num_bdy_sets;
    const int64_t *bdy_nodes_in_set_start; /* num_bdy_sets+1 */
    const int64_t *bdy_nodes_in_set;
    int64_t num_bdy_nodes;
    const int64_t *bdy_nodes;              /* boundary nodes of the mesh, first-seen order */
    const int64_t *nodes_in_node_start;    /* N+1 */
    const int64_t *nodes_in_node;          /* edge neighbours */
} orz_mesh;

/* ---- geometry (K/geometry.cpp:314-547) ---- */
void   orz_get_bmatrix2D(double B[8], const double x[4], const double y[4]);
double orz_get_area_quad(const double x[4], const double y[4]);
double orz_get_vol_quad(const double x[4], const double y[4]);
double orz_heron(double x1, double y1, double x2, double y2, double x3, double y3);
void   orz_get_area_weights2D(double A[4], const double x[4], const double y[4]);

/* ---- per-stage operators ---- */
void orz_get_vol(const orz_mesh *m, orc_state *s);
void orz_get_divergence(const orz_mesh *m, orc_state *s);
void orz_get_force(const orz_mesh *m, orc_state *s, const orc_material *mat, double fuzz,
                   double small, double dt, double rk_alpha);
void orz_get_velgrad2D(double L[9], const double u[4], const double v[4], const double B[8],
                       double elem_vol, double elem_area);
void orz_update_velocity(const orz_mesh *m, orc_state *s, double rk_alpha, double dt);
void orz_boundary_velocity(const orz_mesh *m, orc_state *s, const orc_boundary *bc);
void orz_update_energy(const orz_mesh *m, orc_state *s, double rk_alpha, double dt);
void orz_update_position(const orz_mesh *m, orc_state *s, double rk_alpha, double dt);
void orz_update_state(const orz_mesh *m, orc_state *s);
void orz_rk_init(const orz_mesh *m, orc_state *s);
void orz_get_timestep(const orz_mesh *m, orc_state *s, double time_value, double graphics_time,
                      double time_final, double dt_max, double dt_min, double dt_cfl,
                      double *dt, double fuzz);
/* node_extensive_mass = areal node mass x radius (K/sgh_solve.cpp:107-116) */
void orz_extensive_mass(const orz_mesh *m, const orc_state *s, double *ext);
/* areal node mass from the extensive mass and the current radius, with the
 * on-axis fix (K/sgh_solve.cpp:495-581) */
void orz_update_node_mass(const orz_mesh *m, orc_state *s, const double *ext, double tiny);
/* IE = sum M e, KE = 1/2 sum m R |v|^2 (K/sgh_solve.cpp:66-104, 2D branch) */
void orz_energy_tallies(const orz_mesh *m, const orc_state *s, double *ie, double *ke);

/* ---- cycle driver (K/sgh_solve.cpp:63-688, 2D branches) ---- */
void orz_sgh_solve(const orz_mesh *m, orc_state *s, const orc_material *mat,
                   const orc_boundary *bc, const orc_run_params *p, double tiny,
                   orc_run_result *r);

/* ---- mesh input and connectivity ---- */
/* nx x ny quads on [0,Lx]x[0,Ly], L = L_num/L_den: node (i,j) = i + (nx+1) j
 * at ((i Lx_num)/(nx L_den), (j Ly_num)/(ny L_den)); element nodes (i,j),
 * (i+1,j), (i+1,j+1), (i,j+1). */
void orz_gen_rect(int64_t nx, int64_t ny, double lx_num, double ly_num, double l_den,
                  double *coords /* N x 2 */, int64_t *nodes_in_elem /* E x 4 */);
/* EnSight Gold ASCII reader, 2D (K/read_mesh.cpp:12-147 with num_dims = 2:
 * the z block is read and dropped, 4 node ids per element).  Call with
 * coords/nodes NULL to get the sizes. */
int orz_read_ensight(const char *path, int64_t *num_nodes, int64_t *num_elems,
                     double *coords /* N x 2 */, int64_t *nodes_in_elem /* E x 4 */);
void orz_build_corner_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                   int64_t *corners_in_node_start, int64_t *corners_in_node,
                                   int64_t *elems_in_node);
int64_t orz_build_elem_elem_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                         const int64_t *corners_in_node_start,
                                         const int64_t *elems_in_node,
                                         int64_t *elems_in_elem_start, int64_t *elems_in_elem);
/* Patches = element edges (K/mesh.h:342-649, num_dims = 2).  counts[0] =
 * patches, counts[1] = boundary patches, counts[2] = boundary nodes. */
int64_t orz_build_patch_connectivity(int64_t N, int64_t E, const int64_t *nodes_in_elem,
                                     const int64_t *elems_in_elem_start,
                                     const int64_t *elems_in_elem, int64_t *nodes_in_patch,
                                     int64_t *bdy_patches, int64_t *bdy_nodes, int64_t *counts);
/* nodes_in_node from the patches (K/mesh.h:651-796, 2D branch); start has
 * N+1 entries; out capacity 2 x patches. */
void orz_build_node_node_connectivity(int64_t N, int64_t P, const int64_t *nodes_in_patch,
                                      int64_t *start, int64_t *out);
/* Boundary sets over the 2-node patches (surfaces 0, 1 and 3 only: the
 * reference's check_bdy reads an unset third coordinate for 2 and 4). */
int64_t orz_build_bdy_sets(const double *xy, int64_t num_bdy_patches, const int64_t *bdy_patches,
                           const int64_t *nodes_in_patch, int64_t nsets, const orc_boundary *bc,
                           int64_t N, int64_t *out_start, int64_t *out_nodes);

/* ---- initial conditions (K/setup.cpp:106-424, 2D) ---- */
void orz_setup_fills(const orz_mesh *m, orc_state *s, int64_t num_fills, const orc_fill *fills,
                     const double *state_vars, const int64_t *mat_num_state_vars);
/* areal node mass: sum over corners_in_node of corner area x density */
void orz_node_mass(const orz_mesh *m, orc_state *s);

#ifdef __cplusplus
}
#endif
#endif
