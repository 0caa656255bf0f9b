/* abi_check.c -- include/fsgh.h from C11: every entry point is referenced
 * (so each declaration is valid C and resolves when linked against
 * libfsgh.so), and a host-only context (device -1: connectivity and boundary
 * sets, no GPU) is driven through the ABI.
 *
 *   abi_check <mesh.txt>
 *
 * mesh.txt: "N E", coords N x 3, nodes_in_elem E x 8, then "num_bcs" and
 * "surface hydro_bc value" per set.  Prints "entry_points <count>", the
 * corner CSR ("cin_start ...", "cin ...") and one "set <k> <nodes...>" line
 * per boundary set, in the library's order. */
#include <stdio.h>
#include <stdlib.h>

#include "fsgh.h"

typedef void (*entry_fn)(void);
#define ENTRY(f) ((entry_fn)(f))

static const entry_fn entry_points[] = {
    ENTRY(fsgh_create), ENTRY(fsgh_destroy), ENTRY(fsgh_last_error), ENTRY(fsgh_device_count),
    ENTRY(fsgh_get_corners_in_node), ENTRY(fsgh_get_renumbering), ENTRY(fsgh_get_device_csr),
    ENTRY(fsgh_get_stage_info), ENTRY(fsgh_set_materials), ENTRY(fsgh_set_material_models),
    ENTRY(fsgh_set_strength_params), ENTRY(fsgh_set_bcs), ENTRY(fsgh_bc_set_size),
    ENTRY(fsgh_get_bc_set), ENTRY(fsgh_setup_fills), ENTRY(fsgh_upload), ENTRY(fsgh_download),
    ENTRY(fsgh_rk_init), ENTRY(fsgh_get_timestep), ENTRY(fsgh_get_divergence),
    ENTRY(fsgh_get_force_sgh), ENTRY(fsgh_update_velocity_sgh), ENTRY(fsgh_boundary_velocity),
    ENTRY(fsgh_update_energy_sgh), ENTRY(fsgh_update_position_sgh), ENTRY(fsgh_get_vol),
    ENTRY(fsgh_update_state), ENTRY(fsgh_energy), ENTRY(fsgh_energy_noalloc),
    ENTRY(fsgh_rz_extensive_mass), ENTRY(fsgh_get_velgrad), ENTRY(fsgh_rz_update_node_mass),
    ENTRY(fsgh_run_begin), ENTRY(fsgh_run), ENTRY(fsgh_run_status), ENTRY(fsgh_autotune),
    ENTRY(fsgh_set_tuning), ENTRY(fsgh_get_tuning), ENTRY(fsgh_forward), ENTRY(fsgh_dt_history),
    ENTRY(fsgh_backward), ENTRY(fsgh_tangent), ENTRY(fsgh_profile_enable),
    ENTRY(fsgh_profile_read), ENTRY(fsgh_check), ENTRY(fsgh_stream),
};

static fsgh_ctx *ctx = NULL;

static void check(int rc, const char *what)
{
    if (rc == FSGH_OK) return;
    fprintf(stderr, "%s failed: %d (%s)\n", what, rc, ctx ? fsgh_last_error(ctx) : "no context");
    exit(2);
}

static double rd(FILE *f)
{
    double v;
    if (fscanf(f, "%lf", &v) != 1) {
        fprintf(stderr, "mesh file: short read\n");
        exit(2);
    }
    return v;
}

int main(int argc, char **argv)
{
    FILE *f;
    int64_t n, e, k, q, count;
    int32_t s, num_bcs;
    double *coords;
    int64_t *conn, *start, *corners, *nodes;
    fsgh_boundary *bcs;
    fsgh_mesh_desc d;

    if (argc != 2 || !(f = fopen(argv[1], "r"))) {
        fprintf(stderr, "usage: %s mesh.txt\n", argv[0]);
        return 2;
    }
    printf("entry_points %d\n", (int)(sizeof entry_points / sizeof entry_points[0]));
    n = (int64_t)rd(f);
    e = (int64_t)rd(f);
    coords = malloc((size_t)(3 * n) * sizeof *coords);
    conn = malloc((size_t)(8 * e) * sizeof *conn);
    for (k = 0; k < 3 * n; ++k) coords[k] = rd(f);
    for (k = 0; k < 8 * e; ++k) conn[k] = (int64_t)rd(f);
    num_bcs = (int32_t)rd(f);
    bcs = malloc((size_t)(num_bcs > 0 ? num_bcs : 1) * sizeof *bcs);
    for (s = 0; s < num_bcs; ++s) {
        bcs[s].surface = (int32_t)rd(f);
        bcs[s].hydro_bc = (int32_t)rd(f);
        bcs[s].value = rd(f);
    }
    fclose(f);

    d.num_nodes = n;
    d.num_elems = e;
    d.num_instances = 1;
    d.renumber = 1;
    d.instance_node_off = NULL;
    d.instance_elem_off = NULL;
    d.num_dims = 3;
    d.reserved_ = 0;
    check(fsgh_create(&d, conn, coords, -1, &ctx), "fsgh_create");
    start = malloc((size_t)(n + 1) * sizeof *start);
    corners = malloc((size_t)(8 * e) * sizeof *corners);
    check(fsgh_get_corners_in_node(ctx, start, corners), "fsgh_get_corners_in_node");
    printf("cin_start");
    for (k = 0; k <= n; ++k) printf(" %lld", (long long)start[k]);
    printf("\ncin");
    for (k = 0; k < 8 * e; ++k) printf(" %lld", (long long)corners[k]);
    printf("\n");
    check(fsgh_set_bcs(ctx, 0, num_bcs, bcs), "fsgh_set_bcs");
    for (s = 0; s < num_bcs; ++s) {
        check(fsgh_bc_set_size(ctx, 0, s, &count), "fsgh_bc_set_size");
        nodes = malloc((size_t)(count > 0 ? count : 1) * sizeof *nodes);
        check(fsgh_get_bc_set(ctx, 0, s, nodes), "fsgh_get_bc_set");
        printf("set %d", (int)s);
        for (q = 0; q < count; ++q) printf(" %lld", (long long)nodes[q]);
        printf("\n");
        free(nodes);
    }
    /* an operator on a host-only context is refused, not run */
    if (fsgh_rk_init(ctx) != FSGH_E_STATE) {
        fprintf(stderr, "fsgh_rk_init on a host-only context did not return FSGH_E_STATE\n");
        return 2;
    }
    check(fsgh_destroy(ctx), "fsgh_destroy");
    free(start);
    free(corners);
    free(bcs);
    free(conn);
    free(coords);
    return 0;
}
