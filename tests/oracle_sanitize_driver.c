/*
 * Runs the CPU oracle's set-up and cycle loop, 3D and 2D RZ, so that
 * tests/test_oracle_sanitize.py can build the oracle with AddressSanitizer
 * and UndefinedBehaviorSanitizer (gcc) and check that none of its paths
 * reads or writes out of bounds, leaks, or hits undefined behaviour.  Test
 * infrastructure; prints one line per case and exits non-zero on a
 * non-finite result.
 *
 * usage: oracle_sanitize_driver <path of a 3D EnSight .geo file>
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rz_oracle.h"
#include "sgh_oracle.h"

static void *zalloc(size_t n, size_t size)
{
    void *p = calloc(n ? n : 1, size);
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

static int finite_all(const double *a, int64_t n)
{
    for (int64_t i = 0; i < n; ++i)
        if (!isfinite(a[i])) return 0;
    return 1;
}

/* the Sedov-type blast on a given 3D mesh: sphere deposit, 3 reflected planes */
static int run_3d(int64_t N, int64_t E, const double *coords, const int64_t *conn, int cycles,
                  const char *tag)
{
    int64_t *cstart = zalloc((size_t)N + 1, sizeof(int64_t));
    int64_t *cin = zalloc((size_t)(8 * E), sizeof(int64_t));
    int64_t *ein = zalloc((size_t)(8 * E), sizeof(int64_t));
    int64_t *cie = zalloc((size_t)(8 * E), sizeof(int64_t));
    orc_build_corner_connectivity(N, E, conn, cstart, cin, ein, cie);
    int64_t *estart = zalloc((size_t)E + 1, sizeof(int64_t));
    int64_t nb = orc_build_elem_elem_connectivity(N, E, conn, cstart, ein, estart, NULL);
    int64_t *ee = zalloc((size_t)nb, sizeof(int64_t));
    orc_build_elem_elem_connectivity(N, E, conn, cstart, ein, estart, ee);
    int64_t counts[3];
    orc_build_patch_connectivity(N, E, conn, estart, ee, NULL, NULL, NULL, NULL, counts);
    int64_t *pie = zalloc((size_t)(6 * E), sizeof(int64_t));
    int64_t *nip = zalloc((size_t)(4 * counts[0]), sizeof(int64_t));
    int64_t *bp = zalloc((size_t)counts[1], sizeof(int64_t));
    int64_t *bn = zalloc((size_t)counts[2], sizeof(int64_t));
    orc_build_patch_connectivity(N, E, conn, estart, ee, pie, nip, bp, bn, counts);
    orc_boundary bc[3] = {{0, 0.0, 1}, {1, 0.0, 1}, {2, 0.0, 1}};
    int64_t sstart[4];
    int64_t *snodes = zalloc((size_t)(3 * counts[1] * 4), sizeof(int64_t));
    orc_build_bdy_sets(coords, counts[1], bp, nip, 3, bc, N, sstart, snodes);
    orc_mesh m = {N, E, conn, cstart, cin, ein, estart, ee, 3, sstart, snodes};

    const int64_t nsv = 6;
    orc_state s;
    memset(&s, 0, sizeof s);
    s.node_coords = zalloc((size_t)(6 * N), sizeof(double));
    memcpy(s.node_coords, coords, (size_t)(3 * N) * sizeof(double));
    memcpy(s.node_coords + 3 * N, coords, (size_t)(3 * N) * sizeof(double));
    s.node_vel = zalloc((size_t)(6 * N), sizeof(double));
    s.node_mass = zalloc((size_t)N, sizeof(double));
    s.elem_den = zalloc((size_t)E, sizeof(double));
    s.elem_pres = zalloc((size_t)E, sizeof(double));
    s.elem_sspd = zalloc((size_t)E, sizeof(double));
    s.elem_sie = zalloc((size_t)(2 * E), sizeof(double));
    s.elem_vol = zalloc((size_t)E, sizeof(double));
    s.elem_div = zalloc((size_t)E, sizeof(double));
    s.elem_mass = zalloc((size_t)E, sizeof(double));
    s.elem_mat_id = zalloc((size_t)E, sizeof(int64_t));
    s.elem_statev = zalloc((size_t)(E * nsv), sizeof(double));
    s.num_state_vars = nsv;
    s.corner_force = zalloc((size_t)(24 * E), sizeof(double));
    orc_material mat = {1.0, 1.0, 1.3333, 0.0};
    orc_fill fills[2];
    memset(fills, 0, sizeof fills);
    fills[0].den = 1.0;
    fills[0].sie = 1.0e-10;
    fills[1].volume = 3;
    fills[1].den = 1.0;
    fills[1].sie = 10.0;
    fills[1].radius2 = 0.2;
    double sv[6] = {5.0 / 3.0, 1.0e-14, 1.0, 0.0, 0.0, 0.0};
    int64_t mnsv[1] = {3};
    orc_get_vol(&m, &s);
    orc_setup_fills(&m, &s, 2, fills, sv, mnsv);
    orc_boundary_velocity(&m, &s, bc);
    orc_node_mass(&m, &s);
    orc_run_params p;
    memset(&p, 0, sizeof p);
    p.time_final = 1.0e16;
    p.dt_start = 1.0e-5;
    p.dt_max = 1.0e-2;
    p.dt_min = 1.0e-8;
    p.dt_cfl = 0.4;
    p.graphics_time = 1.0e16;
    p.graphics_dt_ival = 1.0e16;
    p.graphics_cyc_ival = 7;
    p.cycle_stop = cycles;
    p.rk_num_stages = 2;
    p.fuzz = 1.0e-16;
    p.small = 1.0e-8;
    double *rec = zalloc((size_t)cycles, sizeof(double));
    int64_t *wrec = zalloc((size_t)cycles, sizeof(int64_t));
    p.dt_record = rec;
    p.write_record = wrec;
    orc_run_result r;
    orc_sgh_solve(&m, &s, &mat, bc, &p, &r);
    int ok = finite_all(s.elem_den, E) && finite_all(s.node_coords, 6 * N) && isfinite(r.te1);
    printf("%s: N=%lld E=%lld cycles=%lld TE0=%.15e TE1=%.15e %s\n", tag, (long long)N,
           (long long)E, (long long)r.cycles, r.te0, r.te1, ok ? "ok" : "NONFINITE");
    free(rec);
    free(wrec);
    free(s.node_coords);
    free(s.node_vel);
    free(s.node_mass);
    free(s.elem_den);
    free(s.elem_pres);
    free(s.elem_sspd);
    free(s.elem_sie);
    free(s.elem_vol);
    free(s.elem_div);
    free(s.elem_mass);
    free(s.elem_mat_id);
    free(s.elem_statev);
    free(s.corner_force);
    free(snodes);
    free(bn);
    free(bp);
    free(nip);
    free(pie);
    free(ee);
    free(estart);
    free(cie);
    free(ein);
    free(cin);
    free(cstart);
    return ok;
}

/* the RZ Sedov blast on nx x ny quads, 4 reflected sides */
static int run_rz(int64_t nx, int64_t ny, int cycles)
{
    const int64_t N = (nx + 1) * (ny + 1), E = nx * ny;
    double *xy = zalloc((size_t)(2 * N), sizeof(double));
    int64_t *conn = zalloc((size_t)(4 * E), sizeof(int64_t));
    orz_gen_rect(nx, ny, 12.0, 12.0, 10.0, xy, conn);
    int64_t *cstart = zalloc((size_t)N + 1, sizeof(int64_t));
    int64_t *cin = zalloc((size_t)(4 * E), sizeof(int64_t));
    int64_t *ein = zalloc((size_t)(4 * E), sizeof(int64_t));
    orz_build_corner_connectivity(N, E, conn, cstart, cin, ein);
    int64_t *estart = zalloc((size_t)E + 1, sizeof(int64_t));
    int64_t nb = orz_build_elem_elem_connectivity(N, E, conn, cstart, ein, estart, NULL);
    int64_t *ee = zalloc((size_t)nb, sizeof(int64_t));
    orz_build_elem_elem_connectivity(N, E, conn, cstart, ein, estart, ee);
    int64_t counts[3];
    orz_build_patch_connectivity(N, E, conn, estart, ee, NULL, NULL, NULL, counts);
    int64_t *nip = zalloc((size_t)(2 * counts[0]), sizeof(int64_t));
    int64_t *bp = zalloc((size_t)counts[1], sizeof(int64_t));
    int64_t *bn = zalloc((size_t)counts[2], sizeof(int64_t));
    orz_build_patch_connectivity(N, E, conn, estart, ee, nip, bp, bn, counts);
    int64_t *nnstart = zalloc((size_t)N + 1, sizeof(int64_t));
    int64_t *nn = zalloc((size_t)(2 * counts[0]), sizeof(int64_t));
    orz_build_node_node_connectivity(N, counts[0], nip, nnstart, nn);
    orc_boundary bc[4] = {{0, 0.0, 1}, {1, 0.0, 1}, {0, 1.2, 1}, {1, 1.2, 1}};
    int64_t sstart[5];
    int64_t *snodes = zalloc((size_t)(4 * counts[1] * 2), sizeof(int64_t));
    orz_build_bdy_sets(xy, counts[1], bp, nip, 4, bc, N, sstart, snodes);
    orz_mesh m = {N, E, conn, cstart, cin, estart, ee, 4, sstart, snodes, counts[2], bn,
                  nnstart, nn};

    const int64_t nsv = 6;
    orc_state s;
    memset(&s, 0, sizeof s);
    s.node_coords = zalloc((size_t)(4 * N), sizeof(double));
    memcpy(s.node_coords, xy, (size_t)(2 * N) * sizeof(double));
    memcpy(s.node_coords + 2 * N, xy, (size_t)(2 * N) * sizeof(double));
    s.node_vel = zalloc((size_t)(4 * N), sizeof(double));
    s.node_mass = zalloc((size_t)N, sizeof(double));
    s.elem_den = zalloc((size_t)E, sizeof(double));
    s.elem_pres = zalloc((size_t)E, sizeof(double));
    s.elem_sspd = zalloc((size_t)E, sizeof(double));
    s.elem_sie = zalloc((size_t)(2 * E), sizeof(double));
    s.elem_vol = zalloc((size_t)E, sizeof(double));
    s.elem_div = zalloc((size_t)E, sizeof(double));
    s.elem_mass = zalloc((size_t)E, sizeof(double));
    s.elem_mat_id = zalloc((size_t)E, sizeof(int64_t));
    s.elem_statev = zalloc((size_t)(E * nsv), sizeof(double));
    s.num_state_vars = nsv;
    s.corner_force = zalloc((size_t)(8 * E), sizeof(double));
    orc_material mat = {1.0, 1.0, 1.3333, 0.0};
    orc_fill fills[2];
    memset(fills, 0, sizeof fills);
    fills[0].den = 1.0;
    fills[0].sie = 1.0e-10;
    fills[1].volume = 3;
    fills[1].den = 1.0;
    fills[1].sie = 50.0;
    fills[1].radius2 = 0.2;
    double sv[6] = {5.0 / 3.0, 1.0e-14, 1.0, 0.0, 0.0, 0.0};
    int64_t mnsv[1] = {3};
    orz_get_vol(&m, &s);
    orz_setup_fills(&m, &s, 2, fills, sv, mnsv);
    orz_boundary_velocity(&m, &s, bc);
    orz_node_mass(&m, &s);
    orc_run_params p;
    memset(&p, 0, sizeof p);
    p.time_final = 1.0e16;
    p.dt_start = 1.0e-5;
    p.dt_max = 1.0e-2;
    p.dt_min = 1.0e-8;
    p.dt_cfl = 0.4;
    p.graphics_time = 1.0e16;
    p.graphics_dt_ival = 1.0e16;
    p.graphics_cyc_ival = 1000;
    p.cycle_stop = cycles;
    p.rk_num_stages = 2;
    p.fuzz = 1.0e-16;
    p.small = 1.0e-8;
    orc_run_result r;
    orz_sgh_solve(&m, &s, &mat, bc, &p, 1.0e-12, &r);
    int ok = finite_all(s.elem_den, E) && finite_all(s.node_coords, 4 * N) && isfinite(r.te1);
    printf("rz %lldx%lld: cycles=%lld TE0=%.15e TE1=%.15e %s\n", (long long)nx, (long long)ny,
           (long long)r.cycles, r.te0, r.te1, ok ? "ok" : "NONFINITE");
    free(s.node_coords);
    free(s.node_vel);
    free(s.node_mass);
    free(s.elem_den);
    free(s.elem_pres);
    free(s.elem_sspd);
    free(s.elem_sie);
    free(s.elem_vol);
    free(s.elem_div);
    free(s.elem_mass);
    free(s.elem_mat_id);
    free(s.elem_statev);
    free(s.corner_force);
    free(snodes);
    free(nn);
    free(nnstart);
    free(bn);
    free(bp);
    free(nip);
    free(ee);
    free(estart);
    free(ein);
    free(cin);
    free(cstart);
    free(conn);
    free(xy);
    return ok;
}

int main(int argc, char **argv)
{
    int ok = 1;
    /* generated box */
    {
        const int64_t n = 6, N = (n + 1) * (n + 1) * (n + 1), E = n * n * n;
        double *xyz = zalloc((size_t)(3 * N), sizeof(double));
        int64_t *conn = zalloc((size_t)(8 * E), sizeof(int64_t));
        orc_gen_box(n, n, n, 12.0, 12.0, 12.0, 10.0, xyz, conn);
        ok &= run_3d(N, E, xyz, conn, 30, "box 6^3");
        free(conn);
        free(xyz);
    }
    /* the reference's mesh file, through the EnSight reader */
    if (argc > 1) {
        int64_t N = 0, E = 0;
        if (orc_read_ensight(argv[1], &N, &E, NULL, NULL) != 0) {
            fprintf(stderr, "cannot read %s\n", argv[1]);
            return 3;
        }
        double *xyz = zalloc((size_t)(3 * N), sizeof(double));
        int64_t *conn = zalloc((size_t)(8 * E), sizeof(int64_t));
        orc_read_ensight(argv[1], &N, &E, xyz, conn);
        ok &= run_3d(N, E, xyz, conn, 20, "geo");
        free(conn);
        free(xyz);
    }
    ok &= run_rz(10, 8, 25);
    return ok ? 0 : 1;
}
