// sgh_caller.cpp -- a C++ caller of include/fsgh.h: the reference driver's
// set-up (K/driver.cpp:94-313) and the cycle loop of sgh_solve
// (K/sgh_solve.cpp:125-634) with the library's entry points in place of the
// per-stage operators (K/mesh.h:938-1406), linked against libfsgh.so.
//
//   sgh_caller <problem.txt> <out.txt> ops|run
//
// ops: the loop operator by operator, host-side dt and output cadence as in
//      sgh_solve (graphics_time starts at graphics_dt_ival, K/driver.cpp:40,
//      and moves to time + graphics_dt_ival after each output cycle);
// run: fsgh_run_begin + fsgh_run (one hipGraph per cycle, dt on the device).
//
// problem.txt (whitespace-separated, written by the test):
//   N E / coords N x 3 / nodes_in_elem E x 8
//   num_materials / q1 q1ex q2 q2ex per material
//   num_state_vars / state_vars num_materials x num_state_vars / mat_num_state_vars
//   num_bcs / surface hydro_bc value per set
//   num_fills / volume mat_id x1 x2 y1 y2 z1 z2 r1 r2 velocity u v w speed sie den per fill
//   time_final dt_start dt_max dt_min dt_cfl graphics_dt_ival graphics_cyc_ival cycle_stop
//   fuzz small rk_num_stages
// out.txt: "cycles time_value ie ke outputs", then the density of every element.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fsgh.h"

namespace {

fsgh_ctx* g_ctx = nullptr;

void check(int rc, const char* what)
{
    if (rc == FSGH_OK) return;
    std::fprintf(stderr, "%s failed: %d (%s)\n", what, rc,
                 g_ctx ? fsgh_last_error(g_ctx) : "no context");
    std::exit(2);
}

struct Reader {
    FILE* f;
    double d()
    {
        double v;
        if (std::fscanf(f, "%lf", &v) != 1) {
            std::fprintf(stderr, "problem file: short read\n");
            std::exit(2);
        }
        return v;
    }
    int64_t i() { return static_cast<int64_t>(d()); }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 4 || (std::strcmp(argv[3], "ops") && std::strcmp(argv[3], "run"))) {
        std::fprintf(stderr, "usage: %s problem.txt out.txt ops|run\n", argv[0]);
        return 2;
    }
    const bool ops = std::strcmp(argv[3], "ops") == 0;
    Reader r{std::fopen(argv[1], "r")};
    if (!r.f) {
        std::perror(argv[1]);
        return 2;
    }
    const int64_t num_nodes = r.i(), num_elems = r.i();
    std::vector<double> coords(static_cast<size_t>(3 * num_nodes));
    std::vector<int64_t> nodes_in_elem(static_cast<size_t>(8 * num_elems));
    for (double& x : coords) x = r.d();
    for (int64_t& n : nodes_in_elem) n = r.i();
    const int32_t num_materials = static_cast<int32_t>(r.i());
    std::vector<fsgh_material> mats(static_cast<size_t>(num_materials));
    for (fsgh_material& m : mats) {
        m.q1 = r.d();
        m.q1ex = r.d();
        m.q2 = r.d();
        m.q2ex = r.d();
    }
    const int32_t num_state_vars = static_cast<int32_t>(r.i());
    std::vector<double> state_vars(static_cast<size_t>(num_materials) * num_state_vars);
    for (double& v : state_vars) v = r.d();
    std::vector<int32_t> mat_num_state_vars(static_cast<size_t>(num_materials));
    for (int32_t& v : mat_num_state_vars) v = static_cast<int32_t>(r.i());
    const int32_t num_bcs = static_cast<int32_t>(r.i());
    std::vector<fsgh_boundary> bcs(static_cast<size_t>(num_bcs));
    for (fsgh_boundary& b : bcs) {
        b.surface = static_cast<int32_t>(r.i());
        b.hydro_bc = static_cast<int32_t>(r.i());
        b.value = r.d();
    }
    const int32_t num_fills = static_cast<int32_t>(r.i());
    std::vector<fsgh_fill> fills(static_cast<size_t>(num_fills));
    for (fsgh_fill& f : fills) {
        std::memset(&f, 0, sizeof f);
        f.volume = static_cast<int32_t>(r.i());
        f.mat_id = static_cast<int32_t>(r.i());
        f.x1 = r.d();
        f.x2 = r.d();
        f.y1 = r.d();
        f.y2 = r.d();
        f.z1 = r.d();
        f.z2 = r.d();
        f.radius1 = r.d();
        f.radius2 = r.d();
        f.velocity = static_cast<int32_t>(r.i());
        f.u = r.d();
        f.v = r.d();
        f.w = r.d();
        f.speed = r.d();
        f.sie = r.d();
        f.den = r.d();
    }
    const double time_final = r.d(), dt_start = r.d(), dt_max = r.d(), dt_min = r.d(),
                 dt_cfl = r.d(), graphics_dt_ival = r.d();
    const int64_t graphics_cyc_ival = r.i(), cycle_stop = r.i();
    const double fuzz = r.d(), small = r.d();
    const int64_t rk_num_stages = r.i();
    std::fclose(r.f);

    // set-up (K/driver.cpp:138-260)
    fsgh_mesh_desc d = {num_nodes, num_elems, 1, /*renumber=*/1, nullptr, nullptr,
                        /*num_dims=*/3, 0};
    check(fsgh_create(&d, nodes_in_elem.data(), coords.data(), 0, &g_ctx), "fsgh_create");
    check(fsgh_set_materials(g_ctx, num_materials, mats.data()), "fsgh_set_materials");
    check(fsgh_set_bcs(g_ctx, 0, num_bcs, bcs.data()), "fsgh_set_bcs");
    check(fsgh_setup_fills(g_ctx, 0, num_fills, fills.data(), num_materials, num_state_vars,
                           state_vars.data(), mat_num_state_vars.data()),
          "fsgh_setup_fills");

    int64_t cycles = 0, outputs = 0;
    double time_value = 0.0;
    if (ops) {
        // the cycle loop, operator by operator (K/sgh_solve.cpp:125-634)
        double dt = dt_start, graphics_time = graphics_dt_ival;
        for (int64_t cycle = 0; cycle < cycle_stop; cycle++) {
            check(fsgh_get_timestep(g_ctx, time_value, graphics_time, time_final, dt_max, dt_min,
                                    dt_cfl, fuzz, &dt),
                  "fsgh_get_timestep");
            check(fsgh_rk_init(g_ctx), "fsgh_rk_init");
            for (int64_t s = 0; s < rk_num_stages; s++) {
                const double a = 1.0 / (static_cast<double>(rk_num_stages) - static_cast<double>(s));
                check(fsgh_get_divergence(g_ctx), "fsgh_get_divergence");
                check(fsgh_get_force_sgh(g_ctx, fuzz, small, dt, a), "fsgh_get_force_sgh");
                check(fsgh_update_velocity_sgh(g_ctx, a, dt), "fsgh_update_velocity_sgh");
                check(fsgh_boundary_velocity(g_ctx), "fsgh_boundary_velocity");
                check(fsgh_update_energy_sgh(g_ctx, a, dt), "fsgh_update_energy_sgh");
                check(fsgh_update_position_sgh(g_ctx, a, dt), "fsgh_update_position_sgh");
                check(fsgh_get_vol(g_ctx), "fsgh_get_vol");
                check(fsgh_update_state(g_ctx, dt, a), "fsgh_update_state");
            }
            time_value += dt;
            cycles = cycle + 1;
            // output cadence (K/sgh_solve.cpp:594-628): the dump itself is the
            // caller's; the next output time follows it
            const bool write = ((cycle + 1) % graphics_cyc_ival == 0 && cycle > 0) ||
                               time_value >= time_final || time_value >= graphics_time;
            if (write) {
                ++outputs;
                graphics_time = time_value + graphics_dt_ival;
            }
            if (time_value >= time_final) break;
        }
    } else {
        // the whole loop, fused and graph-captured, dt on the device
        fsgh_run_params p = {time_final, dt_start, dt_max, dt_min, dt_cfl, graphics_dt_ival,
                             graphics_cyc_ival, cycle_stop, fuzz, small, 1.0e-12};
        check(fsgh_run_begin(g_ctx, &p), "fsgh_run_begin");
        for (;;) {
            check(fsgh_run(g_ctx, cycle_stop, FSGH_RUN_GRAPH | FSGH_RUN_STOP_ON_OUTPUT),
                  "fsgh_run");
            fsgh_instance_status st;
            check(fsgh_run_status(g_ctx, 0, &st), "fsgh_run_status");
            if (st.write_pending) ++outputs;
            cycles = st.cycles;
            time_value = st.time_value;
            if (st.stopped) break;
        }
    }
    double ie = 0.0, ke = 0.0;
    check(fsgh_energy(g_ctx, 0, &ie, &ke), "fsgh_energy");
    std::vector<double> den(static_cast<size_t>(num_elems));
    check(fsgh_download(g_ctx, FSGH_ELEM_DEN, den.data(), den.size() * sizeof(double)),
          "fsgh_download");
    check(fsgh_destroy(g_ctx), "fsgh_destroy");
    g_ctx = nullptr;

    FILE* o = std::fopen(argv[2], "w");
    if (!o) {
        std::perror(argv[2]);
        return 2;
    }
    std::fprintf(o, "%lld %.17g %.17g %.17g %lld\n", static_cast<long long>(cycles), time_value,
                 ie, ke, static_cast<long long>(outputs));
    for (double v : den) std::fprintf(o, "%.17g\n", v);
    std::fclose(o);
    std::printf("sgh_caller %s: %lld cycles, t = %.9e\n", argv[3], static_cast<long long>(cycles),
                time_value);
    return 0;
}
