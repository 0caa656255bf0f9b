// fsgh_tune.hip -- launch configurations of the fused-cycle kernels and the
// autotuner that picks them (fsgh_autotune / fsgh_set_tuning / fsgh_get_tuning).
//
// A configuration is (workgroup size, occupancy floor, XCD dealing).  The
// workgroup size fixes the work tables (one contiguous range of that many
// elements or nodes per workgroup, never straddling an instance); the
// occupancy floor selects a kernel instantiation compiled with
// amdgpu_waves_per_eu; XCD dealing orders the table so that each XCD sweeps
// one contiguous range.  Per-element and per-node arithmetic is the same in
// every variant, so results are bit-identical; only the dt partial minima
// depend on the element table, through their per-instance block ranges.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#define FSGH_HD __host__ __device__
#include "fsgh_ctx.h"

using namespace fsgh;

namespace {

// xcd_order: 0 launch order, 1 contiguous per XCD, 2..kMaxDealChunk chunks
constexpr int32_t kMaxDealChunk = 4096;

int32_t table_key(bool elem, int32_t bs, int32_t xcd)
{
    return (elem ? (1 << 30) : 0) | (bs << 16) | xcd;
}

bool elem_kernel(int k) { return k == kTunedForce || k == kTunedEnergy; }

bool default_table(const KernelCfg& g) { return g.bs == kBlock && g.xcd == kDefaultDeal; }

void drop_cycle_graphs(fsgh_ctx* c)
{
    for (auto& g : c->graph)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
}

DMesh mesh_for(fsgh_ctx* c, int k, const KernelCfg& g)
{
    DMesh m = make_mesh(c);
    m.bs = g.bs;
    m.minw = g.minw;
    // the element table below is the kernel's own partition for an element
    // kernel and the default one otherwise
    attach_node_stage(c, m, elem_kernel(k) ? g.bs : kBlock);
    attach_corner_stage(c, m, elem_kernel(k) ? kBlock : g.bs);
    if (!default_table(g)) {
        const fsgh_ctx::WorkTable& w = c->tables.at(table_key(elem_kernel(k), g.bs, g.xcd));
        if (elem_kernel(k)) {
            m.eblk = w.d;
            m.n_eblk = w.n;
        } else {
            m.nblk = w.d;
            m.n_nblk = w.n;
        }
    }
    return m;
}

// The dt partial minima of a running loop were written per logical block of
// the old energy table: point every instance's block range at the new table
// and recompute them from the current state (the same values the last energy
// kernel wrote: element dt of the current coordinates and sound speed).
int resync_dt(fsgh_ctx* c)
{
    return resync_dt_ranges(c);
}


// the work table of a configuration and the per-block lists its kernel
// reads (node lists: element kernels, with the table; sorted corner lists:
// the velocity kernel only)
int ensure_kernel_data(fsgh_ctx* c, int k, const KernelCfg& g)
{
    if (int rc = ensure_table(c, elem_kernel(k), g.bs, g.xcd)) return rc;
    return k == kTunedVelBcPos ? build_corner_stage(c, g.bs) : FSGH_OK;
}

int apply(fsgh_ctx* c, int k, const KernelCfg& g)
{
    if (int rc = ensure_kernel_data(c, k, g)) return rc;
    const int32_t old_bs = c->tune[kTunedEnergy].bs;
    c->tune[k] = g;
    drop_cycle_graphs(c);
    if (k == kTunedEnergy && g.bs != old_bs) return resync_dt(c);
    return FSGH_OK;
}

bool valid(int k, const KernelCfg& g)
{
    return k >= 0 && k < kNumTuned && g.xcd >= 0 && g.xcd <= kMaxDealChunk &&
           tuned_variant_ok(k, g.bs, g.minw);
}

std::vector<KernelCfg> candidates(int k)
{
    std::vector<KernelCfg> out;
    for (int32_t bs : {256, 128, 64, 512})
        for (int32_t minw : {1, 3})
            for (int32_t xcd : {1, 0, 4, 8, 16, 32}) {
                KernelCfg g;
                g.bs = bs;
                g.minw = minw;
                g.xcd = xcd;
                if (valid(k, g)) out.push_back(g);
            }
    return out;
}

// One device buffer holding a copy of every array a trial launch may write.
struct Snapshot {
    std::vector<std::pair<void*, size_t>> parts;
    char* buf = nullptr;
    size_t bytes = 0;
};

void snapshot_parts(fsgh_ctx* c, Snapshot& S)
{
    const size_t N = static_cast<size_t>(c->N), E = static_cast<size_t>(c->E);
    for (int b = 0; b < 2; ++b) {
        for (int k = 0; k < 6; ++k) S.parts.push_back({c->d_node[b][k], N * 8});
        S.parts.push_back({c->d_sie[b], E * 8});
        S.parts.push_back({c->d_div[b], E * 8});
    }
    S.parts.push_back({c->d_den, E * 8});
    S.parts.push_back({c->d_pres, E * 8});
    S.parts.push_back({c->d_sspd, E * 8});
    S.parts.push_back({c->d_vol, E * 8});
    S.parts.push_back({c->d_ext_lo, N * 16});
    S.parts.push_back({c->d_ext_hi, N * 16});
    S.parts.push_back({c->d_F, E * 24 * 8});
    S.parts.push_back({c->d_ctl, static_cast<size_t>(c->L.ninst) * sizeof(InstCtl)});
    S.parts.push_back({c->d_partial, c->partial_cap * 8});
    for (auto& p : S.parts) S.bytes += (p.second + 255) / 256 * 256;
}

int copy_parts(fsgh_ctx* c, Snapshot& S, bool save)
{
    size_t off = 0;
    for (auto& p : S.parts) {
        if (save)
            HIPCHK(c, hipMemcpyAsync(S.buf + off, p.first, p.second, hipMemcpyDeviceToDevice,
                                     c->stream));
        else
            HIPCHK(c, hipMemcpyAsync(p.first, S.buf + off, p.second, hipMemcpyDeviceToDevice,
                                     c->stream));
        off += (p.second + 255) / 256 * 256;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSGH_OK;
}

void to_public(int k, const KernelCfg& g, double ms, double ms_def, fsgh_tuning* o)
{
    o->kernel = k;
    o->block = g.bs;
    o->min_waves = g.minw;
    o->xcd_order = g.xcd;
    o->ms = ms;
    o->ms_default = ms_def;
}

}  // namespace

int fsgh::ensure_table(fsgh_ctx* c, bool elem, int32_t bs, int32_t xcd, bool lists)
{
    if (bs == kBlock && xcd == kDefaultDeal) return FSGH_OK;
    const int32_t key = table_key(elem, bs, xcd);
    if (c->tables.count(key)) return (elem && lists) ? build_node_stage(c, bs) : FSGH_OK;
    std::vector<BlockItem> tab;
    fsgh_ctx::WorkTable w;
    block_table(elem ? c->L.elem_off : c->L.node_off, tab, w.inst_begin, bs, xcd);
    w.n = static_cast<int32_t>(tab.size());
    if (elem && tab.size() > c->partial_cap) {
        // the dt partial minima are indexed by logical element block
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double* p = nullptr;
        if (hipMalloc(reinterpret_cast<void**>(&p), tab.size() * sizeof(double)) != hipSuccess) {
            set_error(c, "hipMalloc: dt partial minima");
            return FSGH_E_NOMEM;
        }
        HIPCHK(c, hipMemcpyAsync(p, c->d_partial, c->partial_cap * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        hipFree(c->d_partial);
        c->d_partial = p;
        c->partial_cap = tab.size();
        drop_cycle_graphs(c);
    }
    if (hipMalloc(reinterpret_cast<void**>(&w.d), std::max<size_t>(tab.size(), 1) * sizeof(BlockItem)) !=
        hipSuccess) {
        set_error(c, "hipMalloc: work table");
        return FSGH_E_NOMEM;
    }
    HIPCHK(c, hipMemcpyAsync(w.d, tab.data(), tab.size() * sizeof(BlockItem),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->tables[key] = std::move(w);
    return (elem && lists) ? build_node_stage(c, bs) : FSGH_OK;
}

DMesh fsgh::tuned_mesh(fsgh_ctx* c, int kernel) { return mesh_for(c, kernel, c->tune[kernel]); }

int fsgh::small_block_tables(fsgh_ctx* c)
{
    if (int rc = ensure_table(c, true, 64, 1, false)) return rc;
    return ensure_table(c, false, 64, 1, false);
}

DMesh fsgh::small_block_mesh(fsgh_ctx* c)
{
    DMesh m = make_mesh(c);
    const fsgh_ctx::WorkTable& we = c->tables.at(table_key(true, 64, 1));
    const fsgh_ctx::WorkTable& wn = c->tables.at(table_key(false, 64, 1));
    m.bs = 64;
    m.eblk = we.d;
    m.n_eblk = we.n;
    m.nblk = wn.d;
    m.n_nblk = wn.n;
    attach_node_stage(c, m, 64);
    attach_corner_stage(c, m, 64);
    return m;
}

int fsgh::adjoint_node_tables(fsgh_ctx* c)
{
    if (int rc = ensure_table(c, false, kAdjNodeBlock, kAdjNodeDeal, false)) return rc;
    return build_corner_stage(c, kAdjNodeBlock);
}

DMesh fsgh::adjoint_node_mesh(fsgh_ctx* c)
{
    DMesh m = make_mesh(c);
    const fsgh_ctx::WorkTable& wn = c->tables.at(table_key(false, kAdjNodeBlock, kAdjNodeDeal));
    m.bs = kAdjNodeBlock;
    m.nblk = wn.d;
    m.n_nblk = wn.n;
    attach_corner_stage(c, m, kAdjNodeBlock);
    return m;
}

DMesh fsgh::dt_mesh(fsgh_ctx* c)
{
    if (c->dims == 2 || c->d_stress[0]) return make_mesh(c);
    return tuned_mesh(c, kTunedEnergy);
}

int fsgh::resync_dt_ranges(fsgh_ctx* c)
{
    if (!c->run_active) return FSGH_OK;
    std::vector<InstCtl> h(static_cast<size_t>(c->L.ninst));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_ctl, h.size() * sizeof(InstCtl), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const std::vector<int32_t>& beg = dt_block_begin(c);
    for (size_t i = 0; i < h.size(); ++i) {
        h[i].eblk_begin = beg[i];
        h[i].eblk_end = beg[i + 1];
    }
    HIPCHK(c, hipMemcpyAsync(c->d_ctl, h.data(), h.size() * sizeof(InstCtl), hipMemcpyHostToDevice,
                             c->stream));
    const DState s = make_state(c, 1 - c->cur, c->cur);
    if (c->dims == 2)
        launch_rz_dt_partial(dt_mesh(c), s, c->d_ctl, c->d_partial, 0, 0, 0, 0, 1, c->stream);
    else
        launch_dt_partial(dt_mesh(c), s, c->d_ctl, c->d_partial, 0, 0, 0, 0, 1, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSGH_OK;
}


const std::vector<int32_t>& fsgh::dt_block_begin(fsgh_ctx* c)
{
    const KernelCfg& g = c->tune[kTunedEnergy];
    if (default_table(g) || c->dims == 2 || c->d_stress[0]) return c->eblk_begin;
    return c->tables.at(table_key(true, g.bs, g.xcd)).inst_begin;
}

extern "C" {

int fsgh_set_tuning(fsgh_ctx* c, const fsgh_tuning* cfg, int32_t count)
{
    CHECK_CTX(c);
    if (c->host_only) return FSGH_E_STATE;
    if (c->dims == 2 || c->d_stress[0]) {
        set_error(c, "launch tuning covers the 3D strength-free kernels only");
        return FSGH_E_UNSUPPORTED;
    }
    if (count < 0 || (count > 0 && !cfg)) return FSGH_E_INVALID;
    for (int32_t i = 0; i < count; ++i) {
        KernelCfg g;
        g.bs = cfg[i].block;
        g.minw = cfg[i].min_waves;
        g.xcd = cfg[i].xcd_order;
        if (!valid(cfg[i].kernel, g)) {
            set_error(c, "fsgh_set_tuning: no compiled variant for kernel " +
                             std::to_string(cfg[i].kernel) + " block " + std::to_string(g.bs) +
                             " min_waves " + std::to_string(g.minw) + " xcd_order " +
                             std::to_string(g.xcd));
            return FSGH_E_INVALID;
        }
    }
    HIPCHK(c, hipSetDevice(c->device));
    for (int32_t i = 0; i < count; ++i) {
        KernelCfg g;
        g.bs = cfg[i].block;
        g.minw = cfg[i].min_waves;
        g.xcd = cfg[i].xcd_order;
        if (int rc = apply(c, cfg[i].kernel, g)) return rc;
    }
    return FSGH_OK;
}

int fsgh_get_tuning(fsgh_ctx* c, fsgh_tuning* out, int32_t count)
{
    CHECK_CTX(c);
    if (count < 0 || (count > 0 && !out)) return FSGH_E_INVALID;
    for (int32_t k = 0; k < count && k < kNumTuned; ++k) to_public(k, c->tune[k], 0.0, 0.0, out + k);
    return FSGH_OK;
}

int fsgh_autotune(fsgh_ctx* c, int32_t reps, fsgh_tuning* out)
{
    CHECK_CTX(c);
    if (c->host_only) return FSGH_E_STATE;
    if (c->dims == 2 || c->d_stress[0]) {
        set_error(c, "launch tuning covers the 3D strength-free kernels only");
        return FSGH_E_UNSUPPORTED;
    }
    if (c->nmat <= 0) {
        set_error(c, "materials not set (fsgh_set_materials)");
        return FSGH_E_STATE;
    }
    if (reps <= 0) reps = 10;
    reps = std::min(reps, 1000);
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<KernelCfg> cand[kNumTuned];
    for (int k = 0; k < kNumTuned; ++k) {
        cand[k] = candidates(k);
        for (const KernelCfg& g : cand[k])
            if (int rc = ensure_kernel_data(c, k, g)) return rc;
    }
    Snapshot S;
    snapshot_parts(c, S);
    if (hipMalloc(reinterpret_cast<void**>(&S.buf), S.bytes) != hipSuccess) {
        set_error(c, "fsgh_autotune: no device memory for the state copy (" +
                         std::to_string(S.bytes) + " bytes)");
        return FSGH_E_NOMEM;
    }
    struct Guard {
        char* p;
        ~Guard() { hipFree(p); }
    } guard{S.buf};
    if (int rc = copy_parts(c, S, true)) return rc;
    // every exit from here on, an error return included, puts the fields and
    // the control blocks back (the trial launches run on the live state);
    // declared after guard, so it runs before the copy is freed
    struct Restore {
        fsgh_ctx* c;
        Snapshot* S;
        bool done;
        ~Restore() {
            if (done) return;
            size_t off = 0;
            for (auto& p : S->parts) {
                hipMemcpyAsync(p.first, S->buf + off, p.second, hipMemcpyDeviceToDevice,
                               c->stream);
                off += (p.second + 255) / 256 * 256;
            }
            hipStreamSynchronize(c->stream);
        }
    } restore{c, &S, false};

    // trial control block: every instance live, a small positive dt
    std::vector<InstCtl> h(static_cast<size_t>(c->L.ninst));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_ctl, h.size() * sizeof(InstCtl), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (InstCtl& q : h) {
        q.stopped = 0;
        if (!(q.dt > 0.0)) q.dt = 1.0e-9;
        if (!(q.small > 0.0)) q.small = 1.0e-8;
        if (!(q.fuzz > 0.0)) q.fuzz = 1.0e-16;
        if (!(q.dt_cfl > 0.0)) q.dt_cfl = 0.5;
        if (!(q.dt_max > 0.0)) q.dt_max = 1.0e-2;
        if (!(q.dt_min > 0.0)) q.dt_min = 1.0e-8;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_ctl, h.data(), h.size() * sizeof(InstCtl), hipMemcpyHostToDevice,
                             c->stream));

    // the first stage of a cycle from the current state, as the fused loop
    // issues it; its outputs are the inputs of the trial launches
    const int old = c->cur, nw = 1 - old;
    DState st0 = make_state(c, old, old);
    st0.div = c->d_div[0];
    DState st = make_state(c, old, nw);
    DState sd0 = st;
    sd0.div = c->d_div[0];
    sd0.div_out = c->d_div[1];
    const InstCtl* ctl = c->d_ctl;
    const DMesh m0 = make_mesh(c);
    launch_get_divergence(m0, st0, c->stream);
    launch_div_extremes(m0, st0, ctl, c->stream);
    launch_force(m0, st0, ctl, 0.0, 0.0, c->stream);
    launch_vel_bc_pos(m0, st, ctl, 0.5, c->stream);
    launch_energy_state_div(m0, sd0, ctl, 0.5, c->d_partial, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));

    auto run = [&](int k, const DMesh& m) {
        switch (k) {
        case kTunedDivExtremes: launch_div_extremes(m, st0, ctl, c->stream); break;
        case kTunedForce: launch_force(m, st0, ctl, 0.0, 0.0, c->stream); break;
        case kTunedVelBcPos: launch_vel_bc_pos(m, st, ctl, 0.5, c->stream); break;
        default: launch_energy_state_div(m, sd0, ctl, 0.5, c->d_partial, c->stream); break;
        }
    };
    hipEvent_t ea, eb;
    HIPCHK(c, hipEventCreate(&ea));
    HIPCHK(c, hipEventCreate(&eb));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            hipEventDestroy(a);
            hipEventDestroy(b);
        }
    } evguard{ea, eb};
    KernelCfg best[kNumTuned];
    double best_ms[kNumTuned], def_ms[kNumTuned];
    for (int k = 0; k < kNumTuned; ++k) {
        std::vector<double> t(cand[k].size(), 1.0e300);
        for (int pass = 0; pass < 2; ++pass)
            for (size_t v = 0; v < cand[k].size(); ++v) {
                const DMesh m = mesh_for(c, k, cand[k][v]);
                run(k, m);   // warm-up
                HIPCHK(c, hipEventRecord(ea, c->stream));
                for (int r = 0; r < reps; ++r) run(k, m);
                HIPCHK(c, hipEventRecord(eb, c->stream));
                HIPCHK(c, hipEventSynchronize(eb));
                HIPCHK(c, hipGetLastError());
                float ms = 0.f;
                HIPCHK(c, hipEventElapsedTime(&ms, ea, eb));
                t[v] = std::min(t[v], static_cast<double>(ms) / reps);
            }
        size_t bi = 0;
        for (size_t v = 1; v < t.size(); ++v)
            if (t[v] < t[bi]) bi = v;
        best[k] = cand[k][bi];
        best_ms[k] = t[bi];
        def_ms[k] = t[0];   // candidates start with the default (256, 1, 1)
    }
    restore.done = true;
    if (int rc = copy_parts(c, S, false)) return rc;
    for (int k = 0; k < kNumTuned; ++k)
        if (int rc = apply(c, k, best[k])) return rc;
    if (out)
        for (int k = 0; k < kNumTuned; ++k) to_public(k, best[k], best_ms[k], def_ms[k], out + k);
    return FSGH_OK;
}

}  // extern "C"
