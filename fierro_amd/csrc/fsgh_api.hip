// fsgh_api.hip -- the C ABI of include/fsgh.h: context, field transfer,
// per-stage operators and the fused, graph-captured cycle loop.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#define FSGH_HD __host__ __device__
#include "fsgh_ctx.h"
#include "fsgh_physics.h"

using namespace fsgh;

namespace fsgh {
void set_error(fsgh_ctx* ctx, const std::string& msg)
{
    if (ctx) ctx->err = msg;
}

void mark_advanced(fsgh_ctx* c) { c->set_since.assign(static_cast<size_t>(c->L.ninst), 0); }

void mark_set(fsgh_ctx* c, int32_t inst, uint8_t parts)
{
    if (c->set_since.size() != static_cast<size_t>(c->L.ninst))
        c->set_since.assign(static_cast<size_t>(c->L.ninst), 7);   // never advanced
    for (size_t i = 0; i < c->set_since.size(); ++i)
        if (inst < 0 || static_cast<size_t>(inst) == i) c->set_since[i] |= parts;
}

bool any_advanced(const fsgh_ctx* c)
{
    for (uint8_t a : c->set_since)
        if (a != 7) return true;
    return false;
}
}  // namespace fsgh


DState fsgh::make_state(fsgh_ctx* c, int b0, int b1)
{
    DState s;
    s.x0 = c->d_node[b0][0]; s.y0 = c->d_node[b0][1]; s.z0 = c->d_node[b0][2];
    s.u0 = c->d_node[b0][3]; s.v0 = c->d_node[b0][4]; s.w0 = c->d_node[b0][5];
    s.x1 = c->d_node[b1][0]; s.y1 = c->d_node[b1][1]; s.z1 = c->d_node[b1][2];
    s.u1 = c->d_node[b1][3]; s.v1 = c->d_node[b1][4]; s.w1 = c->d_node[b1][5];
    s.nmass = c->d_nmass;
    s.bcmask = c->d_bcmask;
    s.den = c->d_den; s.pres = c->d_pres; s.sspd = c->d_sspd;
    s.sie0 = c->d_sie[b0]; s.sie1 = c->d_sie[b1];
    s.vol = c->d_vol; s.div = c->d_div[c->divcur]; s.div_out = s.div; s.emass = c->d_emass;
    s.gamma = c->d_gamma; s.csmin = c->d_csmin;
    s.mat = c->d_mat; s.F = c->d_F;
    s.mats = c->d_mats; s.nmat = c->nmat;
    s.user_eos = 0;
    for (const Material& m : c->mats) s.user_eos |= m.eos == FSGH_EOS_USER ? 1 : 0;
    s.sig0 = c->d_stress[b0];
    s.sig1 = c->d_stress[b1];
    s.sig1w = c->d_stress[b1];
    s.sparams = c->d_sparams;
    s.ext_lo = c->d_ext_lo; s.ext_hi = c->d_ext_hi;
    return s;
}


DMesh fsgh::make_mesh(fsgh_ctx* c)
{
    DMesh m;
    m.N = static_cast<int32_t>(c->N);
    m.E = static_cast<int32_t>(c->E);
    m.conn = c->d_conn;
    m.cin_start = c->d_cin_start;
    m.cin = c->d_cin;
    m.eblk = c->d_eblk;
    m.nblk = c->d_nblk;
    m.n_eblk = c->n_eblk;
    m.n_nblk = c->n_nblk;
    m.bs = kBlock;
    m.minw = 1;
    m.cshift = c->dims == 2 ? 2 : 3;
    m.pad_ = 0;
    attach_node_stage(c, m, kBlock);
    attach_corner_stage(c, m, kBlock);
    return m;
}


namespace {

DState ref_state(fsgh_ctx* c) { return make_state(c, 1 - c->cur, c->cur); }


template <typename T>
int dalloc(fsgh_ctx* c, T** p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
    if (e != hipSuccess) {
        set_error(c, std::string("hipMalloc: ") + hipGetErrorString(e));
        return FSGH_E_NOMEM;
    }
    e = hipMemsetAsync(*p, 0, n * sizeof(T), c->stream);
    if (e != hipSuccess) {
        set_error(c, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
        return FSGH_E_HIP;
    }
    return FSGH_OK;
}

void resolve_profile(fsgh_ctx* c)
{
    if (c->recs.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto& r : c->recs) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, r.a, r.b);
        auto it = c->prof_acc.find(r.name);
        if (it == c->prof_acc.end()) {
            c->prof_order.push_back(r.name);
            c->prof_acc[r.name] = {ms, 1};
        } else {
            it->second.first += ms;
            it->second.second += 1;
        }
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    c->recs.clear();
}

int ready(fsgh_ctx* c)
{
    if (c->nmat <= 0) {
        set_error(c, "materials not set (fsgh_set_materials)");
        return FSGH_E_STATE;
    }
    return FSGH_OK;
}

// host <-> device helpers in caller numbering
int put_nodes(fsgh_ctx* c, double* dst, const double* src, int stride, int comp)
{
    std::vector<double> tmp(static_cast<size_t>(c->N));
    for (int64_t n = 0; n < c->N; ++n) tmp[static_cast<size_t>(c->nn_of_on[n])] = src[n * stride + comp];
    HIPCHK(c, hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSGH_OK;
}

int get_nodes(fsgh_ctx* c, const double* src, double* dst, int stride, int comp)
{
    std::vector<double> tmp(static_cast<size_t>(c->N));
    HIPCHK(c, hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t n = 0; n < c->N; ++n) dst[n * stride + comp] = tmp[static_cast<size_t>(c->nn_of_on[n])];
    return FSGH_OK;
}

template <typename TD, typename TH>
int put_elems(fsgh_ctx* c, TD* dst, const TH* src, int stride, int comp)
{
    std::vector<TD> tmp(static_cast<size_t>(c->E));
    for (int64_t e = 0; e < c->E; ++e)
        tmp[static_cast<size_t>(c->ne_of_oe[e])] = static_cast<TD>(src[e * stride + comp]);
    HIPCHK(c, hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(TD), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSGH_OK;
}

template <typename TD, typename TH>
int get_elems(fsgh_ctx* c, const TD* src, TH* dst, int stride, int comp)
{
    std::vector<TD> tmp(static_cast<size_t>(c->E));
    HIPCHK(c, hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(TD), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t e = 0; e < c->E; ++e)
        dst[e * stride + comp] = static_cast<TH>(tmp[static_cast<size_t>(c->ne_of_oe[e])]);
    return FSGH_OK;
}

// Bring stopped instances whose newest state sits in the other physical bin
// back to the common mapping (bin 1 = phys[cur]).
int normalize_bins(fsgh_ctx* c)
{
    if (!c->run_active) return FSGH_OK;
    std::vector<InstCtl> h(static_cast<size_t>(c->L.ninst));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_ctl, h.size() * sizeof(InstCtl), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    bool changed = false;
    DState s = make_state(c, 0, 1);
    for (int32_t i = 0; i < c->L.ninst; ++i) {
        if (h[i].stopped && h[i].final_bin != c->cur) {
            launch_swap_bins_range(s, static_cast<int32_t>(c->L.node_off[i]),
                                   static_cast<int32_t>(c->L.node_off[i + 1]),
                                   static_cast<int32_t>(c->L.elem_off[i]),
                                   static_cast<int32_t>(c->L.elem_off[i + 1]),
                                   static_cast<int32_t>(c->E), c->stream);
            h[i].final_bin = c->cur;
            changed = true;
        }
    }
    if (changed) {
        HIPCHK(c, hipMemcpyAsync(c->d_ctl, h.data(), h.size() * sizeof(InstCtl),
                                 hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return FSGH_OK;
}

void drop_graphs(fsgh_ctx* c)
{
    for (auto& g : c->graph)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
}

}  // namespace

extern "C" {

int fsgh_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* fsgh_last_error(const fsgh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* fsgh_stream(fsgh_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int fsgh_create(const fsgh_mesh_desc* d, const int64_t* nodes_in_elem, const double* coords,
                int device, fsgh_ctx** out)
{
    if (!d || !nodes_in_elem || !coords || !out) return FSGH_E_INVALID;
    *out = nullptr;
    if (d->num_nodes <= 0 || d->num_elems <= 0 || d->num_instances < 1) return FSGH_E_INVALID;
    const int dims = d->num_dims == 0 ? 3 : d->num_dims;
    if (dims != 2 && dims != 3) return FSGH_E_INVALID;
    const int npe = dims == 2 ? 4 : 8;
    if (npe * d->num_elems >= (int64_t(1) << 31) || d->num_nodes >= (int64_t(1) << 31))
        return FSGH_E_UNSUPPORTED;
    auto* c = new fsgh_ctx();
    c->N = d->num_nodes;
    c->E = d->num_elems;
    c->dims = dims;
    c->npe = npe;
    c->device = device;
    c->renumber = d->renumber;
    c->L.ninst = d->num_instances;
    if (d->num_instances == 1) {
        c->L.node_off = {0, c->N};
        c->L.elem_off = {0, c->E};
    } else {
        if (!d->instance_node_off || !d->instance_elem_off) {
            delete c;
            return FSGH_E_INVALID;
        }
        c->L.node_off.assign(d->instance_node_off, d->instance_node_off + d->num_instances + 1);
        c->L.elem_off.assign(d->instance_elem_off, d->instance_elem_off + d->num_instances + 1);
    }
    for (int32_t i = 0; i < c->L.ninst; ++i)
        if (c->L.elem_off[i + 1] <= c->L.elem_off[i] || c->L.node_off[i + 1] <= c->L.node_off[i]) {
            delete c;
            return FSGH_E_INVALID;
        }
    if (c->L.elem_off[0] != 0 || c->L.node_off[0] != 0 || c->L.elem_off.back() != c->E ||
        c->L.node_off.back() != c->N) {
        delete c;
        return FSGH_E_INVALID;
    }
    c->conn.assign(nodes_in_elem, nodes_in_elem + npe * c->E);
    if (dims == 3) {
        c->coords.assign(coords, coords + 3 * c->N);
    } else {
        c->coords.assign(static_cast<size_t>(3 * c->N), 0.0);   // (axial, radius, 0)
        for (int64_t n = 0; n < c->N; ++n) {
            c->coords[static_cast<size_t>(3 * n)] = coords[2 * n];
            c->coords[static_cast<size_t>(3 * n + 1)] = coords[2 * n + 1];
        }
    }
    for (int32_t i = 0; i < c->L.ninst; ++i)
        for (int64_t e = c->L.elem_off[i]; e < c->L.elem_off[i + 1]; ++e)
            for (int a = 0; a < npe; ++a) {
                const int64_t n = c->conn[static_cast<size_t>(npe * e + a)];
                if (n < c->L.node_off[i] || n >= c->L.node_off[i + 1]) {
                    delete c;
                    return FSGH_E_INVALID;
                }
            }
    // Host-only context: numbering and the caller-numbered corner CSR on the
    // CPU.  A device context builds the numbering and the internal
    // connectivity / CSR on the GPU (fsgh_setup.hip) once its buffers exist.
    if (device < 0) {
        // numbering
        if (c->renumber) {
            morton_renumber(c->L, c->conn.data(), c->coords.data(), c->ne_of_oe, c->nn_of_on, npe);
        } else {
            c->ne_of_oe.resize(static_cast<size_t>(c->E));
            c->nn_of_on.resize(static_cast<size_t>(c->N));
            std::iota(c->ne_of_oe.begin(), c->ne_of_oe.end(), 0);
            std::iota(c->nn_of_on.begin(), c->nn_of_on.end(), 0);
        }
        c->oe_of_ne.resize(static_cast<size_t>(c->E));
        c->on_of_nn.resize(static_cast<size_t>(c->N));
        for (int64_t e = 0; e < c->E; ++e) c->oe_of_ne[static_cast<size_t>(c->ne_of_oe[e])] = static_cast<int32_t>(e);
        for (int64_t n = 0; n < c->N; ++n) c->on_of_nn[static_cast<size_t>(c->nn_of_on[n])] = static_cast<int32_t>(n);
        corner_csr(c->N, c->E, c->conn.data(), c->cstart, c->corners, npe);
    }
    std::vector<BlockItem> etab, ntab;
    block_table(c->L.elem_off, etab, c->eblk_begin, kBlock, kDefaultDeal);
    block_table(c->L.node_off, ntab, c->nblk_begin, kBlock, kDefaultDeal);
    c->n_eblk = static_cast<int32_t>(etab.size());
    c->n_nblk = static_cast<int32_t>(ntab.size());

    if (device < 0) {
        c->host_only = true;
        c->bcmask.assign(static_cast<size_t>(c->N), 0);
        c->bcsets.assign(static_cast<size_t>(c->L.ninst), {});
        *out = c;
        return FSGH_OK;
    }
    hipError_t he = hipSetDevice(device);
    if (he != hipSuccess) {
        set_error(c, std::string("hipSetDevice: ") + hipGetErrorString(he));
        *out = c;
        return FSGH_E_HIP;
    }
    he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        set_error(c, std::string("hipStreamCreate: ") + hipGetErrorString(he));
        *out = c;
        return FSGH_E_HIP;
    }
    *out = c;
    const size_t N = static_cast<size_t>(c->N), E = static_cast<size_t>(c->E);
    int rc = FSGH_OK;
    rc |= dalloc(c, &c->d_conn, 8 * E);
    rc |= dalloc(c, &c->d_cin_start, N + 1);
    rc |= dalloc(c, &c->d_cin, 8 * E);
    rc |= dalloc(c, &c->d_eblk, etab.size());
    rc |= dalloc(c, &c->d_nblk, ntab.size());
    for (int b = 0; b < 2; ++b) {
        for (int k = 0; k < 6; ++k) rc |= dalloc(c, &c->d_node[b][k], N);
        rc |= dalloc(c, &c->d_sie[b], E);
    }
    rc |= dalloc(c, &c->d_nmass, N);
    rc |= dalloc(c, &c->d_den, E);
    rc |= dalloc(c, &c->d_pres, E);
    rc |= dalloc(c, &c->d_sspd, E);
    rc |= dalloc(c, &c->d_vol, E);
    rc |= dalloc(c, &c->d_div[0], E);
    rc |= dalloc(c, &c->d_div[1], E);
    rc |= dalloc(c, &c->d_ext_lo, N);
    rc |= dalloc(c, &c->d_ext_hi, N);
    rc |= dalloc(c, &c->d_emass, E);
    rc |= dalloc(c, &c->d_gamma, E);
    rc |= dalloc(c, &c->d_csmin, E);
    rc |= dalloc(c, &c->d_F, 24 * E);
    rc |= dalloc(c, &c->d_mat, E);
    rc |= dalloc(c, &c->d_bcmask, N);
    rc |= dalloc(c, &c->d_ctl, static_cast<size_t>(c->L.ninst));
    rc |= dalloc(c, &c->d_partial, etab.size());
    c->partial_cap = std::max<size_t>(etab.size(), 1);
    rc |= dalloc(c, &c->d_scalar, 4);
    rc |= dalloc(c, &c->d_scratch, 2 * static_cast<size_t>(c->nscratch));
    rc |= dalloc(c, &c->d_flag, 1);
    if (dims == 2) {
        rc |= dalloc(c, &c->d_ext, N);
        rc |= dalloc(c, &c->d_rin, N);
    }
    if (rc != FSGH_OK) return FSGH_E_NOMEM;
    if ((rc = dims == 2 ? rz_build_maps(c) : gpu_build_maps(c)) != FSGH_OK) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_eblk, etab.data(), etab.size() * sizeof(BlockItem), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_nblk, ntab.data(), ntab.size() * sizeof(BlockItem), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = build_node_stage(c, kBlock)) != FSGH_OK) return rc;
    if ((rc = build_corner_stage(c, kBlock)) != FSGH_OK) return rc;
    // coordinates into both bins (K/read_mesh.cpp:129-137)
    for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 3; ++d)
            if ((rc = put_nodes(c, c->d_node[b][d], c->coords.data(), 3, d)) != FSGH_OK) return rc;
    c->bcmask.assign(N, 0);
    c->bcsets.assign(static_cast<size_t>(c->L.ninst), {});
    return FSGH_OK;
}

int fsgh_destroy(fsgh_ctx* c)
{
    CHECK_CTX(c);
    if (c->stream) hipStreamSynchronize(c->stream);
    resolve_profile(c);
    drop_graphs(c);
    free_adjoint(c);
    void* ptrs[] = {c->d_conn, c->d_cin_start, c->d_cin, c->d_eblk, c->d_nblk, c->d_nmass,
                    c->d_den, c->d_pres, c->d_sspd, c->d_vol, c->d_div[0], c->d_div[1], c->d_ext_lo, c->d_ext_hi, c->d_emass, c->d_gamma,
                    c->d_csmin, c->d_F, c->d_mat, c->d_bcmask, c->d_mats, c->d_ctl, c->d_partial,
                    c->d_scalar, c->d_scratch, c->d_flag, c->d_bcnodes, c->d_bcmasks,
                    c->d_sie[0], c->d_sie[1], c->d_on_of_nn, c->d_oe_of_ne, c->d_ext, c->d_rin,
                    c->d_axis};
    for (void* p : ptrs)
        if (p) hipFree(p);
    for (auto& kv : c->tables)
        if (kv.second.d) hipFree(kv.second.d);
    free_node_stages(c);
    for (double* p : {c->d_stress[0], c->d_stress[1], c->d_sparams})
        if (p) hipFree(p);
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 6; ++k)
            if (c->d_node[b][k]) hipFree(c->d_node[b][k]);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return FSGH_OK;
}

int fsgh_get_corners_in_node(fsgh_ctx* c, int64_t* start, int64_t* corners)
{
    CHECK_CTX(c);
    if (!start || !corners) return FSGH_E_INVALID;
    // a device context keeps only the internal CSR; the caller-numbered one
    // is built on first request
    if (c->cstart.empty()) corner_csr(c->N, c->E, c->conn.data(), c->cstart, c->corners, c->npe);
    std::copy(c->cstart.begin(), c->cstart.end(), start);
    std::copy(c->corners.begin(), c->corners.end(), corners);
    return FSGH_OK;
}

int fsgh_get_device_csr(fsgh_ctx* c, int32_t* start, int32_t* corners)
{
    CHECK_CTX(c);
    if (c->host_only) return FSGH_E_STATE;
    if (!start || !corners) return FSGH_E_INVALID;
    HIPCHK(c, hipMemcpyAsync(start, c->d_cin_start, static_cast<size_t>(c->N + 1) * 4,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(corners, c->d_cin, static_cast<size_t>(c->npe * c->E) * 4,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSGH_OK;
}

int fsgh_get_stage_info(fsgh_ctx* c, int32_t elem_block, int64_t* list_entries,
                        int32_t* max_block_nodes)
{
    CHECK_CTX(c);
    if (!list_entries || !max_block_nodes) return FSGH_E_INVALID;
    auto it = c->stages.find(elem_block);
    *list_entries = it == c->stages.end() ? 0 : it->second.total;
    *max_block_nodes = it == c->stages.end() ? 0 : it->second.max_nodes;
    return FSGH_OK;
}

int fsgh_get_renumbering(fsgh_ctx* c, int64_t* e_map, int64_t* n_map)
{
    CHECK_CTX(c);
    if (e_map) std::copy(c->ne_of_oe.begin(), c->ne_of_oe.end(), e_map);
    if (n_map) std::copy(c->nn_of_on.begin(), c->nn_of_on.end(), n_map);
    return FSGH_OK;
}

int fsgh_set_materials(fsgh_ctx* c, int32_t nmat, const fsgh_material* mats)
{
    CHECK_CTX(c);
    if (c->host_only) return FSGH_E_STATE;
    if (nmat <= 0 || !mats) return FSGH_E_INVALID;
    if (c->d_mats) hipFree(c->d_mats);
    c->d_mats = nullptr;
    std::vector<Material> h(static_cast<size_t>(nmat));
    for (int32_t i = 0; i < nmat; ++i)
        h[i] = {mats[i].q1, mats[i].q1ex, mats[i].q2, mats[i].q2ex, FSGH_EOS_IDEAL_GAS, 0};
    int rc = dalloc(c, &c->d_mats, h.size());
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_mats, h.data(), h.size() * sizeof(Material), hipMemcpyHostToDevice,
                             c->stream));
    c->mats = h;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nmat = nmat;
    drop_graphs(c);
    return FSGH_OK;
}

// the stress field (both bins, zero as K/setup.cpp:214-220 sets it) and the
// strength parameter table, on first need
static int ensure_stress(fsgh_ctx* c)
{
    if (c->d_stress[0]) return FSGH_OK;
    const size_t n = 9 * static_cast<size_t>(c->E);
    for (int b = 0; b < 2; ++b) {
        if (hipMalloc(reinterpret_cast<void**>(&c->d_stress[b]), n * sizeof(double)) != hipSuccess) {
            set_error(c, "hipMalloc: stress field");
            return FSGH_E_NOMEM;
        }
        HIPCHK(c, hipMemsetAsync(c->d_stress[b], 0, n * sizeof(double), c->stream));
    }
    const size_t np = static_cast<size_t>(kStrengthParams) * std::max<int32_t>(c->nmat, 1);
    if (hipMalloc(reinterpret_cast<void**>(&c->d_sparams), np * sizeof(double)) != hipSuccess) {
        set_error(c, "hipMalloc: strength parameters");
        return FSGH_E_NOMEM;
    }
    HIPCHK(c, hipMemsetAsync(c->d_sparams, 0, np * sizeof(double), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    // the stress kernels run on the default element table: a running loop's
    // dt partial minima follow it from now on
    return c->run_active ? resync_dt_ranges(c) : FSGH_OK;
}

int fsgh_set_material_models(fsgh_ctx* c, int32_t mat, int32_t eos_kind, int32_t strength)
{
    CHECK_CTX(c);
    if (c->host_only) return FSGH_E_STATE;
    if (c->nmat <= 0 || c->mats.empty()) {
        set_error(c, "materials not set (fsgh_set_materials)");
        return FSGH_E_STATE;
    }
    if (mat < 0 || mat >= c->nmat) return FSGH_E_INVALID;
    if (strength != FSGH_STRENGTH_NONE && strength != FSGH_STRENGTH_HYPO &&
        strength != FSGH_STRENGTH_HYPER)
        return FSGH_E_INVALID;
    if (c->dims == 2 && strength == FSGH_STRENGTH_HYPER) {
        // the reference's 2D update_state evaluates the 3D B matrix on a quad
        set_error(c, "hyper-elastic strength in 2D RZ is not supported");
        return FSGH_E_UNSUPPORTED;
    }
    if (eos_kind != FSGH_EOS_IDEAL_GAS && eos_kind != FSGH_EOS_USER) return FSGH_E_INVALID;
    if (strength != FSGH_STRENGTH_NONE)
        if (int rc = ensure_stress(c)) return rc;
    c->mats[static_cast<size_t>(mat)].eos = eos_kind;
    c->mats[static_cast<size_t>(mat)].strength = strength;
    drop_graphs(c);
    HIPCHK(c, hipMemcpyAsync(c->d_mats + mat, &c->mats[static_cast<size_t>(mat)], sizeof(Material),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSGH_OK;
}

int fsgh_set_strength_params(fsgh_ctx* c, int32_t mat, int32_t n, const double* params)
{
    CHECK_CTX(c);
    if (c->host_only) return FSGH_E_STATE;
    if (c->nmat <= 0) {
        set_error(c, "materials not set (fsgh_set_materials)");
        return FSGH_E_STATE;
    }
    if (mat < 0 || mat >= c->nmat || n < 0 || n > kStrengthParams || (n > 0 && !params))
        return FSGH_E_INVALID;
    if (int rc = ensure_stress(c)) return rc;
    double p[kStrengthParams] = {};
    for (int32_t i = 0; i < n; ++i) p[i] = params[i];
    HIPCHK(c, hipMemcpyAsync(c->d_sparams + static_cast<size_t>(kStrengthParams) * mat, p,
                             sizeof(p), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSGH_OK;
}

static int upload_bc_tables(fsgh_ctx* c)
{
    std::vector<uint8_t> m(static_cast<size_t>(c->N));
    std::vector<int32_t> nodes;
    std::vector<uint8_t> masks;
    for (int64_t n = 0; n < c->N; ++n) m[static_cast<size_t>(c->nn_of_on[n])] = c->bcmask[static_cast<size_t>(n)];
    for (int64_t nn = 0; nn < c->N; ++nn)
        if (m[static_cast<size_t>(nn)]) {
            nodes.push_back(static_cast<int32_t>(nn));
            masks.push_back(m[static_cast<size_t>(nn)]);
        }
    HIPCHK(c, hipMemcpyAsync(c->d_bcmask, m.data(), m.size(), hipMemcpyHostToDevice, c->stream));
    if (c->d_bcnodes) hipFree(c->d_bcnodes);
    if (c->d_bcmasks) hipFree(c->d_bcmasks);
    c->d_bcnodes = nullptr;
    c->d_bcmasks = nullptr;
    int rc = dalloc(c, &c->d_bcnodes, nodes.size());
    rc |= dalloc(c, &c->d_bcmasks, masks.size());
    if (rc) return FSGH_E_NOMEM;
    if (!nodes.empty()) {
        HIPCHK(c, hipMemcpyAsync(c->d_bcnodes, nodes.data(), nodes.size() * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_bcmasks, masks.data(), masks.size(), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n_bcnodes = static_cast<int32_t>(nodes.size());
    return FSGH_OK;
}

int fsgh_set_bcs(fsgh_ctx* c, int32_t inst, int32_t nsets, const fsgh_boundary* sets)
{
    CHECK_CTX(c);
    if (inst < 0 || inst >= c->L.ninst || nsets < 0 || (nsets > 0 && !sets)) return FSGH_E_INVALID;
    for (int32_t s = 0; s < nsets; ++s) {
        if (sets[s].surface < 0 || sets[s].surface > 4) return FSGH_E_INVALID;
        if (c->dims == 2 && (sets[s].surface == 2 || sets[s].surface == 4)) {
            set_error(c, "2D context: boundary surfaces are x-plane, y-plane or cylinder");
            return FSGH_E_UNSUPPORTED;
        }
        // the reference indexes velocity component `surface` for reflected sets,
        // which is only meaningful for planes (K/boundary.cpp:130-135)
        if (sets[s].hydro_bc == 1 && sets[s].surface > 2) {
            set_error(c, "reflected boundary on a curved surface is not supported");
            return FSGH_E_UNSUPPORTED;
        }
    }
    auto& out = c->bcsets[static_cast<size_t>(inst)];
    if (c->host_only || c->dims == 2) {
        boundary_sets(c->L, inst, c->conn.data(), c->coords.data(), c->cstart, c->corners, nsets,
                      sets, out, c->npe);
    } else if (int rc = gpu_boundary_sets(c, inst, nsets, sets, out)) {
        return rc;
    }
    for (int64_t n = c->L.node_off[inst]; n < c->L.node_off[inst + 1]; ++n) c->bcmask[static_cast<size_t>(n)] = 0;
    for (int32_t s = 0; s < nsets; ++s) {
        uint8_t bits = 0;
        if (sets[s].hydro_bc == 1) bits = static_cast<uint8_t>(1u << sets[s].surface);
        else if (sets[s].hydro_bc == 0) bits = 7;
        for (int64_t n : out[static_cast<size_t>(s)]) c->bcmask[static_cast<size_t>(n)] |= bits;
    }
    drop_graphs(c);
    return c->host_only ? FSGH_OK : upload_bc_tables(c);
}

int fsgh_get_bc_set(fsgh_ctx* c, int32_t inst, int32_t set, int64_t* nodes)
{
    CHECK_CTX(c);
    if (inst < 0 || inst >= c->L.ninst || !nodes) return FSGH_E_INVALID;
    const auto& v = c->bcsets[static_cast<size_t>(inst)];
    if (set < 0 || set >= static_cast<int32_t>(v.size())) return FSGH_E_INVALID;
    std::copy(v[static_cast<size_t>(set)].begin(), v[static_cast<size_t>(set)].end(), nodes);
    return FSGH_OK;
}

int fsgh_bc_set_size(fsgh_ctx* c, int32_t inst, int32_t set, int64_t* count)
{
    CHECK_CTX(c);
    if (inst < 0 || inst >= c->L.ninst || !count) return FSGH_E_INVALID;
    const auto& v = c->bcsets[static_cast<size_t>(inst)];
    if (set < 0 || set >= static_cast<int32_t>(v.size())) return FSGH_E_INVALID;
    *count = static_cast<int64_t>(v[static_cast<size_t>(set)].size());
    return FSGH_OK;
}

// K/setup.cpp:106-424 for one instance, evaluated on the host in caller numbering.

static int setup_fills_body(fsgh_ctx* c, int32_t inst, int32_t nfills, const fsgh_fill* fills,
                            int32_t nmat, int32_t nsv, const double* state_vars,
                            const int32_t* mat_nsv);

int fsgh_setup_fills(fsgh_ctx* c, int32_t inst, int32_t nfills, const fsgh_fill* fills,
                     int32_t nmat, int32_t nsv, const double* state_vars, const int32_t* mat_nsv)
{
    const int rc = setup_fills_body(c, inst, nfills, fills, nmat, nsv, state_vars, mat_nsv);
    if (rc == FSGH_OK) mark_set(c, inst, 2 | 4);   // velocities and energies
    return rc;
}

static int setup_fills_body(fsgh_ctx* c, int32_t inst, int32_t nfills, const fsgh_fill* fills,
                            int32_t nmat, int32_t nsv, const double* state_vars,
                            const int32_t* mat_nsv)
{
    CHECK_CTX(c);
    if (c->host_only) return FSGH_E_STATE;
    if (inst < 0 || inst >= c->L.ninst || nfills < 0 || (nfills && !fills) || nmat <= 0 ||
        nsv < 2 || !state_vars || !mat_nsv)
        return FSGH_E_INVALID;
    if (c->nmat <= 0 || nmat > c->nmat) {
        set_error(c, "fsgh_setup_fills: materials not set or fewer than num_materials");
        return FSGH_E_STATE;
    }
    for (int32_t f = 0; f < nfills; ++f) {
        if (fills[f].mat_id < 0 || fills[f].mat_id >= nmat) return FSGH_E_INVALID;
        if (fills[f].velocity < 0 || fills[f].velocity > 5 || fills[f].volume < 0 || fills[f].volume > 3)
            return FSGH_E_UNSUPPORTED;
    }
    if (c->dims == 2) {
        drop_graphs(c);
        return rz_setup_fills(c, inst, nfills, fills, nmat, nsv, state_vars, mat_nsv);
    }
    if (nfills <= 64) {
        drop_graphs(c);
        return gpu_setup_fills(c, inst, nfills, fills, nmat, nsv, state_vars, mat_nsv);
    }
    // more than 64 fills: the same computation on the host
    if (c->cstart.empty()) corner_csr(c->N, c->E, c->conn.data(), c->cstart, c->corners);
    const int64_t e0 = c->L.elem_off[inst], e1 = c->L.elem_off[inst + 1];
    const int64_t n0 = c->L.node_off[inst], n1 = c->L.node_off[inst + 1];
    const size_t NE = static_cast<size_t>(e1 - e0), NN = static_cast<size_t>(n1 - n0);
    std::vector<double> vol(NE), den(NE), mass(NE), sie(NE), pres(NE), sspd(NE), gam(NE), csm(NE);
    std::vector<double> sv4(NE, 0.0);
    std::vector<int32_t> mat(NE);
    std::vector<double> vel(3 * NN, 0.0), nmass(NN, 0.0);
    const double* X = c->coords.data();
    for (int64_t e = e0; e < e1; ++e) {
        double x[8], y[8], z[8];
        for (int a = 0; a < 8; ++a) {
            const int64_t n = c->conn[static_cast<size_t>(8 * e + a)];
            x[a] = X[3 * n];
            y[a] = X[3 * n + 1];
            z[a] = X[3 * n + 2];
        }
        vol[static_cast<size_t>(e - e0)] = hex_volume(x, y, z);
    }
    for (int32_t f = 0; f < nfills; ++f) {
        const fsgh_fill& F = fills[f];
        for (int64_t e = e0; e < e1; ++e) {
            const size_t k = static_cast<size_t>(e - e0);
            double ec[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < 8; ++a) {
                const int64_t n = c->conn[static_cast<size_t>(8 * e + a)];
                ec[0] += X[3 * n];
                ec[1] += X[3 * n + 1];
                ec[2] += X[3 * n + 2];
            }
            ec[0] = ec[0] / 8;
            ec[1] = ec[1] / 8;
            ec[2] = ec[2] / 8;
            const double radius = std::sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
            const double radius_cyl = std::sqrt(ec[0] * ec[0] + ec[1] * ec[1]);
            bool fill = false;
            switch (F.volume) {
            case 0: fill = true; break;
            case 1:
                fill = ec[0] >= F.x1 && ec[0] <= F.x2 && ec[1] >= F.y1 && ec[1] <= F.y2 &&
                       ec[2] >= F.z1 && ec[2] <= F.z2;
                break;
            case 2: fill = radius_cyl >= F.radius1 && radius_cyl <= F.radius2; break;
            case 3: fill = radius >= F.radius1 && radius <= F.radius2; break;
            }
            if (!fill) continue;
            den[k] = F.den;
            mass[k] = den[k] * vol[k];
            sie[k] = F.sie;
            mat[k] = F.mat_id;
            const int32_t m = F.mat_id;
            if (mat_nsv[m] > 0) gam[k] = state_vars[static_cast<size_t>(m) * nsv + 0];
            if (mat_nsv[m] > 1) csm[k] = state_vars[static_cast<size_t>(m) * nsv + 1];
            if (mat_nsv[m] > 4) sv4[k] = state_vars[static_cast<size_t>(m) * nsv + 4];
            // the material's EOS (K/setup.cpp:225-232)
            eos(c->mats[static_cast<size_t>(m)].eos, gam[k], csm[k], den[k], sie[k], pres[k],
                sspd[k]);
            if (F.velocity == 5) tg_vortex_state(ec, F.den, sv4[k], pres[k], sie[k]);
            for (int a = 0; a < 8; ++a) {
                const int64_t n = c->conn[static_cast<size_t>(8 * e + a)];
                double* v = &vel[3 * static_cast<size_t>(n - n0)];
                if (F.velocity == 0) {
                    v[0] = F.u;
                    v[1] = F.v;
                    v[2] = F.w;
                } else if (F.velocity == 5) {
                    tg_vortex_velocity(X + 3 * n, true, v);
                } else if (F.velocity == 1 || F.velocity == 2) {
                    const int nd = F.velocity == 1 ? 2 : 3;
                    double dir[3] = {0.0, 0.0, 0.0}, rv = 0.0;
                    for (int d = 0; d < nd; ++d) {
                        dir[d] = X[3 * n + d];
                        rv += X[3 * n + d] * X[3 * n + d];
                    }
                    rv = std::sqrt(rv);
                    for (int d = 0; d < nd; ++d) dir[d] = rv > 1.0e-14 ? dir[d] / rv : 0.0;
                    v[0] = F.speed * dir[0];
                    v[1] = F.speed * dir[1];
                    v[2] = nd == 3 ? F.speed * dir[2] : 0.0;
                }
            }
        }
    }
    // boundary velocity (K/setup.cpp:362), then node mass (K/setup.cpp:393-404)
    for (int64_t n = n0; n < n1; ++n) {
        const uint8_t mk = c->bcmask[static_cast<size_t>(n)];
        for (int d = 0; d < 3; ++d)
            if (mk & (1u << d)) vel[3 * static_cast<size_t>(n - n0) + d] = 0.0;
        double acc = 0.0;
        for (int64_t q = c->cstart[static_cast<size_t>(n)]; q < c->cstart[static_cast<size_t>(n + 1)]; ++q)
            acc += 1.0 / 8.0 * mass[static_cast<size_t>((c->corners[static_cast<size_t>(q)] >> 3) - e0)];
        nmass[static_cast<size_t>(n - n0)] = acc;
    }
    // upload the instance ranges (renumbering stays inside an instance)
    auto upe = [&](double* dst, const std::vector<double>& src) -> int {
        std::vector<double> t(NE);
        for (int64_t e = e0; e < e1; ++e) t[static_cast<size_t>(c->ne_of_oe[e] - e0)] = src[static_cast<size_t>(e - e0)];
        HIPCHK(c, hipMemcpyAsync(dst + e0, t.data(), NE * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return FSGH_OK;
    };
    auto upn = [&](double* dst, const std::vector<double>& src, int stride, int comp) -> int {
        std::vector<double> t(NN);
        for (int64_t n = n0; n < n1; ++n)
            t[static_cast<size_t>(c->nn_of_on[n] - n0)] = src[static_cast<size_t>(n - n0) * stride + comp];
        HIPCHK(c, hipMemcpyAsync(dst + n0, t.data(), NN * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return FSGH_OK;
    };
    int rc = 0;
    const std::vector<double> zeroE(NE, 0.0), zeroN(3 * NN, 0.0);
    const int b1 = c->cur, b0 = 1 - c->cur;
    rc |= upe(c->d_vol, vol);
    rc |= upe(c->d_den, den);
    rc |= upe(c->d_emass, mass);
    rc |= upe(c->d_sie[b1], sie);
    rc |= upe(c->d_sie[b0], zeroE);
    rc |= upe(c->d_pres, pres);
    rc |= upe(c->d_sspd, sspd);
    rc |= upe(c->d_gamma, gam);
    rc |= upe(c->d_csmin, csm);
    rc |= upe(c->d_div[c->divcur], zeroE);
    for (int d = 0; d < 3; ++d) {
        rc |= upn(c->d_node[b1][3 + d], vel, 3, d);
        rc |= upn(c->d_node[b0][3 + d], zeroN, 3, d);
    }
    rc |= upn(c->d_nmass, nmass, 1, 0);
    {
        std::vector<int32_t> t(NE);
        for (int64_t e = e0; e < e1; ++e) t[static_cast<size_t>(c->ne_of_oe[e] - e0)] = mat[static_cast<size_t>(e - e0)];
        HIPCHK(c, hipMemcpyAsync(c->d_mat + e0, t.data(), NE * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return rc ? FSGH_E_HIP : FSGH_OK;
}

static int field_io(fsgh_ctx* c, int32_t field, void* host, size_t bytes, bool up)
{
    CHECK_CTX(c);
    if (c->host_only) return FSGH_E_STATE;
    if (!host) return FSGH_E_INVALID;
    const size_t N = static_cast<size_t>(c->N), E = static_cast<size_t>(c->E);
    int rc = normalize_bins(c);
    if (rc) return rc;
    const int phys[2] = {1 - c->cur, c->cur};
    switch (field) {
    case FSGH_NODE_COORDS:
    case FSGH_NODE_VEL: {
        const int D = c->dims;
        if (bytes != 2 * N * D * 8) return FSGH_E_INVALID;
        const int k0 = field == FSGH_NODE_COORDS ? 0 : 3;
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < D; ++d) {
                double* h = static_cast<double*>(host) + b * N * D;
                rc = up ? put_nodes(c, c->d_node[phys[b]][k0 + d], h, D, d)
                        : get_nodes(c, c->d_node[phys[b]][k0 + d], h, D, d);
                if (rc) return rc;
            }
        if (up) drop_graphs(c);
        return FSGH_OK;
    }
    case FSGH_NODE_MASS:
        if (bytes != N * 8) return FSGH_E_INVALID;
        return up ? put_nodes(c, c->d_nmass, static_cast<double*>(host), 1, 0)
                  : get_nodes(c, c->d_nmass, static_cast<double*>(host), 1, 0);
    case FSGH_ELEM_SIE: {
        if (bytes != 2 * E * 8) return FSGH_E_INVALID;
        for (int b = 0; b < 2; ++b) {
            double* h = static_cast<double*>(host) + b * E;
            rc = up ? put_elems(c, c->d_sie[phys[b]], h, 1, 0) : get_elems(c, c->d_sie[phys[b]], h, 1, 0);
            if (rc) return rc;
        }
        return FSGH_OK;
    }
    case FSGH_ELEM_MAT_ID:
        if (bytes != E * 8) return FSGH_E_INVALID;
        return up ? put_elems(c, c->d_mat, static_cast<int64_t*>(host), 1, 0)
                  : get_elems(c, c->d_mat, static_cast<int64_t*>(host), 1, 0);
    case FSGH_ELEM_STRESS: {
        // [2 bins][E][3][3] at the ABI, [bin][9][E] inside
        if (bytes != 2 * E * 9 * 8) return FSGH_E_INVALID;
        if (up) {
            if ((rc = ensure_stress(c))) return rc;
        } else if (!c->d_stress[0]) {
            std::memset(host, 0, bytes);
            return FSGH_OK;
        }
        std::vector<double> t(9 * E);
        for (int b = 0; b < 2; ++b) {
            double* h = static_cast<double*>(host) + b * 9 * E;
            if (up) {
                for (size_t e = 0; e < E; ++e)
                    for (int k = 0; k < 9; ++k)
                        t[k * E + static_cast<size_t>(c->ne_of_oe[e])] = h[9 * e + k];
                HIPCHK(c, hipMemcpyAsync(c->d_stress[phys[b]], t.data(), t.size() * 8,
                                         hipMemcpyHostToDevice, c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
            } else {
                HIPCHK(c, hipMemcpyAsync(t.data(), c->d_stress[phys[b]], t.size() * 8,
                                         hipMemcpyDeviceToHost, c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
                for (size_t e = 0; e < E; ++e)
                    for (int k = 0; k < 9; ++k)
                        h[9 * e + k] = t[k * E + static_cast<size_t>(c->ne_of_oe[e])];
            }
        }
        if (up) drop_graphs(c);
        return FSGH_OK;
    }
    case FSGH_CORNER_FORCE: {
        // per element: npe corners x num_dims components, contiguous
        const size_t K = static_cast<size_t>(c->npe * c->dims);
        if (bytes != K * E * 8) return FSGH_E_INVALID;
        std::vector<double> t(K * E);
        double* h = static_cast<double*>(host);
        if (up) {
            for (size_t e = 0; e < E; ++e)
                std::memcpy(&t[K * static_cast<size_t>(c->ne_of_oe[e])], h + K * e, K * 8);
            HIPCHK(c, hipMemcpyAsync(c->d_F, t.data(), t.size() * 8, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            HIPCHK(c, hipMemcpyAsync(t.data(), c->d_F, t.size() * 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (size_t e = 0; e < E; ++e)
                std::memcpy(h + K * e, &t[K * static_cast<size_t>(c->ne_of_oe[e])], K * 8);
        }
        return FSGH_OK;
    }
    default: break;
    }
    double* dst = nullptr;
    switch (field) {
    case FSGH_ELEM_DEN: dst = c->d_den; break;
    case FSGH_ELEM_PRES: dst = c->d_pres; break;
    case FSGH_ELEM_SSPD: dst = c->d_sspd; break;
    case FSGH_ELEM_VOL: dst = c->d_vol; break;
    case FSGH_ELEM_DIV: dst = c->d_div[c->divcur]; break;
    case FSGH_ELEM_MASS: dst = c->d_emass; break;
    case FSGH_ELEM_GAMMA: dst = c->d_gamma; break;
    case FSGH_ELEM_CSMIN: dst = c->d_csmin; break;
    default: return FSGH_E_INVALID;
    }
    if (bytes != E * 8) return FSGH_E_INVALID;
    return up ? put_elems(c, dst, static_cast<double*>(host), 1, 0)
              : get_elems(c, dst, static_cast<double*>(host), 1, 0);
}

int fsgh_upload(fsgh_ctx* c, int32_t field, const void* host, size_t bytes)
{
    const int rc = field_io(c, field, const_cast<void*>(host), bytes, true);
    if (rc == FSGH_OK)
        mark_set(c, -1, field == FSGH_NODE_COORDS ? 1 : field == FSGH_NODE_VEL ? 2
                        : field == FSGH_ELEM_SIE ? 4 : 0);
    return rc;
}

int fsgh_download(fsgh_ctx* c, int32_t field, void* host, size_t bytes)
{
    return field_io(c, field, host, bytes, false);
}

// ------------------------------------------------------------ operators --
#define OP_PROLOGUE(c)                   \
    CHECK_CTX(c);                        \
    if (c->host_only) {                  \
        set_error(c, "host-only context"); \
        return FSGH_E_STATE;             \
    }                                    \
    {                                    \
        int rc_ = normalize_bins(c);     \
        if (rc_) return rc_;             \
    }                                    \
    HIPCHK(c, hipSetDevice(c->device));

#define OP_EPILOGUE(c)                   \
    HIPCHK(c, hipGetLastError());        \
    return FSGH_OK;

int fsgh_rk_init(fsgh_ctx* c)
{
    OP_PROLOGUE(c);
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    timed(c, "rk_init", [&] { launch_rk_init(m, s, c->stream); });
    OP_EPILOGUE(c);
}

int fsgh_get_timestep(fsgh_ctx* c, double time_value, double graphics_time, double time_final,
                      double dt_max, double dt_min, double dt_cfl, double fuzz, double* dt)
{
    OP_PROLOGUE(c);
    if (!dt) return FSGH_E_INVALID;
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    double d = *dt * 1.1;
    timed(c, "dt_partial", [&] {
        if (c->dims == 2)
            launch_rz_dt_partial(m, s, nullptr, c->d_partial, dt_cfl, dt_max, dt_min, fuzz, 0,
                                 c->stream);
        else
            launch_dt_partial(m, s, nullptr, c->d_partial, dt_cfl, dt_max, dt_min, fuzz, 0,
                              c->stream);
    });
    timed(c, "dt_reduce", [&] { launch_reduce_min(c->d_partial, c->n_eblk, c->d_scalar, c->stream); });
    double min_dt_calc = 0.0;
    HIPCHK(c, hipMemcpyAsync(&min_dt_calc, c->d_scalar, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (min_dt_calc < d) d = min_dt_calc;
    d = std::fmin(d, (graphics_time - time_value) + fuzz);
    d = std::fmin(d, time_final - time_value);
    *dt = d;
    OP_EPILOGUE(c);
}

int fsgh_get_divergence(fsgh_ctx* c)
{
    OP_PROLOGUE(c);
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    timed(c, "get_divergence", [&] {
        if (c->dims == 2) launch_rz_get_divergence(m, s, c->stream);   // K/momentum.cpp:298-378
        // nodes staged in LDS, in the force kernel's tuned partition
        else launch_get_divergence_staged(tuned_mesh(c, kTunedForce), s, c->stream);
    });
    OP_EPILOGUE(c);
}

int fsgh_get_force_sgh(fsgh_ctx* c, double fuzz, double small, double dt, double rk_alpha)
{
    OP_PROLOGUE(c);
    if (int rc = ready(c)) return rc;
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    // the 3D strength-free kernels run in the context's tuned launch
    // configuration, as in the fused loop (bit-identical in any)
    const bool tuned = c->dims == 3 && !s.sig1;
    const DMesh md = tuned ? tuned_mesh(c, kTunedDivExtremes) : m;
    const DMesh mf = tuned ? tuned_mesh(c, kTunedForce) : m;
    timed(c, "div_extremes", [&] { launch_div_extremes(md, s, nullptr, c->stream); });
    timed(c, "force", [&] {
        if (c->dims == 2) launch_rz_force(m, s, nullptr, fuzz, small, dt, rk_alpha, c->stream);   // K/force_sgh.cpp:407-808
        else if (s.sig1) launch_force_stress(m, s, nullptr, fuzz, small, dt, rk_alpha, c->stream);
        else launch_force(mf, s, nullptr, fuzz, small, c->stream);
    });
    OP_EPILOGUE(c);
}

int fsgh_update_velocity_sgh(fsgh_ctx* c, double rk_alpha, double dt)
{
    OP_PROLOGUE(c);
    mark_advanced(c);
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    timed(c, "update_velocity", [&] {
        if (c->dims == 2) launch_rz_update_velocity(m, s, rk_alpha, dt, c->stream);
        else if (s.sig1) launch_update_velocity(m, s, rk_alpha, dt, c->stream);
        // the node kernel of the fused loop with its corner forces staged in
        // LDS, in the tuned launch configuration, without boundary and position
        else launch_update_velocity_staged(tuned_mesh(c, kTunedVelBcPos), s, rk_alpha, dt,
                                           c->stream);
    });
    OP_EPILOGUE(c);
}

int fsgh_boundary_velocity(fsgh_ctx* c)
{
    OP_PROLOGUE(c);
    mark_advanced(c);
    DState s = ref_state(c);
    timed(c, "boundary_velocity",
          [&] { launch_boundary_velocity(c->d_bcnodes, c->d_bcmasks, c->n_bcnodes, s, c->stream); });
    OP_EPILOGUE(c);
}

int fsgh_update_energy_sgh(fsgh_ctx* c, double rk_alpha, double dt)
{
    OP_PROLOGUE(c);
    mark_advanced(c);
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    timed(c, "update_energy", [&] {
        if (c->dims == 2) launch_rz_update_energy(m, s, rk_alpha, dt, c->stream);
        else if (s.sig1) launch_update_energy(m, s, rk_alpha, dt, c->stream);
        // staged like the fused loop's energy kernel, in its tuned partition
        else launch_update_energy_staged(tuned_mesh(c, kTunedEnergy), s, rk_alpha, dt,
                                         c->stream);
    });
    OP_EPILOGUE(c);
}

int fsgh_update_position_sgh(fsgh_ctx* c, double rk_alpha, double dt)
{
    OP_PROLOGUE(c);
    mark_advanced(c);
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    timed(c, "update_position", [&] { launch_update_position(m, s, rk_alpha, dt, c->stream); });
    OP_EPILOGUE(c);
}

int fsgh_get_vol(fsgh_ctx* c)
{
    OP_PROLOGUE(c);
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    timed(c, "get_vol", [&] {
        if (c->dims == 2) launch_rz_get_vol(m, s, c->stream);   // K/geometry.cpp:381-431
        else launch_get_vol(m, s, c->stream);
    });
    OP_EPILOGUE(c);
}

int fsgh_update_state(fsgh_ctx* c, double dt, double rk_alpha)
{
    OP_PROLOGUE(c);
    mark_advanced(c);
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    timed(c, "update_state", [&] {
        if (s.sig1) launch_update_state_stress(m, s, dt, rk_alpha, c->stream);
        else launch_update_state(m, s, c->stream);
    });
    OP_EPILOGUE(c);
}

int fsgh_get_velgrad(fsgh_ctx* c, double* out)
{
    OP_PROLOGUE(c);
    if (!out) return FSGH_E_INVALID;
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    const size_t E = static_cast<size_t>(c->E);
    double* d = nullptr;
    HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&d), 9 * E * sizeof(double)));
    if (c->dims == 2) launch_rz_velgrad(m, s, d, c->stream);
    else launch_velgrad(m, s, d, c->stream);
    std::vector<double> t(9 * E);
    const hipError_t e1 = hipMemcpyAsync(t.data(), d, t.size() * 8, hipMemcpyDeviceToHost, c->stream);
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    hipFree(d);
    HIPCHK(c, e1);
    HIPCHK(c, e2);
    for (size_t e = 0; e < E; ++e)
        std::memcpy(out + 9 * e, &t[9 * static_cast<size_t>(c->ne_of_oe[e])], 9 * 8);
    OP_EPILOGUE(c);
}

int fsgh_energy(fsgh_ctx* c, int32_t inst, double* ie, double* ke)
{
    OP_PROLOGUE(c);
    if (inst < 0 || inst >= c->L.ninst || !ie || !ke) return FSGH_E_INVALID;
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    std::vector<int32_t> eo(c->L.elem_off.begin(), c->L.elem_off.end());
    std::vector<int32_t> no(c->L.node_off.begin(), c->L.node_off.end());
    launch_energy_sums(m, s, eo.data(), no.data(), inst, c->d_scalar, c->d_scratch, c->nscratch,
                       c->dims == 2 ? 1 : 0, c->stream);
    double out[2];
    HIPCHK(c, hipMemcpyAsync(out, c->d_scalar, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *ie = out[0];
    *ke = out[1];
    OP_EPILOGUE(c);
}

int fsgh_rz_extensive_mass(fsgh_ctx* c)
{
    OP_PROLOGUE(c);
    if (c->dims != 2) return FSGH_E_STATE;
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    timed(c, "rz_extensive_mass", [&] { launch_rz_ext(m, s, c->d_ext, c->stream); });
    OP_EPILOGUE(c);
}

int fsgh_rz_update_node_mass(fsgh_ctx* c, double tiny)
{
    OP_PROLOGUE(c);
    if (c->dims != 2) return FSGH_E_STATE;
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    timed(c, "rz_node_mass", [&] {
        launch_rz_node_mass(m, s, c->d_ext, tiny, c->d_axis, c->n_axis, c->stream);
    });
    OP_EPILOGUE(c);
}

// ------------------------------------------------------ fused cycle loop --
int fsgh_run_begin(fsgh_ctx* c, const fsgh_run_params* p)
{
    OP_PROLOGUE(c);
    if (!p) return FSGH_E_INVALID;
    if (int rc = ready(c)) return rc;
    std::vector<InstCtl> h(static_cast<size_t>(c->L.ninst));
    for (int32_t i = 0; i < c->L.ninst; ++i) {
        const fsgh_run_params& q = p[i];
        if (q.graphics_cyc_ival <= 0 || q.cycle_stop < 0) return FSGH_E_INVALID;
        InstCtl& k = h[static_cast<size_t>(i)];
        std::memset(&k, 0, sizeof(k));
        k.dt = q.dt_start;
        k.time = 0.0;
        k.graphics_time = q.graphics_dt_ival;   // K/driver.cpp:274
        k.time_final = q.time_final;
        k.dt_max = q.dt_max;
        k.dt_min = q.dt_min;
        k.dt_cfl = q.dt_cfl;
        k.graphics_dt_ival = q.graphics_dt_ival;
        k.fuzz = q.fuzz;
        k.small = q.small;
        k.tiny = q.tiny;
        k.cycle = 0;
        k.cycle_stop = q.cycle_stop;
        k.graphics_cyc_ival = q.graphics_cyc_ival;
        k.stopped = q.cycle_stop == 0 ? 1 : 0;
        k.final_bin = c->cur;
        k.eblk_begin = dt_block_begin(c)[static_cast<size_t>(i)];
        k.eblk_end = dt_block_begin(c)[static_cast<size_t>(i) + 1];
    }
    HIPCHK(c, hipMemcpyAsync(c->d_ctl, h.data(), h.size() * sizeof(InstCtl), hipMemcpyHostToDevice,
                             c->stream));
    c->divcur = 0;
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    // the first stage's divergence comes from the current state
    // (K/sgh_solve.cpp:364); later stages take it from the fused energy kernel
    if (c->dims == 2) {
        // node_extensive_mass at the start of sgh_solve (K/sgh_solve.cpp:107-116)
        timed(c, "rz_extensive_mass", [&] { launch_rz_ext(m, s, c->d_ext, c->stream); });
        timed(c, "get_divergence", [&] { launch_rz_get_divergence(m, s, c->stream); });
        timed(c, "dt_partial", [&] {
            launch_rz_dt_partial(m, s, c->d_ctl, c->d_partial, 0, 0, 0, 0, 1, c->stream);
        });
    } else {
        timed(c, "get_divergence", [&] { launch_get_divergence(m, s, c->stream); });
        timed(c, "dt_partial", [&] {
            launch_dt_partial(dt_mesh(c), s, c->d_ctl, c->d_partial, 0, 0, 0, 0, 1, c->stream);
        });
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->run_active = true;
    OP_EPILOGUE(c);
}

}  // extern "C"

// one cycle: old = c->cur holds t_n; results go to the other physical bin.
void fsgh::issue_cycle(fsgh_ctx* c, int old, const double* dt_replay, double* dt_record)
{
    const int nw = 1 - old;
    const DMesh mx = tuned_mesh(c, kTunedDivExtremes), mf = tuned_mesh(c, kTunedForce);
    const DMesh mv = tuned_mesh(c, kTunedVelBcPos), me = tuned_mesh(c, kTunedEnergy);
    DState st0 = make_state(c, old, old);   // stage 0 force reads the t_n state
    DState st = make_state(c, old, nw);
    DState sd[2] = {st, st};
    for (int k = 0; k < 2; ++k) {
        sd[k].div = c->d_div[k];
        sd[k].div_out = c->d_div[1 - k];
    }
    st0.div = c->d_div[0];
    // with a stress field the first stage reads the t_n stress and its force
    // kernel writes the stage's stress into the new bin
    const bool stress = c->d_stress[0] != nullptr;
    st0.sig1w = c->d_stress[nw];
    const DMesh m0 = make_mesh(c);
    timed(c, "cycle_dt", [&] {
        launch_cycle_dt(c->d_ctl, c->L.ninst, c->d_partial, dt_replay, dt_record, c->stream);
    });
    for (int stage = 0; stage < 2; ++stage) {
        const double alpha = 1.0 / (2.0 - static_cast<double>(stage));   // K/sgh_solve.cpp:353
        const InstCtl* ctl = c->d_ctl;
        const DState& sf = stage == 0 ? st0 : sd[1];
        double* partial = stage == 1 ? c->d_partial : nullptr;
        timed(c, "div_extremes", [&] { launch_div_extremes(mx, sf, ctl, c->stream); });
        timed(c, "force", [&] {
            if (stress) launch_force_stress(m0, sf, ctl, 0.0, 0.0, 0.0, alpha, c->stream);
            else launch_force(mf, sf, ctl, 0.0, 0.0, c->stream);
        });
        timed(c, "vel_bc_pos", [&] { launch_vel_bc_pos(mv, st, ctl, alpha, c->stream); });
        timed(c, "energy_state_div", [&] {
            if (stress) launch_energy_state_div_stress(m0, sd[stage], ctl, alpha, partial, c->stream);
            else launch_energy_state_div(me, sd[stage], ctl, alpha, partial, c->stream);
        });
    }
    timed(c, "time_advance", [&] { launch_time_advance(c->d_ctl, c->L.ninst, nw, c->stream); });
}

// one fused RZ cycle (the 2D branches of K/sgh_solve.cpp:125-634): divergence
// extremes -> force -> node (velocity, boundary, position, stage-input
// radius, RZ mass) + on-axis mass -> energy + volume + EOS + next divergence
void fsgh::issue_cycle_rz(fsgh_ctx* c, int old, const double* dt_replay, double* dt_record)
{
    const int nw = 1 - old;
    const DMesh m = make_mesh(c);
    DState st0 = make_state(c, old, old);
    DState st = make_state(c, old, nw);
    DState sd[2] = {st, st};
    for (int k = 0; k < 2; ++k) {
        sd[k].div = c->d_div[k];
        sd[k].div_out = c->d_div[1 - k];
    }
    st0.div = c->d_div[0];
    st0.sig1w = c->d_stress[nw];
    timed(c, "cycle_dt", [&] {
        launch_cycle_dt(c->d_ctl, c->L.ninst, c->d_partial, dt_replay, dt_record, c->stream);
    });
    for (int stage = 0; stage < 2; ++stage) {
        const double alpha = 1.0 / (2.0 - static_cast<double>(stage));   // K/sgh_solve.cpp:353
        const DState& sf = stage == 0 ? st0 : sd[1];
        timed(c, "rz_div_extremes", [&] { launch_div_extremes(m, sf, c->d_ctl, c->stream); });
        timed(c, "rz_force", [&] { launch_rz_force(m, sf, c->d_ctl, 0.0, 0.0, 0.0, alpha, c->stream); });
        // the stage-input radius: t_n (bin old) on stage 0, the stage-0 result on stage 1
        const double* rsrc = stage == 0 ? st.y0 : st.y1;
        timed(c, "rz_node", [&] {
            launch_rz_node(m, st, c->d_ctl, alpha, rsrc, c->d_rin, c->d_ext, c->d_axis, c->n_axis,
                           c->stream);
        });
        timed(c, "rz_energy_state_div", [&] {
            launch_rz_energy_state_div(m, sd[stage], c->d_ctl, alpha, c->d_rin,
                                       stage == 1 ? c->d_partial : nullptr, c->stream);
        });
    }
    timed(c, "time_advance", [&] { launch_time_advance(c->d_ctl, c->L.ninst, nw, c->stream); });
}

extern "C" {

int fsgh_run(fsgh_ctx* c, int64_t ncycles, int32_t flags)
{
    CHECK_CTX(c);
    if (flags & FSGH_RUN_PERSISTENT) {
        // switched off: the first cooperative launch of a one-launch cycle
        // ended the process with SIGABRT and no runtime message on the
        // MI355X, and no cause was found; the implementation was taken out
        set_error(c, "FSGH_RUN_PERSISTENT is switched off in this build (a one-launch cycle's "
                     "first cooperative launch aborted on the MI355X with no runtime message; "
                     "cause not found)");
        return FSGH_E_UNSUPPORTED;
    }
    OP_PROLOGUE(c);
    if (!c->run_active) {
        set_error(c, "fsgh_run before fsgh_run_begin");
        return FSGH_E_STATE;
    }
    if (ncycles < 0) return FSGH_E_INVALID;
    if (ncycles > 0) mark_advanced(c);
    const bool graph = (flags & FSGH_RUN_GRAPH) != 0 && !c->prof;
    const bool stop_on_output = (flags & FSGH_RUN_STOP_ON_OUTPUT) != 0;
    const int64_t poll = stop_on_output ? 1 : 32;
    for (int64_t k = 0; k < ncycles; ++k) {
        const int old = c->cur;
        if (graph) {
            if (!c->graph[old]) {
                hipGraph_t g;
                HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                if (c->dims == 2) issue_cycle_rz(c, old, nullptr, nullptr);
                else issue_cycle(c, old, nullptr, nullptr);
                HIPCHK(c, hipStreamEndCapture(c->stream, &g));
                HIPCHK(c, hipGraphInstantiate(&c->graph[old], g, nullptr, nullptr, 0));
                HIPCHK(c, hipGraphDestroy(g));
            }
            HIPCHK(c, hipGraphLaunch(c->graph[old], c->stream));
        } else if (c->dims == 2) {
            issue_cycle_rz(c, old, nullptr, nullptr);
        } else {
            issue_cycle(c, old, nullptr, nullptr);
        }
        c->cur = 1 - old;
        if ((k + 1) % poll == 0 || k + 1 == ncycles) {
            std::vector<InstCtl> h(static_cast<size_t>(c->L.ninst));
            HIPCHK(c, hipMemcpyAsync(h.data(), c->d_ctl, h.size() * sizeof(InstCtl),
                                     hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            bool all = true, out = false;
            for (auto& q : h) {
                all = all && q.stopped;
                out = out || q.write_pending;
            }
            if (all || (stop_on_output && out)) break;
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the reference's elem_div after a cycle is the last stage's input
    c->divcur = 1;
    OP_EPILOGUE(c);
}

int fsgh_run_status(fsgh_ctx* c, int32_t inst, fsgh_instance_status* out)
{
    CHECK_CTX(c);
    if (c->host_only) return FSGH_E_STATE;
    if (inst < 0 || inst >= c->L.ninst || !out) return FSGH_E_INVALID;
    InstCtl h;
    HIPCHK(c, hipMemcpyAsync(&h, c->d_ctl + inst, sizeof(InstCtl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->cycles = h.cycle;
    out->time_value = h.time;
    out->dt = h.dt;
    out->graphics_time = h.graphics_time;
    out->stopped = h.stopped;
    out->write_pending = h.write_pending;
    return FSGH_OK;
}

// ---------------------------------------------------------- diagnostics --
int fsgh_profile_enable(fsgh_ctx* c, int32_t enable)
{
    CHECK_CTX(c);
    resolve_profile(c);
    c->prof = enable != 0;
    if (enable) {
        c->prof_acc.clear();
        c->prof_order.clear();
    }
    return FSGH_OK;
}

int fsgh_profile_read(fsgh_ctx* c, int32_t max, char* names, size_t names_bytes, double* ms,
                      int64_t* calls, int32_t* count)
{
    CHECK_CTX(c);
    resolve_profile(c);
    std::string all;
    int32_t k = 0;
    for (const auto& nm : c->prof_order) {
        if (k >= max) break;
        const auto& v = c->prof_acc[nm];
        if (ms) ms[k] = v.first;
        if (calls) calls[k] = v.second;
        all += nm;
        all += '\n';
        ++k;
    }
    if (names && names_bytes) {
        const size_t n = std::min(names_bytes - 1, all.size());
        std::memcpy(names, all.data(), n);
        names[n] = 0;
    }
    if (count) *count = k;
    return FSGH_OK;
}

int fsgh_check(fsgh_ctx* c)
{
    OP_PROLOGUE(c);
    DMesh m = make_mesh(c);
    DState s = ref_state(c);
    int32_t flag = 0;
    HIPCHK(c, hipMemcpyAsync(c->d_flag, &flag, 4, hipMemcpyHostToDevice, c->stream));
    launch_check(m, s, c->d_flag, c->stream);
    HIPCHK(c, hipMemcpyAsync(&flag, c->d_flag, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (flag) {
        set_error(c, "non-finite or non-positive state");
        return FSGH_E_NONFINITE;
    }
    return FSGH_OK;
}

}  // extern "C"
