// fsgh_setup.hip -- mesh preparation on the GPU: Morton renumbering, internal
// connectivity and corner CSR, boundary node sets, fills / EOS / masses.
//
// Each function produces exactly what the host path of fsgh_mesh.hip and
// fsgh_api.hip produces (same orders, same FP64 operations), with the
// serial loops replaced by element-, node- and face-parallel kernels and
// rocPRIM sorts and scans:
//   * element order: stable radix sort of the Morton codes per instance
//     (ties keep ascending original id, as sorting (code, id) pairs does);
//   * node order: first touch, i.e. the node's smallest (new element, local
//     node) key (an atomic minimum, order-independent), then a stable sort;
//   * corner CSR: stable radix sort of the corners (in ascending original
//     corner id) by internal node id, so each node's list keeps the
//     reference's ascending original-element order;
//   * boundary sets: per node surface test, per element face boundary test
//     (no other element holds the face's four nodes), then a stream
//     compaction in ascending original node id;
//   * fills: each element applies the fills in order (the last match wins);
//     each node takes its velocity from the last fill, in fill order, that
//     matches one of its elements and sets velocities (K/setup.cpp:106-424).
#include <algorithm>
#include <cmath>
#include <cstring>   // before rocPRIM, whose host code calls memset
#include <vector>

#include <rocprim/rocprim.hpp>

#define FSGH_HD __host__ __device__
#include "fsgh_ctx.h"
#include "fsgh_physics.h"

namespace fsgh {

namespace {

constexpr int kB = kBlock;

inline unsigned grid(size_t n) { return static_cast<unsigned>((n + kB - 1) / kB); }

// RAII device buffer for set-up temporaries
template <typename T>
struct DBuf {
    T* p = nullptr;
    hipError_t alloc(size_t n)
    {
        return hipMalloc(reinterpret_cast<void**>(&p), (n ? n : 1) * sizeof(T));
    }
    ~DBuf()
    {
        if (p) hipFree(p);
    }
};

#define SETUP_HIP(expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            set_error(c, std::string("setup: ") + #expr + ": " + hipGetErrorString(e_));  \
            return FSGH_E_HIP;                                                            \
        }                                                                                 \
    } while (0)

__device__ __forceinline__ int find_segment(const int64_t* off, int nseg, int64_t i)
{
    int lo = 0, hi = nseg;   // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_conn32(const int64_t* in, int32_t* out, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) out[i] = static_cast<int32_t>(in[i]);
}

// Morton code of each element's centroid (morton_renumber's quantisation),
// value = original element id
__global__ void k_morton(const int32_t* conn, const double* xyz, const int64_t* eoff, int ninst,
                         const double* box, uint64_t* code, int32_t* id, int64_t E)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (e >= E) return;
    const int inst = find_segment(eoff, ninst, e);
    const double* lo = box + 6 * inst;
    const double* hi = lo + 3;
    double cc[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 8; ++a)
        for (int d = 0; d < 3; ++d) cc[d] += xyz[3 * static_cast<int64_t>(conn[8 * e + a]) + d];
    uint64_t k = 0;
    for (int d = 0; d < 3; ++d) k |= morton_spread3(morton_quantize(cc[d] / 8.0, lo[d], hi[d])) << d;
    code[e] = k;
    id[e] = static_cast<int32_t>(e);
}

// rank scatter: map[sorted[k]] = k
// instance of every id in `ids` (off: ninst+1 ascending offsets)
__global__ void k_inst_of(const int32_t* ids, const int64_t* off, int ninst, uint32_t* out,
                          int64_t n)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (k >= n) return;
    const int64_t id = ids[k];
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= id) lo = mid;
        else hi = mid - 1;
    }
    out[k] = static_cast<uint32_t>(lo);
}

__global__ void k_rank(const int32_t* sorted, int32_t* map, int64_t n)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (k < n) map[sorted[k]] = static_cast<int32_t>(k);
}

__global__ void k_iota(int32_t* a, int64_t n)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (k < n) a[k] = static_cast<int32_t>(k);
}

__global__ void k_fill_u64(unsigned long long* a, int64_t n, unsigned long long v)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (k < n) a[k] = v;
}

// first touch of each node: smallest (new element id, local node) key
__global__ void k_first_touch(const int32_t* conn, const int32_t* ne_of_oe, unsigned long long* ft,
                              int64_t E)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (e >= E) return;
    const unsigned long long base = 8ull * static_cast<unsigned long long>(ne_of_oe[e]);
    for (int a = 0; a < 8; ++a) atomicMin(&ft[conn[8 * e + a]], base + static_cast<unsigned>(a));
}

// internal connectivity, and the CSR sort input: per original corner, the
// internal node (key) and the internal corner id (value)
__global__ void k_internal(const int32_t* conn, const int32_t* ne_of_oe, const int32_t* nn_of_on,
                           int32_t* iconn, int32_t* ckey, int32_t* cval, int64_t E)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (e >= E) return;
    const int32_t ne = ne_of_oe[e];
    for (int a = 0; a < 8; ++a) {
        const int32_t nn = nn_of_on[conn[8 * e + a]];
        iconn[8 * static_cast<int64_t>(ne) + a] = nn;
        ckey[8 * e + a] = nn;
        cval[8 * e + a] = 8 * ne + a;
    }
}

__global__ void k_count(const int32_t* key, int32_t* cnt, int64_t n)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (k < n) atomicAdd(&cnt[key[k]], 1);
}

__global__ void k_invert(const int32_t* map, int32_t* inv, int64_t n)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (k < n) inv[map[k]] = static_cast<int32_t>(k);
}

// ---------------------------------------------------------------- BC sets --
// surface test of check_bdy (K/setup.cpp:514-556), internal node numbering,
// on the initial coordinates in caller numbering
__global__ void k_bc_on(const double* xyz, const int32_t* on_of_nn, int64_t nn0, int64_t nn1,
                        int tag, double val, uint8_t* on)
{
    const int64_t nn = nn0 + static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (nn >= nn1) return;
    const double* c = xyz + 3 * static_cast<int64_t>(on_of_nn[nn]);
    bool ok = false;
    if (tag >= 0 && tag <= 2) ok = fabs(c[tag] - val) <= 1.0e-7;
    else if (tag == 3) ok = fabs(sqrt(c[0] * c[0] + c[1] * c[1]) - val) <= 1.0e-7;
    else if (tag == 4) ok = fabs(sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) - val) <= 1.0e-7;
    on[nn] = ok ? 1 : 0;
}

// faces with all four nodes on the surface and held by no other element
// mark their nodes (node lids of the faces: K/mesh.h:353-358).  The mark is
// the smallest position at which the reference's walk meets the node,
// (caller element - first caller element of the instance) * 24 + face * 4 +
// face node, so that sorting a set by it gives the reference's first-seen
// order (K/setup.cpp:591-642); unmarked nodes keep ~0.
__global__ void k_bc_faces(const int32_t* iconn, const int32_t* cin_start, const int32_t* cin,
                           const uint8_t* on, const int32_t* oe_of_ne, int64_t e0,
                           int64_t ne0, int64_t ne1, unsigned long long* mark)
{
    const int64_t e = ne0 + static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (e >= ne1) return;
    const int face[6][4] = {{0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4},
                            {2, 3, 7, 6}, {0, 3, 2, 1}, {4, 5, 6, 7}};
    const int32_t* en = iconn + 8 * e;
    for (int f = 0; f < 6; ++f) {
        int32_t fn[4];
        bool all = true;
        for (int q = 0; q < 4; ++q) {
            fn[q] = en[face[f][q]];
            all = all && on[fn[q]];
        }
        if (!all) continue;
        bool shared = false;
        for (int q = cin_start[fn[0]]; q < cin_start[fn[0] + 1] && !shared; ++q) {
            const int32_t e2 = cin[q] >> 3;
            if (e2 == e) continue;
            int hit = 0;
            for (int k = 1; k < 4; ++k)
                for (int a = 0; a < 8; ++a)
                    if (iconn[8 * static_cast<int64_t>(e2) + a] == fn[k]) {
                        ++hit;
                        break;
                    }
            shared = hit == 3;
        }
        if (shared) continue;
        const unsigned long long pos =
            static_cast<unsigned long long>(oe_of_ne[e] - e0) * 24ull + 4ull * f;
        for (int q = 0; q < 4; ++q) atomicMin(mark + fn[q], pos + q);
    }
}

// flag per original node id of one instance (scattered from the internal
// ids), and the original ids themselves
__global__ void k_bc_flag(const unsigned long long* mark, const int32_t* on_of_nn, int64_t n0,
                          int64_t n1, uint8_t* flag, int64_t* ids, unsigned long long* pos)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    const int64_t nn = n0 + k;
    if (nn >= n1) return;
    const int64_t o = on_of_nn[nn] - n0;
    flag[o] = mark[nn] != ~0ull;
    pos[o] = mark[nn];
    ids[k] = nn;
}

// ------------------------------------------------------------------ fills --
__device__ __forceinline__ bool fill_matches(const fsgh_fill& F, const double ec[3])
{
    const double radius = sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
    const double radius_cyl = sqrt(ec[0] * ec[0] + ec[1] * ec[1]);
    switch (F.volume) {
    case 0: return true;
    case 1:
        return ec[0] >= F.x1 && ec[0] <= F.x2 && ec[1] >= F.y1 && ec[1] <= F.y2 && ec[2] >= F.z1 &&
               ec[2] <= F.z2;
    case 2: return radius_cyl >= F.radius1 && radius_cyl <= F.radius2;
    case 3: return radius >= F.radius1 && radius <= F.radius2;
    }
    return false;
}

// element fields of one instance: fills in order, the last match wins
// (K/setup.cpp:117-232); vel_mask: the matching fills that set velocities
__global__ void k_fill_elem(const fsgh_fill* fills, int nfills, const int32_t* iconn, const double* xyz,
                            const int32_t* on_of_nn, const int32_t* oe_of_ne, int64_t ne0,
                            int64_t ne1, const double* state_vars, const int32_t* mat_nsv, int nsv,
                            const Material* mats, double* den, double* mass, double* sie_new,
                            double* sie_old, double* pres, double* sspd, double* vol, double* gam,
                            double* csm, int32_t* mat, double* div, unsigned long long* vel_mask)
{
    const int64_t e = ne0 + static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (e >= ne1) return;
    (void)oe_of_ne;
    double x[8], y[8], z[8];
    for (int a = 0; a < 8; ++a) {
        const double* p = xyz + 3 * static_cast<int64_t>(on_of_nn[iconn[8 * e + a]]);
        x[a] = p[0];
        y[a] = p[1];
        z[a] = p[2];
    }
    const double v = hex_volume(x, y, z);
    double ec[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 8; ++a) {
        ec[0] += x[a];
        ec[1] += y[a];
        ec[2] += z[a];
    }
    ec[0] = ec[0] / 8;
    ec[1] = ec[1] / 8;
    ec[2] = ec[2] / 8;
    double d = 0.0, ms = 0.0, ie = 0.0, g = 0.0, cm = 0.0, p = 0.0, cs = 0.0, sv4 = 0.0;
    int32_t mt = 0;
    unsigned long long vm = 0;
    for (int f = 0; f < nfills; ++f) {
        const fsgh_fill F = fills[f];
        if (!fill_matches(F, ec)) continue;
        d = F.den;
        ms = d * v;
        ie = F.sie;
        mt = F.mat_id;
        if (mat_nsv[mt] > 0) g = state_vars[static_cast<int64_t>(mt) * nsv + 0];
        if (mat_nsv[mt] > 1) cm = state_vars[static_cast<int64_t>(mt) * nsv + 1];
        if (mat_nsv[mt] > 4) sv4 = state_vars[static_cast<int64_t>(mt) * nsv + 4];
        eos(mats[mt].eos, g, cm, d, ie, p, cs);
        if (F.velocity == 5) tg_vortex_state(ec, F.den, sv4, p, ie);
        if ((F.velocity >= 0 && F.velocity <= 2) || F.velocity == 5) vm |= 1ull << f;
    }
    den[e] = d;
    mass[e] = ms;
    sie_new[e] = ie;   // the current bin; the other bin starts at zero as in the reference
    sie_old[e] = 0.0;
    div[e] = 0.0;
    pres[e] = p;
    sspd[e] = cs;
    vol[e] = v;
    gam[e] = g;
    csm[e] = cm;
    mat[e] = mt;
    vel_mask[e] = vm;
}

// node velocity from the last velocity-setting fill (in fill order) that
// matches one of the node's elements, the boundary velocity, and the node
// mass sum_{elems of n} M/8 in corner-list order (K/setup.cpp:236-404)
__global__ void k_fill_node(const fsgh_fill* fills, const int32_t* cin_start, const int32_t* cin,
                            const unsigned long long* vel_mask, const double* xyz,
                            const int32_t* on_of_nn, const uint8_t* bcmask, const double* mass,
                            int64_t nn0, int64_t nn1, double* u_new, double* v_new,
                            double* w_new, double* u_old, double* v_old, double* w_old,
                            double* nmass)
{
    const int64_t n = nn0 + static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (n >= nn1) return;
    unsigned long long any = 0;
    double acc = 0.0;
    for (int q = cin_start[n]; q < cin_start[n + 1]; ++q) {
        const int32_t e = cin[q] >> 3;
        any |= vel_mask[e];
        acc += 1.0 / 8.0 * mass[e];
    }
    double vel[3] = {0.0, 0.0, 0.0};
    if (any) {
        const fsgh_fill F = fills[63 - __builtin_clzll(any)];
        const double* X = xyz + 3 * static_cast<int64_t>(on_of_nn[n]);
        if (F.velocity == 0) {
            vel[0] = F.u;
            vel[1] = F.v;
            vel[2] = F.w;
        } else if (F.velocity == 5) {
            tg_vortex_velocity(X, true, vel);
        } else {
            const int nd = F.velocity == 1 ? 2 : 3;
            double dir[3] = {0.0, 0.0, 0.0}, rv = 0.0;
            for (int d = 0; d < nd; ++d) {
                dir[d] = X[d];
                rv += X[d] * X[d];
            }
            rv = sqrt(rv);
            for (int d = 0; d < nd; ++d) dir[d] = rv > 1.0e-14 ? dir[d] / rv : 0.0;
            vel[0] = F.speed * dir[0];
            vel[1] = F.speed * dir[1];
            vel[2] = nd == 3 ? F.speed * dir[2] : 0.0;
        }
    }
    const uint8_t mk = bcmask[n];
    for (int d = 0; d < 3; ++d)
        if (mk & (1u << d)) vel[d] = 0.0;
    u_new[n] = vel[0];
    v_new[n] = vel[1];
    w_new[n] = vel[2];
    u_old[n] = 0.0;
    v_old[n] = 0.0;
    w_old[n] = 0.0;
    nmass[n] = acc;
}

}  // namespace

// Morton renumbering (or identity), internal connectivity and corner CSR on
// the device; fills c->d_conn, d_cin_start, d_cin and the host maps.
// Stable sort of (key, id) pairs inside the instance ranges `off` (the ids
// of range i are off[i]..off[i+1]-1 and arrive in ascending order): one
// global stable radix sort by key, then, for several instances, a stable
// sort by instance.  Same result as a segmented sort, but every pass is a
// device-wide sort (a segmented sort gives a large segment to one block).
template <typename K>
hipError_t sort_in_instances(K* key, K* key2, int32_t* id, int32_t* id2, int64_t n,
                             const int64_t* off_dev, int ninst, int end_bit, hipStream_t st)
{
    size_t tmp_bytes = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, key, key2, id, id2,
                                             static_cast<size_t>(n), 0, end_bit, st);
    if (e != hipSuccess) return e;
    {
        DBuf<char> tmp;
        if ((e = tmp.alloc(tmp_bytes)) != hipSuccess) return e;
        e = rocprim::radix_sort_pairs(tmp.p, tmp_bytes, key, key2, id, id2,
                                      static_cast<size_t>(n), 0, end_bit, st);
        if (e != hipSuccess) return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    }
    if (ninst == 1) return hipSuccess;
    DBuf<uint32_t> ik, ik2;
    if ((e = ik.alloc(static_cast<size_t>(n))) != hipSuccess) return e;
    if ((e = ik2.alloc(static_cast<size_t>(n))) != hipSuccess) return e;
    hipLaunchKernelGGL(k_inst_of, dim3(grid(static_cast<size_t>(n))), dim3(kB), 0, st, id2, off_dev,
                       ninst, ik.p, n);
    int bits = 1;
    while ((1 << bits) < ninst) ++bits;
    // the result goes back to id2: sort (ik, id2) -> (ik2, id)
    tmp_bytes = 0;
    e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, ik.p, ik2.p, id2, id, static_cast<size_t>(n),
                                  0, bits, st);
    if (e != hipSuccess) return e;
    DBuf<char> tmp;
    if ((e = tmp.alloc(tmp_bytes)) != hipSuccess) return e;
    e = rocprim::radix_sort_pairs(tmp.p, tmp_bytes, ik.p, ik2.p, id2, id, static_cast<size_t>(n), 0,
                                  bits, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(id2, id, static_cast<size_t>(n) * sizeof(int32_t), hipMemcpyDeviceToDevice,
                       st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

int gpu_build_maps(fsgh_ctx* c)
{
    const int64_t N = c->N, E = c->E, C8 = 8 * E;
    const int ninst = c->L.ninst;
    hipStream_t st = c->stream;
    DBuf<int32_t> conn, ne, nn, ckey, cval, ckey2, cnt;
    DBuf<double> xyz;
    DBuf<int64_t> eoff, noff;
    SETUP_HIP(conn.alloc(static_cast<size_t>(C8)));
    SETUP_HIP(ne.alloc(static_cast<size_t>(E)));
    SETUP_HIP(nn.alloc(static_cast<size_t>(N)));
    {
        DBuf<int64_t> conn64;
        SETUP_HIP(conn64.alloc(static_cast<size_t>(C8)));
        SETUP_HIP(hipMemcpyAsync(conn64.p, c->conn.data(), static_cast<size_t>(C8) * 8,
                                 hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_conn32, dim3(grid(static_cast<size_t>(C8))), dim3(kB), 0, st, conn64.p,
                           conn.p, static_cast<size_t>(C8));
        SETUP_HIP(hipStreamSynchronize(st));
    }
    if (c->renumber) {
        SETUP_HIP(xyz.alloc(static_cast<size_t>(3 * N)));
        SETUP_HIP(eoff.alloc(static_cast<size_t>(ninst + 1)));
        SETUP_HIP(noff.alloc(static_cast<size_t>(ninst + 1)));
        SETUP_HIP(hipMemcpyAsync(xyz.p, c->coords.data(), static_cast<size_t>(3 * N) * 8,
                                 hipMemcpyHostToDevice, st));
        SETUP_HIP(hipMemcpyAsync(eoff.p, c->L.elem_off.data(), (ninst + 1) * 8,
                                 hipMemcpyHostToDevice, st));
        SETUP_HIP(hipMemcpyAsync(noff.p, c->L.node_off.data(), (ninst + 1) * 8,
                                 hipMemcpyHostToDevice, st));
        // per-instance bounding boxes (host: one pass over the coordinates)
        std::vector<double> box(static_cast<size_t>(6 * ninst));
        for (int i = 0; i < ninst; ++i) {
            double* lo = &box[static_cast<size_t>(6 * i)];
            double* hi = lo + 3;
            for (int d = 0; d < 3; ++d) {
                lo[d] = HUGE_VAL;
                hi[d] = -HUGE_VAL;
            }
            for (int64_t n = c->L.node_off[i]; n < c->L.node_off[i + 1]; ++n)
                for (int d = 0; d < 3; ++d) {
                    lo[d] = std::min(lo[d], c->coords[static_cast<size_t>(3 * n + d)]);
                    hi[d] = std::max(hi[d], c->coords[static_cast<size_t>(3 * n + d)]);
                }
        }
        DBuf<double> dbox;
        SETUP_HIP(dbox.alloc(box.size()));
        SETUP_HIP(hipMemcpyAsync(dbox.p, box.data(), box.size() * 8, hipMemcpyHostToDevice, st));
        // elements: sort by Morton code within each instance
        DBuf<uint64_t> code, code2;
        DBuf<int32_t> id, id2;
        SETUP_HIP(code.alloc(static_cast<size_t>(E)));
        SETUP_HIP(code2.alloc(static_cast<size_t>(E)));
        SETUP_HIP(id.alloc(static_cast<size_t>(E)));
        SETUP_HIP(id2.alloc(static_cast<size_t>(E)));
        hipLaunchKernelGGL(k_morton, dim3(grid(static_cast<size_t>(E))), dim3(kB), 0, st, conn.p,
                           xyz.p, eoff.p, ninst, dbox.p, code.p, id.p, E);
        SETUP_HIP(sort_in_instances(code.p, code2.p, id.p, id2.p, E, eoff.p, ninst, 63, st));
        hipLaunchKernelGGL(k_rank, dim3(grid(static_cast<size_t>(E))), dim3(kB), 0, st, id2.p, ne.p,
                           E);
        // nodes: first touch in the new element order, within each instance
        DBuf<unsigned long long> ft, ft2;
        DBuf<int32_t> nid, nid2;
        SETUP_HIP(ft.alloc(static_cast<size_t>(N)));
        SETUP_HIP(ft2.alloc(static_cast<size_t>(N)));
        SETUP_HIP(nid.alloc(static_cast<size_t>(N)));
        SETUP_HIP(nid2.alloc(static_cast<size_t>(N)));
        hipLaunchKernelGGL(k_fill_u64, dim3(grid(static_cast<size_t>(N))), dim3(kB), 0, st, ft.p, N,
                           ~0ull);
        hipLaunchKernelGGL(k_iota, dim3(grid(static_cast<size_t>(N))), dim3(kB), 0, st, nid.p, N);
        hipLaunchKernelGGL(k_first_touch, dim3(grid(static_cast<size_t>(E))), dim3(kB), 0, st, conn.p,
                           ne.p, ft.p, E);
        SETUP_HIP(sort_in_instances(ft.p, ft2.p, nid.p, nid2.p, N, noff.p, ninst, 64, st));
        hipLaunchKernelGGL(k_rank, dim3(grid(static_cast<size_t>(N))), dim3(kB), 0, st, nid2.p, nn.p,
                           N);
        SETUP_HIP(hipStreamSynchronize(st));   // before the sort buffers go out of scope
    } else {
        hipLaunchKernelGGL(k_iota, dim3(grid(static_cast<size_t>(E))), dim3(kB), 0, st, ne.p, E);
        hipLaunchKernelGGL(k_iota, dim3(grid(static_cast<size_t>(N))), dim3(kB), 0, st, nn.p, N);
    }
    // internal connectivity and the corner CSR
    SETUP_HIP(ckey.alloc(static_cast<size_t>(C8)));
    SETUP_HIP(cval.alloc(static_cast<size_t>(C8)));
    SETUP_HIP(ckey2.alloc(static_cast<size_t>(C8)));
    SETUP_HIP(cnt.alloc(static_cast<size_t>(N + 1)));
    hipLaunchKernelGGL(k_internal, dim3(grid(static_cast<size_t>(E))), dim3(kB), 0, st, conn.p, ne.p,
                       nn.p, c->d_conn, ckey.p, cval.p, E);
    int end_bit = 1;
    while ((int64_t(1) << end_bit) < N) ++end_bit;
    size_t tmp_bytes = 0;
    SETUP_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, ckey.p, ckey2.p, cval.p, c->d_cin,
                                        static_cast<size_t>(C8), 0, end_bit, st));
    {
        DBuf<char> tmp;
        SETUP_HIP(tmp.alloc(tmp_bytes));
        SETUP_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, ckey.p, ckey2.p, cval.p, c->d_cin,
                                            static_cast<size_t>(C8), 0, end_bit, st));
        SETUP_HIP(hipMemsetAsync(cnt.p, 0, static_cast<size_t>(N + 1) * 4, st));
        hipLaunchKernelGGL(k_count, dim3(grid(static_cast<size_t>(C8))), dim3(kB), 0, st, ckey.p,
                           cnt.p, C8);
        size_t sb = 0;
        SETUP_HIP(rocprim::exclusive_scan(nullptr, sb, cnt.p, c->d_cin_start, 0,
                                          static_cast<size_t>(N + 1), rocprim::plus<int32_t>(), st));
        DBuf<char> tmp2;
        SETUP_HIP(tmp2.alloc(sb));
        SETUP_HIP(rocprim::exclusive_scan(tmp2.p, sb, cnt.p, c->d_cin_start, 0,
                                          static_cast<size_t>(N + 1), rocprim::plus<int32_t>(), st));
        SETUP_HIP(hipStreamSynchronize(st));
    }
    // maps to the host (field transfers permute there) and kept on the device
    SETUP_HIP(hipMalloc(reinterpret_cast<void**>(&c->d_on_of_nn), static_cast<size_t>(N) * 4));
    SETUP_HIP(hipMalloc(reinterpret_cast<void**>(&c->d_oe_of_ne), static_cast<size_t>(E) * 4));
    hipLaunchKernelGGL(k_invert, dim3(grid(static_cast<size_t>(N))), dim3(kB), 0, st, nn.p,
                       c->d_on_of_nn, N);
    hipLaunchKernelGGL(k_invert, dim3(grid(static_cast<size_t>(E))), dim3(kB), 0, st, ne.p,
                       c->d_oe_of_ne, E);
    c->ne_of_oe.resize(static_cast<size_t>(E));
    c->nn_of_on.resize(static_cast<size_t>(N));
    c->oe_of_ne.resize(static_cast<size_t>(E));
    c->on_of_nn.resize(static_cast<size_t>(N));
    SETUP_HIP(hipMemcpyAsync(c->ne_of_oe.data(), ne.p, static_cast<size_t>(E) * 4,
                             hipMemcpyDeviceToHost, st));
    SETUP_HIP(hipMemcpyAsync(c->nn_of_on.data(), nn.p, static_cast<size_t>(N) * 4,
                             hipMemcpyDeviceToHost, st));
    SETUP_HIP(hipMemcpyAsync(c->oe_of_ne.data(), c->d_oe_of_ne, static_cast<size_t>(E) * 4,
                             hipMemcpyDeviceToHost, st));
    SETUP_HIP(hipMemcpyAsync(c->on_of_nn.data(), c->d_on_of_nn, static_cast<size_t>(N) * 4,
                             hipMemcpyDeviceToHost, st));
    SETUP_HIP(hipStreamSynchronize(st));
    return FSGH_OK;
}

// Boundary node sets of one instance (boundary_sets of fsgh_mesh.hip).
int gpu_boundary_sets(fsgh_ctx* c, int32_t inst, int32_t nsets, const fsgh_boundary* bc,
                      std::vector<std::vector<int64_t>>& out)
{
    const int64_t N = c->N;
    const int64_t n0 = c->L.node_off[inst], n1 = c->L.node_off[inst + 1];
    const int64_t e0 = c->L.elem_off[inst], e1 = c->L.elem_off[inst + 1];
    // internal ids of the instance: contiguous ranges (renumbering stays inside instances)
    const int64_t nn0 = n0, nn1 = n1, ne0 = e0, ne1 = e1;
    const int64_t NI = n1 - n0;
    hipStream_t st = c->stream;
    out.assign(static_cast<size_t>(nsets), {});
    if (nsets == 0) return FSGH_OK;
    DBuf<double> xyz;
    DBuf<uint8_t> on, flag;
    DBuf<unsigned long long> mark, pos, selpos;
    DBuf<int64_t> ids, sel;
    DBuf<unsigned> nsel;
    SETUP_HIP(xyz.alloc(static_cast<size_t>(3 * N)));
    SETUP_HIP(on.alloc(static_cast<size_t>(N)));
    SETUP_HIP(mark.alloc(static_cast<size_t>(N)));
    SETUP_HIP(flag.alloc(static_cast<size_t>(NI)));
    SETUP_HIP(ids.alloc(static_cast<size_t>(NI)));
    SETUP_HIP(sel.alloc(static_cast<size_t>(NI)));
    SETUP_HIP(pos.alloc(static_cast<size_t>(NI)));
    SETUP_HIP(selpos.alloc(static_cast<size_t>(NI)));
    SETUP_HIP(nsel.alloc(1));
    SETUP_HIP(hipMemcpyAsync(xyz.p, c->coords.data(), static_cast<size_t>(3 * N) * 8,
                             hipMemcpyHostToDevice, st));
    size_t sb = 0, sb2 = 0;
    SETUP_HIP(rocprim::select(nullptr, sb, ids.p, flag.p, sel.p, nsel.p, static_cast<size_t>(NI),
                              st));
    SETUP_HIP(rocprim::select(nullptr, sb2, pos.p, flag.p, selpos.p, nsel.p,
                              static_cast<size_t>(NI), st));
    sb = std::max(sb, sb2);
    DBuf<char> tmp;
    SETUP_HIP(tmp.alloc(sb));
    for (int32_t s = 0; s < nsets; ++s) {
        hipLaunchKernelGGL(k_bc_on, dim3(grid(static_cast<size_t>(NI))), dim3(kB), 0, st, xyz.p,
                           c->d_on_of_nn, nn0, nn1, bc[s].surface, bc[s].value, on.p);
        SETUP_HIP(hipMemsetAsync(mark.p, 0xff, static_cast<size_t>(N) * 8, st));
        hipLaunchKernelGGL(k_bc_faces, dim3(grid(static_cast<size_t>(ne1 - ne0))), dim3(kB), 0, st,
                           c->d_conn, c->d_cin_start, c->d_cin, on.p, c->d_oe_of_ne, e0, ne0, ne1,
                           mark.p);
        hipLaunchKernelGGL(k_bc_flag, dim3(grid(static_cast<size_t>(NI))), dim3(kB), 0, st, mark.p,
                           c->d_on_of_nn, n0, n1, flag.p, ids.p, pos.p);
        SETUP_HIP(rocprim::select(tmp.p, sb, ids.p, flag.p, sel.p, nsel.p, static_cast<size_t>(NI),
                                  st));
        SETUP_HIP(rocprim::select(tmp.p, sb, pos.p, flag.p, selpos.p, nsel.p,
                                  static_cast<size_t>(NI), st));
        unsigned count = 0;
        SETUP_HIP(hipMemcpyAsync(&count, nsel.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        SETUP_HIP(hipStreamSynchronize(st));
        std::vector<int64_t> node(count);
        std::vector<unsigned long long> at(count);
        SETUP_HIP(hipMemcpyAsync(node.data(), sel.p, count * 8, hipMemcpyDeviceToHost, st));
        SETUP_HIP(hipMemcpyAsync(at.data(), selpos.p, count * 8, hipMemcpyDeviceToHost, st));
        SETUP_HIP(hipStreamSynchronize(st));
        // first-seen order: by the position of each node's first appearance
        std::vector<uint32_t> order(count);
        for (uint32_t k = 0; k < count; ++k) order[k] = k;
        std::sort(order.begin(), order.end(),
                  [&](uint32_t a, uint32_t b) { return at[a] < at[b]; });
        auto& v = out[static_cast<size_t>(s)];
        v.resize(count);
        for (uint32_t k = 0; k < count; ++k) v[k] = node[order[k]];
    }
    return FSGH_OK;
}

// Fills, EOS, velocities, boundary velocity and node masses of one instance
// (fsgh_setup_fills' host loop), written to the device state.
int gpu_setup_fills(fsgh_ctx* c, int32_t inst, int32_t nfills, const fsgh_fill* fills,
                    int32_t nmat, int32_t nsv, const double* state_vars, const int32_t* mat_nsv)
{
    if (nfills > 64) return FSGH_E_UNSUPPORTED;
    const int64_t N = c->N;
    const int64_t ne0 = c->L.elem_off[inst], ne1 = c->L.elem_off[inst + 1];
    const int64_t nn0 = c->L.node_off[inst], nn1 = c->L.node_off[inst + 1];
    hipStream_t st = c->stream;
    DBuf<fsgh_fill> df;
    SETUP_HIP(df.alloc(static_cast<size_t>(nfills)));
    if (nfills > 0)
        SETUP_HIP(hipMemcpyAsync(df.p, fills, static_cast<size_t>(nfills) * sizeof(fsgh_fill),
                                 hipMemcpyHostToDevice, st));
    DBuf<double> xyz, sv;
    DBuf<int32_t> msv;
    DBuf<unsigned long long> vmask;
    SETUP_HIP(xyz.alloc(static_cast<size_t>(3 * N)));
    SETUP_HIP(sv.alloc(static_cast<size_t>(nmat) * nsv));
    SETUP_HIP(msv.alloc(static_cast<size_t>(nmat)));
    SETUP_HIP(vmask.alloc(static_cast<size_t>(c->E)));
    SETUP_HIP(hipMemcpyAsync(xyz.p, c->coords.data(), static_cast<size_t>(3 * N) * 8,
                             hipMemcpyHostToDevice, st));
    SETUP_HIP(hipMemcpyAsync(sv.p, state_vars, static_cast<size_t>(nmat) * nsv * 8,
                             hipMemcpyHostToDevice, st));
    SETUP_HIP(hipMemcpyAsync(msv.p, mat_nsv, static_cast<size_t>(nmat) * 4, hipMemcpyHostToDevice,
                             st));
    const int b1 = c->cur, b0 = 1 - c->cur;
    hipLaunchKernelGGL(k_fill_elem, dim3(grid(static_cast<size_t>(ne1 - ne0))), dim3(kB), 0, st,
                       static_cast<const fsgh_fill*>(df.p), nfills,
                       c->d_conn, xyz.p, c->d_on_of_nn, c->d_oe_of_ne, ne0, ne1, sv.p, msv.p, nsv,
                       c->d_mats, c->d_den, c->d_emass, c->d_sie[b1], c->d_sie[b0], c->d_pres,
                       c->d_sspd, c->d_vol, c->d_gamma, c->d_csmin, c->d_mat, c->d_div[c->divcur],
                       vmask.p);
    hipLaunchKernelGGL(k_fill_node, dim3(grid(static_cast<size_t>(nn1 - nn0))), dim3(kB), 0, st,
                       static_cast<const fsgh_fill*>(df.p),
                       c->d_cin_start, c->d_cin, vmask.p, xyz.p, c->d_on_of_nn, c->d_bcmask,
                       c->d_emass, nn0, nn1, c->d_node[b1][3], c->d_node[b1][4], c->d_node[b1][5],
                       c->d_node[b0][3], c->d_node[b0][4], c->d_node[b0][5], c->d_nmass);
    SETUP_HIP(hipGetLastError());
    SETUP_HIP(hipStreamSynchronize(st));
    return FSGH_OK;
}

}  // namespace fsgh
