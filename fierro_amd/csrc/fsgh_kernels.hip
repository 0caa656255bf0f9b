// fsgh_kernels.hip -- CDNA4 (gfx950) FP64 kernels of the SGH cycle.
//
// Every kernel walks a table of BlockItems: one workgroup = one contiguous
// element (or node) range inside one instance, dealt so that consecutive
// ranges of the Morton-ordered mesh land on the same XCD (shared L2).
//
// Floating point: the library is compiled with -ffp-contract=off and every
// expression keeps the reference's operation order (file:line cited at each
// kernel), so results match the reference arithmetic bit for bit except for
// pow() (ocml vs. the host libm) and reductions whose order the reference
// leaves to Kokkos.
#include <cfloat>
#include <cmath>

#define FSGH_HD __host__ __device__
#include "fsgh_internal.h"
#include "fsgh_lds.h"
#include "fsgh_physics.h"

namespace fsgh {
namespace {

__device__ __forceinline__ void load_conn(const int32_t* conn, int e, int32_t nd[8])
{
    const int4* p = reinterpret_cast<const int4*>(conn + 8 * static_cast<size_t>(e));
    const int4 a = p[0];
    const int4 b = p[1];
    nd[0] = a.x; nd[1] = a.y; nd[2] = a.z; nd[3] = a.w;
    nd[4] = b.x; nd[5] = b.y; nd[6] = b.z; nd[7] = b.w;
}

__device__ __forceinline__ void gather3(const double* __restrict__ a, const double* __restrict__ b,
                                        const double* __restrict__ c, const int32_t nd[8],
                                        double x[8], double y[8], double z[8])
{
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        x[i] = a[nd[i]];
        y[i] = b[nd[i]];
        z[i] = c[nd[i]];
    }
}

// LDS staging of a block's nodes (lists built in fsgh_stage.hip): capacity
// in nodes per element block of BS elements (a Morton brick of BS elements
// has (a+1)(b+1)(c+1) nodes, 405 for 256; irregular blocks above the
// capacity read from global memory instead)
template <int BS>
constexpr int stage_cap()
{
    return BS <= 64 ? 192 : BS <= 128 ? 320 : BS <= 256 ? 640 : 1024;
}

__device__ __forceinline__ void load_lconn(const uint16_t* lconn, int e, int32_t lc[8])
{
    const uint4 p = *reinterpret_cast<const uint4*>(lconn + 8 * static_cast<size_t>(e));
    lc[0] = static_cast<int32_t>(p.x & 0xffffu);
    lc[1] = static_cast<int32_t>(p.x >> 16);
    lc[2] = static_cast<int32_t>(p.y & 0xffffu);
    lc[3] = static_cast<int32_t>(p.y >> 16);
    lc[4] = static_cast<int32_t>(p.z & 0xffffu);
    lc[5] = static_cast<int32_t>(p.z >> 16);
    lc[6] = static_cast<int32_t>(p.w & 0xffffu);
    lc[7] = static_cast<int32_t>(p.w >> 16);
}

// The 24 corner forces of this lane's element, read cooperatively by the
// wave: the wave's elements own one contiguous stretch of F, which the lanes
// load as consecutive 16-byte pieces (coalesced; wave_issue_forces, whose
// loads can be in flight while other work proceeds) and pass through a
// padded LDS region (13 x 16 B per element, two halves of 32 elements) to the
// lane of each element (wave_scatter_forces).  buf: 32 x 13 double2 of this
// wave; e0: the wave's first element; nvalid: elements of the wave inside the
// work item.
__device__ __forceinline__ void wave_issue_forces(const double* __restrict__ F, int e0, int nvalid,
                                                  double2 t[12])
{
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const double2* src = reinterpret_cast<const double2*>(F + 24 * static_cast<size_t>(e0));
#pragma unroll
    for (int r = 0; r < 12; ++r) {
        // pieces past the wave's valid rows re-read piece 0 (nvalid >= 1):
        // unconditional loads, whose values go unused
        const int idx = 64 * r + lane;                // piece within the wave's stretch
        t[r] = stream_load(src + (idx < 12 * nvalid ? idx : 0));
    }
}

__device__ __forceinline__ void wave_scatter_forces(const double2 t[12], double2* buf, double f[24])
{
    const int lane = static_cast<int>(threadIdx.x) & 63;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const int idx = 64 * r + lane;            // piece within the half
            buf[(idx / 12) * 13 + idx % 12] = t[6 * h + r];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if ((lane >> 5) == h) {
            const double2* mine = buf + (lane & 31) * 13;
#pragma unroll
            for (int i = 0; i < 12; ++i) {
                const double2 v = mine[i];
                f[2 * i] = v.x;
                f[2 * i + 1] = v.y;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__device__ __forceinline__ void lds_gather3(const double* __restrict__ a, const double* __restrict__ b,
                                            const double* __restrict__ c, const int32_t lc[8],
                                            double x[8], double y[8], double z[8])
{
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        x[i] = a[lc[i]];
        y[i] = b[lc[i]];
        z[i] = c[lc[i]];
    }
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

// block-wide min over BS threads; result valid in thread 0
template <int BS = kBlock>
__device__ __forceinline__ double block_min(double v)
{
    __shared__ double red[BS / 64];
    v = wave_min(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int i = 1; i < BS / 64; ++i) r = fmin(r, red[i]);
    return r;
}

__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[kBlock / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int i = 1; i < kBlock / 64; ++i) r = r + red[i];
    __syncthreads();
    return r;
}

// shortest of the 28 node-pair distances (K/time_integration.cpp:101-128).
// sqrt is correctly rounded and monotone, so min(sqrt(d2)) == sqrt(min(d2)).
__device__ __forceinline__ double min_pair_distance(const double x[8], const double y[8],
                                                    const double z[8])
{
    double m = DBL_MAX;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = a + 1; b < 8; ++b) {
            const double dx = x[b] - x[a], dy = y[b] - y[a], dz = z[b] - z[a];
            m = fmin(m, dx * dx + dy * dy + dz * dz);
        }
    return sqrt(m);
}

// local CFL dt of one element, clamped (K/time_integration.cpp:131-135)
__device__ __forceinline__ double elem_dt(double dist_min, double sspd, double dt_cfl,
                                          double dt_max, double dt_min, double fuzz)
{
    double d = dt_cfl * dist_min / (sspd + fuzz);
    d = fmin(d, dt_max);
    d = fmax(d, dt_min);
    return d;
}

// --------------------------------------------------------------------------
// corner force, 3D, strength-free (K/force_sgh.cpp:35-359)
// --------------------------------------------------------------------------
// BS threads per workgroup (one element each); MINW = waves-per-SIMD floor
// requested from the register allocator (1 = the compiler's choice)
template <int BS, int MINW>
__global__ __launch_bounds__(BS, MINW) void k_force(DMesh m, DState s, const InstCtl* ctl,
                                                    double fuzz, double small)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    if (ctl != nullptr) {
        if (ctl[bi.inst].stopped) return;
        fuzz = ctl[bi.inst].fuzz;
        small = ctl[bi.inst].small;
    }
    const int e = bi.start + static_cast<int>(threadIdx.x);
    // One LDS buffer, used twice: first the block's distinct nodes (x, y, z,
    // u, v, w as [6][CAP] doubles, then the divergence extremes as [2][CAP]
    // double2; lists from fsgh_stage.hip), then the block's corner forces
    // (row stride 13 x 16 B, free of bank conflicts), written out as
    // contiguous 16-byte stores.  The element's own inputs and connectivity
    // are requested before the staging barrier.  Blocks over the capacity
    // gather from global memory (block-uniform branch, same results).
    constexpr int CAP = stage_cap<BS>();
    constexpr int kWords = BS * 13 > CAP * 5 ? BS * 13 : CAP * 5;
    __shared__ double2 smem[kWords];
    double* sd = reinterpret_cast<double*>(smem);
    double2* sx = smem + 3 * CAP;
    const bool live = e < bi.end;
    // lanes past the work item read the item's first element (in bounds,
    // results unused), so the loads need no divergent branches
    const int ec = live ? e : bi.start;
    // Staging: every list index is requested first (the load counter
    // retires in issue order, so the node loads then wait for nothing
    // else), then the element's inputs and connectivity, then every node
    // value, and only then are the values stored to LDS (R = CAP / BS
    // rounded up, list entries per lane).
    int q0, nn;
    stage_range(m, bi, CAP, q0, nn);
    const bool staged = nn > 0 && nn <= CAP;   // lists are never empty; the test keeps nn - 1 in range
    constexpr int R = (CAP + BS - 1) / BS;
    uint32_t nid[R];   // unsigned: its address arithmetic needs no sign extension
    double g[R][10];   // x, y, z, u, v, w, ext_lo pair, ext_hi pair
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // clamped, unconditional loads: nothing computed from them before
            // the node loads, so those wait for these alone
            const int i = r * BS + static_cast<int>(threadIdx.x);
            nid[r] = static_cast<uint32_t>(m.sn_nodes[q0 + (i < nn ? i : nn - 1)]);
        }
    }
    // the divergence and sound speed select the impedance coefficients
    // now; every element value the arithmetic needs is read again after the
    // barrier (cache hits or a short wait behind the B matrix), so none of
    // them occupies a register across the staging
    const double div0 = s.div[ec];
    const double sspd0 = s.sspd[ec];
    // connectivity, kept packed until after the barrier (unpacking it here
    // would make the staging loads wait for it)
    uint4 craw0, craw1;
    if (staged) {
        craw0 = *reinterpret_cast<const uint4*>(m.sn_lconn + 8 * static_cast<size_t>(ec));
    } else {
        const uint4* p = reinterpret_cast<const uint4*>(m.conn + 8 * static_cast<size_t>(ec));
        craw0 = p[0];
        craw1 = p[1];
    }
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (wave_in_list(r * BS, nn)) {
                const uint32_t n = nid[r];
                g[r][0] = s.x1[n];
                g[r][1] = s.y1[n];
                g[r][2] = s.z1[n];
                g[r][3] = s.u1[n];
                g[r][4] = s.v1[n];
                g[r][5] = s.w1[n];
                const double2 lo = s.ext_lo[n], hi = s.ext_hi[n];
                g[r][6] = lo.x;
                g[r][7] = lo.y;
                g[r][8] = hi.x;
                g[r][9] = hi.y;
            }
        }
    }
    // the element-only terms, formed while the node values are in flight
    // (this also shortens what stays live across the staging barrier); one
    // material is read uniformly, without a per-element index load
    Material mt;
    if (s.nmat > 1) mt = s.mats[s.mat[ec]];
    else mt = s.mats[0];
    ForceCoef<double> kc =
        force_coef(0.0, 0.0, sspd0, 1.0, div0, mt.q1, mt.q1ex, mt.q2, mt.q2ex, fuzz);
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * BS + static_cast<int>(threadIdx.x);
            if (i < nn) {
#pragma unroll
                for (int k = 0; k < 6; ++k) sd[k * CAP + i] = g[r][k];
                sx[i] = make_double2(g[r][6], g[r][7]);
                sx[CAP + i] = make_double2(g[r][8], g[r][9]);
            }
        }
        __syncthreads();
    }
    int32_t nd[8];
    if (staged) {
        const uint4 p = craw0;
        nd[0] = static_cast<int32_t>(p.x & 0xffffu);
        nd[1] = static_cast<int32_t>(p.x >> 16);
        nd[2] = static_cast<int32_t>(p.y & 0xffffu);
        nd[3] = static_cast<int32_t>(p.y >> 16);
        nd[4] = static_cast<int32_t>(p.z & 0xffffu);
        nd[5] = static_cast<int32_t>(p.z >> 16);
        nd[6] = static_cast<int32_t>(p.w & 0xffffu);
        nd[7] = static_cast<int32_t>(p.w >> 16);
    } else {
        nd[0] = static_cast<int32_t>(craw0.x);
        nd[1] = static_cast<int32_t>(craw0.y);
        nd[2] = static_cast<int32_t>(craw0.z);
        nd[3] = static_cast<int32_t>(craw0.w);
        nd[4] = static_cast<int32_t>(craw1.x);
        nd[5] = static_cast<int32_t>(craw1.y);
        nd[6] = static_cast<int32_t>(craw1.z);
        nd[7] = static_cast<int32_t>(craw1.w);
    }
    // the element values, requested after the barrier: their latency hides
    // behind the B matrix
    kc.pres = s.pres[ec];
    kc.den = s.den[ec];
    kc.div = s.div[ec];
    kc.sspd = s.sspd[ec];
    kc.vol = s.vol[ec];
    const double div_e = kc.div;
    double f[24];
    if (live) {
        double x[8], y[8], z[8], u[8], v[8], w[8];
        if (staged) {
            lds_gather3(sd, sd + CAP, sd + 2 * CAP, nd, x, y, z);
            lds_gather3(sd + 3 * CAP, sd + 4 * CAP, sd + 5 * CAP, nd, u, v, w);
        } else {
            gather3(s.x1, s.y1, s.z1, nd, x, y, z);
            gather3(s.u1, s.v1, s.w1, nd, u, v, w);
        }
        // Shock detector neighbour extreme (K/force_sgh.cpp:297-310).  r_face =
        // 0.9 (d_nb + small)/(div + small) is monotone in d_nb (every rounding
        // step is monotone), so the minimum over neighbours is reached at min d_nb
        // when div+small >= 0 and at max d_nb otherwise.  Neighbours = every other
        // element sharing a node (the reference's elems_in_elem).  k_div_extremes
        // stores per node the two smallest and two largest divergences of its
        // elements (with multiplicity); removing one copy of this element's own
        // value leaves the extreme over the other elements of that node.
        const bool lo = div_e + small >= 0.0;
        double ext = lo ? HUGE_VAL : -HUGE_VAL;
        if (staged) {
            const double2* xs = lo ? sx : sx + CAP;
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const double2 t = xs[nd[a]];
                const double other = (t.x == div_e) ? t.y : t.x;
                ext = lo ? fmin(ext, other) : fmax(ext, other);
            }
        } else {
            const double2* __restrict__ xs = lo ? s.ext_lo : s.ext_hi;
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const double2 t = xs[nd[a]];
                const double other = (t.x == div_e) ? t.y : t.x;
                ext = lo ? fmin(ext, other) : fmax(ext, other);
            }
        }
        force_elem_coef(x, y, z, u, v, w, kc, ext, fuzz, small, f);
    }
    if (staged) __syncthreads();   // every node read is done before smem is reused
    if (live) {
#pragma unroll
        for (int i = 0; i < 12; ++i)
            smem[threadIdx.x * 13 + i] = make_double2(f[2 * i], f[2 * i + 1]);
    }
    __syncthreads();
    const int cnt = 12 * (bi.end - bi.start);
    double2* out = reinterpret_cast<double2*>(s.F + 24 * static_cast<size_t>(bi.start));
#pragma unroll
    for (int r = 0; r < 12; ++r) {
        const int idx = r * BS + static_cast<int>(threadIdx.x);
        if (idx < cnt) stream_store(out + idx, smem[(idx / 12) * 13 + idx % 12]);
    }
}

// --------------------------------------------------------------------------
// operator kernels with the reference's per-operator semantics
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_get_vol(DMesh m, DState s)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[8];
    load_conn(m.conn, e, nd);
    double x[8], y[8], z[8];
    gather3(s.x1, s.y1, s.z1, nd, x, y, z);
    s.vol[e] = hex_volume(x, y, z);
}

__global__ __launch_bounds__(kBlock) void k_get_divergence(DMesh m, DState s)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[8];
    load_conn(m.conn, e, nd);
    double x[8], y[8], z[8], u[8], v[8], w[8], B[24];
    gather3(s.x1, s.y1, s.z1, nd, x, y, z);
    hex_bmatrix(x, y, z, B);
    gather3(s.u1, s.v1, s.w1, nd, u, v, w);
    s.div[e] = divergence(B, u, v, w, s.vol[e]);
}

// get_divergence with the block's nodes (x, y, z, u, v, w) staged in LDS,
// list indices requested first; the same arithmetic as k_get_divergence.
// Blocks over the LDS capacity gather.
template <int BS>
__global__ __launch_bounds__(BS) void k_get_divergence_staged(DMesh m, DState s)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    constexpr int CAP = stage_cap<BS>();
    __shared__ double sd[6 * CAP];
    int q0, nn;
    stage_range(m, bi, CAP, q0, nn);
    const bool staged = nn > 0 && nn <= CAP;
    if (staged) {
        constexpr int R = (CAP + BS - 1) / BS;
        uint32_t nid[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * BS + static_cast<int>(threadIdx.x);
            nid[r] = static_cast<uint32_t>(m.sn_nodes[q0 + (i < nn ? i : nn - 1)]);
        }
        double g[R][6];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (wave_in_list(r * BS, nn)) {
                const uint32_t n = nid[r];
                g[r][0] = s.x1[n];
                g[r][1] = s.y1[n];
                g[r][2] = s.z1[n];
                g[r][3] = s.u1[n];
                g[r][4] = s.v1[n];
                g[r][5] = s.w1[n];
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * BS + static_cast<int>(threadIdx.x);
            if (i < nn) {
#pragma unroll
                for (int k = 0; k < 6; ++k) sd[k * CAP + i] = g[r][k];
            }
        }
        __syncthreads();
    }
    if (e >= bi.end) return;
    int32_t nd[8];
    double x[8], y[8], z[8], u[8], v[8], w[8], B[24];
    if (staged) {
        load_lconn(m.sn_lconn, e, nd);
        lds_gather3(sd, sd + CAP, sd + 2 * CAP, nd, x, y, z);
        hex_bmatrix(x, y, z, B);
        lds_gather3(sd + 3 * CAP, sd + 4 * CAP, sd + 5 * CAP, nd, u, v, w);
    } else {
        load_conn(m.conn, e, nd);
        gather3(s.x1, s.y1, s.z1, nd, x, y, z);
        hex_bmatrix(x, y, z, B);
        gather3(s.u1, s.v1, s.w1, nd, u, v, w);
    }
    s.div[e] = divergence(B, u, v, w, s.vol[e]);
}

// K/momentum.cpp:23-49: CSR gather in corners_in_node order, no atomics.
// The first eight entries are fetched with independent loads (indices clamped
// to the list, so every load is in bounds and none waits on a branch), then
// summed in list order; longer lists (irregular meshes) continue in a loop.
__device__ __forceinline__ void gather_node_force(const DMesh& m, const double* __restrict__ F,
                                                  int n, double& f0, double& f1, double& f2)
{
    f0 = 0.0;
    f1 = 0.0;
    f2 = 0.0;
    const int q0 = m.cin_start[n], cnt = m.cin_start[n + 1] - q0;
    if (cnt <= 0) return;
    const int last = cnt - 1;
    double g0[8], g1[8], g2[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const size_t k = 3 * static_cast<size_t>(m.cin[q0 + (j < cnt ? j : last)]);
        g0[j] = F[k];
        g1[j] = F[k + 1];
        g2[j] = F[k + 2];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool ok = j < cnt;
        f0 = ok ? f0 + g0[j] : f0;
        f1 = ok ? f1 + g1[j] : f1;
        f2 = ok ? f2 + g2[j] : f2;
    }
    for (int q = q0 + 8; q < q0 + cnt; ++q) {
        const size_t k = 3 * static_cast<size_t>(m.cin[q]);
        f0 += F[k];
        f1 += F[k + 1];
        f2 += F[k + 2];
    }
}

// gather_node_force from the block's LDS copy of its corner forces (the
// entries' positions in the block's sorted corner list), same order and sums
__device__ __forceinline__ void lds_node_force(const DMesh& m, const double* sF, int n, double& f0,
                                               double& f1, double& f2)
{
    f0 = 0.0;
    f1 = 0.0;
    f2 = 0.0;
    const int q0 = m.cin_start[n], cnt = m.cin_start[n + 1] - q0;
    if (cnt <= 0) return;
    const int last = cnt - 1;
    int p[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) p[j] = 3 * m.cs_pos[q0 + (j < cnt ? j : last)];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool ok = j < cnt;
        const double g0 = sF[p[j]], g1 = sF[p[j] + 1], g2 = sF[p[j] + 2];
        f0 = ok ? f0 + g0 : f0;
        f1 = ok ? f1 + g1 : f1;
        f2 = ok ? f2 + g2 : f2;
    }
    for (int q = q0 + 8; q < q0 + cnt; ++q) {
        const int k = 3 * m.cs_pos[q];
        f0 += sF[k];
        f1 += sF[k + 1];
        f2 += sF[k + 2];
    }
}

// insert v into the pair (smallest, second smallest), multiplicity counted
__device__ __forceinline__ void insert_min2(double& a, double& b, double v)
{
    b = fmin(b, fmax(a, v));
    a = fmin(a, v);
}
__device__ __forceinline__ void insert_max2(double& a, double& b, double v)
{
    b = fmax(b, fmin(a, v));
    a = fmax(a, v);
}

// Per node, the two smallest and two largest divergences of the elements
// around it, counted with multiplicity (feeds the shock detector).  A NaN
// divergence is left out, as fmin/fmax leave it out of the reference's
// neighbour minimum.  NC: corners read unconditionally (clamped to the list)
// before the loop over longer lists: 8 for hexes, 4 for RZ quads (a node of
// a quad mesh has at most 4 corners unless elements are repeated).
template <int BS, int NC = 8>
__global__ __launch_bounds__(BS) void k_div_extremes(DMesh m, DState s, const InstCtl* ctl)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    if (ctl != nullptr && ctl[bi.inst].stopped) return;
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    const int q0 = m.cin_start[n], cnt = m.cin_start[n + 1] - q0;
    double mn1 = HUGE_VAL, mn2 = HUGE_VAL, mx1 = -HUGE_VAL, mx2 = -HUGE_VAL;
    if (cnt > 0) {
        const int last = cnt - 1;
        double d[NC];
#pragma unroll
        for (int j = 0; j < NC; ++j) d[j] = s.div[m.cin[q0 + (j < cnt ? j : last)] >> m.cshift];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            const bool ok = j < cnt && !isnan(d[j]);
            insert_min2(mn1, mn2, ok ? d[j] : HUGE_VAL);
            insert_max2(mx1, mx2, ok ? d[j] : -HUGE_VAL);
        }
        for (int q = q0 + NC; q < q0 + cnt; ++q) {
            const double v = s.div[m.cin[q] >> m.cshift];
            if (!isnan(v)) {
                insert_min2(mn1, mn2, v);
                insert_max2(mx1, mx2, v);
            }
        }
    }
    s.ext_lo[n] = make_double2(mn1, mn2);
    s.ext_hi[n] = make_double2(mx1, mx2);
}

__global__ __launch_bounds__(kBlock) void k_update_velocity(DMesh m, DState s, double alpha,
                                                            double dt)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    double f0, f1, f2;
    gather_node_force(m, s.F, n, f0, f1, f2);
    const double mass = s.nmass[n];
    s.u1[n] = s.u0[n] + alpha * dt * f0 / mass;
    s.v1[n] = s.v0[n] + alpha * dt * f1 / mass;
    s.w1[n] = s.w0[n] + alpha * dt * f2 / mass;
}

// K/boundary.cpp:16-55, all sets merged into one mask per boundary node
__global__ __launch_bounds__(kBlock) void k_boundary_velocity(const int32_t* nodes,
                                                              const uint8_t* masks, int32_t count,
                                                              DState s)
{
    const int i = blockIdx.x * kBlock + static_cast<int>(threadIdx.x);
    if (i >= count) return;
    const int n = nodes[i];
    const uint8_t mk = masks[i];
    if (mk & 1) s.u1[n] = 0.0;
    if (mk & 2) s.v1[n] = 0.0;
    if (mk & 4) s.w1[n] = 0.0;
}

__device__ __forceinline__ void load_forces(const DState& s, int e, double f[24])
{
    const double2* Fp = reinterpret_cast<const double2*>(s.F + 24 * static_cast<size_t>(e));
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        const double2 t = Fp[i];
        f[2 * i] = t.x;
        f[2 * i + 1] = t.y;
    }
}

__global__ __launch_bounds__(kBlock) void k_update_energy(DMesh m, DState s, double alpha,
                                                          double dt)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[8];
    load_conn(m.conn, e, nd);
    double f[24], u1[8], v1[8], w1[8], u0[8], v0[8], w0[8];
    load_forces(s, e, f);
    gather3(s.u1, s.v1, s.w1, nd, u1, v1, w1);
    gather3(s.u0, s.v0, s.w0, nd, u0, v0, w0);
    s.sie1[e] = energy_update(f, u1, v1, w1, u0, v0, w0, s.sie0[e], s.emass[e], alpha, dt);
}

// update_energy_sgh (K/energy_sgh.cpp:16-52) as the fused loop's energy
// kernel reads its inputs: the block's nodes staged in LDS as their mean
// stage velocities half_sum(v1, v0) (list indices requested first), the
// wave's corner-force rows read as one coalesced stretch; the same
// arithmetic as k_update_energy.  Blocks over the LDS capacity gather.
template <int BS>
__global__ __launch_bounds__(BS) void k_update_energy_staged(DMesh m, DState s, double alpha,
                                                             double dt)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    constexpr int CAP = stage_cap<BS>();
    __shared__ double sd[3 * CAP];
    __shared__ double2 sfw[(BS / 64) * 32 * 13];
    int q0, nn;
    stage_range(m, bi, CAP, q0, nn);
    const bool staged = nn > 0 && nn <= CAP;
    constexpr int R = (CAP + BS - 1) / BS;
    uint32_t nid[R];
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * BS + static_cast<int>(threadIdx.x);
            nid[r] = static_cast<uint32_t>(m.sn_nodes[q0 + (i < nn ? i : nn - 1)]);
        }
    }
    const int w0e = bi.start + (static_cast<int>(threadIdx.x) & ~63);
    const bool wave_live = w0e < bi.end;   // wave-uniform
    double2 ft[12];
    wave_issue_forces(s.F, wave_live ? w0e : bi.start, wave_live ? bi.end - w0e : 1, ft);
    const bool live = e < bi.end;
    const int ec = live ? e : bi.start;    // in bounds; unused past the item
    const double emass = s.emass[ec];
    const double sie0 = s.sie0[ec];
    if (staged) {
        double g[R][6];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (wave_in_list(r * BS, nn)) {
                const uint32_t n = nid[r];
                g[r][0] = s.u1[n];
                g[r][1] = s.v1[n];
                g[r][2] = s.w1[n];
                g[r][3] = s.u0[n];
                g[r][4] = s.v0[n];
                g[r][5] = s.w0[n];
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * BS + static_cast<int>(threadIdx.x);
            if (i < nn) {
#pragma unroll
                for (int k = 0; k < 3; ++k) sd[k * CAP + i] = half_sum(g[r][k], g[r][3 + k]);
            }
        }
        __syncthreads();
    }
    double f[24];
    if (wave_live) wave_scatter_forces(ft, sfw + (threadIdx.x >> 6) * 32 * 13, f);
    if (!live) return;
    double sie;
    if (staged) {
        int32_t nd[8];
        load_lconn(m.sn_lconn, e, nd);
        double hu[8], hv[8], hw[8];
        lds_gather3(sd, sd + CAP, sd + 2 * CAP, nd, hu, hv, hw);
        sie = energy_update_h(f, hu, hv, hw, sie0, emass, alpha, dt);
    } else {
        int32_t nd[8];
        load_conn(m.conn, e, nd);
        double u1[8], v1[8], w1[8], u0[8], v0[8], w0[8];
        gather3(s.u1, s.v1, s.w1, nd, u1, v1, w1);
        gather3(s.u0, s.v0, s.w0, nd, u0, v0, w0);
        sie = energy_update(f, u1, v1, w1, u0, v0, w0, sie0, emass, alpha, dt);
    }
    s.sie1[e] = sie;
}

__global__ __launch_bounds__(kBlock) void k_update_position(DMesh m, DState s, double alpha,
                                                            double dt)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    s.x1[n] = s.x0[n] + alpha * dt * ((s.u1[n] + s.u0[n]) * 0.5);
    s.y1[n] = s.y0[n] + alpha * dt * ((s.v1[n] + s.v0[n]) * 0.5);
    s.z1[n] = s.z0[n] + alpha * dt * ((s.w1[n] + s.w0[n]) * 0.5);
}

// K/properties.cpp:30-109 with the ideal gas
__global__ __launch_bounds__(kBlock) void k_update_state(DMesh m, DState s)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    const double den = s.emass[e] / s.vol[e];
    s.den[e] = den;
    double p, c;
    eos(s.mats[s.nmat > 1 ? s.mat[e] : 0].eos, s.gamma[e], s.csmin[e], den, s.sie1[e], p, c);
    s.pres[e] = p;
    s.sspd[e] = c;
}

__global__ __launch_bounds__(kBlock) void k_rk_init_nodes(DMesh m, DState s)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    s.x0[n] = s.x1[n];
    s.y0[n] = s.y1[n];
    s.z0[n] = s.z1[n];
    s.u0[n] = s.u1[n];
    s.v0[n] = s.v1[n];
    s.w0[n] = s.w1[n];
}

__global__ __launch_bounds__(kBlock) void k_rk_init_elems(DMesh m, DState s)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    s.sie0[e] = s.sie1[e];
    if (s.sig0 != nullptr) {
        // stress is always 3 x 3 (K/time_integration.cpp:19-24)
        const size_t E = static_cast<size_t>(m.E);
#pragma unroll
        for (int k = 0; k < 9; ++k) s.sig0[k * E + e] = s.sig1[k * E + e];
    }
}

// per-element CFL dt, block minimum into partial[logical block]
template <int BS>
__global__ __launch_bounds__(BS) void k_dt_partial(DMesh m, DState s, const InstCtl* ctl,
                                                   double* partial, double dt_cfl,
                                                   double dt_max, double dt_min, double fuzz,
                                                   int use_ctl)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    double cfl = dt_cfl, dmax = dt_max, dmin = dt_min, fz = fuzz;
    if (use_ctl) {
        const InstCtl& c = ctl[bi.inst];
        cfl = c.dt_cfl;
        dmax = c.dt_max;
        dmin = c.dt_min;
        fz = c.fuzz;
    }
    double d = DBL_MAX;
    if (e < bi.end) {
        int32_t nd[8];
        load_conn(m.conn, e, nd);
        double x[8], y[8], z[8];
        gather3(s.x1, s.y1, s.z1, nd, x, y, z);
        d = elem_dt(min_pair_distance(x, y, z), s.sspd[e], cfl, dmax, dmin, fz);
    }
    d = block_min<BS>(d);
    if (threadIdx.x == 0) partial[bi.lid] = d;
}

__global__ __launch_bounds__(1024) void k_reduce_min(const double* partial, int32_t n, double* out)
{
    double v = DBL_MAX;
    for (int i = threadIdx.x; i < n; i += 1024) v = fmin(v, partial[i]);
    __shared__ double red[16];
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int i = 1; i < 16; ++i) r = fmin(r, red[i]);
        *out = r;
    }
}

// --------------------------------------------------------------------------
// fused cycle kernels (dt, time and stop state stay on the device)
// --------------------------------------------------------------------------

// start of a cycle: get_timestep's host part (K/time_integration.cpp:69-150)
__global__ __launch_bounds__(1024) void k_cycle_dt(InstCtl* ctl, int32_t ninst,
                                                    const double* partial,
                                                    const double* dt_replay, double* dt_record)
{
    InstCtl& c = ctl[blockIdx.x];
    if (c.stopped) return;
    const size_t slot = static_cast<size_t>(c.cycle) * ninst + blockIdx.x;
    if (dt_replay != nullptr) {
        // frozen time steps (derivative checks): the recorded dt of this cycle
        if (threadIdx.x == 0) {
            c.dt = dt_replay[slot];
            if (dt_record != nullptr) dt_record[slot] = c.dt;
        }
        return;
    }
    double v0 = DBL_MAX, v1 = DBL_MAX, v2 = DBL_MAX, v3 = DBL_MAX;
    int i = c.eblk_begin + static_cast<int>(threadIdx.x);
    for (; i + 3 * 1024 < c.eblk_end; i += 4 * 1024) {
        v0 = fmin(v0, partial[i]);
        v1 = fmin(v1, partial[i + 1024]);
        v2 = fmin(v2, partial[i + 2048]);
        v3 = fmin(v3, partial[i + 3072]);
    }
    for (; i < c.eblk_end; i += 1024) v0 = fmin(v0, partial[i]);
    double v = fmin(fmin(v0, v1), fmin(v2, v3));
    __shared__ double red[16];
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < 16; ++k) r = fmin(r, red[k]);
        double dt = c.dt * 1.1;
        if (r < dt) dt = r;
        dt = fmin(dt, (c.graphics_time - c.time) + c.fuzz);
        dt = fmin(dt, c.time_final - c.time);
        c.dt = dt;
        if (dt_record != nullptr) dt_record[slot] = dt;
    }
}

// velocity + boundary + position, node-centric (K/momentum.cpp:23-49,
// K/boundary.cpp:16-55, K/geometry.cpp:17-24)
// FULL = false: the velocity update alone (the update_velocity_sgh operator,
// K/momentum.cpp:23-49) with dt given by the caller (ctl == nullptr); the
// boundary and position operators then run separately.
template <int BS, bool FULL = true>
__global__ __launch_bounds__(BS) void k_vel_bc_pos(DMesh m, DState s, const InstCtl* ctl,
                                                   double alpha, double dt_arg = 0.0)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    double dt = dt_arg;
    if (ctl != nullptr) {
        if (ctl[bi.inst].stopped) return;
        dt = ctl[bi.inst].dt;
    }
    const int n = bi.start + static_cast<int>(threadIdx.x);
    // The block's corner forces in LDS, loaded in ascending corner order (the
    // block's sorted corner list): consecutive lanes read consecutive doubles
    // of few elements' force rows.  Blocks over the capacity gather from
    // global memory.
    constexpr int CAPC = corner_stage_cap(BS);
    __shared__ double sF[3 * CAPC];
    const int q0 = m.cin_start[bi.start];
    const int nc = m.cin_start[bi.end] - q0;
    const bool staged = m.cs_cor != nullptr && nc <= CAPC;
    if (staged) {
        // CAPC / BS corners per lane, every load issued before any store
        constexpr int R = CAPC / BS;
        int cid[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = r * BS + static_cast<int>(threadIdx.x);
            cid[r] = k < nc ? m.cs_cor[q0 + k] : -1;
        }
        double g[R][3];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double* src = s.F + 3 * static_cast<size_t>(cid[r] < 0 ? 0 : cid[r]);
            g[r][0] = src[0];
            g[r][1] = src[1];
            g[r][2] = src[2];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = r * BS + static_cast<int>(threadIdx.x);
            if (cid[r] >= 0) {
                sF[3 * k] = g[r][0];
                sF[3 * k + 1] = g[r][1];
                sF[3 * k + 2] = g[r][2];
            }
        }
        __syncthreads();
    }
    if (n >= bi.end) return;
    double f[3], v0[3], x0[3], v1[3], x1[3];
    if (staged) lds_node_force(m, sF, n, f[0], f[1], f[2]);
    else gather_node_force(m, s.F, n, f[0], f[1], f[2]);
    v0[0] = s.u0[n];
    v0[1] = s.v0[n];
    v0[2] = s.w0[n];
    if (!FULL) {
        // v1 = v0 + alpha dt F / m, as node_update evaluates it
        const double mass = s.nmass[n];
        s.u1[n] = v0[0] + alpha * dt * f[0] / mass;
        s.v1[n] = v0[1] + alpha * dt * f[1] / mass;
        s.w1[n] = v0[2] + alpha * dt * f[2] / mass;
        return;
    }
    x0[0] = s.x0[n];
    x0[1] = s.y0[n];
    x0[2] = s.z0[n];
    node_update(f, s.nmass[n], v0, x0, s.bcmask[n], alpha, dt, v1, x1);
    // streaming stores: the next reader is another kernel, megabytes later
    __builtin_nontemporal_store(v1[0], &s.u1[n]);
    __builtin_nontemporal_store(v1[1], &s.v1[n]);
    __builtin_nontemporal_store(v1[2], &s.w1[n]);
    __builtin_nontemporal_store(x1[0], &s.x1[n]);
    __builtin_nontemporal_store(x1[1], &s.y1[n]);
    __builtin_nontemporal_store(x1[2], &s.z1[n]);
}

// energy + volume + EOS + next divergence (+ next cycle's dt candidates),
// element-centric (K/energy_sgh.cpp:16-52, K/geometry.cpp:255-306,
// K/properties.cpp:30-106, K/momentum.cpp:229-287, K/time_integration.cpp:74-140)
// USER = false: every material is ideal gas, the EOS selector folds away
template <int BS, int MINW, bool USER>
__global__ __launch_bounds__(BS, MINW) void k_energy_state_div(DMesh m, DState s,
                                                               const InstCtl* ctl, double alpha,
                                                               double* dt_partial)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const InstCtl& c = ctl[bi.inst];
    if (c.stopped) return;
    const int e = bi.start + static_cast<int>(threadIdx.x);
    // the block's distinct nodes in LDS: u1, v1, w1, u0, v0, w0, x1, y1, z1;
    // per wave a region for its corner forces (wave_scatter_forces)
    constexpr int CAP = stage_cap<BS>();
    __shared__ double sd[9 * CAP];
    __shared__ double2 sfw[(BS / 64) * 32 * 13];
    // The load counter retires in issue order, so what the staging needs
    // first is requested first: the block's node-list indices, then this
    // wave's corner forces (they arrive while the block stages its nodes),
    // then the element's own inputs.
    int q0, nn;
    stage_range(m, bi, CAP, q0, nn);
    const bool staged = nn > 0 && nn <= CAP;   // lists are never empty; the test keeps nn - 1 in range
    constexpr int R = (CAP + BS - 1) / BS;
    uint32_t nid[R];   // unsigned: its address arithmetic needs no sign extension
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // clamped, unconditional loads: nothing computed from them before
            // the node loads, so those wait for these alone
            const int i = r * BS + static_cast<int>(threadIdx.x);
            nid[r] = static_cast<uint32_t>(m.sn_nodes[q0 + (i < nn ? i : nn - 1)]);
        }
    }
    const int w0e = bi.start + (static_cast<int>(threadIdx.x) & ~63);
    const bool wave_live = w0e < bi.end;   // wave-uniform
    double2 ft[12];
    // a wave past the item reads the item's first row (values unused)
    wave_issue_forces(s.F, wave_live ? w0e : bi.start, wave_live ? bi.end - w0e : 1, ft);
    const bool live = e < bi.end;
    const double emass = live ? s.emass[e] : 1.0;
    const double sie0 = live ? s.sie0[e] : 0.0;
    const double gam = live ? s.gamma[e] : 0.0;
    const double csmin = live ? s.csmin[e] : 0.0;
    if (staged) {
        // every node value, then the LDS stores
        double g[R][9];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (wave_in_list(r * BS, nn)) {
                const uint32_t n = nid[r];
                g[r][0] = s.u1[n];
                g[r][1] = s.v1[n];
                g[r][2] = s.w1[n];
                g[r][3] = s.u0[n];
                g[r][4] = s.v0[n];
                g[r][5] = s.w0[n];
                g[r][6] = s.x1[n];
                g[r][7] = s.y1[n];
                g[r][8] = s.z1[n];
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * BS + static_cast<int>(threadIdx.x);
            if (i < nn) {
                // slots 3..5 hold the node's mean velocity (v1 + v0) / 2
#pragma unroll
                for (int k = 0; k < 3; ++k) g[r][3 + k] = half_sum(g[r][k], g[r][3 + k]);
#pragma unroll
                for (int k = 0; k < 9; ++k) sd[k * CAP + i] = g[r][k];
            }
        }
        __syncthreads();
    }
    double dtl = DBL_MAX;
    double f[24];
    if (wave_live) wave_scatter_forces(ft, sfw + (threadIdx.x >> 6) * 32 * 13, f);
    if (e < bi.end) {
        int32_t nd[8];
        double sie;
        if (staged) {
            load_lconn(m.sn_lconn, e, nd);
            double hu[8], hv[8], hw[8];
            lds_gather3(sd + 3 * CAP, sd + 4 * CAP, sd + 5 * CAP, nd, hu, hv, hw);
            sie = energy_update_h(f, hu, hv, hw, sie0, emass, alpha, c.dt);
        } else {
            load_conn(m.conn, e, nd);
            double u1[8], v1[8], w1[8], u0[8], v0[8], w0[8];
            gather3(s.u1, s.v1, s.w1, nd, u1, v1, w1);
            gather3(s.u0, s.v0, s.w0, nd, u0, v0, w0);
            sie = energy_update(f, u1, v1, w1, u0, v0, w0, sie0, emass, alpha, c.dt);
        }
        s.sie1[e] = sie;
        double x[8], y[8], z[8], u1[8], v1[8], w1[8];
        if (staged) {
            lds_gather3(sd, sd + CAP, sd + 2 * CAP, nd, u1, v1, w1);
            lds_gather3(sd + 6 * CAP, sd + 7 * CAP, sd + 8 * CAP, nd, x, y, z);
        } else {
            gather3(s.u1, s.v1, s.w1, nd, u1, v1, w1);
            gather3(s.x1, s.y1, s.z1, nd, x, y, z);
        }
        double vol, den, p, cs, div;
        const int32_t kind = USER ? s.mats[s.nmat > 1 ? s.mat[e] : 0].eos : kEosIdealGas;
        derived_state(x, y, z, u1, v1, w1, emass, sie, gam, csmin, kind, vol, den, p, cs, div);
        s.vol[e] = vol;
        s.den[e] = den;
        s.pres[e] = p;
        s.sspd[e] = cs;
        s.div_out[e] = div;
        if (dt_partial != nullptr)
            dtl = elem_dt(min_pair_distance(x, y, z), cs, c.dt_cfl, c.dt_max, c.dt_min, c.fuzz);
    }
    if (dt_partial != nullptr) {
        dtl = block_min<BS>(dtl);
        if (threadIdx.x == 0) dt_partial[bi.lid] = dtl;
    }
}

// --------------------------------------------------------------------------
// kernels with a stress field (strength-model materials)
// --------------------------------------------------------------------------
__device__ __forceinline__ void load_stress(const double* sig, size_t E, int e, double st[9])
{
#pragma unroll
    for (int k = 0; k < 9; ++k) st[k] = sig[static_cast<size_t>(k) * E + e];
}
__device__ __forceinline__ void store_stress(double* sig, size_t E, int e, const double st[9])
{
#pragma unroll
    for (int k = 0; k < 9; ++k) sig[static_cast<size_t>(k) * E + e] = st[k];
}

// get_force_sgh with tau = stress - p I and the hypo-elastic hook after the
// forces (K/force_sgh.cpp:84-99, 127-137, 338-391)
__global__ __launch_bounds__(kBlock) void k_force_stress(DMesh m, DState s, const InstCtl* ctl,
                                                         double fuzz, double small, double dt,
                                                         double alpha)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    if (ctl != nullptr) {
        if (ctl[bi.inst].stopped) return;
        fuzz = ctl[bi.inst].fuzz;
        small = ctl[bi.inst].small;
        dt = ctl[bi.inst].dt;
    }
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    const size_t E = static_cast<size_t>(m.E);
    int32_t nd[8];
    load_conn(m.conn, e, nd);
    double x[8], y[8], z[8], u[8], v[8], w[8];
    gather3(s.x1, s.y1, s.z1, nd, x, y, z);
    gather3(s.u1, s.v1, s.w1, nd, u, v, w);
    const double vol = s.vol[e];
    const double div = s.div[e];
    const double den = s.den[e];
    const double sspd = s.sspd[e];
    const double pres = s.pres[e];
    const int32_t mat = s.nmat > 1 ? s.mat[e] : 0;
    const Material mt = s.mats[mat];
    const bool lo = div + small >= 0.0;
    const double2* __restrict__ xs = lo ? s.ext_lo : s.ext_hi;
    double ext = lo ? HUGE_VAL : -HUGE_VAL;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const double2 t = xs[nd[a]];
        const double other = (t.x == div) ? t.y : t.x;
        ext = lo ? fmin(ext, other) : fmax(ext, other);
    }
    double st[9], B[24], f[24];
    load_stress(s.sig1, E, e, st);
    force_elem_stressed(x, y, z, u, v, w, den, pres, sspd, vol, div, ext, mt.q1, mt.q1ex, mt.q2,
                        mt.q2ex, fuzz, small, st, B, f);
    double2* out = reinterpret_cast<double2*>(s.F + 24 * static_cast<size_t>(e));
#pragma unroll
    for (int i = 0; i < 12; ++i) out[i] = make_double2(f[2 * i], f[2 * i + 1]);
    if (mt.strength == kStrengthHypo) {
        double L[9];
        velgrad(B, u, v, w, vol, L);
        double p = pres, c = sspd, stn[9];
        load_stress(s.sig0, E, e, stn);
        user_strength(mat, s.sparams + kStrengthParams * mat, den, s.sie0[e], L, vol, dt, alpha,
                      stn, st, p, c);
        s.pres[e] = p;
        s.sspd[e] = c;
    }
    store_stress(s.sig1w, E, e, st);
}

// the fused energy/state kernel with a stress field: the hyper-elastic hook
// before the EOS (K/properties.cpp:39-106), the user EOS seeing the stress
__global__ __launch_bounds__(kBlock) void k_energy_state_div_stress(DMesh m, DState s,
                                                                    const InstCtl* ctl,
                                                                    double alpha,
                                                                    double* dt_partial)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const InstCtl& c = ctl[bi.inst];
    if (c.stopped) return;
    const int e = bi.start + static_cast<int>(threadIdx.x);
    double dtl = DBL_MAX;
    if (e < bi.end) {
        const size_t E = static_cast<size_t>(m.E);
        int32_t nd[8];
        load_conn(m.conn, e, nd);
        double f[24], x[8], y[8], z[8], u1[8], v1[8], w1[8], u0[8], v0[8], w0[8];
        load_forces(s, e, f);
        gather3(s.u1, s.v1, s.w1, nd, u1, v1, w1);
        gather3(s.u0, s.v0, s.w0, nd, u0, v0, w0);
        const double emass = s.emass[e];
        const double sie =
            energy_update(f, u1, v1, w1, u0, v0, w0, s.sie0[e], emass, alpha, c.dt);
        s.sie1[e] = sie;
        gather3(s.x1, s.y1, s.z1, nd, x, y, z);
        double B[24];
        const double vol = hex_bmatrix_volume(x, y, z, B);
        const double den = emass / vol;
        const double div = divergence(B, u1, v1, w1, vol);
        const int32_t mat = s.nmat > 1 ? s.mat[e] : 0;
        const Material mt = s.mats[mat];
        double st[9];
        load_stress(s.sig1, E, e, st);
        double p = s.pres[e], cs = s.sspd[e];
        if (mt.strength == kStrengthHyper) {
            double L[9];
            velgrad(B, u1, v1, w1, vol, L);
            double stn[9];
            load_stress(s.sig0, E, e, stn);
            user_strength(mat, s.sparams + kStrengthParams * mat, den, s.sie0[e], L, vol, c.dt,
                          alpha, stn, st, p, cs);
            store_stress(s.sig1, E, e, st);
        }
        eos(mt.eos, s.gamma[e], s.csmin[e], den, sie, p, cs, st);
        s.vol[e] = vol;
        s.den[e] = den;
        s.pres[e] = p;
        s.sspd[e] = cs;
        s.div_out[e] = div;
        if (dt_partial != nullptr)
            dtl = elem_dt(min_pair_distance(x, y, z), cs, c.dt_cfl, c.dt_max, c.dt_min, c.fuzz);
    }
    if (dt_partial != nullptr) {
        dtl = block_min<kBlock>(dtl);
        if (threadIdx.x == 0) dt_partial[bi.lid] = dtl;
    }
}

// update_state with a stress field (K/properties.cpp:9-114)
__global__ __launch_bounds__(kBlock) void k_update_state_stress(DMesh m, DState s, double dt,
                                                                double alpha)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    const size_t E = static_cast<size_t>(m.E);
    const double den = s.emass[e] / s.vol[e];
    s.den[e] = den;
    const int32_t mat = s.nmat > 1 ? s.mat[e] : 0;
    const Material mt = s.mats[mat];
    double st[9];
    load_stress(s.sig1, E, e, st);
    double p = s.pres[e], c = s.sspd[e];
    if (mt.strength == kStrengthHyper) {
        int32_t nd[8];
        load_conn(m.conn, e, nd);
        double x[8], y[8], z[8], u[8], v[8], w[8], B[24], L[9];
        gather3(s.x1, s.y1, s.z1, nd, x, y, z);
        gather3(s.u1, s.v1, s.w1, nd, u, v, w);
        hex_bmatrix(x, y, z, B);
        velgrad(B, u, v, w, s.vol[e], L);
        double stn[9];
        load_stress(s.sig0, E, e, stn);
        user_strength(mat, s.sparams + kStrengthParams * mat, den, s.sie0[e], L, s.vol[e], dt,
                      alpha, stn, st, p, c);
        store_stress(s.sig1, E, e, st);
    }
    eos(mt.eos, s.gamma[e], s.csmin[e], den, s.sie1[e], p, c, st);
    s.pres[e] = p;
    s.sspd[e] = c;
}

// velocity gradient of the current state (K/momentum.cpp:60-148), E x 9
__global__ __launch_bounds__(kBlock) void k_velgrad(DMesh m, DState s, double* out)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[8];
    load_conn(m.conn, e, nd);
    double x[8], y[8], z[8], u[8], v[8], w[8], B[24], L[9];
    gather3(s.x1, s.y1, s.z1, nd, x, y, z);
    gather3(s.u1, s.v1, s.w1, nd, u, v, w);
    hex_bmatrix(x, y, z, B);
    velgrad(B, u, v, w, s.vol[e], L);
#pragma unroll
    for (int k = 0; k < 9; ++k) out[9 * static_cast<size_t>(e) + k] = L[k];
}

// end of a cycle (K/sgh_solve.cpp:590-632)
__global__ void k_time_advance(InstCtl* ctl, int32_t ninst, int32_t newest_bin)
{
    const int i = blockIdx.x * blockDim.x + static_cast<int>(threadIdx.x);
    if (i >= ninst) return;
    InstCtl& c = ctl[i];
    if (c.stopped) return;
    const long long cycle = c.cycle;
    c.time += c.dt;
    int write = 0;
    if ((cycle + 1) % c.graphics_cyc_ival == 0 && cycle > 0) write = 1;
    else if (c.time >= c.time_final) write = 1;
    else if (c.time >= c.graphics_time) write = 1;
    if (write) c.graphics_time = c.time + c.graphics_dt_ival;
    c.write_pending = write;
    c.cycle = cycle + 1;
    if (c.time >= c.time_final || c.cycle >= c.cycle_stop) {
        c.stopped = 1;
        c.final_bin = newest_bin;
    }
}

// per-instance energy tallies, fixed-order two-level sums (K/sgh_solve.cpp:77-101)
__global__ __launch_bounds__(kBlock) void k_energy_partial(DState s, int32_t e0, int32_t e1,
                                                           int32_t n0, int32_t n1, double* scratch,
                                                           int32_t nscratch, int rz)
{
    double ie = 0.0, ke = 0.0;
    const int stride = nscratch * kBlock;
    for (int e = e0 + blockIdx.x * kBlock + threadIdx.x; e < e1; e += stride)
        ie += s.emass[e] * s.sie1[e];
    for (int n = n0 + blockIdx.x * kBlock + threadIdx.x; n < n1; n += stride) {
        double k = 0;
        k += s.u1[n] * s.u1[n];
        k += s.v1[n] * s.v1[n];
        if (rz) {
            ke += s.nmass[n] * s.y1[n] * k;
        } else {
            k += s.w1[n] * s.w1[n];
            ke += s.nmass[n] * k;
        }
    }
    ie = block_sum(ie);
    ke = block_sum(ke);
    if (threadIdx.x == 0) {
        scratch[2 * blockIdx.x] = ie;
        scratch[2 * blockIdx.x + 1] = ke;
    }
}

__global__ __launch_bounds__(kBlock) void k_energy_final(const double* scratch, int32_t nscratch,
                                                         double* out2)
{
    double ie = 0.0, ke = 0.0;
    for (int i = threadIdx.x; i < nscratch; i += kBlock) {
        ie += scratch[2 * i];
        ke += scratch[2 * i + 1];
    }
    ie = block_sum(ie);
    ke = block_sum(ke);
    if (threadIdx.x == 0) {
        out2[0] = ie;
        out2[1] = 0.5 * ke;
    }
}

__global__ __launch_bounds__(kBlock) void k_swap_bins(DState s, int32_t n0, int32_t n1, int32_t e0,
                                                      int32_t e1, int32_t Etot)
{
    const int i = blockIdx.x * kBlock + static_cast<int>(threadIdx.x);
    const int n = n0 + i;
    if (n < n1) {
        double t;
        t = s.x0[n]; s.x0[n] = s.x1[n]; s.x1[n] = t;
        t = s.y0[n]; s.y0[n] = s.y1[n]; s.y1[n] = t;
        t = s.z0[n]; s.z0[n] = s.z1[n]; s.z1[n] = t;
        t = s.u0[n]; s.u0[n] = s.u1[n]; s.u1[n] = t;
        t = s.v0[n]; s.v0[n] = s.v1[n]; s.v1[n] = t;
        t = s.w0[n]; s.w0[n] = s.w1[n]; s.w1[n] = t;
    }
    const int e = e0 + i;
    if (e < e1) {
        const double t = s.sie0[e];
        s.sie0[e] = s.sie1[e];
        s.sie1[e] = t;
        if (s.sig0 != nullptr)
            for (int k = 0; k < 9; ++k) {
                const size_t q = static_cast<size_t>(k) * Etot + e;
                const double u = s.sig0[q];
                s.sig0[q] = s.sig1[q];
                s.sig1[q] = u;
            }
    }
}

__global__ __launch_bounds__(kBlock) void k_check(DMesh m, DState s, int32_t* flag)
{
    const int i = blockIdx.x * kBlock + static_cast<int>(threadIdx.x);
    int bad = 0;
    if (i < m.E) {
        const double vals[6] = {s.sie1[i], s.den[i], s.pres[i], s.sspd[i], s.vol[i], s.emass[i]};
        for (double v : vals) bad |= !isfinite(v);
        bad |= !(s.vol[i] > 0.0) || !(s.emass[i] > 0.0);
    }
    if (i < m.N) {
        const double vals[7] = {s.x1[i], s.y1[i], s.z1[i], s.u1[i], s.v1[i], s.w1[i], s.nmass[i]};
        for (double v : vals) bad |= !isfinite(v);
        bad |= !(s.nmass[i] > 0.0);
    }
    if (bad) atomicOr(flag, 1);
}

inline unsigned grid_for(int32_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

}  // namespace

// ---------------------------------------------------------------- launchers --
void launch_get_vol(const DMesh& m, const DState& s, hipStream_t st)
{
    if (m.n_eblk) hipLaunchKernelGGL(k_get_vol, dim3(m.n_eblk), dim3(kBlock), 0, st, m, s);
}
void launch_get_divergence(const DMesh& m, const DState& s, hipStream_t st)
{
    if (m.n_eblk) hipLaunchKernelGGL(k_get_divergence, dim3(m.n_eblk), dim3(kBlock), 0, st, m, s);
}
void launch_get_divergence_staged(const DMesh& m, const DState& s, hipStream_t st)
{
    if (!m.n_eblk) return;
    if (m.bs == 64)
        hipLaunchKernelGGL((k_get_divergence_staged<64>), dim3(m.n_eblk), dim3(64), 0, st, m, s);
    else if (m.bs == 128)
        hipLaunchKernelGGL((k_get_divergence_staged<128>), dim3(m.n_eblk), dim3(128), 0, st, m,
                           s);
    else if (m.bs == 256)
        hipLaunchKernelGGL((k_get_divergence_staged<256>), dim3(m.n_eblk), dim3(256), 0, st, m,
                           s);
    else
        launch_get_divergence(m, s, st);
}
// Variant dispatch of the tuned kernels: the work tables in `m` were built
// for workgroups of m.bs entries; m.minw selects the occupancy floor.
#define FSGH_HEAVY_VARIANTS(X) X(64, 1) X(64, 3) X(128, 1) X(128, 3) X(256, 1) X(256, 3)
#define FSGH_LIGHT_VARIANTS(X) X(64) X(128) X(256) X(512)

bool tuned_variant_ok(int kernel, int32_t bs, int32_t minw)
{
    const bool heavy = kernel == kTunedForce || kernel == kTunedEnergy;
#define FSGH_VH(B, W) if (bs == B && minw == W) return true;
#define FSGH_VL(B) if (bs == B && minw == 1) return true;
    if (heavy) { FSGH_HEAVY_VARIANTS(FSGH_VH) }
    else { FSGH_LIGHT_VARIANTS(FSGH_VL) }
#undef FSGH_VH
#undef FSGH_VL
    return false;
}

void launch_force(const DMesh& m, const DState& s, const InstCtl* ctl, double fuzz, double small,
                  hipStream_t st)
{
    if (!m.n_eblk) return;
#define FSGH_VH(B, W)                                                                       \
    if (m.bs == B && m.minw == W) {                                                        \
        hipLaunchKernelGGL((k_force<B, W>), dim3(m.n_eblk), dim3(B), 0, st, m, s, ctl, fuzz, \
                           small);                                                         \
        return;                                                                            \
    }
    FSGH_HEAVY_VARIANTS(FSGH_VH)
#undef FSGH_VH
}
void launch_div_extremes(const DMesh& m, const DState& s, const InstCtl* ctl, hipStream_t st)
{
    if (!m.n_nblk) return;
    if (m.cshift == 2 && m.bs == kBlock) {   // RZ quads: at most 4 corners per node
        hipLaunchKernelGGL((k_div_extremes<kBlock, 4>), dim3(m.n_nblk), dim3(kBlock), 0, st, m,
                           s, ctl);
        return;
    }
#define FSGH_VL(B)                                                                          \
    if (m.bs == B) {                                                                       \
        hipLaunchKernelGGL((k_div_extremes<B>), dim3(m.n_nblk), dim3(B), 0, st, m, s, ctl);  \
        return;                                                                            \
    }
    FSGH_LIGHT_VARIANTS(FSGH_VL)
#undef FSGH_VL
}
void launch_update_velocity(const DMesh& m, const DState& s, double alpha, double dt,
                            hipStream_t st)
{
    if (m.n_nblk)
        hipLaunchKernelGGL(k_update_velocity, dim3(m.n_nblk), dim3(kBlock), 0, st, m, s, alpha, dt);
}
void launch_update_velocity_staged(const DMesh& m, const DState& s, double alpha, double dt,
                                   hipStream_t st)
{
    if (!m.n_nblk) return;
#define FSGH_VO(B)                                                                          \
    if (m.bs == B) {                                                                       \
        hipLaunchKernelGGL((k_vel_bc_pos<B, false>), dim3(m.n_nblk), dim3(B), 0, st, m, s,   \
                           nullptr, alpha, dt);                                            \
        return;                                                                            \
    }
    FSGH_LIGHT_VARIANTS(FSGH_VO)
#undef FSGH_VO
}
void launch_boundary_velocity(const int32_t* nodes, const uint8_t* masks, int32_t count,
                              const DState& s, hipStream_t st)
{
    if (count > 0)
        hipLaunchKernelGGL(k_boundary_velocity, dim3(grid_for(count)), dim3(kBlock), 0, st, nodes,
                           masks, count, s);
}
void launch_update_energy(const DMesh& m, const DState& s, double alpha, double dt, hipStream_t st)
{
    if (m.n_eblk)
        hipLaunchKernelGGL(k_update_energy, dim3(m.n_eblk), dim3(kBlock), 0, st, m, s, alpha, dt);
}
void launch_update_energy_staged(const DMesh& m, const DState& s, double alpha, double dt,
                                 hipStream_t st)
{
    if (!m.n_eblk) return;
    if (m.bs == 64)
        hipLaunchKernelGGL((k_update_energy_staged<64>), dim3(m.n_eblk), dim3(64), 0, st, m, s,
                           alpha, dt);
    else if (m.bs == 128)
        hipLaunchKernelGGL((k_update_energy_staged<128>), dim3(m.n_eblk), dim3(128), 0, st, m,
                           s, alpha, dt);
    else if (m.bs == 256)
        hipLaunchKernelGGL((k_update_energy_staged<256>), dim3(m.n_eblk), dim3(256), 0, st, m,
                           s, alpha, dt);
    else
        launch_update_energy(m, s, alpha, dt, st);
}
void launch_update_position(const DMesh& m, const DState& s, double alpha, double dt,
                            hipStream_t st)
{
    if (m.n_nblk)
        hipLaunchKernelGGL(k_update_position, dim3(m.n_nblk), dim3(kBlock), 0, st, m, s, alpha, dt);
}
void launch_update_state(const DMesh& m, const DState& s, hipStream_t st)
{
    if (m.n_eblk) hipLaunchKernelGGL(k_update_state, dim3(m.n_eblk), dim3(kBlock), 0, st, m, s);
}
void launch_rk_init(const DMesh& m, const DState& s, hipStream_t st)
{
    if (m.n_eblk) hipLaunchKernelGGL(k_rk_init_elems, dim3(m.n_eblk), dim3(kBlock), 0, st, m, s);
    if (m.n_nblk) hipLaunchKernelGGL(k_rk_init_nodes, dim3(m.n_nblk), dim3(kBlock), 0, st, m, s);
}
void launch_dt_partial(const DMesh& m, const DState& s, const InstCtl* ctl, double* partial,
                       double dt_cfl, double dt_max, double dt_min, double fuzz, int use_ctl,
                       hipStream_t st)
{
    if (!m.n_eblk) return;
#define FSGH_VL(B)                                                                          \
    if (m.bs == B) {                                                                       \
        hipLaunchKernelGGL((k_dt_partial<B>), dim3(m.n_eblk), dim3(B), 0, st, m, s, ctl,     \
                           partial, dt_cfl, dt_max, dt_min, fuzz, use_ctl);                \
        return;                                                                            \
    }
    FSGH_LIGHT_VARIANTS(FSGH_VL)
#undef FSGH_VL
}
void launch_reduce_min(const double* partial, int32_t n, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_min, dim3(1), dim3(1024), 0, st, partial, n, out);
}
void launch_cycle_dt(InstCtl* ctl, int32_t ninst, const double* partial, const double* dt_replay,
                     double* dt_record, hipStream_t st)
{
    hipLaunchKernelGGL(k_cycle_dt, dim3(ninst), dim3(1024), 0, st, ctl, ninst, partial, dt_replay,
                       dt_record);
}
void launch_vel_bc_pos(const DMesh& m, const DState& s, const InstCtl* ctl, double alpha,
                       hipStream_t st)
{
    if (!m.n_nblk) return;
#define FSGH_VL(B)                                                                          \
    if (m.bs == B) {                                                                       \
        hipLaunchKernelGGL((k_vel_bc_pos<B, true>), dim3(m.n_nblk), dim3(B), 0, st, m, s,    \
                           ctl, alpha, 0.0);                                               \
        return;                                                                            \
    }
    FSGH_LIGHT_VARIANTS(FSGH_VL)
#undef FSGH_VL
}
void launch_energy_state_div(const DMesh& m, const DState& s, const InstCtl* ctl, double alpha,
                             double* dt_partial, hipStream_t st)
{
    if (!m.n_eblk) return;
#define FSGH_VH(B, W)                                                                       \
    if (m.bs == B && m.minw == W) {                                                        \
        if (s.user_eos)                                                                    \
            hipLaunchKernelGGL((k_energy_state_div<B, W, true>), dim3(m.n_eblk), dim3(B), 0, \
                               st, m, s, ctl, alpha, dt_partial);                          \
        else                                                                               \
            hipLaunchKernelGGL((k_energy_state_div<B, W, false>), dim3(m.n_eblk), dim3(B), 0, \
                               st, m, s, ctl, alpha, dt_partial);                          \
        return;                                                                            \
    }
    FSGH_HEAVY_VARIANTS(FSGH_VH)
#undef FSGH_VH
}
void launch_force_stress(const DMesh& m, const DState& s, const InstCtl* ctl, double fuzz,
                         double small, double dt, double alpha, hipStream_t st)
{
    if (m.n_eblk)
        hipLaunchKernelGGL(k_force_stress, dim3(m.n_eblk), dim3(kBlock), 0, st, m, s, ctl, fuzz,
                           small, dt, alpha);
}
void launch_energy_state_div_stress(const DMesh& m, const DState& s, const InstCtl* ctl,
                                    double alpha, double* dt_partial, hipStream_t st)
{
    if (m.n_eblk)
        hipLaunchKernelGGL(k_energy_state_div_stress, dim3(m.n_eblk), dim3(kBlock), 0, st, m, s,
                           ctl, alpha, dt_partial);
}
void launch_update_state_stress(const DMesh& m, const DState& s, double dt, double alpha,
                                hipStream_t st)
{
    if (m.n_eblk)
        hipLaunchKernelGGL(k_update_state_stress, dim3(m.n_eblk), dim3(kBlock), 0, st, m, s, dt,
                           alpha);
}
void launch_velgrad(const DMesh& m, const DState& s, double* out, hipStream_t st)
{
    if (m.n_eblk) hipLaunchKernelGGL(k_velgrad, dim3(m.n_eblk), dim3(kBlock), 0, st, m, s, out);
}
void launch_time_advance(InstCtl* ctl, int32_t ninst, int32_t newest_bin, hipStream_t st)
{
    hipLaunchKernelGGL(k_time_advance, dim3(grid_for(ninst)), dim3(kBlock), 0, st, ctl, ninst,
                       newest_bin);
}
void launch_energy_sums(const DMesh& m, const DState& s, const int32_t* eoff, const int32_t* noff,
                        int32_t inst, double* out2, double* scratch, int32_t nscratch, int rz,
                        hipStream_t st)
{
    (void)m;
    hipLaunchKernelGGL(k_energy_partial, dim3(nscratch), dim3(kBlock), 0, st, s, eoff[inst],
                       eoff[inst + 1], noff[inst], noff[inst + 1], scratch, nscratch, rz);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, st, scratch, nscratch, out2);
}
void launch_swap_bins_range(const DState& s, int32_t n0, int32_t n1, int32_t e0, int32_t e1,
                            int32_t Etot, hipStream_t st)
{
    const int32_t cnt = (n1 - n0) > (e1 - e0) ? (n1 - n0) : (e1 - e0);
    if (cnt > 0)
        hipLaunchKernelGGL(k_swap_bins, dim3(grid_for(cnt)), dim3(kBlock), 0, st, s, n0, n1, e0, e1,
                           Etot);
}
void launch_check(const DMesh& m, const DState& s, int32_t* flag, hipStream_t st)
{
    const int32_t cnt = m.N > m.E ? m.N : m.E;
    if (cnt > 0) hipLaunchKernelGGL(k_check, dim3(grid_for(cnt)), dim3(kBlock), 0, st, m, s, flag);
}

}  // namespace fsgh
