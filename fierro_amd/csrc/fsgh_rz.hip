// fsgh_rz.hip -- the 2D axisymmetric (RZ) SGH path: quad meshes in the
// (axial, radius) half plane, the reference's num_dims == 2 branches.
//
// Operators: get_vol (RZ volume), get_divergence2D, get_force_sgh2D (MARS +
// shock detector + hoop terms), velocity / energy (radius-weighted power) /
// timestep2D, and the RZ node-mass update with its on-axis fix
// (K/sgh_solve.cpp:495-581).  The fused cycle runs per RK stage: divergence
// extremes (shared with 3D) -> force -> node (velocity, boundary, position,
// stage-input radius, RZ mass) -> on-axis mass -> energy + volume + EOS +
// next divergence (+ dt candidates).  Mesh maps, boundary sets and fills of
// a 2D context are built on the host (set-up, untimed) and uploaded.
//
// Layout: connectivity E x 4 int32 (int4 per element), corner forces
// [4E][2], node coordinates (x, y) = (axial, radius) in the x and y arrays
// (z and w stay zero).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>

#define FSGH_HD __host__ __device__
#include "fsgh_ctx.h"
#include "fsgh_lds.h"
#include "fsgh_physics.h"
#include "fsgh_rz_geom.h"

namespace fsgh {
namespace {

constexpr int kB = kBlock;

// the 4 edges of a quad as node lids (K/mesh.h:366-377)
const int kEdge[4][2] = {{0, 3}, {1, 2}, {0, 1}, {3, 2}};

__device__ __forceinline__ void load_quad(const DMesh& m, int e, int32_t nd[4])
{
    const int4 a = reinterpret_cast<const int4*>(m.conn)[e];
    nd[0] = a.x;
    nd[1] = a.y;
    nd[2] = a.z;
    nd[3] = a.w;
}

__device__ __forceinline__ void gather2(const double* __restrict__ a, const double* __restrict__ b,
                                        const int32_t nd[4], double x[4], double y[4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        x[i] = a[nd[i]];
        y[i] = b[nd[i]];
    }
}

__device__ __forceinline__ double rz_elem_dt(double dist_min, double sspd, double dt_cfl,
                                             double dt_max, double dt_min, double fuzz)
{
    double d = dt_cfl * dist_min / (sspd + fuzz);
    d = fmin(d, dt_max);
    d = fmax(d, dt_min);
    return d;
}

__device__ __forceinline__ double rz_block_min(double v)
{
    __shared__ double red[kB / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int i = 1; i < kB / 64; ++i) r = fmin(r, red[i]);
    return r;
}

// corner forces of one quad (K/force_sgh.cpp:407-808).  Strength-free
// (stress == nullptr): tau = -p I, so a_i . tau_d = B_id p.  With a stress
// field: tau = stress - p I and a_i . tau_d = a_i0 tau_0d + a_i1 tau_1d with
// a = -B, term by term.  Hoop terms tau_10 A_i / R and (tau_11 - tau_22)
// A_i / R at nodes off the axis (:755-769).  B (outward) and the area are
// returned for the velocity gradient.
__device__ __forceinline__ void rz_force_elem(const double x[4], const double y[4],
                                              const double u[4], const double v[4], double den,
                                              double pres, double sspd, double div, double ext,
                                              const Material& mt, double fuzz, double small,
                                              const double* stress, double B[8], double& area,
                                              double f[8])
{
    double A[4];
    rz::bmatrix(x, y, B);
    const double elem_area = rz::area(x, y);
    area = elem_area;
    rz::corner_areas(x, y, A);
    double vs0 = 0.0, vs1 = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        vs0 += 0.25 * u[a];
        vs1 += 0.25 * v[a];
    }
    const double qa = div < 0.0 ? mt.q1 : mt.q1ex;
    const double qb = div < 0.0 ? mt.q2 : mt.q2ex;
    double sum0 = 0.0, sum1 = 0.0, sum3 = 0.0, mu[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double mag_vel =
            sqrt((u[a] - vs0) * (u[a] - vs0) + (v[a] - vs1) * (v[a] - vs1));
        const double muc = den * (qa * sspd + qb * mag_vel);
        const double bx = B[2 * a], by = B[2 * a + 1];
        const double mu_term = muc * sqrt(bx * bx + by * by);
        sum0 += mu_term * u[a];
        sum1 += mu_term * v[a];
        sum3 += mu_term;
        mu[a] = mu_term;
    }
    if (sum3 > fuzz) {
        vs0 = sum0 / sum3;
        vs1 = sum1 / sum3;
    } else {
        vs0 = 0.0;
        vs1 = 0.0;
    }
    // shock detector: the neighbour extreme gives the smallest face ratio
    const double r_face = 0.9 * (ext + small) / (div + small);
    const double r_min = fmin(r_face, 1.0);
    double phi = 1.0 - fmax(0.0, r_min);
    const double c_length = sqrt(elem_area);
    const double alpha = fmin(1.0, 20.0 * (c_length * fabs(div)) / (sspd + fuzz));
    phi = alpha * phi;
    double tau00 = 0.0 - pres, tau01 = 0.0, tau10 = 0.0;
    double tau11 = 0.0 - pres, tau22 = 0.0 - pres;
    if (stress != nullptr) {
        tau00 = stress[0] - pres;
        tau01 = stress[1];
        tau10 = stress[3];
        tau11 = stress[4] - pres;
        tau22 = stress[8] - pres;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double pm = phi * mu[a];
        double f0, f1;
        if (stress != nullptr) {
            const double ax = (-1.0) * B[2 * a], ay = (-1.0) * B[2 * a + 1];
            f0 = ax * tau00 + ay * tau10 + pm * (vs0 - u[a]);
            f1 = ax * tau01 + ay * tau11 + pm * (vs1 - v[a]);
        } else {
            f0 = B[2 * a + 0] * pres + pm * (vs0 - u[a]);
            f1 = B[2 * a + 1] * pres + pm * (vs1 - v[a]);
        }
        const double node_radius = y[a];
        if (node_radius > 1e-14) {
            if (stress != nullptr) {
                f0 += tau10 * A[a] / node_radius;
                f1 += (tau11 - tau22) * A[a] / node_radius;
            } else {
                // tau10 and tau11 - tau22 are exactly zero without a stress
                // field, and 0 A / R = 0 (A / R) for every finite A / R (the
                // same signed zero, NaN alike): one division for both terms
                const double q = A[a] / node_radius;
                f0 += tau10 * q;
                f1 += (tau11 - tau22) * q;
            }
        }
        f[2 * a] = f0;
        f[2 * a + 1] = f1;
    }
}

// neighbour extreme of the shock detector from the per-node divergence pairs
// (see k_force: every other element sharing a node)
__device__ __forceinline__ double rz_neighbour_extreme(const DState& s, const int32_t nd[4],
                                                       double div, double small)
{
    const bool lo = div + small >= 0.0;
    const double2* __restrict__ xs = lo ? s.ext_lo : s.ext_hi;
    double ext = lo ? HUGE_VAL : -HUGE_VAL;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double2 t = xs[nd[a]];
        const double other = (t.x == div) ? t.y : t.x;
        ext = lo ? fmin(ext, other) : fmax(ext, other);
    }
    return ext;
}

// velocity gradient in RZ with the outward B (K/momentum.cpp:151-213)
__device__ __forceinline__ void rz_velgrad(const double B[8], const double u[4], const double v[4],
                                           double vol, double area, double L[9])
{
#pragma unroll
    for (int k = 0; k < 9; ++k) L[k] = 0.0;
    const double mean_radius = vol / area;
    const double elem_vel_r = 0.25 * (v[0] + v[1] + v[2] + v[3]);
    const double inverse_area = 1.0 / area;
    L[0] = (u[0] * B[0] + u[1] * B[2] + u[2] * B[4] + u[3] * B[6]) * inverse_area;
    L[1] = (u[0] * B[1] + u[1] * B[3] + u[2] * B[5] + u[3] * B[7]) * inverse_area;
    L[3] = (v[0] * B[0] + v[1] * B[2] + v[2] * B[4] + v[3] * B[6]) * inverse_area;
    L[4] = (v[0] * B[1] + v[1] * B[3] + v[2] * B[5] + v[3] * B[7]) * inverse_area;
    L[8] = elem_vel_r / mean_radius;
}

template <bool STRESS>
__global__ __launch_bounds__(kB) void k_rz_force(DMesh m, DState s, const InstCtl* ctl,
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
    int32_t nd[4];
    load_quad(m, e, nd);
    double x[4], y[4], u[4], v[4], f[8];
    gather2(s.x1, s.y1, nd, x, y);
    gather2(s.u1, s.v1, nd, u, v);
    const double div = s.div[e];
    const double ext = rz_neighbour_extreme(s, nd, div, small);
    const int32_t mat = s.nmat > 1 ? s.mat[e] : 0;
    const Material mt = s.mats[mat];
    double B[8], area, st[9];
    const size_t E = static_cast<size_t>(m.E);
    if (STRESS)
#pragma unroll
        for (int k = 0; k < 9; ++k) st[k] = s.sig1[static_cast<size_t>(k) * E + e];
    rz_force_elem(x, y, u, v, s.den[e], s.pres[e], s.sspd[e], div, ext, mt, fuzz, small,
                  STRESS ? st : nullptr, B, area, f);
    double2* out = reinterpret_cast<double2*>(s.F + 8 * static_cast<size_t>(e));
#pragma unroll
    for (int a = 0; a < 4; ++a) out[a] = make_double2(f[2 * a], f[2 * a + 1]);
    if (STRESS) {
        if (mt.strength == kStrengthHypo) {
            double L[9];
            rz_velgrad(B, u, v, s.vol[e], area, L);
            double p = s.pres[e], c = s.sspd[e], stn[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) stn[k] = s.sig0[static_cast<size_t>(k) * E + e];
            user_strength(mat, s.sparams + kStrengthParams * mat, s.den[e], s.sie0[e], L, s.vol[e],
                          dt, alpha, stn, st, p, c);
            s.pres[e] = p;
            s.sspd[e] = c;
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) s.sig1w[static_cast<size_t>(k) * E + e] = st[k];
    }
}

// Strength-free RZ corner forces with the block's nodes staged in LDS.
// Each workgroup of kB quads (a 16 x 16 Morton square: 289 distinct nodes)
// loads its distinct nodes' x, y, u, v and both divergence-extreme pairs
// once, consecutive lanes on consecutive list entries (lists from
// fsgh_stage.hip), then every quad reads its four nodes from LDS through its
// 16-bit local connectivity.  The list indices are requested first, then the
// quad's own inputs, then every node value, and only then is LDS written
// (vector loads retire in issue order).  The corner forces go through the
// same LDS buffer (row stride 5 x 16 B) and leave as contiguous 16-byte
// streaming stores.  A block whose list exceeds the capacity (e.g. a long
// row of a lexicographic order) gathers from global memory instead: a
// block-uniform branch with the same arithmetic, so the same bits.
constexpr int kRzStageCap = 2 * kB;   // distinct nodes per block held in LDS

__global__ __launch_bounds__(kB) void k_rz_force_staged(DMesh m, DState s, const InstCtl* ctl,
                                                        double fuzz, double small)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    if (ctl != nullptr) {
        if (ctl[bi.inst].stopped) return;
        fuzz = ctl[bi.inst].fuzz;
        small = ctl[bi.inst].small;
    }
    constexpr int CAP = kRzStageCap;
    constexpr int R = CAP / kB;
    // x, y, u, v as [4][CAP] doubles, then the extreme pairs as [2][CAP]
    // double2; reused for the corner-force rows
    constexpr int kWords = 4 * CAP / 2 + 2 * CAP > kB * 5 ? 4 * CAP / 2 + 2 * CAP : kB * 5;
    __shared__ double2 smem[kWords];
    double* sd = reinterpret_cast<double*>(smem);
    double2* sx = smem + 2 * CAP;
    const int e = bi.start + static_cast<int>(threadIdx.x);
    const bool live = e < bi.end;
    const int ec = live ? e : bi.start;   // in bounds; results of idle lanes unused
    int q0, nn;
    stage_range(m, bi, CAP, q0, nn);
    const bool staged = nn > 0 && nn <= CAP;
    uint32_t nid[R];
    double g[R][8];   // x, y, u, v, ext_lo pair, ext_hi pair
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * kB + static_cast<int>(threadIdx.x);
            nid[r] = static_cast<uint32_t>(m.sn_nodes[q0 + (i < nn ? i : nn - 1)]);
        }
    }
    const double div = s.div[ec];
    const double den = s.den[ec];
    const double pres = s.pres[ec];
    const double sspd = s.sspd[ec];
    uint2 lraw;
    int4 craw;
    if (staged) lraw = *reinterpret_cast<const uint2*>(m.sn_lconn + 4 * static_cast<size_t>(ec));
    else craw = reinterpret_cast<const int4*>(m.conn)[ec];
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (wave_in_list(r * kB, nn)) {
                const uint32_t n = nid[r];
                g[r][0] = s.x1[n];
                g[r][1] = s.y1[n];
                g[r][2] = s.u1[n];
                g[r][3] = s.v1[n];
                const double2 lo = s.ext_lo[n], hi = s.ext_hi[n];
                g[r][4] = lo.x;
                g[r][5] = lo.y;
                g[r][6] = hi.x;
                g[r][7] = hi.y;
            }
        }
    }
    const Material mt = s.nmat > 1 ? s.mats[s.mat[ec]] : s.mats[0];
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * kB + static_cast<int>(threadIdx.x);
            if (i < nn) {
#pragma unroll
                for (int k = 0; k < 4; ++k) sd[k * CAP + i] = g[r][k];
                sx[i] = make_double2(g[r][4], g[r][5]);
                sx[CAP + i] = make_double2(g[r][6], g[r][7]);
            }
        }
        __syncthreads();
    }
    double f[8];
    if (live) {
        int32_t nd[4];
        double x[4], y[4], u[4], v[4];
        if (staged) {
            nd[0] = static_cast<int32_t>(lraw.x & 0xffffu);
            nd[1] = static_cast<int32_t>(lraw.x >> 16);
            nd[2] = static_cast<int32_t>(lraw.y & 0xffffu);
            nd[3] = static_cast<int32_t>(lraw.y >> 16);
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                x[a] = sd[nd[a]];
                y[a] = sd[CAP + nd[a]];
                u[a] = sd[2 * CAP + nd[a]];
                v[a] = sd[3 * CAP + nd[a]];
            }
        } else {
            nd[0] = craw.x;
            nd[1] = craw.y;
            nd[2] = craw.z;
            nd[3] = craw.w;
            gather2(s.x1, s.y1, nd, x, y);
            gather2(s.u1, s.v1, nd, u, v);
        }
        // neighbour extreme of the shock detector (as rz_neighbour_extreme)
        const bool lo = div + small >= 0.0;
        double ext = lo ? HUGE_VAL : -HUGE_VAL;
        const double2* __restrict__ xs = staged ? (lo ? sx : sx + CAP) : (lo ? s.ext_lo : s.ext_hi);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const double2 t = xs[nd[a]];
            const double other = (t.x == div) ? t.y : t.x;
            ext = lo ? fmin(ext, other) : fmax(ext, other);
        }
        double B[8], area;
        rz_force_elem(x, y, u, v, den, pres, sspd, div, ext, mt, fuzz, small, nullptr, B, area, f);
    }
    if (staged) __syncthreads();   // every node read is done before smem is reused
    if (live) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
            smem[threadIdx.x * 5 + a] = make_double2(f[2 * a], f[2 * a + 1]);
    }
    __syncthreads();
    const int cnt = 4 * (bi.end - bi.start);
    double2* out = reinterpret_cast<double2*>(s.F + 8 * static_cast<size_t>(bi.start));
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int idx = r * kB + static_cast<int>(threadIdx.x);
        if (idx < cnt) stream_store(out + idx, smem[(idx >> 2) * 5 + (idx & 3)]);
    }
}

__global__ __launch_bounds__(kB) void k_rz_velgrad(DMesh m, DState s, double* out)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[4];
    load_quad(m, e, nd);
    double x[4], y[4], u[4], v[4], B[8], L[9];
    gather2(s.x1, s.y1, nd, x, y);
    gather2(s.u1, s.v1, nd, u, v);
    rz::bmatrix(x, y, B);
    rz_velgrad(B, u, v, s.vol[e], rz::area(x, y), L);
#pragma unroll
    for (int k = 0; k < 9; ++k) out[9 * static_cast<size_t>(e) + k] = L[k];
}

// corner -> node force sum in corners_in_node order (K/momentum.cpp:31-41)
__device__ __forceinline__ void rz_node_force(const DMesh& m, const double* __restrict__ F, int n,
                                              double& f0, double& f1)
{
    f0 = 0.0;
    f1 = 0.0;
    const int q0 = m.cin_start[n], q1 = m.cin_start[n + 1];
    for (int q = q0; q < q1; ++q) {
        const double2 t = reinterpret_cast<const double2*>(F)[m.cin[q]];
        f0 += t.x;
        f1 += t.y;
    }
}

__global__ __launch_bounds__(kB) void k_rz_update_velocity(DMesh m, DState s, double alpha,
                                                           double dt)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    double f0, f1;
    rz_node_force(m, s.F, n, f0, f1);
    const double mass = s.nmass[n];
    s.u1[n] = s.u0[n] + alpha * dt * f0 / mass;
    s.v1[n] = s.v0[n] + alpha * dt * f1 / mass;
}

// power with the node radius of the stage input (K/energy_sgh.cpp:16-52)
__device__ __forceinline__ double rz_power(const double f[8], const double u1[4],
                                           const double v1[4], const double u0[4],
                                           const double v0[4], const double r[4])
{
    double p = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double hx = (u1[a] + u0[a]) * 0.5;
        p += f[2 * a] * r[a] * hx;
        const double hy = (v1[a] + v0[a]) * 0.5;
        p += f[2 * a + 1] * r[a] * hy;
    }
    return p;
}

__device__ __forceinline__ void rz_load_forces(const DState& s, int e, double f[8])
{
    const double2* Fp = reinterpret_cast<const double2*>(s.F + 8 * static_cast<size_t>(e));
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double2 t = Fp[a];
        f[2 * a] = t.x;
        f[2 * a + 1] = t.y;
    }
}

__global__ __launch_bounds__(kB) void k_rz_update_energy(DMesh m, DState s, double alpha,
                                                         double dt)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[4];
    load_quad(m, e, nd);
    double f[8], u1[4], v1[4], u0[4], v0[4], r[4];
    rz_load_forces(s, e, f);
    gather2(s.u1, s.v1, nd, u1, v1);
    gather2(s.u0, s.v0, nd, u0, v0);
#pragma unroll
    for (int a = 0; a < 4; ++a) r[a] = s.y1[nd[a]];
    const double p = rz_power(f, u1, v1, u0, v0, r);
    s.sie1[e] = s.sie0[e] - alpha * dt / s.emass[e] * p;
}

__global__ __launch_bounds__(kB) void k_rz_get_vol(DMesh m, DState s)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[4];
    load_quad(m, e, nd);
    double x[4], y[4];
    gather2(s.x1, s.y1, nd, x, y);
    s.vol[e] = rz::volume(x, y);
}

__global__ __launch_bounds__(kB) void k_rz_get_divergence(DMesh m, DState s)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[4];
    load_quad(m, e, nd);
    double x[4], y[4], u[4], v[4], B[8];
    gather2(s.x1, s.y1, nd, x, y);
    gather2(s.u1, s.v1, nd, u, v);
    rz::bmatrix(x, y, B);
    s.div[e] = rz::divergence(B, u, v, rz::area(x, y), s.vol[e]);
}

__global__ __launch_bounds__(kB) void k_rz_dt_partial(DMesh m, DState s, const InstCtl* ctl,
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
        int32_t nd[4];
        load_quad(m, e, nd);
        double x[4], y[4];
        gather2(s.x1, s.y1, nd, x, y);
        d = rz_elem_dt(rz::min_pair_distance(x, y), s.sspd[e], cfl, dmax, dmin, fz);
    }
    d = rz_block_min(d);
    if (threadIdx.x == 0) partial[bi.lid] = d;
}

// node_extensive_mass = areal mass x radius (K/sgh_solve.cpp:107-116)
__global__ __launch_bounds__(kB) void k_rz_ext(DMesh m, DState s, double* ext)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    ext[n] = s.nmass[n] * s.y1[n];
}

// areal node mass from the current radius (K/sgh_solve.cpp:498-505)
__global__ __launch_bounds__(kB) void k_rz_node_mass(DMesh m, DState s, const double* ext,
                                                     double tiny)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    const double r = s.y1[n];
    double nm = 0.0;
    if (r > tiny) nm = ext[n] / r;
    s.nmass[n] = nm;
}

// on-axis boundary nodes take half the largest areal mass of their off-axis
// edge neighbours (K/sgh_solve.cpp:561-577).  The edge neighbours of node n
// are the two adjacent nodes of every element corner at n; fmax is
// order-free and repeats do not change it.
__global__ __launch_bounds__(kB) void k_rz_axis(DMesh m, DState s, const int2* list, int32_t count,
                                                const InstCtl* ctl, double tiny)
{
    const int i = blockIdx.x * kB + static_cast<int>(threadIdx.x);
    if (i >= count) return;
    const int2 it = list[i];
    const int n = it.x;
    if (ctl != nullptr) {
        if (ctl[it.y].stopped) return;
        tiny = ctl[it.y].tiny;
    }
    if (!(s.y1[n] < tiny)) return;
    double nm = s.nmass[n];
    for (int q = m.cin_start[n]; q < m.cin_start[n + 1]; ++q) {
        const int cid = m.cin[q];
        const int e = cid >> 2, a = cid & 3;
        const int4 c4 = reinterpret_cast<const int4*>(m.conn)[e];
        const int nd[4] = {c4.x, c4.y, c4.z, c4.w};
        const int nb0 = nd[(a + 1) & 3], nb1 = nd[(a + 3) & 3];
        if (s.y1[nb0] > tiny) nm = fmax(nm, s.nmass[nb0] / 2.0);
        if (s.y1[nb1] > tiny) nm = fmax(nm, s.nmass[nb1] / 2.0);
    }
    s.nmass[n] = nm;
}

// corners of a node block staged in LDS by the RZ node kernel: 6 per node
// (4 on a structured quad mesh); blocks with more gather from global memory
constexpr int kRzCornerCap = 6 * kB;

// fused stage, node part: velocity (K/momentum.cpp:23-49), boundary
// (K/boundary.cpp:16-55), position (K/geometry.cpp:17-24), the stage-input
// radius for the energy kernel and the RZ areal mass of the new position.
// The block's corner forces are loaded into LDS in ascending corner order
// (its sorted corner list: consecutive lanes read consecutive 16-byte
// pieces of few elements' rows) and each node sums its own in
// corners_in_node order, as rz_node_force does from global memory.
__global__ __launch_bounds__(kB) void k_rz_node(DMesh m, DState s, const InstCtl* ctl,
                                                double alpha, const double* rsrc, double* rin,
                                                const double* ext)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const InstCtl& c = ctl[bi.inst];
    if (c.stopped) return;
    __shared__ double2 sF[kRzCornerCap];
    const int qb = m.cin_start[bi.start];
    const int nc = m.cin_start[bi.end] - qb;
    const bool staged = m.cs_cor != nullptr && nc <= kRzCornerCap;   // block-uniform
    const double2* __restrict__ F2 = reinterpret_cast<const double2*>(s.F);
    if (staged) {
        constexpr int R = kRzCornerCap / kB;
        int cid[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = r * kB + static_cast<int>(threadIdx.x);
            cid[r] = k < nc ? m.cs_cor[qb + k] : -1;
        }
        double2 g[R];
#pragma unroll
        for (int r = 0; r < R; ++r) g[r] = F2[cid[r] < 0 ? 0 : cid[r]];
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (cid[r] >= 0) sF[r * kB + static_cast<int>(threadIdx.x)] = g[r];
        __syncthreads();
    }
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    const double dt = c.dt;
    double f0, f1;
    if (staged) {
        f0 = 0.0;
        f1 = 0.0;
        const int q1 = m.cin_start[n + 1];
        for (int q = m.cin_start[n]; q < q1; ++q) {
            const double2 t = sF[m.cs_pos[q]];
            f0 += t.x;
            f1 += t.y;
        }
    } else {
        rz_node_force(m, s.F, n, f0, f1);
    }
    const double mass = s.nmass[n];
    const uint8_t mk = s.bcmask[n];
    const double u0 = s.u0[n], v0 = s.v0[n];
    double un = u0 + alpha * dt * f0 / mass;
    double vn = v0 + alpha * dt * f1 / mass;
    if (mk & 1u) un = 0.0;
    if (mk & 2u) vn = 0.0;
    const double r_in = rsrc[n];
    const double xn = s.x0[n] + alpha * dt * ((un + u0) * 0.5);
    const double yn = s.y0[n] + alpha * dt * ((vn + v0) * 0.5);
    s.u1[n] = un;
    s.v1[n] = vn;
    s.x1[n] = xn;
    s.y1[n] = yn;
    rin[n] = r_in;
    double nm = 0.0;
    if (yn > c.tiny) nm = ext[n] / yn;
    s.nmass[n] = nm;
}

// fused stage, element part: energy (K/energy_sgh.cpp:16-52), RZ volume,
// density and EOS (K/properties.cpp:118-234), the next stage's divergence
// (K/momentum.cpp:298-378) and, on the second stage, the next cycle's dt
// candidates (K/time_integration.cpp:161-246)
__global__ __launch_bounds__(kB) void k_rz_energy_state_div(DMesh m, DState s, const InstCtl* ctl,
                                                            double alpha, const double* rin,
                                                            double* dt_partial)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const InstCtl& c = ctl[bi.inst];
    if (c.stopped) return;
    const int e = bi.start + static_cast<int>(threadIdx.x);
    double dtl = DBL_MAX;
    if (e < bi.end) {
        int32_t nd[4];
        load_quad(m, e, nd);
        double f[8], u1[4], v1[4], u0[4], v0[4], r[4], x[4], y[4], B[8];
        rz_load_forces(s, e, f);
        gather2(s.u1, s.v1, nd, u1, v1);
        gather2(s.u0, s.v0, nd, u0, v0);
#pragma unroll
        for (int a = 0; a < 4; ++a) r[a] = rin[nd[a]];
        const double emass = s.emass[e];
        const double sie = s.sie0[e] - alpha * c.dt / emass * rz_power(f, u1, v1, u0, v0, r);
        s.sie1[e] = sie;
        gather2(s.x1, s.y1, nd, x, y);
        const double vol = rz::volume(x, y);
        const double den = emass / vol;
        double p, cs, st[9];
        const double* stp = nullptr;
        if (s.sig1 != nullptr) {
            const size_t E = static_cast<size_t>(m.E);
#pragma unroll
            for (int k = 0; k < 9; ++k) st[k] = s.sig1[static_cast<size_t>(k) * E + e];
            stp = st;
        }
        eos(s.mats[s.nmat > 1 ? s.mat[e] : 0].eos, s.gamma[e], s.csmin[e], den, sie, p, cs, stp);
        rz::bmatrix(x, y, B);
        const double div = rz::divergence(B, u1, v1, rz::area(x, y), vol);
        s.vol[e] = vol;
        s.den[e] = den;
        s.pres[e] = p;
        s.sspd[e] = cs;
        s.div_out[e] = div;
        if (dt_partial != nullptr)
            dtl = rz_elem_dt(rz::min_pair_distance(x, y), cs, c.dt_cfl, c.dt_max, c.dt_min, c.fuzz);
    }
    if (dt_partial != nullptr) {
        dtl = rz_block_min(dtl);
        if (threadIdx.x == 0) dt_partial[bi.lid] = dtl;
    }
}

// The same fused element pass for a context without a stress field, with the
// block's nodes staged in LDS (lists as in k_rz_force_staged): per distinct
// node the stage-output coordinates and velocities, the mean stage velocity
// (v1 + v0) * 0.5 of each component (formed once per node, the same
// operations the power takes per corner) and the stage-input radius.  The
// list indices are requested first, then the quad's force row and inputs,
// then the node values; LDS is written after every load is issued.  Blocks
// over the capacity gather from global memory (block-uniform, same bits).
__global__ __launch_bounds__(kB) void k_rz_energy_state_div_staged(DMesh m, DState s,
                                                                   const InstCtl* ctl, double alpha,
                                                                   const double* rin,
                                                                   double* dt_partial)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const InstCtl& c = ctl[bi.inst];
    if (c.stopped) return;
    constexpr int CAP = kRzStageCap;
    constexpr int R = CAP / kB;
    constexpr int NV = 7;   // x1, y1, u1, v1, mean u, mean v, radius
    __shared__ double sd[NV * CAP];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    const bool live = e < bi.end;
    const int ec = live ? e : bi.start;
    int q0, nn;
    stage_range(m, bi, CAP, q0, nn);
    const bool staged = nn > 0 && nn <= CAP;
    uint32_t nid[R];
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * kB + static_cast<int>(threadIdx.x);
            nid[r] = static_cast<uint32_t>(m.sn_nodes[q0 + (i < nn ? i : nn - 1)]);
        }
    }
    double f[8];
    rz_load_forces(s, ec, f);
    const double emass = s.emass[ec];
    const double sie0 = s.sie0[ec];
    const double gamma = s.gamma[ec];
    const double csmin = s.csmin[ec];
    uint2 lraw;
    int4 craw;
    if (staged) lraw = *reinterpret_cast<const uint2*>(m.sn_lconn + 4 * static_cast<size_t>(ec));
    else craw = reinterpret_cast<const int4*>(m.conn)[ec];
    if (staged) {
        double g[R][NV];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (wave_in_list(r * kB, nn)) {
                const uint32_t n = nid[r];
                g[r][0] = s.x1[n];
                g[r][1] = s.y1[n];
                g[r][2] = s.u1[n];
                g[r][3] = s.v1[n];
                g[r][4] = s.u0[n];
                g[r][5] = s.v0[n];
                g[r][6] = rin[n];
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = r * kB + static_cast<int>(threadIdx.x);
            if (i < nn) {
                sd[i] = g[r][0];
                sd[CAP + i] = g[r][1];
                sd[2 * CAP + i] = g[r][2];
                sd[3 * CAP + i] = g[r][3];
                sd[4 * CAP + i] = (g[r][2] + g[r][4]) * 0.5;
                sd[5 * CAP + i] = (g[r][3] + g[r][5]) * 0.5;
                sd[6 * CAP + i] = g[r][6];
            }
        }
        __syncthreads();
    }
    double dtl = DBL_MAX;
    if (live) {
        double x[4], y[4], u1[4], v1[4], hu[4], hv[4], r[4], B[8];
        if (staged) {
            int32_t nd[4];
            nd[0] = static_cast<int32_t>(lraw.x & 0xffffu);
            nd[1] = static_cast<int32_t>(lraw.x >> 16);
            nd[2] = static_cast<int32_t>(lraw.y & 0xffffu);
            nd[3] = static_cast<int32_t>(lraw.y >> 16);
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                x[a] = sd[nd[a]];
                y[a] = sd[CAP + nd[a]];
                u1[a] = sd[2 * CAP + nd[a]];
                v1[a] = sd[3 * CAP + nd[a]];
                hu[a] = sd[4 * CAP + nd[a]];
                hv[a] = sd[5 * CAP + nd[a]];
                r[a] = sd[6 * CAP + nd[a]];
            }
        } else {
            const int32_t nd[4] = {craw.x, craw.y, craw.z, craw.w};
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                x[a] = s.x1[nd[a]];
                y[a] = s.y1[nd[a]];
                u1[a] = s.u1[nd[a]];
                v1[a] = s.v1[nd[a]];
                hu[a] = (u1[a] + s.u0[nd[a]]) * 0.5;
                hv[a] = (v1[a] + s.v0[nd[a]]) * 0.5;
                r[a] = rin[nd[a]];
            }
        }
        // power in rz_power's order: per corner the axial then the radial term
        double pw = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            pw += f[2 * a] * r[a] * hu[a];
            pw += f[2 * a + 1] * r[a] * hv[a];
        }
        const double sie = sie0 - alpha * c.dt / emass * pw;
        const double vol = rz::volume(x, y);
        const double den = emass / vol;
        double p, cs;
        eos(s.mats[s.nmat > 1 ? s.mat[ec] : 0].eos, gamma, csmin, den, sie, p, cs);
        rz::bmatrix(x, y, B);
        const double div = rz::divergence(B, u1, v1, rz::area(x, y), vol);
        s.sie1[e] = sie;
        s.vol[e] = vol;
        s.den[e] = den;
        s.pres[e] = p;
        s.sspd[e] = cs;
        s.div_out[e] = div;
        if (dt_partial != nullptr)
            dtl = rz_elem_dt(rz::min_pair_distance(x, y), cs, c.dt_cfl, c.dt_max, c.dt_min, c.fuzz);
    }
    if (dt_partial != nullptr) {
        dtl = rz_block_min(dtl);
        if (threadIdx.x == 0) dt_partial[bi.lid] = dtl;
    }
}

unsigned grid_of(int32_t n) { return static_cast<unsigned>((n + kB - 1) / kB); }

}  // namespace

// ---------------------------------------------------------------- launchers --
void launch_rz_get_vol(const DMesh& m, const DState& s, hipStream_t st)
{
    if (m.n_eblk) hipLaunchKernelGGL(k_rz_get_vol, dim3(m.n_eblk), dim3(kB), 0, st, m, s);
}
void launch_rz_get_divergence(const DMesh& m, const DState& s, hipStream_t st)
{
    if (m.n_eblk) hipLaunchKernelGGL(k_rz_get_divergence, dim3(m.n_eblk), dim3(kB), 0, st, m, s);
}
void launch_rz_force(const DMesh& m, const DState& s, const InstCtl* ctl, double fuzz,
                     double small, double dt, double alpha, hipStream_t st)
{
    if (!m.n_eblk) return;
    if (s.sig1 != nullptr)
        hipLaunchKernelGGL(k_rz_force<true>, dim3(m.n_eblk), dim3(kB), 0, st, m, s, ctl, fuzz, small,
                           dt, alpha);
    else if (m.bs == kB)
        hipLaunchKernelGGL(k_rz_force_staged, dim3(m.n_eblk), dim3(kB), 0, st, m, s, ctl, fuzz,
                           small);
    else
        hipLaunchKernelGGL(k_rz_force<false>, dim3(m.n_eblk), dim3(kB), 0, st, m, s, ctl, fuzz,
                           small, dt, alpha);
}
void launch_rz_velgrad(const DMesh& m, const DState& s, double* out, hipStream_t st)
{
    if (m.n_eblk) hipLaunchKernelGGL(k_rz_velgrad, dim3(m.n_eblk), dim3(kB), 0, st, m, s, out);
}
void launch_rz_update_velocity(const DMesh& m, const DState& s, double alpha, double dt,
                               hipStream_t st)
{
    if (m.n_nblk)
        hipLaunchKernelGGL(k_rz_update_velocity, dim3(m.n_nblk), dim3(kB), 0, st, m, s, alpha, dt);
}
void launch_rz_update_energy(const DMesh& m, const DState& s, double alpha, double dt,
                             hipStream_t st)
{
    if (m.n_eblk)
        hipLaunchKernelGGL(k_rz_update_energy, dim3(m.n_eblk), dim3(kB), 0, st, m, s, alpha, dt);
}
void launch_rz_dt_partial(const DMesh& m, const DState& s, const InstCtl* ctl, double* partial,
                          double dt_cfl, double dt_max, double dt_min, double fuzz, int use_ctl,
                          hipStream_t st)
{
    if (m.n_eblk)
        hipLaunchKernelGGL(k_rz_dt_partial, dim3(m.n_eblk), dim3(kB), 0, st, m, s, ctl, partial,
                           dt_cfl, dt_max, dt_min, fuzz, use_ctl);
}
void launch_rz_ext(const DMesh& m, const DState& s, double* ext, hipStream_t st)
{
    if (m.n_nblk) hipLaunchKernelGGL(k_rz_ext, dim3(m.n_nblk), dim3(kB), 0, st, m, s, ext);
}
void launch_rz_node_mass(const DMesh& m, const DState& s, const double* ext, double tiny,
                         const int2* axis, int32_t naxis, hipStream_t st)
{
    if (m.n_nblk)
        hipLaunchKernelGGL(k_rz_node_mass, dim3(m.n_nblk), dim3(kB), 0, st, m, s, ext, tiny);
    if (naxis > 0)
        hipLaunchKernelGGL(k_rz_axis, dim3(grid_of(naxis)), dim3(kB), 0, st, m, s, axis, naxis,
                           static_cast<const InstCtl*>(nullptr), tiny);
}
void launch_rz_node(const DMesh& m, const DState& s, const InstCtl* ctl, double alpha,
                    const double* rsrc, double* rin, const double* ext, const int2* axis,
                    int32_t naxis, hipStream_t st)
{
    if (m.n_nblk)
        hipLaunchKernelGGL(k_rz_node, dim3(m.n_nblk), dim3(kB), 0, st, m, s, ctl, alpha, rsrc, rin,
                           ext);
    if (naxis > 0)
        hipLaunchKernelGGL(k_rz_axis, dim3(grid_of(naxis)), dim3(kB), 0, st, m, s, axis, naxis, ctl,
                           0.0);
}
void launch_rz_energy_state_div(const DMesh& m, const DState& s, const InstCtl* ctl, double alpha,
                                const double* rin, double* dt_partial, hipStream_t st)
{
    if (!m.n_eblk) return;
    if (s.sig1 == nullptr && m.bs == kB)
        hipLaunchKernelGGL(k_rz_energy_state_div_staged, dim3(m.n_eblk), dim3(kB), 0, st, m, s, ctl,
                           alpha, rin, dt_partial);
    else
        hipLaunchKernelGGL(k_rz_energy_state_div, dim3(m.n_eblk), dim3(kB), 0, st, m, s, ctl, alpha,
                           rin, dt_partial);
}

// ------------------------------------------------------------ host set-up --
namespace {

template <typename T>
int up(fsgh_ctx* c, T* dst, const T* src, size_t n)
{
    HIPCHK(c, hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FSGH_OK;
}

}  // namespace

// Numbering, internal connectivity, corner CSR and the on-axis candidate
// list of a 2D context, on the host (the reference's serial set-up,
// K/mesh.h:180-649), then uploaded.
int rz_build_maps(fsgh_ctx* c)
{
    const int64_t N = c->N, E = c->E;
    if (c->renumber) {
        morton_renumber(c->L, c->conn.data(), c->coords.data(), c->ne_of_oe, c->nn_of_on, 4);
    } else {
        c->ne_of_oe.resize(static_cast<size_t>(E));
        c->nn_of_on.resize(static_cast<size_t>(N));
        std::iota(c->ne_of_oe.begin(), c->ne_of_oe.end(), 0);
        std::iota(c->nn_of_on.begin(), c->nn_of_on.end(), 0);
    }
    c->oe_of_ne.resize(static_cast<size_t>(E));
    c->on_of_nn.resize(static_cast<size_t>(N));
    for (int64_t e = 0; e < E; ++e) c->oe_of_ne[static_cast<size_t>(c->ne_of_oe[e])] = static_cast<int32_t>(e);
    for (int64_t n = 0; n < N; ++n) c->on_of_nn[static_cast<size_t>(c->nn_of_on[n])] = static_cast<int32_t>(n);
    corner_csr(N, E, c->conn.data(), c->cstart, c->corners, 4);
    std::vector<int32_t> iconn(static_cast<size_t>(4 * E)), istart(static_cast<size_t>(N + 1)),
        icin(static_cast<size_t>(4 * E));
    for (int64_t oe = 0; oe < E; ++oe)
        for (int a = 0; a < 4; ++a)
            iconn[static_cast<size_t>(4 * c->ne_of_oe[oe] + a)] =
                c->nn_of_on[static_cast<size_t>(c->conn[static_cast<size_t>(4 * oe + a)])];
    // each internal node keeps its caller node's corner list (ascending caller element)
    int64_t pos = 0;
    for (int64_t k = 0; k < N; ++k) {
        istart[static_cast<size_t>(k)] = static_cast<int32_t>(pos);
        const int64_t on = c->on_of_nn[static_cast<size_t>(k)];
        for (int64_t q = c->cstart[static_cast<size_t>(on)]; q < c->cstart[static_cast<size_t>(on + 1)]; ++q) {
            const int64_t oc = c->corners[static_cast<size_t>(q)];
            icin[static_cast<size_t>(pos++)] = static_cast<int32_t>(4 * c->ne_of_oe[static_cast<size_t>(oc / 4)] + oc % 4);
        }
    }
    istart[static_cast<size_t>(N)] = static_cast<int32_t>(pos);
    // boundary nodes of the mesh: both ends of every edge that no other element holds
    std::vector<int2> axis;
    std::vector<char> mark(static_cast<size_t>(N), 0);
    for (int32_t i = 0; i < c->L.ninst; ++i) {
        for (int64_t e = c->L.elem_off[i]; e < c->L.elem_off[i + 1]; ++e)
            for (int f = 0; f < 4; ++f) {
                const int64_t a = c->conn[static_cast<size_t>(4 * e + kEdge[f][0])];
                const int64_t b = c->conn[static_cast<size_t>(4 * e + kEdge[f][1])];
                bool shared = false;
                for (int64_t q = c->cstart[static_cast<size_t>(a)]; q < c->cstart[static_cast<size_t>(a + 1)] && !shared; ++q) {
                    const int64_t e2 = c->corners[static_cast<size_t>(q)] / 4;
                    if (e2 == e) continue;
                    for (int k = 0; k < 4; ++k)
                        if (c->conn[static_cast<size_t>(4 * e2 + k)] == b) shared = true;
                }
                if (!shared) {
                    mark[static_cast<size_t>(a)] = 1;
                    mark[static_cast<size_t>(b)] = 1;
                }
            }
        for (int64_t n = c->L.node_off[i]; n < c->L.node_off[i + 1]; ++n)
            if (mark[static_cast<size_t>(n)]) axis.push_back(make_int2(c->nn_of_on[static_cast<size_t>(n)], i));
    }
    std::sort(axis.begin(), axis.end(), [](const int2& p, const int2& q) { return p.x < q.x; });
    c->n_axis = static_cast<int32_t>(axis.size());
    if (hipMalloc(reinterpret_cast<void**>(&c->d_axis), std::max<size_t>(axis.size(), 1) * sizeof(int2)) != hipSuccess)
        return FSGH_E_NOMEM;
    int rc = FSGH_OK;
    rc |= up(c, c->d_conn, iconn.data(), iconn.size());
    rc |= up(c, c->d_cin_start, istart.data(), istart.size());
    rc |= up(c, c->d_cin, icin.data(), icin.size());
    if (!axis.empty()) rc |= up(c, c->d_axis, axis.data(), axis.size());
    if (c->d_on_of_nn) rc |= up(c, c->d_on_of_nn, c->on_of_nn.data(), c->on_of_nn.size());
    if (c->d_oe_of_ne) rc |= up(c, c->d_oe_of_ne, c->oe_of_ne.data(), c->oe_of_ne.size());
    return rc ? FSGH_E_HIP : FSGH_OK;
}

// K/setup.cpp:106-424 for one instance of a 2D context, on the host in
// caller numbering: RZ volume, fills (centroid of the 4 nodes), EOS,
// velocities (cartesian, radial), boundary velocity, corner-area node mass.
int rz_setup_fills(fsgh_ctx* c, int32_t inst, int32_t nfills, const fsgh_fill* fills,
                   int32_t nmat, int32_t nsv, const double* state_vars, const int32_t* mat_nsv)
{
    (void)nmat;
    for (int32_t f = 0; f < nfills; ++f)
        if (fills[f].velocity == 2) {
            set_error(c, "spherical velocity fill needs a third coordinate (2D context)");
            return FSGH_E_UNSUPPORTED;
        }
    const int64_t e0 = c->L.elem_off[inst], e1 = c->L.elem_off[inst + 1];
    const int64_t n0 = c->L.node_off[inst], n1 = c->L.node_off[inst + 1];
    const size_t NE = static_cast<size_t>(e1 - e0), NN = static_cast<size_t>(n1 - n0);
    std::vector<double> vol(NE), den(NE), mass(NE), sie(NE), pres(NE), sspd(NE), gam(NE), csm(NE);
    std::vector<double> sv4(NE, 0.0);
    std::vector<int32_t> mat(NE);
    std::vector<double> vel(2 * NN, 0.0), nmass(NN, 0.0);
    const double* X = c->coords.data();   // N x 3, z = 0
    auto quad = [&](int64_t e, double x[4], double y[4]) {
        for (int a = 0; a < 4; ++a) {
            const int64_t n = c->conn[static_cast<size_t>(4 * e + a)];
            x[a] = X[3 * n];
            y[a] = X[3 * n + 1];
        }
    };
    for (int64_t e = e0; e < e1; ++e) {
        double x[4], y[4];
        quad(e, x, y);
        vol[static_cast<size_t>(e - e0)] = rz::volume(x, y);
    }
    for (int32_t f = 0; f < nfills; ++f) {
        const fsgh_fill& F = fills[f];
        for (int64_t e = e0; e < e1; ++e) {
            const size_t k = static_cast<size_t>(e - e0);
            double ec[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < 4; ++a) {
                const int64_t n = c->conn[static_cast<size_t>(4 * e + a)];
                ec[0] += X[3 * n];
                ec[1] += X[3 * n + 1];
                ec[2] = 0.0;
            }
            ec[0] = ec[0] / 4;
            ec[1] = ec[1] / 4;
            ec[2] = ec[2] / 4;
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
            eos(c->mats[static_cast<size_t>(m)].eos, gam[k], csm[k], den[k], sie[k], pres[k],
                sspd[k]);
            if (F.velocity == 5) tg_vortex_state(ec, F.den, sv4[k], pres[k], sie[k]);
            for (int a = 0; a < 4; ++a) {
                const int64_t n = c->conn[static_cast<size_t>(4 * e + a)];
                double* v = &vel[2 * static_cast<size_t>(n - n0)];
                if (F.velocity == 0) {
                    v[0] = F.u;
                    v[1] = F.v;
                } else if (F.velocity == 5) {
                    tg_vortex_velocity(X + 3 * n, false, v);
                } else if (F.velocity == 1) {
                    double dir[2] = {0.0, 0.0}, rv = 0.0;
                    for (int d = 0; d < 2; ++d) {
                        dir[d] = X[3 * n + d];
                        rv += X[3 * n + d] * X[3 * n + d];
                    }
                    rv = std::sqrt(rv);
                    for (int d = 0; d < 2; ++d) dir[d] = rv > 1.0e-14 ? dir[d] / rv : 0.0;
                    v[0] = F.speed * dir[0];
                    v[1] = F.speed * dir[1];
                }
            }
        }
    }
    // boundary velocity (K/setup.cpp:362), corner masses and the areal node
    // mass (K/setup.cpp:365-420)
    std::vector<double> cmass(4 * NE);
    for (int64_t e = e0; e < e1; ++e) {
        double x[4], y[4], A[4];
        quad(e, x, y);
        rz::corner_areas(x, y, A);
        for (int a = 0; a < 4; ++a)
            cmass[4 * static_cast<size_t>(e - e0) + static_cast<size_t>(a)] = A[a] * den[static_cast<size_t>(e - e0)];
    }
    for (int64_t n = n0; n < n1; ++n) {
        const uint8_t mk = c->bcmask[static_cast<size_t>(n)];
        for (int d = 0; d < 2; ++d)
            if (mk & (1u << d)) vel[2 * static_cast<size_t>(n - n0) + d] = 0.0;
        double acc = 0.0;
        for (int64_t q = c->cstart[static_cast<size_t>(n)]; q < c->cstart[static_cast<size_t>(n + 1)]; ++q)
            acc += cmass[static_cast<size_t>(c->corners[static_cast<size_t>(q)] - 4 * e0)];
        nmass[static_cast<size_t>(n - n0)] = acc;
    }
    auto upe = [&](double* dst, const std::vector<double>& src) -> int {
        std::vector<double> t(NE);
        for (int64_t e = e0; e < e1; ++e) t[static_cast<size_t>(c->ne_of_oe[e] - e0)] = src[static_cast<size_t>(e - e0)];
        return up(c, dst + e0, t.data(), NE);
    };
    auto upn = [&](double* dst, const std::vector<double>& src, int stride, int comp) -> int {
        std::vector<double> t(NN);
        for (int64_t n = n0; n < n1; ++n)
            t[static_cast<size_t>(c->nn_of_on[n] - n0)] = src[static_cast<size_t>(n - n0) * stride + comp];
        return up(c, dst + n0, t.data(), NN);
    };
    int rc = 0;
    const std::vector<double> zeroE(NE, 0.0), zeroN(2 * NN, 0.0);
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
    for (int d = 0; d < 2; ++d) {
        rc |= upn(c->d_node[b1][3 + d], vel, 2, d);
        rc |= upn(c->d_node[b0][3 + d], zeroN, 2, d);
    }
    rc |= upn(c->d_nmass, nmass, 1, 0);
    {
        std::vector<int32_t> t(NE);
        for (int64_t e = e0; e < e1; ++e) t[static_cast<size_t>(c->ne_of_oe[e] - e0)] = mat[static_cast<size_t>(e - e0)];
        rc |= up(c, c->d_mat + e0, t.data(), NE);
    }
    return rc ? FSGH_E_HIP : FSGH_OK;
}

}  // namespace fsgh
