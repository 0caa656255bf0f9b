// fsgh_adjoint.hip -- discrete adjoint and tangent-linear model of the SGH cycle.
//
// Model: the fused RK2 cycle with the time steps frozen to the ones recorded
// by the forward run (the CFL minimum, its x1.1 growth and the clamps are not
// differentiated).  Controls: initial element density rho0 (through the
// element masses M = rho0 V0, the node masses m = sum M/8 and the setup
// density/pressure/sound speed), initial specific internal energy e0, the
// per-element gamma, and the initial node velocity.
//
// Reverse pass, per RK stage (alpha, dt), hand-derived from the forward
// operators in fsgh_physics.h:
//   k_adj_energy   e1 = e0 - alpha dt/M sum_c F_c.(v1+v0)/2: e0-bar, M-bar and
//                  the element scalar g = -alpha dt/M e1-bar
//   k_adj_node     v1 = mask(v0 + alpha dt sum F/m), x1 = x0 + alpha dt (v1+v0)/2,
//                  with the energy's corner terms g (v1+v0)/2 and g F_c formed at
//                  the node (CSR order)                  -> x0-bar, v0-bar, F-bar, m-bar
//   k_arg_extremes per node, the two smallest / largest divergences with the
//                  elements holding them
//   k_adj_force_pair<true>  the shock detector's neighbour-extreme adjoint ext-bar of
//                  each element and the node slot its extreme came from
//   k_adj_ext_node per node and extreme slot, the ext-adjoints of the elements
//                  that chose it (each element then adds the sums of the slots
//                  it occupies to its divergence adjoint, in k_adj_force_pair<false>)
//   k_adj_force_pair<false> corner forces (MARS + shock detector), then vol / den /
//                  EOS / div of the stage input -> e-bar, M-bar, gamma-bar,
//                  corner x-bar, v-bar
//   k_adj_gather   corner adjoints -> node adjoints, in corner order (no atomics)
// Branches (compression test, min/max, clamps, the sound-speed floor) follow
// the forward values; ties of the neighbour extreme go to the first element
// found, the same choice the tangent-linear model makes.
//
// Checkpointing: the state (x, v, e) is stored every `checkpoint_every`
// cycles.  The reverse pass recomputes one segment at a time, keeping every
// cycle's intermediates (stage-0 output, both stages' corner forces and
// input element states) in per-cycle slots written in place by the forward
// kernels, then reverses the segment's cycles from those slots.  Device
// memory of the reverse pass: (K + 1)(6N + 6E) + K(6N + 54E) doubles for
// K = checkpoint_every, plus (ncycles / K)(6N + E) for the checkpoints.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#define FSGH_HD __host__ __device__
#include "fsgh_ctx.h"
#include "fsgh_physics.h"

using namespace fsgh;

namespace {

constexpr int kB = kBlock;

__device__ __forceinline__ void ld_conn(const int32_t* conn, int e, int32_t nd[8])
{
    const int4* p = reinterpret_cast<const int4*>(conn + 8 * static_cast<size_t>(e));
    const int4 a = p[0], b = p[1];
    nd[0] = a.x; nd[1] = a.y; nd[2] = a.z; nd[3] = a.w;
    nd[4] = b.x; nd[5] = b.y; nd[6] = b.z; nd[7] = b.w;
}

template <typename T>
__device__ __forceinline__ void g3(const double* a, const double* b, const double* c,
                                   const int32_t nd[8], T x[8], T y[8], T z[8])
{
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        x[i] = a[nd[i]];
        y[i] = b[nd[i]];
        z[i] = c[nd[i]];
    }
}

// (value, tangent) pair of device arrays
struct TA {
    double* v;
    double* t;
    __device__ Dual ld(size_t i) const { return Dual(v[i], t[i]); }
    __device__ void st(size_t i, Dual a) const
    {
        v[i] = a.v;
        t[i] = a.d;
    }
};

// -------------------------------------------------------------------------
// divergence extremes with the element that holds them (node-centric)
// -------------------------------------------------------------------------
// insert (v, e) into the two smallest / largest so far; ties keep the
// earlier entry (strict comparisons), as a sequential scan would
__device__ __forceinline__ void ins_lo(double& a, double& b, int& ia, int& ib, double v, int e)
{
    const bool la = v < a, lb = v < b;
    b = la ? a : (lb ? v : b);
    ib = la ? ia : (lb ? e : ib);
    a = la ? v : a;
    ia = la ? e : ia;
}
__device__ __forceinline__ void ins_hi(double& a, double& b, int& ia, int& ib, double v, int e)
{
    const bool ga = v > a, gb = v > b;
    b = ga ? a : (gb ? v : b);
    ib = ga ? ia : (gb ? e : ib);
    a = ga ? v : a;
    ia = ga ? e : ia;
}

// The first eight corner ids of node n's CSR list, with indices clamped to
// the list (to entry 0 of the array for an empty list) so every load is in
// bounds and none waits on a branch; returns the list length.  Callers use
// entry j only when j < length.
__device__ __forceinline__ int corners8(const DMesh& m, int n, int q0, int c8[8])
{
    const int cnt = m.cin_start[n + 1] - q0;
    const int base = cnt > 0 ? q0 : 0;
    const int last = cnt > 0 ? cnt - 1 : 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) c8[j] = m.cin[base + (j < last ? j : last)];
    return cnt;
}

// div: element divergence values; dtan (may be null): their tangents
__global__ __launch_bounds__(kB) void k_arg_extremes(DMesh m, const double* div, const double* dtan,
                                                     double2* lo, double2* hi, int2* alo,
                                                     int2* ahi, double2* tlo, double2* thi)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    double a1 = HUGE_VAL, a2 = HUGE_VAL, b1 = -HUGE_VAL, b2 = -HUGE_VAL;
    int ia1 = -1, ia2 = -1, ib1 = -1, ib2 = -1;
    const int q0 = m.cin_start[n];
    int c8[8];
    const int cnt = corners8(m, n, q0, c8);
    if (cnt > 0) {
        double d[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) d[j] = div[c8[j] >> 3];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool ok = j < cnt && !isnan(d[j]);
            const int e = c8[j] >> 3;
            ins_lo(a1, a2, ia1, ia2, ok ? d[j] : HUGE_VAL, e);
            ins_hi(b1, b2, ib1, ib2, ok ? d[j] : -HUGE_VAL, e);
        }
        for (int q = q0 + 8; q < q0 + cnt; ++q) {
            const int e = m.cin[q] >> 3;
            const double v = div[e];
            if (isnan(v)) continue;
            ins_lo(a1, a2, ia1, ia2, v, e);
            ins_hi(b1, b2, ib1, ib2, v, e);
        }
    }
    lo[n] = make_double2(a1, a2);
    hi[n] = make_double2(b1, b2);
    alo[n] = make_int2(ia1, ia2);
    ahi[n] = make_int2(ib1, ib2);
    if (dtan != nullptr) {
        tlo[n] = make_double2(ia1 >= 0 ? dtan[ia1] : 0.0, ia2 >= 0 ? dtan[ia2] : 0.0);
        thi[n] = make_double2(ib1 >= 0 ? dtan[ib1] : 0.0, ib2 >= 0 ? dtan[ib2] : 0.0);
    }
}

// The neighbour extreme of element e and the element holding it: drop e
// itself from each node's pair (by id), keep the first strict improvement.
// code (may be null): 4 * node + slot of the chosen entry (slot 0/1 = the
// node's smallest / second smallest, 2/3 = largest / second largest), -1
// when there is no neighbour.
__device__ __forceinline__ double neighbour_extreme(const int32_t nd[8], int e, bool low,
                                                    const double2* lo, const double2* hi,
                                                    const int2* alo, const int2* ahi, int& arg,
                                                    const double2* tlo, const double2* thi,
                                                    double* tan, int* code = nullptr)
{
    double ext = low ? HUGE_VAL : -HUGE_VAL, te = 0.0;
    arg = -1;
    int cd = -1;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const double2 t = low ? lo[nd[a]] : hi[nd[a]];
        const int2 id = low ? alo[nd[a]] : ahi[nd[a]];
        const bool skip = id.x == e;
        const double o = skip ? t.y : t.x;
        const int oid = skip ? id.y : id.x;
        const bool better = low ? (o < ext) : (o > ext);
        if (better) {
            ext = o;
            arg = oid;
            cd = 4 * nd[a] + (low ? 0 : 2) + (skip ? 1 : 0);
            if (tan != nullptr) {
                const double2 tt = low ? tlo[nd[a]] : thi[nd[a]];
                te = skip ? tt.y : tt.x;
            }
        }
    }
    if (tan != nullptr) *tan = te;
    if (code != nullptr) *code = arg >= 0 ? cd : -1;
    return ext;
}

// -------------------------------------------------------------------------
// tangent-linear kernels (Dual instantiation of the forward physics)
// -------------------------------------------------------------------------
struct TLMState {
    TA x0[3], v0[3], x1[3], v1[3];      // node bins (t_n, stage output)
    TA e0, e1;                           // element energy bins
    TA den, pres, sspd, vol, div_in, div_out;
    TA F;                                // 24E
    TA M, m, g;                          // element mass, node mass, gamma
    const double* csmin;
    const uint8_t* mask;
    const Material* mats;
    const int32_t* mat;
    int32_t nmat;
    double2 *lo, *hi, *tlo, *thi;
    int2 *alo, *ahi;
};

__global__ __launch_bounds__(kB) void k_tlm_force(DMesh m, TLMState s, const InstCtl* ctl)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    const double fuzz = ctl[bi.inst].fuzz, small = ctl[bi.inst].small;
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    Dual x[8], y[8], z[8], u[8], v[8], w[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        x[a] = s.x1[0].ld(nd[a]);
        y[a] = s.x1[1].ld(nd[a]);
        z[a] = s.x1[2].ld(nd[a]);
        u[a] = s.v1[0].ld(nd[a]);
        v[a] = s.v1[1].ld(nd[a]);
        w[a] = s.v1[2].ld(nd[a]);
    }
    const Dual div = s.div_in.ld(e);
    const bool low = div.v + small >= 0.0;
    int arg;
    double te;
    const double ext = neighbour_extreme(nd, e, low, s.lo, s.hi, s.alo, s.ahi, arg, s.tlo, s.thi,
                                         &te);
    const Material mt = s.mats[s.nmat > 1 ? s.mat[e] : 0];
    Dual f[24];
    force_elem(x, y, z, u, v, w, s.den.ld(e), s.pres.ld(e), s.sspd.ld(e), s.vol.ld(e), div,
               Dual(ext, te), mt.q1, mt.q1ex, mt.q2, mt.q2ex, fuzz, small, f);
#pragma unroll
    for (int k = 0; k < 24; ++k) s.F.st(24 * static_cast<size_t>(e) + k, f[k]);
}

__global__ __launch_bounds__(kB) void k_tlm_node(DMesh m, TLMState s, const double* dts,
                                                 double alpha)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    const double dt = dts[bi.inst];
    Dual f[3] = {Dual(0.0), Dual(0.0), Dual(0.0)};
    for (int q = m.cin_start[n]; q < m.cin_start[n + 1]; ++q) {
        const size_t c = static_cast<size_t>(m.cin[q]);
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] += s.F.ld(3 * c + d);
    }
    Dual v0[3], x0[3], v1[3], x1[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        v0[d] = s.v0[d].ld(n);
        x0[d] = s.x0[d].ld(n);
    }
    node_update(f, s.m.ld(n), v0, x0, s.mask[n], alpha, dt, v1, x1);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        s.v1[d].st(n, v1[d]);
        s.x1[d].st(n, x1[d]);
    }
}

__global__ __launch_bounds__(kB) void k_tlm_energy_state(DMesh m, TLMState s, const double* dts,
                                                         double alpha)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    const double dt = dts[bi.inst];
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    Dual f[24], x[8], y[8], z[8], u1[8], v1[8], w1[8], u0[8], v0[8], w0[8];
#pragma unroll
    for (int k = 0; k < 24; ++k) f[k] = s.F.ld(24 * static_cast<size_t>(e) + k);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        x[a] = s.x1[0].ld(nd[a]);
        y[a] = s.x1[1].ld(nd[a]);
        z[a] = s.x1[2].ld(nd[a]);
        u1[a] = s.v1[0].ld(nd[a]);
        v1[a] = s.v1[1].ld(nd[a]);
        w1[a] = s.v1[2].ld(nd[a]);
        u0[a] = s.v0[0].ld(nd[a]);
        v0[a] = s.v0[1].ld(nd[a]);
        w0[a] = s.v0[2].ld(nd[a]);
    }
    const Dual M = s.M.ld(e);
    const Dual sie = energy_update(f, u1, v1, w1, u0, v0, w0, s.e0.ld(e), M, alpha, dt);
    s.e1.st(e, sie);
    Dual vol, den, p, c, div;
    const int32_t kind = s.mats[s.nmat > 1 ? s.mat[e] : 0].eos;
    derived_state(x, y, z, u1, v1, w1, M, sie, s.g.ld(e), s.csmin[e], kind, vol, den, p, c, div);
    s.vol.st(e, vol);
    s.den.st(e, den);
    s.pres.st(e, p);
    s.sspd.st(e, c);
    s.div_out.st(e, div);
}

// initial tangents: den0 = rho0, p0, c0 by the EOS, div0 from v0 (x fixed)
__global__ __launch_bounds__(kB) void k_tlm_setup_elem(DMesh m, TLMState s)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    double x[8], y[8], z[8];
    g3(s.x1[0].v, s.x1[1].v, s.x1[2].v, nd, x, y, z);
    Dual u[8], v[8], w[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        u[a] = s.v1[0].ld(nd[a]);
        v[a] = s.v1[1].ld(nd[a]);
        w[a] = s.v1[2].ld(nd[a]);
    }
    Dual p, c;
    eos(s.mats[s.nmat > 1 ? s.mat[e] : 0].eos, s.g.ld(e), s.csmin[e], s.den.ld(e), s.e1.ld(e), p,
        c);
    s.pres.st(e, p);
    s.sspd.st(e, c);
    double B[24];
    hex_bmatrix(x, y, z, B);
    Dual Bd[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) Bd[k] = B[k];
    s.div_in.st(e, divergence(Bd, u, v, w, Dual(s.vol.v[e], 0.0)));
}

// node mass tangent: m = sum over the node's corners of M/8 (K/setup.cpp:398-404)
__global__ __launch_bounds__(kB) void k_tlm_setup_node(DMesh m, TLMState s)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    double acc = 0.0;
    for (int q = m.cin_start[n]; q < m.cin_start[n + 1]; ++q) acc += 1.0 / 8.0 * s.M.t[m.cin[q] >> 3];
    s.m.t[n] = acc;
}

// -------------------------------------------------------------------------
// reverse kernels
// -------------------------------------------------------------------------
struct AdjView {
    // adjoints of the stage outputs (node x, v; element e) and of the cycle
    // input state, accumulated over both stages
    double *xo[3], *vo[3], *eo;
    double *x0b[3], *v0b[3], *e0b;
    double *Mb, *mb, *gb, *rhob;
    double* ge;                 // element: -alpha dt / M * e-bar of the stage output energy
    // node: the stage's mean velocity h = (v1 + v0)/2 and kv = alpha dt / m
    // times the masked velocity adjoint, one 48-byte record per node
    // (h0, h1, h2, kv0, kv1, kv2); a corner's force adjoint is
    // F-bar = g_e h + kv at its node.  Every h it holds is finite (zeroed at
    // allocation; a quiet node block of k_adj_node keeps the earlier one)
    double* hk;
    // corners, element order: position (cbx) and velocity (cbv) adjoints,
    // rows of 24 doubles per element (8 corners x 3)
    double *cbx, *cbv;
    uint8_t* rowset;            // element: 1 if its rows were written this stage, 0 if zero
    // node, this stage: kv is nonzero (k_adj_node); a slot sum is nonzero (k_adj_ext_node)
    uint8_t *kvflag, *accflag;
    double* extb;               // element: adjoint of its neighbour extreme
    int32_t* argext;            // element: 4 * node + slot of its chosen extreme, or -1
    double* nacc;               // node: per-slot sums of ext-bar (4 x N), where accflag
    double2 *lo, *hi;
    int2 *alo, *ahi;
};

// Wave-local synchronisation around the per-wave LDS transposes of corner
// rows (release/acquire at wavefront scope and a wave barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// e1 = e0 - alpha dt / M * P,  P = sum_c F_c . (v1+v0)/2: the element part
// (e0-bar, M-bar and the scalar g = -alpha dt / M * e1-bar that the node
// kernel turns into the corner terms g (v1+v0)/2 and g F_c).
// assign != 0: the accumulators (e0-bar here, x0-bar / v0-bar in the node
// kernel, the input adjoints in k_adj_force_pair<false> / k_adj_gather) are written
// instead of added to (the first stage reversed in a cycle)
__global__ __launch_bounds__(kB) void k_adj_energy(DMesh m, DState s, const double* F,
                                                   AdjView A, const double* dts, double alpha,
                                                   int assign)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    const double k = alpha * dts[bi.inst];
    const double eb = A.eo[e];
    A.e0b[e] = assign ? eb : A.e0b[e] + eb;
    const double M = s.emass[e];
    A.ge[e] = -k / M * eb;
    if (eb == 0.0) return;
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    const double* f = F + 24 * static_cast<size_t>(e);
    double P = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        P += f[3 * a + 0] * ((s.u1[nd[a]] + s.u0[nd[a]]) * 0.5);
        P += f[3 * a + 1] * ((s.v1[nd[a]] + s.v0[nd[a]]) * 0.5);
        P += f[3 * a + 2] * ((s.w1[nd[a]] + s.w0[nd[a]]) * 0.5);
    }
    A.Mb[e] += k / (M * M) * eb * P;
}

// v1 = mask(v0 + alpha dt F/m), x1 = x0 + alpha dt (v1 + v0)/2, and the
// energy update's corner terms: (v1+v0)/2-bar = sum_c g_e F_c over the
// node's corners (list order).  Writes the node's h = (v1+v0)/2 and
// kv = alpha dt/m v1-bar (masked), from which k_adj_force_pair forms each
// corner's F_c-bar = g_e h + kv.  The block's corner forces and their
// elements' g are staged in LDS in sorted-corner order (coalesced loads, as
// the forward velocity kernel stages its forces); a block over the capacity
// gathers from global memory.  Sums in list order either way.  (Every
// thread reaches the block's barriers: threads past the block's nodes return
// after them.)
template <int BS>
__global__ __launch_bounds__(BS) void k_adj_node(DMesh m, DState s, const double* F, AdjView A,
                                                 const double* dts, double alpha, int assign)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const double k = alpha * dts[bi.inst];
    constexpr int CAPC = corner_stage_cap(BS);
    __shared__ double sF[4 * CAPC];              // per sorted corner: F (3), g of its element
    const int qb = m.cin_start[bi.start];
    const int nc = m.cin_start[bi.end] - qb;
    const bool staged = m.cs_cor != nullptr && nc <= CAPC;
    const int n = bi.start + static_cast<int>(threadIdx.x);
    const bool live = n < bi.end;
    // the stage output's adjoints
    double xo[3] = {0.0, 0.0, 0.0}, vo[3] = {0.0, 0.0, 0.0};
    if (live) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xo[d] = A.xo[d][n];
            vo[d] = A.vo[d][n];
        }
    }
    // A staged block whose corners' g and nodes' stage-output adjoints are
    // all zero is quiet (block-uniform): its outputs are zero, so it reads
    // no corner forces or velocities, leaves its accumulators as they are
    // (writes zeros where it assigns) and writes kv = 0 with its flag; the
    // records keep their h (finite: zero from the allocation or an earlier
    // stage's mean velocity), which enters a corner's force adjoint only as
    // g_e h with g_e = 0 at every corner of these nodes.
    bool quiet = false;
    if (staged) {
        constexpr int R = CAPC / BS;
        int cid[R];
        double ge[R];
        bool nz = false;
#pragma unroll
        for (int d = 0; d < 3; ++d) nz = nz || xo[d] != 0.0 || vo[d] != 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = r * BS + static_cast<int>(threadIdx.x);
            cid[r] = j < nc ? m.cs_cor[qb + j] : -1;
            ge[r] = cid[r] >= 0 ? A.ge[cid[r] >> 3] : 0.0;
            nz = nz || ge[r] != 0.0;
        }
        quiet = !__syncthreads_or(nz);
        if (!quiet) {
            double g[R][3];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int c = cid[r] < 0 ? 0 : cid[r];
                const double* f = F + 3 * static_cast<size_t>(c);
                g[r][0] = f[0];
                g[r][1] = f[1];
                g[r][2] = f[2];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = r * BS + static_cast<int>(threadIdx.x);
                if (cid[r] >= 0) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) sF[4 * j + d] = g[r][d];
                    sF[4 * j + 3] = ge[r];
                }
            }
            __syncthreads();
        }
    }
    if (!live) return;
    if (quiet) {
        if (assign) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                A.x0b[d][n] = 0.0;
                A.v0b[d][n] = 0.0;
            }
        }
        double* rec = A.hk + 6 * static_cast<size_t>(n);
        rec[3] = 0.0;
        reinterpret_cast<double2*>(rec)[2] = make_double2(0.0, 0.0);
        A.kvflag[n] = 0;
        return;
    }
    if (!live) return;
    const int q0 = m.cin_start[n];
    const int cnt = m.cin_start[n + 1] - q0;
    double sh[3] = {0.0, 0.0, 0.0}, sf[3] = {0.0, 0.0, 0.0};
    for (int q = q0; q < q0 + cnt; ++q) {
        double f[3], g;
        if (staged) {
            const double* p = sF + 4 * m.cs_pos[q];
            f[0] = p[0];
            f[1] = p[1];
            f[2] = p[2];
            g = p[3];
        } else {
            const int c = m.cin[q];
            const double* p = F + 3 * static_cast<size_t>(c);
            f[0] = p[0];
            f[1] = p[1];
            f[2] = p[2];
            g = A.ge[c >> 3];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            sh[d] += g * f[d];
            sf[d] += f[d];
        }
    }
    const double mass = s.nmass[n];
    const uint8_t mk = s.bcmask[n];
    const double h[3] = {(s.u1[n] + s.u0[n]) * 0.5, (s.v1[n] + s.v0[n]) * 0.5,
                         (s.w1[n] + s.w0[n]) * 0.5};
    double vm[3];
    double dot = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double xb = xo[d];
        A.x0b[d][n] = assign ? xb : A.x0b[d][n] + xb;
        const double vt = vo[d] + 0.5 * k * xb + 0.5 * sh[d];
        vm[d] = (mk & (1u << d)) ? 0.0 : vt;
        const double v0b = 0.5 * k * xb + 0.5 * sh[d] + vm[d];
        A.v0b[d][n] = assign ? v0b : A.v0b[d][n] + v0b;
        dot += vm[d] * sf[d];
    }
    A.mb[n] += -k * dot / (mass * mass);
    const double km = k / mass;
    const double kv[3] = {km * vm[0], km * vm[1], km * vm[2]};
    double2* rec = reinterpret_cast<double2*>(A.hk + 6 * static_cast<size_t>(n));
    rec[0] = make_double2(h[0], h[1]);
    rec[1] = make_double2(h[2], kv[0]);
    rec[2] = make_double2(kv[1], kv[2]);
    A.kvflag[n] = kv[0] != 0.0 || kv[1] != 0.0 || kv[2] != 0.0;
}

// Shock-detector neighbour term: add to each element's divergence adjoint
// the ext-adjoints of the elements whose extreme neighbour it is.  A
// neighbour sharing several nodes is counted once, at the first shared node
// in local order.
// Shock-detector neighbour term, in two deterministic passes.  Element e2
// took its extreme from slot s of node n (argext = 4 n + s); the element in
// that slot gets ext-bar(e2) added to its divergence adjoint.  Node pass
// (here): per node and slot, the sum of ext-bar over the elements (in CSR
// order) whose choice was that node and slot.  Element pass (start of
// k_adj_force_pair<false>): each element adds the sums of the slots it occupies at its
// 8 nodes.  Every e2 contributes once.
__global__ __launch_bounds__(kB) void k_adj_ext_node(DMesh m, AdjView A)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    const int q0 = m.cin_start[n];
    int c8[8];
    const int cnt = corners8(m, n, q0, c8);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    int code[8];
    double xb[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        code[j] = A.argext[c8[j] >> 3];
        xb[j] = A.extb[c8[j] >> 3];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool mine = j < cnt && (code[j] >> 2) == n && code[j] >= 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] += (mine && (code[j] & 3) == k) ? xb[j] : 0.0;
    }
    for (int q = q0 + 8; q < q0 + cnt; ++q) {
        const int e2 = m.cin[q] >> 3;
        const int cd = A.argext[e2];
        if (cd >= 0 && (cd >> 2) == n) acc[cd & 3] += A.extb[e2];
    }
    // the sums are written only where one is nonzero (accflag; the element
    // pass reads them only there)
    const bool any = acc[0] != 0.0 || acc[1] != 0.0 || acc[2] != 0.0 || acc[3] != 0.0;
    if (any) {
        double2* out = reinterpret_cast<double2*>(A.nacc + 4 * static_cast<size_t>(n));
        out[0] = make_double2(acc[0], acc[1]);
        out[1] = make_double2(acc[2], acc[3]);
    }
    A.accflag[n] = any;
}

// ---- one element over a pair of lanes ----
// Lanes 2i and 2i + 1 of a wave share element i; lane half h = lane & 1
// holds corners 4h .. 4h + 3.  A sum over the element's corners is formed
// per half and the two halves added (a + b == b + a, so both lanes hold the
// same value).
__device__ __forceinline__ double pair_sum(double v) { return v + __shfl_xor(v, 1, 64); }

// hi when h, else lo, by bit masks: the same value as h ? hi : lo, written so
// the compiler cannot turn a choice between two array elements into a
// lane-dependent index (which would move the array to scratch memory)
__device__ __forceinline__ double pick(int h, double lo, double hi)
{
    const long long m = -static_cast<long long>(h);
    return __longlong_as_double((__double_as_longlong(lo) & ~m) | (__double_as_longlong(hi) & m));
}
__device__ __forceinline__ int32_t pick(int h, int32_t lo, int32_t hi)
{
    const int32_t m = -h;
    return (lo & ~m) | (hi & m);
}

// Half rows (12 doubles: corners 4h .. 4h + 3 of element i) held by the lane
// pair (2i, 2i + 1) of a wave -> memory.  The wave's 32 elements own one
// contiguous stretch of 6 KB of the corner array, of which lane l holds bytes
// [96 l, 96 l + 96); the lanes store it as consecutive 16-byte pieces
// (coalesced) through a padded LDS region of 64 x 7 pieces.  e0: the wave's
// first element; nvalid: its elements inside the work item (rows past it are
// not stored).
__device__ __forceinline__ void wave_store_half_rows(double* dst_, int e0, int nvalid,
                                                     const double f[12], double2* buf)
{
    const int lane = static_cast<int>(threadIdx.x) & 63;
    double2* dst = reinterpret_cast<double2*>(dst_ + 24 * static_cast<size_t>(e0));
    wave_sync();
#pragma unroll
    for (int k = 0; k < 6; ++k) buf[lane * 7 + k] = make_double2(f[2 * k], f[2 * k + 1]);
    wave_sync();
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        const int p = 64 * r + lane;   // piece within the wave's stretch
        if (p / 12 < nvalid) dst[p] = buf[(p / 6) * 7 + p % 6];
    }
}

// Whether either lane of the pair holds a true flag.
__device__ __forceinline__ bool pair_or(bool b)
{
    return (static_cast<int>(b) | __shfl_xor(static_cast<int>(b), 1, 64)) != 0;
}

// Reverse of force_elem (K/force_sgh.cpp:35-359) at the stage input state,
// with each element's corners split over a lane pair (two waves of 32
// elements per 64-element work item): a lane holds the per-corner values of
// four corners, so the kernel runs two waves per SIMD (one element per lane
// needed 450 registers, one wave).  The B matrix and the mean velocity are
// formed from all eight nodes in both lanes (the mean velocity in the forward
// kernel's order); the other sums over corners are formed per half and the
// halves added.  Lanes past the work item compute on the item's first
// element and store nothing (every lane takes part in the row transposes).
//
// Two passes over the elements use it:
//   EXT = true  (the "adj_ext" pass): only the shock detector's neighbour-extreme
//     adjoint ext-bar and the chosen extreme, which k_adj_ext_node sums per
//     node slot for the element that holds the extreme.  Zero where the
//     limiter does not depend on the extreme (r_face outside (0, 1]); a wave
//     with no such element skips the recompute.
//   EXT = false (the "adj_force_derived" pass): the whole force reverse, then the
//     reverse of the stage input's derived state (volume, density, EOS,
//     divergence) with the neighbour sums added to the divergence adjoint,
//     the stage's one B-matrix VJP, and the element's corner rows (position
//     and velocity adjoints) written once.  setup != 0: the first stage of
//     the run, whose density, pressure and sound speed come from the initial
//     conditions (den = rho0) and whose coordinates are the fixed initial
//     mesh.  assign != 0: the energy-adjoint accumulator is written, not
//     added to.
template <bool EXT>
__global__ __launch_bounds__(128, 2) void k_adj_force_pair(DMesh m, DState s, AdjView A,
                                                           const InstCtl* ctl, const double* vol_a,
                                                           const double* div_a,
                                                           const double* sie_in, double* sieb_in,
                                                           int setup, int assign)
{
    __shared__ double2 hbuf[2 * 64 * 7];
    const BlockItem bi = m.eblk[blockIdx.x];
    const int h = static_cast<int>(threadIdx.x) & 1;
    const int el = static_cast<int>(threadIdx.x) >> 1;
    const bool live = bi.start + el < bi.end;
    const int e = live ? bi.start + el : bi.start;
    const double fuzz = ctl[bi.inst].fuzz, small = ctl[bi.inst].small;
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    int32_t n4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) n4[k] = pick(h, nd[k], nd[4 + k]);
    const double div = div_a[e];
    const double ge = A.ge[e];
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    const int nvalid = bi.end - bi.start - 32 * wave;   // this wave's elements in the item
    const int e0 = bi.start + 32 * wave;
    double2* buf = hbuf + wave * 64 * 7;
    // The reverse of a zero adjoint is zero.  A corner's force adjoint
    // g_e h + kv is zero where g_e (ge) and its node's kv are (kvflag);
    // ext-bar = -phi-bar alpha 0.9 / (div + small) is zero where every
    // corner's force adjoint is (then phi-bar is) and where alpha is
    // (div = 0); the element's rows and element adjoints are zero where, in
    // addition, its neighbour term is (no nonzero slot sum at its nodes,
    // accflag).  A wave with no element left stops here (wave-uniform
    // branches).
    bool kvs = ge != 0.0, accs = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        kvs = kvs || A.kvflag[n4[k]] != 0;
        if (!EXT) accs = accs || A.accflag[n4[k]] != 0;
    }
    const bool seeded = pair_or(kvs);
    double divnb_all = 0.0;
    if (EXT) {
        if (!__any(live && seeded && div != 0.0)) {
            if (live && h == 0) {
                A.extb[e] = 0.0;
                A.argext[e] = -1;   // no slot to add to
            }
            return;
        }
    } else {
        if (!__any(live && (seeded || pair_or(accs)))) {
            // rows absent (k_adj_gather reads them as zero)
            if (live && h == 0) {
                if (assign) sieb_in[e] = 0.0;
                A.rowset[e] = 0;
            }
            return;
        }
        // the neighbour term of the divergence adjoint: the per-slot sums
        // (k_adj_ext_node) of the slots this element occupies at this lane's
        // four nodes, read where a node has a nonzero one (accflag)
        double divnb = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (A.accflag[n4[k]]) {
                const int2 il = A.alo[n4[k]], ih = A.ahi[n4[k]];
                const double2* p =
                    reinterpret_cast<const double2*>(A.nacc + 4 * static_cast<size_t>(n4[k]));
                const double2 s01 = p[0], s23 = p[1];
                divnb += (il.x == e ? s01.x : 0.0) + (il.y == e ? s01.y : 0.0) +
                         (ih.x == e ? s23.x : 0.0) + (ih.y == e ? s23.y : 0.0);
            }
        }
        divnb_all = pair_sum(divnb);
    }
    const double vol = vol_a[e], den = s.den[e], sspd = s.sspd[e], pres = s.pres[e];
    const Material mt = s.mats[s.nmat > 1 ? s.mat[e] : 0];
    const bool low = div + small >= 0.0;
    int arg;
    int code;
    const double ext = neighbour_extreme(nd, e, low, A.lo, A.hi, A.alo, A.ahi, arg, nullptr,
                                         nullptr, nullptr, &code);
    const double bden = div + small;
    const double r = 0.9 * (ext + small) / bden;
    const double rmin = fmin(r, 1.0);
    // the limiter depends on the neighbour extreme only for r_face in (0, 1]
    const bool rdep = rmin > 0.0 && r <= 1.0 && arg >= 0;
    if (EXT && !__any(live && seeded && rdep)) {
        if (live && h == 0) {
            A.extb[e] = 0.0;
            A.argext[e] = -1;
        }
        return;
    }
    // ---- forward recompute ----
    double B4[12];
    {
        double x[8], y[8], z[8], B[24];
        g3(s.x1, s.y1, s.z1, nd, x, y, z);
        hex_bmatrix(x, y, z, B);
#pragma unroll
        for (int i = 0; i < 12; ++i) B4[i] = pick(h, B[i], B[12 + i]);
    }
    double vbar[3] = {0.0, 0.0, 0.0}, vel4[3][4];
    {
        double vel[3][8];
        g3(s.u1, s.v1, s.w1, nd, vel[0], vel[1], vel[2]);
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) vbar[d] += 0.125 * vel[d][a];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int d = 0; d < 3; ++d) vel4[d][k] = pick(h, vel[d][k], vel[d][4 + k]);
    }
    const double qa = div < 0.0 ? mt.q1 : mt.q1ex;
    const double qb = div < 0.0 ? mt.q2 : mt.q2ex;
    double mv[4], nb[4], Mt[4], S[3] = {0.0, 0.0, 0.0}, W = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double d0 = vel4[0][k] - vbar[0], d1 = vel4[1][k] - vbar[1],
                     d2 = vel4[2][k] - vbar[2];
        mv[k] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        const double muc = den * (qa * sspd + qb * mv[k]);
        nb[k] = sqrt(B4[3 * k] * B4[3 * k] + B4[3 * k + 1] * B4[3 * k + 1] +
                     B4[3 * k + 2] * B4[3 * k + 2]);
        Mt[k] = muc * nb[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) S[d] += Mt[k] * vel4[d][k];
        W += Mt[k];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) S[d] = pair_sum(S[d]);
    W = pair_sum(W);
    const bool riemann = W > fuzz;
    double vs[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) vs[d] = riemann ? S[d] / W : 0.0;
    const double phi0 = 1.0 - fmax(0.0, rmin);
    const double cl = pow(vol, 1.0 / 3.0);
    const double cd = sspd + fuzz;
    const double argv = 20.0 * (cl * fabs(div)) / cd;
    const double alpha = fmin(1.0, argv);
    const double phi = alpha * phi0;
    // ---- reverse ----
    double Bb4[12], velb4[3][4], Mb[4];
    double presb = 0.0, phib = 0.0, vsb[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double g = 0.0;
        const double2* rec = reinterpret_cast<const double2*>(A.hk + 6 * static_cast<size_t>(n4[k]));
        const double2 r0 = rec[0], r1 = rec[1], r2 = rec[2];
        const double hh[3] = {r0.x, r0.y, r1.x}, kv[3] = {r1.y, r2.x, r2.y};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double fb = ge * hh[d] + kv[d];
            presb += B4[3 * k + d] * fb;
            Bb4[3 * k + d] = pres * fb;
            g += fb * (vs[d] - vel4[d][k]);
            vsb[d] += phi * Mt[k] * fb;
            velb4[d][k] = -phi * Mt[k] * fb;
        }
        phib += Mt[k] * g;
        Mb[k] = phi * g;
    }
    presb = pair_sum(presb);
    phib = pair_sum(phib);
    if (EXT) {
        double extb = 0.0;
        if (rmin > 0.0 && r <= 1.0) extb = -(phib * alpha) * 0.9 / bden;
        if (live && h == 0) {
            A.extb[e] = (arg >= 0) ? extb : 0.0;
            A.argext[e] = code;
        }
        return;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) vsb[d] = pair_sum(vsb[d]);
    if (riemann) {
        double Wb = 0.0, Sb[3];
        const double iW = 1.0 / W;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Sb[d] = vsb[d] * iW;
            Wb -= vsb[d] * vs[d] * iW;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                velb4[d][k] += Mt[k] * Sb[d];
                Mb[k] += Sb[d] * vel4[d][k];
            }
            Mb[k] += Wb;
        }
    }
    double denb = 0.0, sspdb = 0.0, vbarb[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double muc = den * (qa * sspd + qb * mv[k]);
        const double mucb = Mb[k] * nb[k];
        const double nbb = Mb[k] * muc;
        if (nb[k] > 0.0) {
            const double q = nbb / nb[k];
#pragma unroll
            for (int d = 0; d < 3; ++d) Bb4[3 * k + d] += q * B4[3 * k + d];
        }
        denb += mucb * (qa * sspd + qb * mv[k]);
        sspdb += mucb * den * qa;
        const double mvb = mucb * den * qb;
        if (mv[k] > 0.0) {
            const double q = mvb / mv[k];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double db = q * (vel4[d][k] - vbar[d]);
                velb4[d][k] += db;
                vbarb[d] -= db;
            }
        }
    }
    denb = pair_sum(denb);
    sspdb = pair_sum(sspdb);
#pragma unroll
    for (int d = 0; d < 3; ++d) vbarb[d] = pair_sum(vbarb[d]);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) velb4[d][k] += 0.125 * vbarb[d];
    double divb = 0.0, volb = 0.0;
    const double alphab = phib * phi0;
    const double phi0b = phib * alpha;
    if (argv < 1.0) {
        const double clb = alphab * 20.0 * fabs(div) / cd;
        divb += alphab * 20.0 * cl * (div < 0.0 ? -1.0 : 1.0) / cd;
        sspdb -= alphab * argv / cd;
        if (vol > 0.0) volb += clb * cl / (3.0 * vol);
    }
    if (rmin > 0.0 && r <= 1.0) {
        // (the neighbour extreme's term, rb 0.9 / bden, is the ext-bar of
        // the EXT pass)
        const double rb = -phi0b;
        divb -= rb * r / bden;
    }
    // ---- the stage input's derived state (K/geometry.cpp:298-306 volume,
    // K/properties.cpp:39,99 density and EOS, K/momentum.cpp:265-286 div) ----
    divb = divb + divnb_all;
    // divergence: div = (sum_a v_a . B_a) / vol
    const double iv = 1.0 / vol;
    {
        double rv[12];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                Bb4[3 * k + d] += divb * iv * vel4[d][k];
                rv[3 * k + d] = velb4[d][k] + divb * iv * B4[3 * k + d];
            }
        if (nvalid > 0) wave_store_half_rows(A.cbv, e0, nvalid, rv, buf);   // wave-uniform
    }
    volb -= divb * div * iv;
    const double gam = s.gamma[e], csmin = s.csmin[e];
    const double sie = sie_in[e];
    // the material's EOS: its partial derivatives in density, energy and gamma
    // by three dual-number evaluations (both lanes; the first stores)
    const int32_t kind = mt.eos;
    Dual pr, cs;
    eos(kind, Dual(gam), csmin, Dual(den, 1.0), Dual(sie), pr, cs);
    const double dnb = denb + presb * pr.d + sspdb * cs.d;
    eos(kind, Dual(gam), csmin, Dual(den), Dual(sie, 1.0), pr, cs);
    const double eb = presb * pr.d + sspdb * cs.d;
    eos(kind, Dual(gam, 1.0), csmin, Dual(den), Dual(sie), pr, cs);
    const double gb = presb * pr.d + sspdb * cs.d;
    if (live && h == 0) {
        sieb_in[e] = assign ? eb : sieb_in[e] + eb;
        A.gb[e] += gb;
        A.rowset[e] = 1;
    }
    double xb[8] = {0}, yb[8] = {0}, zb[8] = {0};
    if (setup) {
        // den = rho0; vol and B belong to the fixed initial mesh (its
        // coordinate adjoint is not a gradient of the run)
        if (live && h == 0) A.rhob[e] += dnb;
    } else {
        // den = M / vol; vol = V(x) with dV/dx_a = B_a, on this lane's nodes
        if (live && h == 0) A.Mb[e] += dnb / vol;
        volb -= dnb * den / vol;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double bx = volb * B4[3 * k], by = volb * B4[3 * k + 1],
                         bz = volb * B4[3 * k + 2];
            xb[k] = pick(h, bx, 0.0);
            yb[k] = pick(h, by, 0.0);
            zb[k] = pick(h, bz, 0.0);
            xb[4 + k] = pick(h, 0.0, bx);
            yb[4 + k] = pick(h, 0.0, by);
            zb[4 + k] = pick(h, 0.0, bz);
        }
    }
    {
        // the one B-matrix VJP of the stage on this lane's rows (the other
        // half zero), on coordinates read again (cache hits) behind an
        // opaque copy, so that none of the B matrix's coordinate differences
        // stay live across the kernel (registers)
        double Bbf[24];
#pragma unroll
        for (int i = 0; i < 12; ++i) {
            Bbf[i] = pick(h, Bb4[i], 0.0);
            Bbf[12 + i] = pick(h, 0.0, Bb4[i]);
        }
        // the element's nodes, too, read again behind an opaque index
        int ei = e;
        asm volatile("" : "+v"(ei));
        ld_conn(m.conn, ei, nd);
        double x[8], y[8], z[8];
        g3(s.x1, s.y1, s.z1, nd, x, y, z);
#pragma unroll
        for (int i = 0; i < 8; ++i) asm volatile("" : "+v"(x[i]), "+v"(y[i]), "+v"(z[i]));
        hex_bmatrix_vjp(x, y, z, Bbf, xb, yb, zb);
    }
    // this lane's nodes: its partial plus the partner's
    double row[12];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double sx = pick(h, xb[4 + k], xb[k]), sy = pick(h, yb[4 + k], yb[k]),
                     sz = pick(h, zb[4 + k], zb[k]);
        row[3 * k] = pick(h, xb[k], xb[4 + k]) + __shfl_xor(sx, 1, 64);
        row[3 * k + 1] = pick(h, yb[k], yb[4 + k]) + __shfl_xor(sy, 1, 64);
        row[3 * k + 2] = pick(h, zb[k], zb[4 + k]) + __shfl_xor(sz, 1, 64);
    }
    if (nvalid > 0) wave_store_half_rows(A.cbx, e0, nvalid, row, buf);
}

// corner adjoints -> node adjoints, summed in corner-list order (no
// atomics).  The block's corner rows are staged in LDS in sorted-corner order
// (coalesced loads), the position part, then the velocity part; a block over
// the capacity gathers from global memory.  rowset (or null: every row
// present): per element, whether its rows were written; an absent row is
// zero and not read.  assign: written, else added to.
template <int BS>
__global__ __launch_bounds__(BS) void k_adj_gather(DMesh m, const double* cbx, const double* cbv,
                                                   const uint8_t* rowset, double* xa0,
                                                   double* xa1, double* xa2, double* va0,
                                                   double* va1, double* va2, int assign)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    constexpr int CAPC = corner_stage_cap(BS);
    constexpr int R = CAPC / BS;
    __shared__ double sC[3 * CAPC];
    const int qb = m.cin_start[bi.start];
    const int nc = m.cin_start[bi.end] - qb;
    const bool staged = m.cs_cor != nullptr && nc <= CAPC;
    const int n = bi.start + static_cast<int>(threadIdx.x);
    const bool live = n < bi.end;
    int cid[R];
    if (staged) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = r * BS + static_cast<int>(threadIdx.x);
            cid[r] = j < nc ? m.cs_cor[qb + j] : -1;
        }
        // a slot whose row is absent is staged as zero, unread (-2)
        bool any = false;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (rowset && cid[r] >= 0 && !rowset[cid[r] >> 3]) cid[r] = -2;
            any = any || cid[r] >= 0;
        }
        // a block with no row present: its nodes' sums are zero (block-uniform)
        if (!__syncthreads_or(any)) {
            if (live && assign) {
                double* out[6] = {xa0, xa1, xa2, va0, va1, va2};
#pragma unroll
                for (int k = 0; k < 6; ++k) out[k][n] = 0.0;
            }
            return;
        }
    }
    int q0 = 0, cnt = 0;
    if (live) {
        q0 = m.cin_start[n];
        cnt = m.cin_start[n + 1] - q0;
    }
    double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        const double* cb = part ? cbv : cbx;
        if (staged) {
            double g[R][3];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                g[r][0] = g[r][1] = g[r][2] = 0.0;
                if (cid[r] >= 0) {
                    const double* p = cb + 3 * static_cast<size_t>(cid[r]);
                    g[r][0] = p[0];
                    g[r][1] = p[1];
                    g[r][2] = p[2];
                }
            }
            if (part) __syncthreads();     // the position part's readers are done
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = r * BS + static_cast<int>(threadIdx.x);
                if (cid[r] != -1) {
                    sC[3 * j] = g[r][0];
                    sC[3 * j + 1] = g[r][1];
                    sC[3 * j + 2] = g[r][2];
                }
            }
            __syncthreads();
        }
        for (int q = q0; q < q0 + cnt; ++q) {
            if (staged) {
                const double* p = sC + 3 * m.cs_pos[q];
                acc[3 * part] += p[0];
                acc[3 * part + 1] += p[1];
                acc[3 * part + 2] += p[2];
            } else if (!rowset || rowset[m.cin[q] >> 3]) {
                const double* p = cb + 3 * static_cast<size_t>(m.cin[q]);
                acc[3 * part] += p[0];
                acc[3 * part + 1] += p[1];
                acc[3 * part + 2] += p[2];
            }
        }
    }
    if (!live) return;
    double* out[6] = {xa0, xa1, xa2, va0, va1, va2};
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k][n] = assign ? acc[k] : acc[k] + out[k][n];
}

void launch_adj_gather(const DMesh& mn, const double* cbx, const double* cbv,
                       const uint8_t* rowset, double* const x[3], double* const v[3], int assign,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_gather<kAdjNodeBlock>, dim3(mn.n_nblk), dim3(kAdjNodeBlock), 0, st,
                       mn, cbx, cbv, rowset, x[0], x[1], x[2], v[0], v[1], v[2], assign);
}

// Final element density den = M / V(x) as an objective term: its seed sd
// gives M-bar += sd / V and, per corner, x-bar = -sd den / V dV/dx_a with
// dV/dx_a = B_a (velocity part zero).
__global__ __launch_bounds__(kB) void k_adj_seed_den(DMesh m, DState s, const double* sd,
                                                     double* Mb, double* cbx, double* cbv)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    double x[8], y[8], z[8], B[24];
    g3(s.x1, s.y1, s.z1, nd, x, y, z);
    const double vol = hex_bmatrix_volume(x, y, z, B);
    const double M = s.emass[e];
    const double g = sd[e];
    Mb[e] += g / vol;
    const double k = -g * (M / vol) / vol;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        double* px = cbx + 24 * static_cast<size_t>(e) + 3 * a;
        double* pv = cbv + 24 * static_cast<size_t>(e) + 3 * a;
        px[0] = k * B[3 * a];
        px[1] = k * B[3 * a + 1];
        px[2] = k * B[3 * a + 2];
        pv[0] = 0.0;
        pv[1] = 0.0;
        pv[2] = 0.0;
    }
}

// d/d rho0 = rhob + V0 (Mb + sum_{nodes} mb / 8); initial-velocity adjoint masked
__global__ __launch_bounds__(kB) void k_adj_finish_elem(DMesh m, AdjView A, const double* vol0,
                                                        double* grad_rho)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) acc += 1.0 / 8.0 * A.mb[nd[a]];
    grad_rho[e] = A.rhob[e] + vol0[e] * (A.Mb[e] + acc);
}


// Up to kMaxSeg (dst, src, n) copies in one launch (src == nullptr: zero fill).
constexpr int kMaxSeg = 16;
struct SegList {
    double* dst[kMaxSeg];
    const double* src[kMaxSeg];
    long long start[kMaxSeg + 1];   // prefix offsets of the segments
    int count;
};

__global__ __launch_bounds__(kB) void k_segcopy(SegList L)
{
    const long long i = static_cast<long long>(blockIdx.x) * kB + threadIdx.x;
    if (i >= L.start[L.count]) return;
    int k = 0;
    while (k + 1 < L.count && i >= L.start[k + 1]) ++k;
    const long long j = i - L.start[k];
    L.dst[k][j] = L.src[k] != nullptr ? L.src[k][j] : 0.0;
}

// cycle counter and replayed dt of every instance; clears `stopped`
__global__ void k_set_cycle_dt(InstCtl* ctl, int32_t ninst, long long cycle, const double* dts)
{
    const int i = blockIdx.x * kB + static_cast<int>(threadIdx.x);
    if (i >= ninst) return;
    ctl[i].cycle = cycle;
    ctl[i].stopped = 0;
    ctl[i].dt = dts[static_cast<size_t>(cycle) * ninst + i];
}

// element state (vol, den, p, c, div) from (x, v, e) at a cycle start
__global__ __launch_bounds__(kB) void k_derived(DMesh m, DState s)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    double x[8], y[8], z[8], u[8], v[8], w[8];
    g3(s.x1, s.y1, s.z1, nd, x, y, z);
    g3(s.u1, s.v1, s.w1, nd, u, v, w);
    double vol, den, p, c, div;
    derived_state(x, y, z, u, v, w, s.emass[e], s.sie1[e], s.gamma[e], s.csmin[e],
                  s.mats[s.nmat > 1 ? s.mat[e] : 0].eos, vol, den, p, c, div);
    s.vol[e] = vol;
    s.den[e] = den;
    s.pres[e] = p;
    s.sspd[e] = c;
    s.div[e] = div;
}

// tangent of the element mass M = rho0 V0
__global__ void k_mul(double* dst, const double* a, const double* b, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) dst[i] = a[i] * b[i];
}

inline unsigned nb_for(size_t n) { return static_cast<unsigned>((n + kB - 1) / kB); }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct AdjointData {
    int64_t ncyc = 0, every = 1;
    int ninst = 1;
    size_t N = 0, E = 0;
    double* dt = nullptr;                 // [ncyc][ninst]
    std::vector<double*> chk;             // states every `every` cycles (x,y,z,u,v,w | e)
    double *den0 = nullptr, *pres0 = nullptr, *sspd0 = nullptr, *vol0 = nullptr;
    std::vector<double*> bufs;            // everything allocated, released together
    bool done = false;
    // reverse pass: one segment of `every` cycles recomputed and kept per
    // bank; with two banks the next segment is recomputed on a second stream
    // while the current one is reversed
    bool rev_ready = false;
    struct Bank {
        std::vector<double*> seg;         // cycle-start states S_c of the segment (every + 1)
        std::vector<double*> sta;         // stage-0 outputs (x,y,z,u,v,w | e)
        std::vector<double*> Fa, Fb1;     // stage-0 / stage-1 corner forces
        std::vector<double*> ea, eb;      // stage-0 / stage-1 input element state (5 x E)
        std::vector<double*> x0, x1;      // stage-0 / stage-1 divergence extremes (6 x N, two banks)
        double* dt = nullptr;             // the segment's time steps [every][instances]
    } bank[2];
    hipGraphExec_t rgraph[4] = {};        // reverse pass of a full segment, per (bank, role parity)
    int nbank = 1;
    hipStream_t rs = nullptr;             // recompute stream (two banks)
    hipEvent_t ready[2] = {}, freed[2] = {};
    double* set[3] = {};                  // adjoint state sets (x,y,z,u,v,w | e), rotating roles
    double* grho = nullptr;
    AdjView A{};                          // per-run accumulators and per-stage scratch
    // tangent-linear model
    bool tlm_ready = false;
    double* tn[2][6][2] = {};             // node bins x (value, tangent)
    double* te[2][2] = {};                // element energy bins
    double* tel[6][2] = {};               // den, pres, sspd, vol, div A, div B
    double* tF[2] = {};
    double *tM = nullptr, *tm = nullptr, *tg = nullptr;
    double2 *lo = nullptr, *hi = nullptr, *tlo = nullptr, *thi = nullptr;
    int2 *alo = nullptr, *ahi = nullptr;
};

namespace {

struct Seg {
    double* dst;
    const double* src;
    size_t n;
};

int adj_alloc(fsgh_ctx* c, double** p, size_t n)
{
    if (hipMalloc(reinterpret_cast<void**>(p), (n ? n : 1) * sizeof(double)) != hipSuccess) {
        *p = nullptr;
        set_error(c, "adjoint: device allocation failed");
        return FSGH_E_NOMEM;
    }
    c->adj->bufs.push_back(*p);
    return FSGH_OK;
}

template <typename T>
int adj_alloc_t(fsgh_ctx* c, T** p, size_t n)
{
    double* d = nullptr;
    const int rc = adj_alloc(c, &d, (n * sizeof(T) + sizeof(double) - 1) / sizeof(double));
    *p = reinterpret_cast<T*>(d);
    return rc;
}

// copies / zero fills, batched kMaxSeg per launch
void segcopy(fsgh_ctx* c, const std::vector<Seg>& v)
{
    for (size_t b = 0; b < v.size(); b += kMaxSeg) {
        SegList L{};
        long long off = 0;
        int k = 0;
        for (size_t i = b; i < v.size() && i < b + kMaxSeg; ++i) {
            if (v[i].n == 0) continue;
            L.dst[k] = v[i].dst;
            L.src[k] = v[i].src;
            L.start[k] = off;
            off += static_cast<long long>(v[i].n);
            ++k;
        }
        if (k == 0) continue;
        L.start[k] = off;
        L.count = k;
        hipLaunchKernelGGL(k_segcopy, dim3(nb_for(static_cast<size_t>(off))), dim3(kB), 0,
                           c->stream, L);
    }
}

// state layout of checkpoint/segment buffers and adjoint sets: x, y, z, u, v, w (N each), e (E)
std::vector<Seg> bin_to_state(fsgh_ctx* c, double* dst, int bin)
{
    const size_t N = static_cast<size_t>(c->N), E = static_cast<size_t>(c->E);
    std::vector<Seg> v;
    for (int k = 0; k < 6; ++k) v.push_back(Seg{dst + k * N, c->d_node[bin][k], N});
    v.push_back(Seg{dst + 6 * N, c->d_sie[bin], E});
    return v;
}

int init_ctl(fsgh_ctx* c, const fsgh_run_params* p)
{
    std::vector<InstCtl> h(static_cast<size_t>(c->L.ninst));
    for (int32_t i = 0; i < c->L.ninst; ++i) {
        const fsgh_run_params& q = p[i];
        if (q.graphics_cyc_ival <= 0) return FSGH_E_INVALID;
        InstCtl& k = h[static_cast<size_t>(i)];
        std::memset(&k, 0, sizeof(k));
        k.dt = q.dt_start;
        k.graphics_time = q.graphics_dt_ival;
        k.time_final = q.time_final;
        k.dt_max = q.dt_max;
        k.dt_min = q.dt_min;
        k.dt_cfl = q.dt_cfl;
        k.graphics_dt_ival = q.graphics_dt_ival;
        k.fuzz = q.fuzz;
        k.small = q.small;
        k.cycle_stop = q.cycle_stop;
        k.graphics_cyc_ival = q.graphics_cyc_ival;
        k.final_bin = c->cur;
        k.eblk_begin = dt_block_begin(c)[static_cast<size_t>(i)];
        k.eblk_end = dt_block_begin(c)[static_cast<size_t>(i) + 1];
    }
    if (hipMemcpyAsync(c->d_ctl, h.data(), h.size() * sizeof(InstCtl), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        return FSGH_E_HIP;
    return FSGH_OK;
}

// host array in caller numbering -> device array in library numbering
int up_elem(fsgh_ctx* c, double* dst, const double* src, std::vector<double>& t)
{
    const size_t E = static_cast<size_t>(c->E);
    t.assign(E, 0.0);
    if (src)
        for (size_t e = 0; e < E; ++e) t[static_cast<size_t>(c->ne_of_oe[e])] = src[e];
    if (hipMemcpyAsync(dst, t.data(), E * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return FSGH_E_HIP;
    return FSGH_OK;
}
int up_node(fsgh_ctx* c, double* dst, const double* src, int comp, std::vector<double>& t)
{
    const size_t N = static_cast<size_t>(c->N);
    t.assign(N, 0.0);
    if (src)
        for (size_t n = 0; n < N; ++n) t[static_cast<size_t>(c->nn_of_on[n])] = src[3 * n + comp];
    if (hipMemcpyAsync(dst, t.data(), N * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return FSGH_E_HIP;
    return FSGH_OK;
}
int down_elem(fsgh_ctx* c, double* dst, const double* src, std::vector<double>& t)
{
    if (!dst) return FSGH_OK;
    const size_t E = static_cast<size_t>(c->E);
    t.resize(E);
    if (hipMemcpy(t.data(), src, E * 8, hipMemcpyDeviceToHost) != hipSuccess) return FSGH_E_HIP;
    for (size_t e = 0; e < E; ++e) dst[e] = t[static_cast<size_t>(c->ne_of_oe[e])];
    return FSGH_OK;
}
int down_node(fsgh_ctx* c, double* dst, const double* src, int comp, std::vector<double>& t)
{
    if (!dst) return FSGH_OK;
    const size_t N = static_cast<size_t>(c->N);
    t.resize(N);
    if (hipMemcpy(t.data(), src, N * 8, hipMemcpyDeviceToHost) != hipSuccess) return FSGH_E_HIP;
    for (size_t n = 0; n < N; ++n) dst[3 * n + comp] = t[static_cast<size_t>(c->nn_of_on[n])];
    return FSGH_OK;
}

// A state view over arbitrary buffers: t_n state `s0` and stage state `s1`
// (x,y,z,u,v,w | e layout), element state `el` (den, pres, sspd, vol, div),
// corner forces F; mesh-level fields from the context.
DState view(fsgh_ctx* c, const double* s0, const double* s1, double* const* el, double* F)
{
    const size_t N = static_cast<size_t>(c->N);
    DState s = make_state(c, 0, 1);
    double* a = const_cast<double*>(s0);
    double* b = const_cast<double*>(s1);
    s.x0 = a; s.y0 = a + N; s.z0 = a + 2 * N; s.u0 = a + 3 * N; s.v0 = a + 4 * N; s.w0 = a + 5 * N;
    s.x1 = b; s.y1 = b + N; s.z1 = b + 2 * N; s.u1 = b + 3 * N; s.v1 = b + 4 * N; s.w1 = b + 5 * N;
    s.sie0 = a + 6 * N;
    s.sie1 = b + 6 * N;
    s.den = el[0];
    s.pres = el[1];
    s.sspd = el[2];
    s.vol = el[3];
    s.div = el[4];
    s.div_out = el[4];
    s.F = F;
    return s;
}

// the stage-0 input element state of cycle `cyc` (segment slot j): the
// set-up state (rho0, EOS pressure and sound speed, V0) for cycle 0
void stage0_set(fsgh_ctx* c, const AdjointData::Bank& B, int64_t cyc, size_t j, double* out[5])
{
    AdjointData& D = *c->adj;
    if (cyc == 0) {
        out[0] = D.den0;
        out[1] = D.pres0;
        out[2] = D.sspd0;
        out[3] = D.vol0;
    } else {
        for (int k = 0; k < 4; ++k) out[k] = B.ea[j] + k * D.E;
    }
    out[4] = B.ea[j] + 4 * D.E;
}
void elem_set(const AdjointData& D, double* base, double* out[5])
{
    for (int k = 0; k < 5; ++k) out[k] = base + k * D.E;
}

// Per-node divergence extremes with their elements (k_arg_extremes output)
// of one stage input, in one block of 6N doubles: lo, hi (double2), alo,
// ahi (int2).
struct ExtPtr {
    double2 *lo, *hi;
    int2 *alo, *ahi;
};
ExtPtr ext_ptr(size_t N, double* base)
{
    ExtPtr p;
    p.lo = reinterpret_cast<double2*>(base);
    p.hi = reinterpret_cast<double2*>(base + 2 * N);
    p.alo = reinterpret_cast<int2*>(base + 4 * N);
    p.ahi = reinterpret_cast<int2*>(base + 5 * N);
    return p;
}

// Recompute the cycles [c0, c1) of one segment from checkpoint `chk` on
// stream `st`, keeping every intermediate the reverse pass reads:
// seg[j] = S_c, sta[j] = stage-0 output, Fa[j] / Fb1[j] = stage-0 / stage-1
// forces, ea[j] / eb[j] = stage-0 / stage-1 input element state.  No
// copies: every kernel reads and writes the slot buffers directly.
void recompute_segment(fsgh_ctx* c, AdjointData::Bank& B, const double* chk, int64_t c0,
                       int64_t c1, hipStream_t st)
{
    AdjointData& D = *c->adj;
    DMesh m = make_mesh(c);
    const DMesh mx = tuned_mesh(c, kTunedDivExtremes), mf = tuned_mesh(c, kTunedForce);
    const DMesh mv = tuned_mesh(c, kTunedVelBcPos), me = tuned_mesh(c, kTunedEnergy);
    {
        // the checkpoint and the segment's time steps into the bank
        SegList L{};
        const long long nst = static_cast<long long>(6 * D.N + D.E);
        const long long ndt = static_cast<long long>(c1 - c0) * D.ninst;
        L.dst[0] = B.seg[0];
        L.src[0] = chk;
        L.dst[1] = B.dt;
        L.src[1] = D.dt + static_cast<size_t>(c0) * D.ninst;
        L.start[0] = 0;
        L.start[1] = nst;
        L.start[2] = nst + ndt;
        L.count = 2;
        hipLaunchKernelGGL(k_segcopy, dim3(nb_for(static_cast<size_t>(nst + ndt))), dim3(kB), 0,
                           st, L);
    }
    for (int64_t cy = c0; cy < c1; ++cy) {
        const size_t j = static_cast<size_t>(cy - c0);
        double* S0[5];
        stage0_set(c, B, cy, j, S0);
        double *EB[5], *EA1[5];
        elem_set(D, B.eb[j], EB);
        elem_set(D, B.ea[j + 1], EA1);
        if (cy == 0) {
            launch_get_divergence(m, view(c, B.seg[0], B.seg[0], S0, nullptr), st);
        } else if (j == 0) {
            hipLaunchKernelGGL(k_derived, dim3(m.n_eblk), dim3(kB), 0, st, m,
                               view(c, B.seg[0], B.seg[0], S0, nullptr));
        }
        hipLaunchKernelGGL(k_set_cycle_dt, dim3(nb_for(static_cast<size_t>(c->L.ninst))),
                           dim3(kB), 0, st, c->d_ctl, c->L.ninst, static_cast<long long>(cy),
                           static_cast<const double*>(D.dt));
        // stage 0 (alpha 1/2): S_c -> sta[j]; element state -> eb[j]
        DState f0 = view(c, B.seg[j], B.seg[j], S0, B.Fa[j]);
        launch_div_extremes(mx, f0, c->d_ctl, st);
        launch_force(mf, f0, c->d_ctl, 0.0, 0.0, st);
        if (!B.x0.empty()) {
            const ExtPtr x = ext_ptr(D.N, B.x0[j]);
            hipLaunchKernelGGL(k_arg_extremes, dim3(m.n_nblk), dim3(kB), 0, st, m,
                               static_cast<const double*>(S0[4]),
                               static_cast<const double*>(nullptr), x.lo, x.hi, x.alo, x.ahi,
                               static_cast<double2*>(nullptr), static_cast<double2*>(nullptr));
        }
        DState u0 = view(c, B.seg[j], B.sta[j], EB, B.Fa[j]);
        launch_vel_bc_pos(mv, u0, c->d_ctl, 0.5, st);
        launch_energy_state_div(me, u0, c->d_ctl, 0.5, nullptr, st);
        // stage 1 (alpha 1): sta[j] -> seg[j+1]; element state -> ea[j+1]
        DState f1 = view(c, B.sta[j], B.sta[j], EB, B.Fb1[j]);
        launch_div_extremes(mx, f1, c->d_ctl, st);
        launch_force(mf, f1, c->d_ctl, 0.0, 0.0, st);
        if (!B.x1.empty()) {
            const ExtPtr x = ext_ptr(D.N, B.x1[j]);
            hipLaunchKernelGGL(k_arg_extremes, dim3(m.n_nblk), dim3(kB), 0, st, m,
                               static_cast<const double*>(EB[4]),
                               static_cast<const double*>(nullptr), x.lo, x.hi, x.alo, x.ahi,
                               static_cast<double2*>(nullptr), static_cast<double2*>(nullptr));
        }
        DState u1 = view(c, B.seg[j], B.seg[j + 1], EA1, B.Fb1[j]);
        launch_vel_bc_pos(mv, u1, c->d_ctl, 1.0, st);
        launch_energy_state_div(me, u1, c->d_ctl, 1.0, nullptr, st);
    }
}

// adjoint set: pointers to its x,y,z / u,v,w / e parts
struct SetPtr {
    double *x[3], *v[3], *e;
};
SetPtr set_ptr(const AdjointData& D, double* base)
{
    SetPtr p;
    for (int d = 0; d < 3; ++d) {
        p.x[d] = base + d * D.N;
        p.v[d] = base + (3 + d) * D.N;
    }
    p.e = base + 6 * D.N;
    return p;
}

// Reverse one RK stage.  s_in: stage input (x1..w1, element state, its
// energy in sie1); s_out: stage output velocities in u1..w1, cycle state in
// u0..w0; F: the stage's corner forces.  out: adjoints of the stage outputs;
// acc: adjoints of the cycle state; in: adjoints of the stage input; assign:
// acc and in are written, not added to.  ext: the stage input's divergence
// extremes if already computed (else computed here).
void reverse_stage(fsgh_ctx* c, int64_t cyc, int stage, const DState& s_in, const DState& s_out,
                   const double* F, double alpha, const SetPtr& out, const SetPtr& acc,
                   const SetPtr& in, int assign, const ExtPtr* ext, const double* dts)
{
    AdjointData& D = *c->adj;
    // 64 elements / nodes per work item: at up to 512 registers per lane these
    // kernels run one wave per SIMD, and small work items spread a mesh of a
    // few hundred blocks over every CU
    const DMesh m = small_block_mesh(c);
    const dim3 tb(static_cast<unsigned>(m.bs));
    AdjView A = D.A;
    for (int d = 0; d < 3; ++d) {
        A.xo[d] = out.x[d];
        A.vo[d] = out.v[d];
        A.x0b[d] = acc.x[d];
        A.v0b[d] = acc.v[d];
    }
    A.eo = out.e;
    A.e0b = acc.e;
    if (ext) {
        A.lo = ext->lo;
        A.hi = ext->hi;
        A.alo = ext->alo;
        A.ahi = ext->ahi;
    }
    timed(c, "adj_energy", [&] {
        hipLaunchKernelGGL(k_adj_energy, dim3(m.n_eblk), tb, 0, c->stream, m, s_out, F, A, dts,
                           alpha, assign);
    });
    const DMesh mn = adjoint_node_mesh(c);
    timed(c, "adj_node", [&] {
        hipLaunchKernelGGL(k_adj_node<kAdjNodeBlock>, dim3(mn.n_nblk), dim3(kAdjNodeBlock), 0,
                           c->stream, mn, s_out, F, A, dts, alpha, assign);
    });
    if (!ext)
        timed(c, "arg_extremes", [&] {
            hipLaunchKernelGGL(k_arg_extremes, dim3(m.n_nblk), tb, 0, c->stream, m,
                               static_cast<const double*>(s_in.div),
                               static_cast<const double*>(nullptr), A.lo, A.hi, A.alo, A.ahi,
                               static_cast<double2*>(nullptr), static_cast<double2*>(nullptr));
        });
    // two lanes per element of the 64-element work items (small_block_mesh)
    const int setup = (cyc == 0 && stage == 0) ? 1 : 0;
    timed(c, "adj_ext", [&] {
        hipLaunchKernelGGL(k_adj_force_pair<true>, dim3(m.n_eblk), dim3(128), 0, c->stream, m,
                           s_in, A, static_cast<const InstCtl*>(c->d_ctl),
                           static_cast<const double*>(s_in.vol),
                           static_cast<const double*>(s_in.div),
                           static_cast<const double*>(s_in.sie1), in.e, setup, assign);
    });
    timed(c, "adj_ext_node", [&] {
        hipLaunchKernelGGL(k_adj_ext_node, dim3(m.n_nblk), tb, 0, c->stream, m, A);
    });
    timed(c, "adj_force_derived", [&] {
        hipLaunchKernelGGL(k_adj_force_pair<false>, dim3(m.n_eblk), dim3(128), 0, c->stream, m,
                           s_in, A, static_cast<const InstCtl*>(c->d_ctl),
                           static_cast<const double*>(s_in.vol),
                           static_cast<const double*>(s_in.div),
                           static_cast<const double*>(s_in.sie1), in.e, setup, assign);
    });
    timed(c, "adj_gather", [&] {
        launch_adj_gather(mn, A.cbx, A.cbv, A.rowset, in.x, in.v, assign, c->stream);
    });
}

int ensure_extreme_buffers(fsgh_ctx* c)
{
    AdjointData& D = *c->adj;
    if (D.lo) return FSGH_OK;
    const size_t N = D.N;
    int rc = 0;
    rc |= adj_alloc_t(c, &D.lo, N);
    rc |= adj_alloc_t(c, &D.hi, N);
    rc |= adj_alloc_t(c, &D.alo, N);
    rc |= adj_alloc_t(c, &D.ahi, N);
    rc |= adj_alloc_t(c, &D.tlo, N);
    rc |= adj_alloc_t(c, &D.thi, N);
    return rc ? FSGH_E_NOMEM : FSGH_OK;
}

int alloc_bank(fsgh_ctx* c, AdjointData::Bank& B, size_t K, bool ext)
{
    AdjointData& D = *c->adj;
    const size_t N = D.N, E = D.E, C = 8 * E, SW = 6 * N + E;
    int rc = adj_alloc(c, &B.dt, K * static_cast<size_t>(D.ninst));
    for (size_t k = 0; k <= K && !rc; ++k) {
        double *p = nullptr, *q = nullptr;
        rc |= adj_alloc(c, &p, SW);
        rc |= adj_alloc(c, &q, 5 * E);
        B.seg.push_back(p);
        B.ea.push_back(q);
    }
    for (size_t k = 0; k < K && !rc; ++k) {
        double *a = nullptr, *f0 = nullptr, *f1 = nullptr, *b = nullptr;
        rc |= adj_alloc(c, &a, SW);
        rc |= adj_alloc(c, &f0, 3 * C);
        rc |= adj_alloc(c, &f1, 3 * C);
        rc |= adj_alloc(c, &b, 5 * E);
        B.sta.push_back(a);
        B.Fa.push_back(f0);
        B.Fb1.push_back(f1);
        B.eb.push_back(b);
        if (ext) {
            double *g0 = nullptr, *g1 = nullptr;
            rc |= adj_alloc(c, &g0, 6 * N);
            rc |= adj_alloc(c, &g1, 6 * N);
            B.x0.push_back(g0);
            B.x1.push_back(g1);
        }
    }
    return rc;
}

int ensure_reverse_buffers(fsgh_ctx* c)
{
    AdjointData& D = *c->adj;
    if (D.rev_ready) return FSGH_OK;
    const size_t N = D.N, E = D.E, C = 8 * E;
    const size_t SW = 6 * N + E;
    const size_t K = static_cast<size_t>(std::min<int64_t>(D.every, D.ncyc));
    // a second bank (recompute overlapped with the reverse pass) when there
    // is more than one segment and both banks, with the reverse pass's other
    // buffers (three adjoint sets; per element the corner rows, 48 doubles,
    // and ~8 element arrays; per node ~20 doubles of records, slot sums,
    // extremes and flags), take under three quarters of the free device
    // memory; with per-kernel timing on (fsgh_profile_enable) the
    // pass keeps one bank on the context stream, so that each reverse kernel's
    // events time it alone
    const size_t bank_bytes = 8 * ((K + 1) * (6 * N + 6 * E) + K * (18 * N + 54 * E));
    const size_t other_bytes = 8 * (3 * (6 * N + E) + 20 * N + 56 * E);
    size_t mfree = 0, mtotal = 0;
    const bool two = !c->prof && D.ncyc > D.every &&
                     hipMemGetInfo(&mfree, &mtotal) == hipSuccess &&
                     2 * bank_bytes + other_bytes < mfree / 4 * 3;
    D.nbank = two ? 2 : 1;
    int rc = small_block_tables(c);
    if (!rc) rc = adjoint_node_tables(c);
    if (!rc) rc = ensure_extreme_buffers(c);
    for (int b = 0; b < D.nbank && !rc; ++b) rc |= alloc_bank(c, D.bank[b], K, two);
    if (!rc && two) {
        if (hipStreamCreateWithFlags(&D.rs, hipStreamNonBlocking) != hipSuccess) rc = FSGH_E_HIP;
        for (int b = 0; b < 2 && !rc; ++b)
            if (hipEventCreateWithFlags(&D.ready[b], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&D.freed[b], hipEventDisableTiming) != hipSuccess)
                rc = FSGH_E_HIP;
    }
    for (int k = 0; k < 3; ++k) rc |= adj_alloc(c, &D.set[k], SW);
    rc |= adj_alloc(c, &D.A.Mb, E);
    rc |= adj_alloc(c, &D.A.mb, N);
    rc |= adj_alloc(c, &D.A.gb, E);
    rc |= adj_alloc(c, &D.A.rhob, E);
    rc |= adj_alloc(c, &D.A.ge, E);
    rc |= adj_alloc(c, &D.grho, E);
    rc |= adj_alloc(c, &D.A.hk, 6 * N);
    rc |= adj_alloc(c, &D.A.cbx, 3 * C);
    rc |= adj_alloc(c, &D.A.cbv, 3 * C);
    rc |= adj_alloc(c, &D.A.extb, E);
    rc |= adj_alloc_t(c, &D.A.rowset, E);
    rc |= adj_alloc_t(c, &D.A.kvflag, N);
    rc |= adj_alloc_t(c, &D.A.accflag, N);
    rc |= adj_alloc_t(c, &D.A.argext, E);
    rc |= adj_alloc(c, &D.A.nacc, 4 * N);
    if (rc) return rc == FSGH_E_HIP ? FSGH_E_HIP : FSGH_E_NOMEM;
    D.A.lo = D.lo;
    D.A.hi = D.hi;
    D.A.alo = D.alo;
    D.A.ahi = D.ahi;
    // the node records start finite (k_adj_node's quiet blocks keep their h)
    segcopy(c, {Seg{D.A.hk, nullptr, 6 * N}});
    D.rev_ready = true;
    return FSGH_OK;
}

int ensure_tlm_buffers(fsgh_ctx* c)
{
    AdjointData& D = *c->adj;
    if (D.tlm_ready) return FSGH_OK;
    const size_t N = D.N, E = D.E;
    int rc = ensure_extreme_buffers(c);
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 6; ++k)
            for (int j = 0; j < 2; ++j) rc |= adj_alloc(c, &D.tn[b][k][j], N);
    for (int b = 0; b < 2; ++b)
        for (int j = 0; j < 2; ++j) rc |= adj_alloc(c, &D.te[b][j], E);
    for (int k = 0; k < 6; ++k)
        for (int j = 0; j < 2; ++j) rc |= adj_alloc(c, &D.tel[k][j], E);
    for (int j = 0; j < 2; ++j) rc |= adj_alloc(c, &D.tF[j], 24 * E);
    rc |= adj_alloc(c, &D.tM, E);
    rc |= adj_alloc(c, &D.tm, N);
    rc |= adj_alloc(c, &D.tg, E);
    if (rc) return FSGH_E_NOMEM;
    D.tlm_ready = true;
    return FSGH_OK;
}

// TLM view: node bins 0 (t_n) / 1 (stage input and output); element
// divergence input slot `din`
TLMState tlm_state(fsgh_ctx* c, int din)
{
    AdjointData& D = *c->adj;
    TLMState s;
    for (int d = 0; d < 3; ++d) {
        s.x0[d] = TA{D.tn[0][d][0], D.tn[0][d][1]};
        s.v0[d] = TA{D.tn[0][3 + d][0], D.tn[0][3 + d][1]};
        s.x1[d] = TA{D.tn[1][d][0], D.tn[1][d][1]};
        s.v1[d] = TA{D.tn[1][3 + d][0], D.tn[1][3 + d][1]};
    }
    s.e0 = TA{D.te[0][0], D.te[0][1]};
    s.e1 = TA{D.te[1][0], D.te[1][1]};
    s.den = TA{D.tel[0][0], D.tel[0][1]};
    s.pres = TA{D.tel[1][0], D.tel[1][1]};
    s.sspd = TA{D.tel[2][0], D.tel[2][1]};
    s.vol = TA{D.tel[3][0], D.tel[3][1]};
    s.div_in = TA{D.tel[4 + din][0], D.tel[4 + din][1]};
    s.div_out = TA{D.tel[5 - din][0], D.tel[5 - din][1]};
    s.F = TA{D.tF[0], D.tF[1]};
    s.M = TA{c->d_emass, D.tM};
    s.m = TA{c->d_nmass, D.tm};
    s.g = TA{c->d_gamma, D.tg};
    s.csmin = c->d_csmin;
    s.mask = c->d_bcmask;
    s.mats = c->d_mats;
    s.mat = c->d_mat;
    s.nmat = c->nmat;
    s.lo = D.lo;
    s.hi = D.hi;
    s.tlo = D.tlo;
    s.thi = D.thi;
    s.alo = D.alo;
    s.ahi = D.ahi;
    return s;
}

}  // namespace

void fsgh::free_adjoint(fsgh_ctx* c)
{
    if (!c->adj) return;
    AdjointData& D = *c->adj;
    if (D.rs) {
        hipStreamSynchronize(D.rs);
        hipStreamDestroy(D.rs);
    }
    for (int b = 0; b < 2; ++b) {
        if (D.ready[b]) hipEventDestroy(D.ready[b]);
        if (D.freed[b]) hipEventDestroy(D.freed[b]);
    }
    for (hipGraphExec_t g : D.rgraph)
        if (g) hipGraphExecDestroy(g);
    for (double* p : c->adj->bufs) hipFree(p);
    delete c->adj;
    c->adj = nullptr;
}

extern "C" {

int fsgh_forward(fsgh_ctx* c, const fsgh_run_params* params, int64_t ncycles,
                 int64_t checkpoint_every, const double* dt_replay)
{
    if (!c) return FSGH_E_INVALID;
    if (c->dims == 2 || c->d_stress[0]) {
        set_error(c, "the adjoint covers the 3D strength-free cycle only");
        return FSGH_E_UNSUPPORTED;
    }
    if (c->host_only) {
        set_error(c, "host-only context");
        return FSGH_E_STATE;
    }
    if (!params || ncycles <= 0 || checkpoint_every <= 0) return FSGH_E_INVALID;
    if (c->nmat <= 0) {
        set_error(c, "fsgh_forward: materials not set");
        return FSGH_E_STATE;
    }
    if (any_advanced(c)) {
        set_error(c, "fsgh_forward: the state was advanced (fsgh_run, fsgh_forward or a "
                     "state-update operator) and its coordinates, velocities and energies "
                     "have not all been set again since (uploads, or fsgh_setup_fills for "
                     "velocities and energies); set the initial state again");
        return FSGH_E_STATE;
    }
    free_adjoint(c);
    c->adj = new AdjointData();
    AdjointData& D = *c->adj;
    D.ncyc = ncycles;
    D.every = checkpoint_every;
    D.ninst = c->L.ninst;
    D.N = static_cast<size_t>(c->N);
    D.E = static_cast<size_t>(c->E);
    const size_t E = D.E, SW = 6 * D.N + E;
    const size_t nchk = static_cast<size_t>((ncycles + checkpoint_every - 1) / checkpoint_every);
    int rc = adj_alloc(c, &D.dt, static_cast<size_t>(ncycles) * D.ninst);
    for (size_t k = 0; k < nchk && !rc; ++k) {
        double* p = nullptr;
        rc = adj_alloc(c, &p, SW);
        D.chk.push_back(p);
    }
    if (!rc) rc |= adj_alloc(c, &D.den0, E);
    if (!rc) rc |= adj_alloc(c, &D.pres0, E);
    if (!rc) rc |= adj_alloc(c, &D.sspd0, E);
    if (!rc) rc |= adj_alloc(c, &D.vol0, E);
    if (rc) {
        free_adjoint(c);
        return FSGH_E_NOMEM;
    }
    HIPCHK(c, hipGetLastError());
    if (dt_replay)
        HIPCHK(c, hipMemcpyAsync(D.dt, dt_replay, static_cast<size_t>(ncycles) * D.ninst * 8,
                                 hipMemcpyHostToDevice, c->stream));
    // the set-up element state: density rho0, EOS pressure and sound speed, volume V0
    segcopy(c, {Seg{D.den0, c->d_den, E}, Seg{D.pres0, c->d_pres, E}, Seg{D.sspd0, c->d_sspd, E},
                Seg{D.vol0, c->d_vol, E}});
    if ((rc = init_ctl(c, params))) return rc;
    c->divcur = 0;
    DMesh m = make_mesh(c);
    {
        DState s = make_state(c, 1 - c->cur, c->cur);
        launch_get_divergence(m, s, c->stream);
        if (!dt_replay)
            launch_dt_partial(tuned_mesh(c, kTunedEnergy), s, c->d_ctl, c->d_partial, 0, 0, 0, 0,
                              1, c->stream);
    }
    for (int64_t k = 0; k < ncycles; ++k) {
        if (k % checkpoint_every == 0)
            segcopy(c, bin_to_state(c, D.chk[static_cast<size_t>(k / checkpoint_every)], c->cur));
        const int old = c->cur;
        issue_cycle(c, old, dt_replay ? D.dt : nullptr, dt_replay ? nullptr : D.dt);
        c->cur = 1 - old;
    }
    c->divcur = 1;
    c->run_active = true;
    std::vector<InstCtl> h(static_cast<size_t>(c->L.ninst));
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_ctl, h.size() * sizeof(InstCtl), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto& q : h)
        if (q.cycle != ncycles) {
            set_error(c, "fsgh_forward: an instance stopped before ncycles (time_final or "
                         "cycle_stop); every instance must run all cycles");
            return FSGH_E_INVALID;
        }
    D.done = true;
    mark_advanced(c);
    return FSGH_OK;
}

int fsgh_dt_history(fsgh_ctx* c, double* out, int64_t n)
{
    if (!c || !out) return FSGH_E_INVALID;
    if (!c->adj || !c->adj->done) {
        set_error(c, "fsgh_dt_history before fsgh_forward");
        return FSGH_E_STATE;
    }
    if (n != c->adj->ncyc * c->adj->ninst) return FSGH_E_INVALID;
    HIPCHK(c, hipMemcpy(out, c->adj->dt, static_cast<size_t>(n) * 8, hipMemcpyDeviceToHost));
    return FSGH_OK;
}

int fsgh_backward(fsgh_ctx* c, const double* seed_x, const double* seed_v, const double* seed_e,
                  const double* seed_den, double* grad_den0, double* grad_sie0,
                  double* grad_gamma, double* grad_vel0)
{
    if (!c) return FSGH_E_INVALID;
    if (!c->adj || !c->adj->done) {
        set_error(c, "fsgh_backward before fsgh_forward");
        return FSGH_E_STATE;
    }
    AdjointData& D = *c->adj;
    if (D.N >= (size_t(1) << 29)) {
        set_error(c, "fsgh_backward: more than 2^29 nodes");
        return FSGH_E_UNSUPPORTED;
    }
    int rc = ensure_reverse_buffers(c);
    if (rc) return rc;
    const size_t N = D.N, E = D.E;
    DMesh m = make_mesh(c);
    // the recomputation rewrites the control blocks (cycle, dt): keep the forward's
    std::vector<InstCtl> ctl(static_cast<size_t>(c->L.ninst));
    HIPCHK(c, hipMemcpyAsync(ctl.data(), c->d_ctl, ctl.size() * sizeof(InstCtl),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // roles of the three adjoint sets: O = output of the cycle being reversed,
    // P = its input (accumulated), Q = input of its second stage (accumulated)
    int iO = 0, iP = 1, iQ = 2;
    std::vector<double> t;
    {
        const SetPtr O = set_ptr(D, D.set[iO]);
        for (int d = 0; d < 3 && !rc; ++d) {
            rc |= up_node(c, O.x[d], seed_x, d, t);
            rc |= up_node(c, O.v[d], seed_v, d, t);
        }
        if (!rc) rc = up_elem(c, O.e, seed_e, t);
        if (!rc && seed_den) rc = up_elem(c, D.grho, seed_den, t);
        if (rc) {
            set_error(c, "fsgh_backward: seed upload failed");
            return FSGH_E_HIP;
        }
    }
    segcopy(c, {Seg{D.A.Mb, nullptr, E}, Seg{D.A.mb, nullptr, N}, Seg{D.A.gb, nullptr, E},
                Seg{D.A.rhob, nullptr, E}});
    if (seed_den) {
        // the final density's seed, through M and the final coordinates
        const SetPtr O = set_ptr(D, D.set[iO]);
        DState sf = make_state(c, 1 - c->cur, c->cur);
        hipLaunchKernelGGL(k_adj_seed_den, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, sf,
                           static_cast<const double*>(D.grho), D.A.Mb, D.A.cbx, D.A.cbv);
        launch_adj_gather(adjoint_node_mesh(c), D.A.cbx, D.A.cbv, nullptr, O.x, O.v, 0,
                          c->stream);
    }
    const int64_t K = D.every;
    const int64_t nseg = (D.ncyc + K - 1) / K;
    auto seg_range = [&](int64_t sg, int64_t& c0, int64_t& c1) {
        c0 = sg * K;
        c1 = std::min(D.ncyc, c0 + K);
    };
    auto recompute = [&](int64_t sg) {
        int64_t c0, c1;
        seg_range(sg, c0, c1);
        const int b = D.nbank == 2 ? static_cast<int>(sg & 1) : 0;
        hipStream_t st = D.nbank == 2 ? D.rs : c->stream;
        recompute_segment(c, D.bank[b], D.chk[static_cast<size_t>(sg)], c0, c1, st);
        if (D.nbank == 2) HIPCHK(c, hipEventRecord(D.ready[b], D.rs));
        return FSGH_OK;
    };
    if (D.nbank == 2) {
        // the recompute stream starts after the seeds and accumulators are set
        HIPCHK(c, hipEventRecord(D.freed[0], c->stream));
        HIPCHK(c, hipEventRecord(D.freed[1], c->stream));
        HIPCHK(c, hipStreamWaitEvent(D.rs, D.freed[(nseg - 1) & 1], 0));
    }
    if ((rc = recompute(nseg - 1))) return rc;
    for (int64_t sg = nseg - 1; sg >= 0; --sg) {
        int64_t c0, c1;
        seg_range(sg, c0, c1);
        const int b = D.nbank == 2 ? static_cast<int>(sg & 1) : 0;
        const AdjointData::Bank& B = D.bank[b];
        if (D.nbank == 2) {
            HIPCHK(c, hipStreamWaitEvent(c->stream, D.ready[b], 0));
            if (sg > 0) {
                // the previous segment goes into the other bank once its last reader is done
                HIPCHK(c, hipStreamWaitEvent(D.rs, D.freed[(sg - 1) & 1], 0));
                if ((rc = recompute(sg - 1))) return rc;
            }
        }
        // A full segment other than the first is the same launch sequence
        // for a given bank and role parity: captured once, replayed.
        const bool graph = D.nbank == 2 && sg > 0 && c1 - c0 == K;
        const int key = 2 * b + iO;
        const bool replay = graph && D.rgraph[key] != nullptr;
        if (graph && !replay)
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int64_t cy = c1 - 1; cy >= c0; --cy) {
            const size_t j = static_cast<size_t>(cy - c0);
            if (!replay) {
                const SetPtr O = set_ptr(D, D.set[iO]), P = set_ptr(D, D.set[iP]),
                             Q = set_ptr(D, D.set[iQ]);
                double *S0[5], *EB[5];
                stage0_set(c, B, cy, j, S0);
                elem_set(D, B.eb[j], EB);
                ExtPtr x0, x1;
                if (!B.x0.empty()) {
                    x0 = ext_ptr(N, B.x0[j]);
                    x1 = ext_ptr(N, B.x1[j]);
                }
                const bool pre = !B.x0.empty();
                const double* dts = B.dt + j * static_cast<size_t>(D.ninst);
                // stage 1 (alpha 1): input sta[j] + eb[j]; output velocities of
                // seg[j+1]; writes P (cycle input) and Q (stage-1 input)
                reverse_stage(c, cy, 1, view(c, B.seg[j], B.sta[j], EB, nullptr),
                              view(c, B.seg[j], B.seg[j + 1], EB, nullptr), B.Fb1[j], 1.0, O, P,
                              Q, 1, pre ? &x1 : nullptr, dts);
                // stage 0 (alpha 1/2): input seg[j] + its element state; output
                // sta[j]; adds to P
                reverse_stage(c, cy, 0, view(c, B.seg[j], B.seg[j], S0, nullptr),
                              view(c, B.seg[j], B.sta[j], S0, nullptr), B.Fa[j], 0.5, Q, P, P, 0,
                              pre ? &x0 : nullptr, dts);
            }
            // the input adjoint of this cycle is the output adjoint of the previous one
            std::swap(iO, iP);
        }
        if (graph && !replay) {
            hipGraph_t gr;
            HIPCHK(c, hipStreamEndCapture(c->stream, &gr));
            const hipError_t ie = hipGraphInstantiate(&D.rgraph[key], gr, nullptr, nullptr, 0);
            hipGraphDestroy(gr);
            HIPCHK(c, ie);
        }
        if (graph) HIPCHK(c, hipGraphLaunch(D.rgraph[key], c->stream));
        if (D.nbank == 2) HIPCHK(c, hipEventRecord(D.freed[b], c->stream));
        else if (sg > 0 && (rc = recompute(sg - 1))) return rc;
    }
    const SetPtr G = set_ptr(D, D.set[iO]);
    {
        AdjView A = D.A;
        hipLaunchKernelGGL(k_adj_finish_elem, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, A,
                           static_cast<const double*>(D.vol0), D.grho);
    }
    HIPCHK(c, hipMemcpyAsync(c->d_ctl, ctl.data(), ctl.size() * sizeof(InstCtl),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    rc |= down_elem(c, grad_den0, D.grho, t);
    rc |= down_elem(c, grad_sie0, G.e, t);
    rc |= down_elem(c, grad_gamma, D.A.gb, t);
    for (int d = 0; d < 3; ++d) rc |= down_node(c, grad_vel0, G.v[d], d, t);
    if (rc) {
        set_error(c, "fsgh_backward: gradient download failed");
        return FSGH_E_HIP;
    }
    return FSGH_OK;
}

int fsgh_tangent(fsgh_ctx* c, const double* d_den0, const double* d_sie0, const double* d_gamma,
                 const double* d_vel0, double* t_x, double* t_v, double* t_e)
{
    if (!c) return FSGH_E_INVALID;
    if (!c->adj || !c->adj->done) {
        set_error(c, "fsgh_tangent before fsgh_forward");
        return FSGH_E_STATE;
    }
    AdjointData& D = *c->adj;
    int rc = ensure_tlm_buffers(c);
    if (rc) return rc;
    const size_t N = D.N, E = D.E;
    DMesh m = make_mesh(c);
    std::vector<double> t;
    // initial values (the forward's first checkpoint and set-up state) and tangents
    const double* S0 = D.chk[0];
    {
        std::vector<Seg> v;
        for (int k = 0; k < 6; ++k) v.push_back(Seg{D.tn[1][k][0], S0 + k * N, N});
        v.push_back(Seg{D.te[1][0], S0 + 6 * N, E});
        for (int d = 0; d < 3; ++d) v.push_back(Seg{D.tn[1][d][1], nullptr, N});
        v.push_back(Seg{D.tel[0][0], D.den0, E});
        v.push_back(Seg{D.tel[1][0], D.pres0, E});
        v.push_back(Seg{D.tel[2][0], D.sspd0, E});
        v.push_back(Seg{D.tel[3][0], D.vol0, E});
        v.push_back(Seg{D.tel[3][1], nullptr, E});
        segcopy(c, v);
    }
    for (int d = 0; d < 3; ++d) rc |= up_node(c, D.tn[1][3 + d][1], d_vel0, d, t);
    rc |= up_elem(c, D.te[1][1], d_sie0, t);
    rc |= up_elem(c, D.tel[0][1], d_den0, t);
    rc |= up_elem(c, D.tg, d_gamma, t);
    if (rc) {
        set_error(c, "fsgh_tangent: direction upload failed");
        return FSGH_E_HIP;
    }
    hipLaunchKernelGGL(k_mul, dim3(nb_for(E)), dim3(kB), 0, c->stream, D.tM,
                       static_cast<const double*>(D.vol0), static_cast<const double*>(D.tel[0][1]),
                       E);
    int din = 0;
    {
        TLMState s = tlm_state(c, din);
        hipLaunchKernelGGL(k_tlm_setup_node, dim3(m.n_nblk), dim3(kB), 0, c->stream, m, s);
        hipLaunchKernelGGL(k_tlm_setup_elem, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, s);
    }
    for (int64_t cy = 0; cy < D.ncyc; ++cy) {
        const double* dts = D.dt + static_cast<size_t>(cy) * D.ninst;
        {
            std::vector<Seg> v;
            for (int k = 0; k < 6; ++k)
                for (int j = 0; j < 2; ++j) v.push_back(Seg{D.tn[0][k][j], D.tn[1][k][j], N});
            for (int j = 0; j < 2; ++j) v.push_back(Seg{D.te[0][j], D.te[1][j], E});
            segcopy(c, v);
        }
        for (int stage = 0; stage < 2; ++stage) {
            const double alpha = stage == 0 ? 0.5 : 1.0;
            TLMState s = tlm_state(c, din);
            hipLaunchKernelGGL(k_arg_extremes, dim3(m.n_nblk), dim3(kB), 0, c->stream, m,
                               static_cast<const double*>(s.div_in.v),
                               static_cast<const double*>(s.div_in.t), D.lo, D.hi, D.alo, D.ahi,
                               D.tlo, D.thi);
            hipLaunchKernelGGL(k_tlm_force, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, s,
                               static_cast<const InstCtl*>(c->d_ctl));
            hipLaunchKernelGGL(k_tlm_node, dim3(m.n_nblk), dim3(kB), 0, c->stream, m, s, dts,
                               alpha);
            hipLaunchKernelGGL(k_tlm_energy_state, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, s,
                               dts, alpha);
            din = 1 - din;
        }
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int d = 0; d < 3; ++d) {
        rc |= down_node(c, t_x, D.tn[1][d][1], d, t);
        rc |= down_node(c, t_v, D.tn[1][3 + d][1], d, t);
    }
    rc |= down_elem(c, t_e, D.te[1][1], t);
    if (rc) {
        set_error(c, "fsgh_tangent: download failed");
        return FSGH_E_HIP;
    }
    return FSGH_OK;
}

}  // extern "C"
