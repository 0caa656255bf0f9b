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
//   k_adj_energy   e1 = e0 - alpha dt/M sum F.(v1+v0)/2     -> Fb, hb (per corner), Mb, e0b
//   k_adj_node     v1 = mask(v0 + alpha dt sum F/m), x1 = x0 + alpha dt (v1+v0)/2
//                  (gathers hb in corner order)             -> x0b, v0b, Fb, mb
//   k_adj_force    corner forces (MARS + shock detector)    -> per-corner (xb, vb), element
//                  (den, pres, sspd, vol, div, ext) adjoints
//   k_adj_derived  neighbour-divergence adjoint (the extreme neighbour chosen by the
//                  forward), then vol/den/EOS/div of the stage input -> eb, Mb, gb, corner xb, vb
//   k_adj_gather   corner adjoints -> node adjoints, in corner order (no atomics)
// Branches (compression test, min/max, clamps, the sound-speed floor) follow
// the forward values; ties of the neighbour extreme go to the first element
// found, the same choice the tangent-linear model makes.
//
// Checkpointing: the state (x, v, e) is stored every `checkpoint_every`
// cycles; each segment is recomputed forward (storing its states) before it
// is reversed; inside a cycle, stage 0 is recomputed from the cycle's state.
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
__device__ __forceinline__ void ins_lo(double& a, double& b, int& ia, int& ib, double v, int e)
{
    if (v < a) { b = a; ib = ia; a = v; ia = e; }
    else if (v < b) { b = v; ib = e; }
}
__device__ __forceinline__ void ins_hi(double& a, double& b, int& ia, int& ib, double v, int e)
{
    if (v > a) { b = a; ib = ia; a = v; ia = e; }
    else if (v > b) { b = v; ib = e; }
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
    for (int q = m.cin_start[n]; q < m.cin_start[n + 1]; ++q) {
        const int e = m.cin[q] >> 3;
        const double v = div[e];
        if (isnan(v)) continue;
        ins_lo(a1, a2, ia1, ia2, v, e);
        ins_hi(b1, b2, ib1, ib2, v, e);
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
__device__ __forceinline__ double neighbour_extreme(const int32_t nd[8], int e, bool low,
                                                    const double2* lo, const double2* hi,
                                                    const int2* alo, const int2* ahi, int& arg,
                                                    const double2* tlo, const double2* thi,
                                                    double* tan)
{
    double ext = low ? HUGE_VAL : -HUGE_VAL, te = 0.0;
    arg = -1;
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
            if (tan != nullptr) {
                const double2 tt = low ? tlo[nd[a]] : thi[nd[a]];
                te = skip ? tt.y : tt.x;
            }
        }
    }
    if (tan != nullptr) *tan = te;
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
    derived_state(x, y, z, u1, v1, w1, M, sie, s.g.ld(e), s.csmin[e], vol, den, p, c, div);
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
    ideal_gas(s.g.ld(e), s.csmin[e], s.den.ld(e), s.e1.ld(e), p, c);
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
    double *Fb, *hb, *cb;       // corners: 3, 3, 6
    double* eadj;               // element: den, pres, sspd, vol, div, ext (6 x E)
    int32_t* argext;
    double2 *lo, *hi;
    int2 *alo, *ahi;
};

// e1 = e0 - alpha dt / M * sum_c F_c . (v1+v0)/2
__global__ __launch_bounds__(kB) void k_adj_energy(DMesh m, DState s, const double* F,
                                                   AdjView A, const double* dts, double alpha)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    const double k = alpha * dts[bi.inst];
    const double eb = A.eo[e];
    A.e0b[e] += eb;
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    const double M = s.emass[e];
    const double g = -k / M * eb;
    double P = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const size_t c = 8 * static_cast<size_t>(e) + a;
        const double h[3] = {(s.u1[nd[a]] + s.u0[nd[a]]) * 0.5, (s.v1[nd[a]] + s.v0[nd[a]]) * 0.5,
                             (s.w1[nd[a]] + s.w0[nd[a]]) * 0.5};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double f = F[3 * c + d];
            P += f * h[d];
            A.Fb[3 * c + d] = g * h[d];
            A.hb[3 * c + d] = g * f;
        }
    }
    A.Mb[e] += k / (M * M) * eb * P;
}

// v1 = mask(v0 + alpha dt F/m), x1 = x0 + alpha dt (v1 + v0)/2, plus the
// energy update's dependence on v1 and v0 through hb.
__global__ __launch_bounds__(kB) void k_adj_node(DMesh m, DState s, const double* F, AdjView A,
                                                 const double* dts, double alpha)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    const double k = alpha * dts[bi.inst];
    const int q0 = m.cin_start[n], q1 = m.cin_start[n + 1];
    double sh[3] = {0.0, 0.0, 0.0}, sf[3] = {0.0, 0.0, 0.0};
    for (int q = q0; q < q1; ++q) {
        const size_t c = static_cast<size_t>(m.cin[q]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            sh[d] += A.hb[3 * c + d];
            sf[d] += F[3 * c + d];
        }
    }
    const double mass = s.nmass[n];
    const uint8_t mk = s.bcmask[n];
    double vm[3];
    double dot = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double xb = A.xo[d][n];
        A.x0b[d][n] += xb;
        const double vt = A.vo[d][n] + 0.5 * k * xb + 0.5 * sh[d];
        vm[d] = (mk & (1u << d)) ? 0.0 : vt;
        A.v0b[d][n] += 0.5 * k * xb + 0.5 * sh[d] + vm[d];
        dot += vm[d] * sf[d];
    }
    A.mb[n] += -k * dot / (mass * mass);
    for (int q = q0; q < q1; ++q) {
        const size_t c = static_cast<size_t>(m.cin[q]);
#pragma unroll
        for (int d = 0; d < 3; ++d) A.Fb[3 * c + d] += k / mass * vm[d];
    }
}

// Reverse of force_elem (K/force_sgh.cpp:35-359) at the stage input state.
__global__ __launch_bounds__(kB) void k_adj_force(DMesh m, DState s, AdjView A, const InstCtl* ctl,
                                                  const double* vol_a, const double* div_a)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    const double fuzz = ctl[bi.inst].fuzz, small = ctl[bi.inst].small;
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    double x[8], y[8], z[8], vel[3][8];
    g3(s.x1, s.y1, s.z1, nd, x, y, z);
    g3(s.u1, s.v1, s.w1, nd, vel[0], vel[1], vel[2]);
    const double vol = vol_a[e], div = div_a[e], den = s.den[e], sspd = s.sspd[e],
                 pres = s.pres[e];
    const Material mt = s.mats[s.nmat > 1 ? s.mat[e] : 0];
    const bool low = div + small >= 0.0;
    int arg;
    const double ext = neighbour_extreme(nd, e, low, A.lo, A.hi, A.alo, A.ahi, arg, nullptr,
                                         nullptr, nullptr);
    // ---- forward recompute ----
    double B[24];
    hex_bmatrix(x, y, z, B);
    double vbar[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) vbar[d] += 0.125 * vel[d][a];
    const double qa = div < 0.0 ? mt.q1 : mt.q1ex;
    const double qb = div < 0.0 ? mt.q2 : mt.q2ex;
    double dl[8][3], mv[8], muc[8], nb[8], Mt[8], S[3] = {0.0, 0.0, 0.0}, W = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
#pragma unroll
        for (int d = 0; d < 3; ++d) dl[a][d] = vel[d][a] - vbar[d];
        mv[a] = sqrt(dl[a][0] * dl[a][0] + dl[a][1] * dl[a][1] + dl[a][2] * dl[a][2]);
        muc[a] = den * (qa * sspd + qb * mv[a]);
        nb[a] = sqrt(B[3 * a] * B[3 * a] + B[3 * a + 1] * B[3 * a + 1] + B[3 * a + 2] * B[3 * a + 2]);
        Mt[a] = muc[a] * nb[a];
#pragma unroll
        for (int d = 0; d < 3; ++d) S[d] += Mt[a] * vel[d][a];
        W += Mt[a];
    }
    const bool riemann = W > fuzz;
    double vs[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) vs[d] = riemann ? S[d] / W : 0.0;
    const double bden = div + small;
    const double r = 0.9 * (ext + small) / bden;
    const double rmin = fmin(r, 1.0);
    const double phi0 = 1.0 - fmax(0.0, rmin);
    const double cl = pow(vol, 1.0 / 3.0);
    const double cd = sspd + fuzz;
    const double argv = 20.0 * (cl * fabs(div)) / cd;
    const double alpha = fmin(1.0, argv);
    const double phi = alpha * phi0;
    // ---- reverse ----
    double Bb[24], velb[3][8], Mb[8];
    double presb = 0.0, phib = 0.0, vsb[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const size_t c = 8 * static_cast<size_t>(e) + a;
        double g = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double fb = A.Fb[3 * c + d];
            presb += B[3 * a + d] * fb;
            Bb[3 * a + d] = pres * fb;
            g += fb * (vs[d] - vel[d][a]);
            vsb[d] += phi * Mt[a] * fb;
            velb[d][a] = -phi * Mt[a] * fb;
        }
        phib += Mt[a] * g;
        Mb[a] = phi * g;
    }
    if (riemann) {
        double Wb = 0.0, Sb[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Sb[d] = vsb[d] / W;
            Wb -= vsb[d] * vs[d] / W;
        }
#pragma unroll
        for (int a = 0; a < 8; ++a) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                velb[d][a] += Mt[a] * Sb[d];
                Mb[a] += Sb[d] * vel[d][a];
            }
            Mb[a] += Wb;
        }
    }
    double denb = 0.0, sspdb = 0.0, vbarb[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const double mucb = Mb[a] * nb[a];
        const double nbb = Mb[a] * muc[a];
        if (nb[a] > 0.0) {
#pragma unroll
            for (int d = 0; d < 3; ++d) Bb[3 * a + d] += nbb * B[3 * a + d] / nb[a];
        }
        denb += mucb * (qa * sspd + qb * mv[a]);
        sspdb += mucb * den * qa;
        const double mvb = mucb * den * qb;
        if (mv[a] > 0.0) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double db = mvb * dl[a][d] / mv[a];
                velb[d][a] += db;
                vbarb[d] -= db;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) velb[d][a] += 0.125 * vbarb[d];
    double divb = 0.0, volb = 0.0, extb = 0.0;
    const double alphab = phib * phi0;
    const double phi0b = phib * alpha;
    if (argv < 1.0) {
        const double clb = alphab * 20.0 * fabs(div) / cd;
        divb += alphab * 20.0 * cl * (div < 0.0 ? -1.0 : 1.0) / cd;
        sspdb -= alphab * argv / cd;
        if (vol > 0.0) volb += clb * cl / (3.0 * vol);
    }
    if (rmin > 0.0 && r <= 1.0) {
        const double rb = -phi0b;
        extb = rb * 0.9 / bden;
        divb -= rb * r / bden;
    }
    double xb[8] = {0}, yb[8] = {0}, zb[8] = {0};
    hex_bmatrix_vjp(x, y, z, Bb, xb, yb, zb);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        double* cb = A.cb + 6 * (8 * static_cast<size_t>(e) + a);
        cb[0] = xb[a];
        cb[1] = yb[a];
        cb[2] = zb[a];
        cb[3] = velb[0][a];
        cb[4] = velb[1][a];
        cb[5] = velb[2][a];
    }
    const size_t E = static_cast<size_t>(m.E);
    A.eadj[0 * E + e] = denb;
    A.eadj[1 * E + e] = presb;
    A.eadj[2 * E + e] = sspdb;
    A.eadj[3 * E + e] = volb;
    A.eadj[4 * E + e] = divb;
    A.eadj[5 * E + e] = (arg >= 0) ? extb : 0.0;
    A.argext[e] = arg;
}

// Neighbour divergence adjoint, then the stage-input state (vol, den, EOS,
// div).  setup != 0: the first stage of the run, whose density, pressure and
// sound speed come from the initial conditions (den = rho0) and whose
// coordinates are the fixed initial mesh.
__global__ __launch_bounds__(kB) void k_adj_derived(DMesh m, DState s, AdjView A,
                                                    const double* vol_a, const double* div_a,
                                                    const double* sie_in, double* sieb_in,
                                                    int setup)
{
    const BlockItem bi = m.eblk[blockIdx.x];
    const int e = bi.start + static_cast<int>(threadIdx.x);
    if (e >= bi.end) return;
    const size_t E = static_cast<size_t>(m.E);
    int32_t nd[8];
    ld_conn(m.conn, e, nd);
    // elements whose shock detector picked this element's divergence; each
    // neighbour is counted once (at the first shared node in local order)
    double divb = A.eadj[4 * E + e];
    for (int a = 0; a < 8; ++a) {
        for (int q = m.cin_start[nd[a]]; q < m.cin_start[nd[a] + 1]; ++q) {
            const int e2 = m.cin[q] >> 3;
            if (e2 == e || A.argext[e2] != e) continue;
            int32_t n2[8];
            ld_conn(m.conn, e2, n2);
            bool earlier = false;
            for (int b = 0; b < a && !earlier; ++b)
                for (int k = 0; k < 8; ++k) earlier = earlier || n2[k] == nd[b];
            if (!earlier) divb += A.eadj[5 * E + e2];
        }
    }
    const double denb = A.eadj[0 * E + e], presb = A.eadj[1 * E + e], sspdb = A.eadj[2 * E + e];
    double volb = A.eadj[3 * E + e];
    double x[8], y[8], z[8], vel[3][8];
    g3(s.x1, s.y1, s.z1, nd, x, y, z);
    g3(s.u1, s.v1, s.w1, nd, vel[0], vel[1], vel[2]);
    const double vol = vol_a[e], div = div_a[e];
    const double gam = s.gamma[e], csmin = s.csmin[e];
    const double sie = sie_in[e];
    const double den = s.den[e];
    // EOS (K/eos.cpp:35-43)
    double eb = presb * (gam - 1.0) * den;
    double dnb = denb + presb * (gam - 1.0) * sie;
    double gb = presb * sie * den;
    const double c = sqrt(gam * (gam - 1.0) * sie);
    if (!(c < csmin) && c > 0.0) {
        eb += sspdb * gam * (gam - 1.0) / (2.0 * c);
        gb += sspdb * (2.0 * gam - 1.0) * sie / (2.0 * c);
    }
    sieb_in[e] += eb;
    A.gb[e] += gb;
    // divergence (K/momentum.cpp:265-286): div = (sum_a v_a . B_a) / vol
    double B[24];
    hex_bmatrix(x, y, z, B);
    const double iv = 1.0 / vol;
    double Bb[24];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Bb[3 * a + d] = divb * iv * vel[d][a];
            double* cb = A.cb + 6 * (8 * static_cast<size_t>(e) + a);
            cb[3 + d] += divb * iv * B[3 * a + d];
        }
    volb -= divb * div * iv;
    if (setup) {
        // den = rho0; vol and B belong to the fixed initial mesh
        A.rhob[e] += dnb;
        return;
    }
    // den = M / vol
    A.Mb[e] += dnb / vol;
    volb -= dnb * den / vol;
    // vol = V(x): dV/dx_a = B_a
    double xb[8] = {0}, yb[8] = {0}, zb[8] = {0};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        xb[a] = volb * B[3 * a];
        yb[a] = volb * B[3 * a + 1];
        zb[a] = volb * B[3 * a + 2];
    }
    hex_bmatrix_vjp(x, y, z, Bb, xb, yb, zb);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        double* cb = A.cb + 6 * (8 * static_cast<size_t>(e) + a);
        cb[0] += xb[a];
        cb[1] += yb[a];
        cb[2] += zb[a];
    }
}

// corner adjoints -> node adjoints, summed in corner order
__global__ __launch_bounds__(kB) void k_adj_gather(DMesh m, const double* cb, double* xa0,
                                                   double* xa1, double* xa2, double* va0,
                                                   double* va1, double* va2)
{
    const BlockItem bi = m.nblk[blockIdx.x];
    const int n = bi.start + static_cast<int>(threadIdx.x);
    if (n >= bi.end) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int q = m.cin_start[n]; q < m.cin_start[n + 1]; ++q) {
        const double* p = cb + 6 * static_cast<size_t>(m.cin[q]);
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] += p[k];
    }
    xa0[n] += s[0];
    xa1[n] += s[1];
    xa2[n] += s[2];
    va0[n] += s[3];
    va1[n] += s[4];
    va2[n] += s[5];
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

__global__ void k_fill(double* p, size_t n, double v)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy(double* dst, const double* src, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) dst[i] = src[i];
}


// cycle counter of every instance (selects the replayed dt); clears `stopped`
__global__ void k_set_cycle(InstCtl* ctl, int32_t ninst, long long cycle)
{
    const int i = blockIdx.x * kB + static_cast<int>(threadIdx.x);
    if (i >= ninst) return;
    ctl[i].cycle = cycle;
    ctl[i].stopped = 0;
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
    derived_state(x, y, z, u, v, w, s.emass[e], s.sie1[e], s.gamma[e], s.csmin[e], vol, den, p,
                  c, div);
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
    std::vector<double*> seg;             // recomputed states of one segment (every + 1)
    double *den0 = nullptr, *pres0 = nullptr, *sspd0 = nullptr, *vol0 = nullptr;
    double* fin = nullptr;                // final forward state, restored after a backward pass
    std::vector<double*> bufs;            // everything allocated, released together
    bool done = false;
    // reverse pass
    bool rev_ready = false;
    AdjView A{};
    double *Fa = nullptr, *vol_in = nullptr, *div_in = nullptr, *esa = nullptr, *grho = nullptr;
    double *xab[3] = {}, *vab[3] = {};
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

void fill(fsgh_ctx* c, double* p, size_t n, double v)
{
    if (n) hipLaunchKernelGGL(k_fill, dim3(nb_for(n)), dim3(kB), 0, c->stream, p, n, v);
}
void copy(fsgh_ctx* c, double* dst, const double* src, size_t n)
{
    if (n) hipLaunchKernelGGL(k_copy, dim3(nb_for(n)), dim3(kB), 0, c->stream, dst, src, n);
}

// state layout of checkpoint/segment buffers: x, y, z, u, v, w (N each), e (E)
void save_state(fsgh_ctx* c, double* dst, int bin)
{
    const size_t N = static_cast<size_t>(c->N), E = static_cast<size_t>(c->E);
    for (int k = 0; k < 6; ++k) copy(c, dst + k * N, c->d_node[bin][k], N);
    copy(c, dst + 6 * N, c->d_sie[bin], E);
}
void load_state(fsgh_ctx* c, const double* src, int bin)
{
    const size_t N = static_cast<size_t>(c->N), E = static_cast<size_t>(c->E);
    for (int k = 0; k < 6; ++k) copy(c, c->d_node[bin][k], src + k * N, N);
    copy(c, c->d_sie[bin], src + 6 * N, E);
}

// every mutable state array of the context (both bins + element state)
std::vector<std::pair<double*, size_t>> state_arrays(fsgh_ctx* c)
{
    const size_t N = static_cast<size_t>(c->N), E = static_cast<size_t>(c->E);
    std::vector<std::pair<double*, size_t>> v;
    for (int b = 0; b < 2; ++b) {
        for (int k = 0; k < 6; ++k) v.emplace_back(c->d_node[b][k], N);
        v.emplace_back(c->d_sie[b], E);
    }
    for (double* p : {c->d_den, c->d_pres, c->d_sspd, c->d_vol, c->d_div[0], c->d_div[1]})
        v.emplace_back(p, E);
    return v;
}
size_t state_words(fsgh_ctx* c) { return 12 * static_cast<size_t>(c->N) + 8 * static_cast<size_t>(c->E); }

void set_cycle(fsgh_ctx* c, int64_t cycle)
{
    hipLaunchKernelGGL(k_set_cycle, dim3(nb_for(static_cast<size_t>(c->L.ninst))), dim3(kB), 0,
                       c->stream, c->d_ctl, c->L.ninst, static_cast<long long>(cycle));
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
        k.eblk_begin = c->eblk_begin[static_cast<size_t>(i)];
        k.eblk_end = c->eblk_begin[static_cast<size_t>(i) + 1];
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

// Recompute cycle `cyc` from `state` (loaded into both bins) up to the
// stage-1 forces.  Leaves: Fa = stage-0 forces, bin 0 = S_c, bin 1 = stage-0
// output, element arrays = stage-1 input state, F = stage-1 forces,
// vol_in/div_in[s*E] = stage-s input volume and divergence.
void recompute_cycle(fsgh_ctx* c, int64_t cyc, const double* state)
{
    AdjointData& D = *c->adj;
    const size_t E = D.E;
    DMesh m = make_mesh(c);
    load_state(c, state, 0);
    load_state(c, state, 1);
    DState s = make_state(c, 0, 1);
    s.div = c->d_div[0];
    if (cyc == 0) {
        // the set-up element state of the run
        copy(c, c->d_den, D.den0, E);
        copy(c, c->d_pres, D.pres0, E);
        copy(c, c->d_sspd, D.sspd0, E);
        copy(c, c->d_vol, D.vol0, E);
        launch_get_divergence(m, s, c->stream);
    } else {
        hipLaunchKernelGGL(k_derived, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, s);
    }
    copy(c, D.vol_in, c->d_vol, E);
    copy(c, D.div_in, c->d_div[0], E);
    set_cycle(c, cyc);
    launch_cycle_dt(c->d_ctl, c->L.ninst, c->d_partial, D.dt, nullptr, c->stream);
    DState s0 = make_state(c, 0, 0);
    s0.div = c->d_div[0];
    launch_div_extremes(m, s0, c->d_ctl, c->stream);
    launch_force(m, s0, c->d_ctl, 0.0, 0.0, c->stream);
    copy(c, D.Fa, c->d_F, 24 * E);
    DState st = make_state(c, 0, 1);
    st.div = c->d_div[0];
    st.div_out = c->d_div[1];
    launch_vel_bc_pos(m, st, c->d_ctl, 0.5, c->stream);
    launch_energy_state_div(m, st, c->d_ctl, 0.5, nullptr, c->stream);
    copy(c, D.vol_in + E, c->d_vol, E);
    copy(c, D.div_in + E, c->d_div[1], E);
    DState s1 = make_state(c, 0, 1);
    s1.div = c->d_div[1];
    launch_div_extremes(m, s1, c->d_ctl, c->stream);
    launch_force(m, s1, c->d_ctl, 0.0, 0.0, c->stream);
}

// Reverse one RK stage.  s_in: stage input (x1..w1, element arrays);
// s_out: stage output velocities in u1..w1, cycle state in bin 0;
// F: the stage's corner forces.  Output adjoints are in A.xo/vo/eo; the
// adjoints of the cycle state accumulate into A.x0b/v0b/e0b, those of the
// stage input into (xdst, vdst, ein_b).
void reverse_stage(fsgh_ctx* c, int64_t cyc, int stage, const DState& s_in, const DState& s_out,
                   const double* F, double alpha, const double* ein, double* ein_b,
                   double* const* xdst, double* const* vdst)
{
    AdjointData& D = *c->adj;
    const size_t E = D.E;
    DMesh m = make_mesh(c);
    const double* dts = D.dt + static_cast<size_t>(cyc) * D.ninst;
    const AdjView A = D.A;
    const double* vol_a = D.vol_in + stage * E;
    const double* div_a = D.div_in + stage * E;
    hipLaunchKernelGGL(k_adj_energy, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, s_out, F, A, dts,
                       alpha);
    hipLaunchKernelGGL(k_adj_node, dim3(m.n_nblk), dim3(kB), 0, c->stream, m, s_out, F, A, dts,
                       alpha);
    hipLaunchKernelGGL(k_arg_extremes, dim3(m.n_nblk), dim3(kB), 0, c->stream, m, div_a,
                       static_cast<const double*>(nullptr), D.lo, D.hi, D.alo, D.ahi,
                       static_cast<double2*>(nullptr), static_cast<double2*>(nullptr));
    hipLaunchKernelGGL(k_adj_force, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, s_in, A,
                       static_cast<const InstCtl*>(c->d_ctl), vol_a, div_a);
    hipLaunchKernelGGL(k_adj_derived, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, s_in, A, vol_a,
                       div_a, ein, ein_b, (cyc == 0 && stage == 0) ? 1 : 0);
    hipLaunchKernelGGL(k_adj_gather, dim3(m.n_nblk), dim3(kB), 0, c->stream, m,
                       static_cast<const double*>(A.cb), xdst[0], xdst[1], xdst[2], vdst[0],
                       vdst[1], vdst[2]);
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

int ensure_reverse_buffers(fsgh_ctx* c)
{
    AdjointData& D = *c->adj;
    if (D.rev_ready) return FSGH_OK;
    const size_t N = D.N, E = D.E, C = 8 * E;
    const size_t SW = 6 * N + E;
    int rc = ensure_extreme_buffers(c);
    for (int64_t k = 0; k <= D.every && !rc; ++k) {
        double* p = nullptr;
        rc = adj_alloc(c, &p, SW);
        D.seg.push_back(p);
    }
    for (int d = 0; d < 3; ++d) {
        rc |= adj_alloc(c, &D.A.xo[d], N);
        rc |= adj_alloc(c, &D.A.vo[d], N);
        rc |= adj_alloc(c, &D.A.x0b[d], N);
        rc |= adj_alloc(c, &D.A.v0b[d], N);
        rc |= adj_alloc(c, &D.xab[d], N);
        rc |= adj_alloc(c, &D.vab[d], N);
    }
    rc |= adj_alloc(c, &D.A.eo, E);
    rc |= adj_alloc(c, &D.A.e0b, E);
    rc |= adj_alloc(c, &D.esa, E);
    rc |= adj_alloc(c, &D.A.Mb, E);
    rc |= adj_alloc(c, &D.A.mb, N);
    rc |= adj_alloc(c, &D.A.gb, E);
    rc |= adj_alloc(c, &D.A.rhob, E);
    rc |= adj_alloc(c, &D.grho, E);
    rc |= adj_alloc(c, &D.A.Fb, 3 * C);
    rc |= adj_alloc(c, &D.A.hb, 3 * C);
    rc |= adj_alloc(c, &D.A.cb, 6 * C);
    rc |= adj_alloc(c, &D.A.eadj, 6 * E);
    rc |= adj_alloc_t(c, &D.A.argext, E);
    rc |= adj_alloc(c, &D.Fa, 3 * C);
    rc |= adj_alloc(c, &D.vol_in, 2 * E);
    rc |= adj_alloc(c, &D.div_in, 2 * E);
    if (rc) return FSGH_E_NOMEM;
    D.A.lo = D.lo;
    D.A.hi = D.hi;
    D.A.alo = D.alo;
    D.A.ahi = D.ahi;
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

// TLM view: node bins b0 (t_n) / b1 (stage input and output); element
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
    for (double* p : c->adj->bufs) hipFree(p);
    delete c->adj;
    c->adj = nullptr;
}

extern "C" {

int fsgh_forward(fsgh_ctx* c, const fsgh_run_params* params, int64_t ncycles,
                 int64_t checkpoint_every, const double* dt_replay)
{
    if (!c) return FSGH_E_INVALID;
    if (c->host_only) {
        set_error(c, "host-only context");
        return FSGH_E_STATE;
    }
    if (!params || ncycles <= 0 || checkpoint_every <= 0) return FSGH_E_INVALID;
    if (c->nmat <= 0) {
        set_error(c, "fsgh_forward: materials not set");
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
    if (!rc) rc |= adj_alloc(c, &D.fin, state_words(c));
    if (rc) {
        free_adjoint(c);
        return FSGH_E_NOMEM;
    }
    HIPCHK(c, hipGetLastError());
    if (dt_replay)
        HIPCHK(c, hipMemcpyAsync(D.dt, dt_replay, static_cast<size_t>(ncycles) * D.ninst * 8,
                                 hipMemcpyHostToDevice, c->stream));
    // the set-up element state: density rho0, EOS pressure and sound speed, volume V0
    copy(c, D.den0, c->d_den, E);
    copy(c, D.pres0, c->d_pres, E);
    copy(c, D.sspd0, c->d_sspd, E);
    copy(c, D.vol0, c->d_vol, E);
    if ((rc = init_ctl(c, params))) return rc;
    c->divcur = 0;
    DMesh m = make_mesh(c);
    {
        DState s = make_state(c, 1 - c->cur, c->cur);
        launch_get_divergence(m, s, c->stream);
        if (!dt_replay) launch_dt_partial(m, s, c->d_ctl, c->d_partial, 0, 0, 0, 0, 1, c->stream);
    }
    for (int64_t k = 0; k < ncycles; ++k) {
        if (k % checkpoint_every == 0)
            save_state(c, D.chk[static_cast<size_t>(k / checkpoint_every)], c->cur);
        const int old = c->cur;
        issue_cycle(c, old, dt_replay ? D.dt : nullptr, dt_replay ? nullptr : D.dt);
        c->cur = 1 - old;
    }
    c->divcur = 1;
    c->run_active = true;
    size_t off = 0;
    for (auto& a : state_arrays(c)) {
        copy(c, D.fin + off, a.first, a.second);
        off += a.second;
    }
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
                  double* grad_den0, double* grad_sie0, double* grad_gamma, double* grad_vel0)
{
    if (!c) return FSGH_E_INVALID;
    if (!c->adj || !c->adj->done) {
        set_error(c, "fsgh_backward before fsgh_forward");
        return FSGH_E_STATE;
    }
    AdjointData& D = *c->adj;
    int rc = ensure_reverse_buffers(c);
    if (rc) return rc;
    const size_t N = D.N, E = D.E, SW = 6 * N + E;
    DMesh m = make_mesh(c);
    std::vector<double> t;
    for (int d = 0; d < 3 && !rc; ++d) {
        rc |= up_node(c, D.A.xo[d], seed_x, d, t);
        rc |= up_node(c, D.A.vo[d], seed_v, d, t);
    }
    if (!rc) rc = up_elem(c, D.A.eo, seed_e, t);
    if (rc) {
        set_error(c, "fsgh_backward: seed upload failed");
        return FSGH_E_HIP;
    }
    fill(c, D.A.Mb, E, 0.0);
    fill(c, D.A.mb, N, 0.0);
    fill(c, D.A.gb, E, 0.0);
    fill(c, D.A.rhob, E, 0.0);
    const int64_t K = D.every;
    const int64_t nseg = (D.ncyc + K - 1) / K;
    for (int64_t sg = nseg - 1; sg >= 0; --sg) {
        const int64_t c0 = sg * K, c1 = std::min(D.ncyc, c0 + K);
        // recompute the segment, keeping every cycle's state
        copy(c, D.seg[0], D.chk[static_cast<size_t>(sg)], SW);
        for (int64_t cy = c0; cy < c1; ++cy) {
            recompute_cycle(c, cy, D.seg[static_cast<size_t>(cy - c0)]);
            DState st = make_state(c, 0, 1);
            st.div = c->d_div[1];
            st.div_out = c->d_div[0];
            launch_vel_bc_pos(m, st, c->d_ctl, 1.0, c->stream);
            launch_energy_state_div(m, st, c->d_ctl, 1.0, nullptr, c->stream);
            save_state(c, D.seg[static_cast<size_t>(cy - c0 + 1)], 1);
        }
        for (int64_t cy = c1 - 1; cy >= c0; --cy) {
            const double* Sc = D.seg[static_cast<size_t>(cy - c0)];
            const double* Sn = D.seg[static_cast<size_t>(cy - c0 + 1)];
            recompute_cycle(c, cy, Sc);
            for (int d = 0; d < 3; ++d) {
                fill(c, D.A.x0b[d], N, 0.0);
                fill(c, D.A.v0b[d], N, 0.0);
                fill(c, D.xab[d], N, 0.0);
                fill(c, D.vab[d], N, 0.0);
            }
            fill(c, D.A.e0b, E, 0.0);
            fill(c, D.esa, E, 0.0);
            // stage 1 (alpha 1): input = stage-0 output (bin 1) + element arrays
            DState s_in = make_state(c, 0, 1);
            DState s_out = make_state(c, 0, 1);
            s_out.u1 = const_cast<double*>(Sn + 3 * N);
            s_out.v1 = const_cast<double*>(Sn + 4 * N);
            s_out.w1 = const_cast<double*>(Sn + 5 * N);
            double* xa[3] = {D.xab[0], D.xab[1], D.xab[2]};
            double* va[3] = {D.vab[0], D.vab[1], D.vab[2]};
            reverse_stage(c, cy, 1, s_in, s_out, c->d_F, 1.0, c->d_sie[1], D.esa, xa, va);
            // stage 0 (alpha 1/2): outputs = stage-1 inputs
            for (int d = 0; d < 3; ++d) {
                copy(c, D.A.xo[d], D.xab[d], N);
                copy(c, D.A.vo[d], D.vab[d], N);
            }
            copy(c, D.A.eo, D.esa, E);
            if (cy == 0) {
                copy(c, c->d_den, D.den0, E);
                copy(c, c->d_pres, D.pres0, E);
                copy(c, c->d_sspd, D.sspd0, E);
            } else {
                DState s = make_state(c, 0, 0);
                s.div = c->d_div[0];
                hipLaunchKernelGGL(k_derived, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, s);
            }
            DState s0_in = make_state(c, 0, 0);
            DState s0_out = make_state(c, 0, 1);
            double* x0[3] = {D.A.x0b[0], D.A.x0b[1], D.A.x0b[2]};
            double* v0[3] = {D.A.v0b[0], D.A.v0b[1], D.A.v0b[2]};
            reverse_stage(c, cy, 0, s0_in, s0_out, D.Fa, 0.5, c->d_sie[0], D.A.e0b, x0, v0);
            // the cycle input adjoint is the previous cycle's output adjoint
            for (int d = 0; d < 3; ++d) {
                copy(c, D.A.xo[d], D.A.x0b[d], N);
                copy(c, D.A.vo[d], D.A.v0b[d], N);
            }
            copy(c, D.A.eo, D.A.e0b, E);
        }
    }
    hipLaunchKernelGGL(k_adj_finish_elem, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, D.A, D.vol0,
                       D.grho);
    // restore the state the forward pass ended with
    size_t off = 0;
    for (auto& a : state_arrays(c)) {
        copy(c, a.first, D.fin + off, a.second);
        off += a.second;
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    rc |= down_elem(c, grad_den0, D.grho, t);
    rc |= down_elem(c, grad_sie0, D.A.eo, t);
    rc |= down_elem(c, grad_gamma, D.A.gb, t);
    for (int d = 0; d < 3; ++d) rc |= down_node(c, grad_vel0, D.A.vo[d], d, t);
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
    for (int k = 0; k < 6; ++k) copy(c, D.tn[1][k][0], S0 + k * N, N);
    copy(c, D.te[1][0], S0 + 6 * N, E);
    for (int d = 0; d < 3; ++d) {
        fill(c, D.tn[1][d][1], N, 0.0);
        rc |= up_node(c, D.tn[1][3 + d][1], d_vel0, d, t);
    }
    rc |= up_elem(c, D.te[1][1], d_sie0, t);
    rc |= up_elem(c, D.tel[0][1], d_den0, t);
    rc |= up_elem(c, D.tg, d_gamma, t);
    if (rc) {
        set_error(c, "fsgh_tangent: direction upload failed");
        return FSGH_E_HIP;
    }
    copy(c, D.tel[0][0], D.den0, E);
    copy(c, D.tel[1][0], D.pres0, E);
    copy(c, D.tel[2][0], D.sspd0, E);
    copy(c, D.tel[3][0], D.vol0, E);
    fill(c, D.tel[3][1], E, 0.0);
    hipLaunchKernelGGL(k_mul, dim3(nb_for(E)), dim3(kB), 0, c->stream, D.tM, D.vol0,
                       D.tel[0][1], E);
    int din = 0;
    {
        TLMState s = tlm_state(c, din);
        hipLaunchKernelGGL(k_tlm_setup_node, dim3(m.n_nblk), dim3(kB), 0, c->stream, m, s);
        hipLaunchKernelGGL(k_tlm_setup_elem, dim3(m.n_eblk), dim3(kB), 0, c->stream, m, s);
    }
    for (int64_t cy = 0; cy < D.ncyc; ++cy) {
        const double* dts = D.dt + static_cast<size_t>(cy) * D.ninst;
        for (int k = 0; k < 6; ++k)
            for (int j = 0; j < 2; ++j) copy(c, D.tn[0][k][j], D.tn[1][k][j], N);
        for (int j = 0; j < 2; ++j) copy(c, D.te[0][j], D.te[1][j], E);
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
