// fsgh_internal.h -- shared declarations of the SGH HIP library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fsgh.h"

namespace fsgh {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, workgroups dealt round-robin

// One workgroup's work item: a contiguous range [start, end) of elements (or
// nodes) inside one instance.  `lid` is the logical (instance-ordered) block id.
struct BlockItem {
    int32_t inst;
    int32_t start;
    int32_t end;
    int32_t lid;
};

// Per-instance control block, device resident (dt and time never leave the
// device inside the cycle loop).
struct InstCtl {
    double dt;               // dt of the current cycle
    double time;             // time_value
    double graphics_time;
    double time_final, dt_max, dt_min, dt_cfl, graphics_dt_ival;
    double fuzz, small;
    double tiny;             // on-axis radius threshold of the RZ node mass (K/driver.cpp:61)
    long long cycle;         // cycles completed
    long long cycle_stop, graphics_cyc_ival;
    int32_t stopped;
    int32_t final_bin;       // physical bin holding the latest state once stopped
    int32_t eblk_begin, eblk_end;   // logical element-block range of the instance
    int32_t write_pending;   // the last cycle met the output condition (K/sgh_solve.cpp:594-606)
    int32_t pad_;
};

struct Material {
    double q1, q1ex, q2, q2ex;
    int32_t eos;               // FSGH_EOS_IDEAL_GAS or FSGH_EOS_USER
    int32_t strength;          // FSGH_STRENGTH_NONE / _HYPO / _HYPER
};
constexpr int kStrengthParams = FSGH_STRENGTH_PARAMS;   // parameters per material for the user strength model

// Device view of the mesh (new numbering, int32 indices).
struct DMesh {
    int32_t N, E;
    const int32_t* conn;       // E x 8
    const int32_t* cin_start;  // N+1
    const int32_t* cin;        // 8E corner ids, each node's list in ascending original-element order
    const BlockItem* eblk;     // element work items, XCD-ordered
    const BlockItem* nblk;     // node work items, XCD-ordered
    int32_t n_eblk, n_nblk;
    int32_t bs;                // entries per work item = workgroup size of the tables above
    int32_t minw;              // occupancy floor (waves per SIMD) of the kernel variant to run
    int32_t cshift;            // log2(corners per element): 3 for hexes, 2 for RZ quads
    int32_t pad_;
    // per logical element block of this partition (bs elements): its distinct
    // nodes (sn_nodes[sn_start[b] .. sn_start[b+1]), ascending) and each
    // corner's position in that list (sn_lconn, 8 per hex, 4 per RZ quad);
    // null when not built (a partition without lists)
    const int32_t* sn_start;
    const int32_t* sn_nodes;
    const uint16_t* sn_lconn;
    // per logical node block of this partition: its corner-list entries'
    // corners sorted ascending (cs_cor) and each entry's position among them
    // (cs_pos), both indexed like cin; null when not built
    const int32_t* cs_cor;
    const uint16_t* cs_pos;
};

// LDS capacity (corner-list entries) of a node block of bs nodes in the
// corner-staged velocity kernel; larger blocks gather from global memory
constexpr int corner_stage_cap(int bs) { return bs * 9; }

// Launch configuration of one fused-cycle kernel, chosen by fsgh_autotune.
// How the default work tables deal work items to the XCDs: chunks of 16
// consecutive items round-robin (see block_table)
constexpr int32_t kDefaultDeal = 16;

struct KernelCfg {
    int32_t bs = kBlock;       // workgroup size (64, 128, 256; 512 for the node kernels)
    int32_t minw = 1;          // waves-per-SIMD floor asked of the register allocator
    int32_t xcd = kDefaultDeal;  // XCD dealing: 0 launch order, 1 contiguous, k chunks of k
};
enum TunedKernel { kTunedDivExtremes = 0, kTunedForce, kTunedVelBcPos, kTunedEnergy, kNumTuned };

// Device view of the state (structure of arrays); bin 0 = t_n, bin 1 = current.
struct DState {
    double *x0, *y0, *z0, *x1, *y1, *z1;
    double *u0, *v0, *w0, *u1, *v1, *w1;
    double* nmass;
    const uint8_t* bcmask;     // per node: bit d set -> velocity component d forced to 0
    double *den, *pres, *sspd, *sie0, *sie1, *vol, *div, *emass, *gamma, *csmin;
    double* div_out;           // where the fused energy kernel writes the next stage's div
    double2* ext_lo;           // per node: two smallest element divergences (with multiplicity)
    double2* ext_hi;           // per node: two largest element divergences
    const int32_t* mat;        // material id per element
    double* F;                 // 8E x 3 corner forces
    const Material* mats;
    int32_t nmat;
    int32_t user_eos;          // some material uses the user EOS (else all are ideal gas)
    // Cauchy stress, SoA [9][E] row-major components, or all null (zero
    // stress): sig0 = t_n, sig1 = the current stage's input, sig1w = where the
    // force kernel writes the stress after the hypo-elastic model (sig1 except
    // on the first stage of a fused cycle, whose input is t_n)
    double *sig0, *sig1, *sig1w;
    const double* sparams;     // kStrengthParams per material
};

void set_error(fsgh_ctx* ctx, const std::string& msg);

}  // namespace fsgh

// ---- kernel launchers (fsgh_kernels.hip) ----
namespace fsgh {
void launch_get_vol(const DMesh&, const DState&, hipStream_t);
void launch_get_divergence(const DMesh&, const DState&, hipStream_t);
// the (bs, minw) pairs compiled for a tuned kernel (TunedKernel id)
bool tuned_variant_ok(int kernel, int32_t bs, int32_t minw);
void launch_force(const DMesh&, const DState&, const InstCtl*, double fuzz, double small,
                  hipStream_t);
// with a stress field (hypo-elastic hook after the forces); dt / alpha are
// taken from ctl when it is given
void launch_force_stress(const DMesh&, const DState&, const InstCtl*, double fuzz, double small,
                         double dt, double alpha, hipStream_t);
void launch_energy_state_div_stress(const DMesh&, const DState&, const InstCtl*, double alpha,
                                    double* dt_partial, hipStream_t);
// velocity gradient of the current state, E x 9 (row-major)
void launch_velgrad(const DMesh&, const DState&, double* out, hipStream_t);
// update_state with stress (hyper-elastic hook before the EOS)
void launch_update_state_stress(const DMesh&, const DState&, double dt, double alpha,
                                hipStream_t);
void launch_div_extremes(const DMesh&, const DState&, const InstCtl*, hipStream_t);
void launch_update_velocity(const DMesh&, const DState&, double alpha, double dt, hipStream_t);
// the divergence with the block's nodes staged in LDS (mesh with node lists)
void launch_get_divergence_staged(const DMesh&, const DState&, hipStream_t);
// the energy update with the block's nodes staged in LDS (mesh with node lists)
void launch_update_energy_staged(const DMesh&, const DState&, double alpha, double dt,
                                 hipStream_t);
// the same update through the staged node kernel (mesh with sorted corner lists)
void launch_update_velocity_staged(const DMesh&, const DState&, double alpha, double dt,
                                   hipStream_t);
void launch_boundary_velocity(const int32_t* nodes, const uint8_t* masks, int32_t count,
                              const DState&, hipStream_t);
void launch_update_energy(const DMesh&, const DState&, double alpha, double dt, hipStream_t);
void launch_update_position(const DMesh&, const DState&, double alpha, double dt, hipStream_t);
void launch_update_state(const DMesh&, const DState&, hipStream_t);
void launch_rk_init(const DMesh&, const DState&, hipStream_t);
// dt partial minima per logical element block (dt_lcl of K/time_integration.cpp:74-140)
void launch_dt_partial(const DMesh&, const DState&, const InstCtl*, double* partial,
                       double dt_cfl, double dt_max, double dt_min, double fuzz, int use_ctl,
                       hipStream_t);
void launch_reduce_min(const double* partial, int32_t n, double* out, hipStream_t);
// fused cycle kernels
// dt_replay / dt_record (layout [cycle][instance]) may be null
void launch_cycle_dt(InstCtl* ctl, int32_t ninst, const double* partial, const double* dt_replay,
                     double* dt_record, hipStream_t);
void launch_vel_bc_pos(const DMesh&, const DState&, const InstCtl*, double alpha, hipStream_t);
void launch_energy_state_div(const DMesh&, const DState&, const InstCtl*, double alpha,
                             double* dt_partial, hipStream_t);
void launch_time_advance(InstCtl* ctl, int32_t ninst, int32_t newest_bin, hipStream_t);
// rz != 0: kinetic energy weighted by the node radius (K/sgh_solve.cpp:92-97)
void launch_energy_sums(const DMesh&, const DState&, const int32_t* eoff, const int32_t* noff,
                        int32_t inst, double* out2, double* scratch, int32_t nscratch, int rz,
                        hipStream_t);
// ---- 2D axisymmetric (RZ) path (fsgh_rz.hip) ----
void launch_rz_get_vol(const DMesh&, const DState&, hipStream_t);
void launch_rz_get_divergence(const DMesh&, const DState&, hipStream_t);
// with a stress field in s: tau = stress - p I, hoop terms of the stress
// and the hypo-elastic hook; dt / alpha from ctl when given
void launch_rz_force(const DMesh&, const DState&, const InstCtl*, double fuzz, double small,
                     double dt, double alpha, hipStream_t);
void launch_rz_velgrad(const DMesh&, const DState&, double* out, hipStream_t);
void launch_rz_update_velocity(const DMesh&, const DState&, double alpha, double dt, hipStream_t);
void launch_rz_update_energy(const DMesh&, const DState&, double alpha, double dt, hipStream_t);
void launch_rz_dt_partial(const DMesh&, const DState&, const InstCtl*, double* partial,
                          double dt_cfl, double dt_max, double dt_min, double fuzz, int use_ctl,
                          hipStream_t);
void launch_rz_ext(const DMesh&, const DState&, double* ext, hipStream_t);
void launch_rz_node_mass(const DMesh&, const DState&, const double* ext, double tiny,
                         const int2* axis, int32_t naxis, hipStream_t);
void launch_rz_node(const DMesh&, const DState&, const InstCtl*, double alpha, const double* rsrc,
                    double* rin, const double* ext, const int2* axis, int32_t naxis, hipStream_t);
void launch_rz_energy_state_div(const DMesh&, const DState&, const InstCtl*, double alpha,
                                const double* rin, double* dt_partial, hipStream_t);
void launch_swap_bins_range(const DState&, int32_t n0, int32_t n1, int32_t e0, int32_t e1,
                            int32_t Etot, hipStream_t);
void launch_check(const DMesh&, const DState&, int32_t* flag, hipStream_t);
}  // namespace fsgh
