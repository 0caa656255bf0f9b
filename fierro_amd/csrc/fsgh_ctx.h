// fsgh_ctx.h -- the context behind the opaque fsgh_ctx handle (library-internal).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "fsgh_internal.h"
#include "fsgh_mesh.h"

struct AdjointData;  // fsgh_adjoint.hip

struct fsgh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    fsgh::Layout L;
    int64_t N = 0, E = 0;
    int renumber = 0;
    int dims = 3;                           // 3: hexes; 2: axisymmetric RZ quads
    int npe = 8;                            // nodes (= corners) per element
    std::vector<int64_t> conn;              // caller numbering, E x npe
    std::vector<double> coords;             // caller numbering, N x 3 (initial; z = 0 in 2D)
    std::vector<int32_t> ne_of_oe, nn_of_on, oe_of_ne, on_of_nn;
    std::vector<int64_t> cstart, corners;   // caller-numbering corner CSR
    std::vector<int32_t> eblk_begin, nblk_begin;
    // device mesh
    int32_t *d_conn = nullptr, *d_cin_start = nullptr, *d_cin = nullptr;
    fsgh::BlockItem *d_eblk = nullptr, *d_nblk = nullptr;
    int32_t n_eblk = 0, n_nblk = 0;
    // device state: node[phys bin][x,y,z,u,v,w], sie[phys bin]
    double* d_node[2][6] = {};
    double* d_sie[2] = {};
    double *d_nmass = nullptr, *d_den = nullptr, *d_pres = nullptr, *d_sspd = nullptr;
    double *d_vol = nullptr, *d_emass = nullptr, *d_gamma = nullptr;
    // two divergence buffers: stage s's force reads d_div[s] and the fused
    // energy kernel of stage s writes d_div[1-s]; operators use d_div[divcur]
    double* d_div[2] = {nullptr, nullptr};
    int divcur = 0;
    double2 *d_ext_lo = nullptr, *d_ext_hi = nullptr;
    double *d_csmin = nullptr, *d_F = nullptr;
    int32_t* d_mat = nullptr;
    uint8_t* d_bcmask = nullptr;
    int32_t *d_on_of_nn = nullptr, *d_oe_of_ne = nullptr;   // internal -> caller ids
    fsgh::Material* d_mats = nullptr;
    std::vector<fsgh::Material> mats;   // host copy of the material table
    // Cauchy stress per physical bin ([9][E] each), allocated when a material
    // gets a strength model or a stress field is uploaded; strength
    // parameters per material
    double* d_stress[2] = {nullptr, nullptr};
    double* d_sparams = nullptr;
    int32_t nmat = 0;
    fsgh::InstCtl* d_ctl = nullptr;
    double* d_partial = nullptr;
    size_t partial_cap = 0;                 // doubles in d_partial
    // launch configuration of the fused-cycle kernels (fsgh_autotune) and the
    // work tables they need, keyed by (element/node, workgroup size, XCD order)
    fsgh::KernelCfg tune[fsgh::kNumTuned];
    struct WorkTable {
        fsgh::BlockItem* d = nullptr;
        int32_t n = 0;
        std::vector<int32_t> inst_begin;
    };
    std::map<int32_t, WorkTable> tables;
    // per element block size: the blocks' distinct-node lists that the
    // LDS-staged element kernels read (fsgh_stage.hip)
    struct NodeStage {
        int32_t* start = nullptr;
        int32_t* nodes = nullptr;
        uint16_t* lconn = nullptr;
        int64_t total = 0;
        int32_t max_nodes = 0;
    };
    std::map<int32_t, NodeStage> stages;
    // per node block size: the blocks' sorted corner lists that the
    // corner-staged velocity kernel reads (fsgh_stage.hip)
    struct CornerStage {
        int32_t* scor = nullptr;
        uint16_t* spos = nullptr;
    };
    std::map<int32_t, CornerStage> cstages;
    double* d_scalar = nullptr;             // 2 doubles
    double* d_scratch = nullptr;            // energy partial sums
    int32_t nscratch = 1024;
    int32_t* d_flag = nullptr;
    int32_t* d_bcnodes = nullptr;
    uint8_t* d_bcmasks = nullptr;
    int32_t n_bcnodes = 0;
    std::vector<uint8_t> bcmask;            // caller numbering
    std::vector<std::vector<std::vector<int64_t>>> bcsets;  // [inst][set] caller node ids
    std::vector<int64_t> eoff32, noff32;
    int cur = 1;                            // physical bin holding reference bin 1
    bool run_active = false;
    // per instance: which of coordinates (1), velocities (2) and energies (4)
    // were set again (uploads; fsgh_setup_fills sets 2 and 4) since the state
    // was last advanced (fsgh_run, fsgh_forward, the state-update operators);
    // 7 for an instance never advanced
    std::vector<uint8_t> set_since;
    bool host_only = false;                 // device < 0: host-side maps only
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // profiling
    bool prof = false;
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::map<std::string, std::pair<double, int64_t>> prof_acc;
    std::vector<std::string> prof_order;
    // discrete adjoint / tangent state (allocated by fsgh_forward)
    AdjointData* adj = nullptr;
    // 2D RZ: node extensive mass (areal mass x radius), the stage-input node
    // radius of the fused loop, and the mesh boundary nodes (internal id,
    // instance) that the on-axis mass fix visits
    double* d_ext = nullptr;
    double* d_rin = nullptr;
    int2* d_axis = nullptr;
    int32_t n_axis = 0;
};


namespace fsgh {
// shared by the ABI translation units
DState make_state(fsgh_ctx* c, int b0, int b1);
// the per-instance record of what was set since the state last advanced:
// every instance advanced; parts (bits of coordinates 1, velocities 2,
// energies 4) set for one instance (inst < 0: every instance); whether any
// instance is advanced and not wholly set again
void mark_advanced(fsgh_ctx* c);
void mark_set(fsgh_ctx* c, int32_t inst, uint8_t parts);
bool any_advanced(const fsgh_ctx* c);
DMesh make_mesh(fsgh_ctx* c);
// the mesh view with the work tables of a tuned kernel's configuration
DMesh tuned_mesh(fsgh_ctx* c, int kernel);
// per-instance logical block ranges of the element table that the fused
// energy kernel (and the dt partial minima) use
const std::vector<int32_t>& dt_block_begin(fsgh_ctx* c);
// the mesh view whose element table the dt partial minima follow
DMesh dt_mesh(fsgh_ctx* c);
// point a running loop's dt block ranges at dt_block_begin and recompute the
// partial minima from the current state
int resync_dt_ranges(fsgh_ctx* c);
// build (once) the work table of a configuration (and, for an element table
// with `lists`, its per-block node lists); FSGH_OK or an error code
int ensure_table(fsgh_ctx* c, bool elem, int32_t bs, int32_t xcd, bool lists = true);
// element and node tables of 64 entries per work item (XCD-dealt), and the
// mesh view over them: the reverse-pass kernels, which hold up to 512
// registers per lane, spread a small mesh over every CU this way
int small_block_tables(fsgh_ctx* c);
DMesh small_block_mesh(fsgh_ctx* c);
// the node table of the reverse pass's node kernels (kAdjNodeBlock nodes per
// work item, chunks of kAdjNodeDeal items per XCD) with its sorted corner
// lists, and the mesh view over it
constexpr int32_t kAdjNodeBlock = 128;
constexpr int32_t kAdjNodeDeal = 4;
int adjoint_node_tables(fsgh_ctx* c);
DMesh adjoint_node_mesh(fsgh_ctx* c);
// per-block node lists of the element partition with `bs` elements per block
// (3D device contexts; no-op otherwise), and the mesh view's pointers to the
// lists of the partition its element table `m.eblk` has (ebs per block)
int build_node_stage(fsgh_ctx* c, int32_t bs);
void attach_node_stage(fsgh_ctx* c, DMesh& m, int32_t ebs);
// sorted corner lists of the node partition with `bs` nodes per block, and
// the mesh view's pointers to those of its node table's partition (nbs)
int build_corner_stage(fsgh_ctx* c, int32_t bs);
void attach_corner_stage(fsgh_ctx* c, DMesh& m, int32_t nbs);
void free_node_stages(fsgh_ctx* c);
void issue_cycle(fsgh_ctx* c, int old, const double* dt_replay, double* dt_record);
// 2D RZ contexts (fsgh_rz.hip)
int rz_build_maps(fsgh_ctx* c);
int rz_setup_fills(fsgh_ctx* c, int32_t inst, int32_t nfills, const fsgh_fill* fills,
                   int32_t nmat, int32_t nsv, const double* state_vars, const int32_t* mat_nsv);
void issue_cycle_rz(fsgh_ctx* c, int old, const double* dt_replay, double* dt_record);
void free_adjoint(fsgh_ctx* c);
// mesh preparation on the device (fsgh_setup.hip)
int gpu_build_maps(fsgh_ctx* c);
int gpu_boundary_sets(fsgh_ctx* c, int32_t inst, int32_t nsets, const fsgh_boundary* bc,
                      std::vector<std::vector<int64_t>>& out);
int gpu_setup_fills(fsgh_ctx* c, int32_t inst, int32_t nfills, const fsgh_fill* fills,
                    int32_t nmat, int32_t nsv, const double* state_vars, const int32_t* mat_nsv);
// Launch `fn` with optional per-kernel event timing on the context stream
// (fsgh_profile_enable); the records are resolved by fsgh_profile_read.
template <typename Fn>
void timed(fsgh_ctx* c, const char* name, Fn fn)
{
    if (!c->prof) {
        fn();
        return;
    }
    fsgh_ctx::Rec r;
    r.name = name;
    hipEventCreate(&r.a);
    hipEventCreate(&r.b);
    hipEventRecord(r.a, c->stream);
    fn();
    hipEventRecord(r.b, c->stream);
    c->recs.push_back(r);
}

}  // namespace fsgh

#define CHECK_CTX(c)                                   \
    do {                                               \
        if ((c) == nullptr) return FSGH_E_INVALID;     \
    } while (0)

#define HIPCHK(ctx, expr)                                                                 \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            set_error(ctx, std::string(#expr) + ": " + hipGetErrorString(e_));            \
            return FSGH_E_HIP;                                                            \
        }                                                                                 \
    } while (0)
