/*
 * fsgh.h -- C ABI of the MI355X-native SGH (staggered-grid hydrodynamics) cycle.
 *
 * The library replaces the per-stage operators of Fierro's explicit
 * Lagrangian solver (reference: Explicit-Lagrange-Kokkos/SGH_solver/, cited
 * below as K/) -- 3D hexahedra and the 2D axisymmetric (RZ) quadrilateral
 * path -- with hand-written HIP kernels for gfx950, behind plain C entry
 * points: no C++ or torch types cross this boundary.
 *
 * Model of use (mirrors K/driver.cpp:94-313):
 *   fsgh_create            mesh + coordinates in   (K/read_mesh.cpp, K/mesh.h:180-245)
 *   fsgh_set_materials     q1,q1ex,q2,q2ex         (K/state.h:130-178)
 *   fsgh_set_bcs           boundary sets           (K/setup.cpp:431-665)
 *   fsgh_setup_fills       initial conditions      (K/setup.cpp:106-424)
 *     or fsgh_upload of every state field
 *   per-stage operators    exactly the reference's sequence, or
 *   fsgh_run               the fused, graph-captured cycle loop (K/sgh_solve.cpp:125-634)
 *   fsgh_download / fsgh_energy
 *
 * Conventions
 *   - Every call returns FSGH_OK (0) or a negative FSGH_E_* code;
 *     fsgh_last_error() gives the message.  No exception crosses the ABI.
 *   - Field layouts at the ABI are the reference's (row-major MATAR):
 *     coordinates/velocity [2 bins][N][3], sie [2 bins][E], corner force
 *     [8E][3] with corner = 8e+i.  Indices are int64 (reference: size_t).
 *     A 2D RZ context (fsgh_mesh_desc.num_dims = 2) has num_dims = 2 in
 *     these shapes: coordinates/velocity [2 bins][N][2] with (x, y) =
 *     (axial, radius), corner force [4E][2], nodes_in_elem E x 4.
 *     Internally the library keeps SoA FP64 fields and int32 connectivity,
 *     optionally renumbered along a Morton curve; uploads and downloads
 *     convert, so callers always see their own numbering.
 *   - Operator calls are asynchronous on the context's HIP stream; download,
 *     energy and any call that returns a host value synchronise.
 *   - A context must not be used by two host threads at once.
 *   - A batched context holds many independent instances packed as one ragged
 *     mesh (instance i owns nodes [node_off[i], node_off[i+1]) and elements
 *     [elem_off[i], elem_off[i+1])); every instance has its own dt, time,
 *     boundary sets and stop state.  A plain mesh is a batch of one.
 */
#ifndef FSGH_H
#define FSGH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSGH_OK 0
#define FSGH_E_INVALID (-1)     /* bad argument, size mismatch, NULL handle */
#define FSGH_E_HIP (-2)         /* HIP runtime error */
#define FSGH_E_NOMEM (-3)       /* device or host allocation failed */
#define FSGH_E_NONFINITE (-4)   /* non-finite or non-physical state found by fsgh_check */
#define FSGH_E_STATE (-5)       /* call out of order (e.g. run before setup) */
#define FSGH_E_UNSUPPORTED (-6) /* feature outside this build (e.g. strength models) */
#define FSGH_E_NOTBUILT (-7)    /* library built without GPU code (never in a product build) */

typedef struct fsgh_ctx fsgh_ctx;

/* ---------------------------------------------------------------- mesh --- */
typedef struct fsgh_mesh_desc {
    int64_t num_nodes;                 /* N, total over instances */
    int64_t num_elems;                 /* E, total over instances */
    int32_t num_instances;             /* 1 for a plain mesh */
    int32_t renumber;                  /* 0: keep caller numbering, 1: Morton order */
    const int64_t *instance_node_off;  /* num_instances+1 entries, or NULL if 1 instance */
    const int64_t *instance_elem_off;  /* num_instances+1 entries, or NULL if 1 instance */
    int32_t num_dims;                  /* 3 (or 0): hexahedra; 2: axisymmetric RZ quads */
    int32_t reserved_;
} fsgh_mesh_desc;

/* Creates a context on HIP device `device`: builds the corner connectivity
 * (corners_in_node in ascending element order, K/mesh.h:180-245), copies the
 * coordinates into both RK bins (K/read_mesh.cpp:129-137), allocates device
 * state.  nodes_in_elem: E x 8 (E x 4 counter-clockwise for num_dims = 2),
 * 0-based, node ids global over the batch; coords N x num_dims.  Replaces:
 * read_mesh_ensight + mesh.build_corner_connectivity (K/driver.cpp:138-139).
 * A 2D context builds its maps, boundary sets and fills on the host (the
 * reference's set-up, untimed); its cycle runs on the GPU. */
int fsgh_create(const fsgh_mesh_desc *desc, const int64_t *nodes_in_elem,
                const double *coords, int device, fsgh_ctx **out);
/* device < 0 creates a host-only context: connectivity, renumbering and
 * boundary sets are built (and can be read back) but nothing touches a GPU;
 * operators and field transfers return FSGH_E_STATE. */
int fsgh_destroy(fsgh_ctx *ctx);
const char *fsgh_last_error(const fsgh_ctx *ctx);
/* 1 when the library holds gfx950 device code and a device is usable. */
int fsgh_device_count(void);

/* corners_in_node as CSR in the caller's numbering (start: N+1, data: 8E).
 * Replaces the mesh_t ragged array read by K/momentum.cpp:31-34. */
int fsgh_get_corners_in_node(fsgh_ctx *ctx, int64_t *start, int64_t *corners);
/* Morton permutations: new_of_old_elem[E], new_of_old_node[N] (identity when
 * renumber == 0). */
int fsgh_get_renumbering(fsgh_ctx *ctx, int64_t *new_of_old_elem, int64_t *new_of_old_node);
/* Diagnostic: the device's corner CSR in internal numbering (start [N+1],
 * corner ids 8 * internal element + local node [8E]); each node's list is in
 * ascending caller element order.  Device contexts only. */
int fsgh_get_device_csr(fsgh_ctx *ctx, int32_t *start, int32_t *corners);
/* Diagnostic (no counterpart in the reference): the per-block node lists of
 * the LDS-staged element kernels for element blocks of `elem_block`
 * elements (64, 128, 256 or 512): total list entries (sum over blocks of
 * their distinct nodes) and the largest block's count; zeros when not built
 * (a block size no kernel configuration uses yet; 2D contexts build 256). */
int fsgh_get_stage_info(fsgh_ctx *ctx, int32_t elem_block, int64_t *list_entries,
                        int32_t *max_block_nodes);

/* ------------------------------------------------------------ materials --- */
typedef struct fsgh_material {
    double q1, q1ex, q2, q2ex;         /* K/state.h:174-177 */
} fsgh_material;
/* EOS is the ideal gas (K/eos.cpp:12-46) with gamma = statev[e][0] and
 * min sound speed = statev[e][1], unless fsgh_set_material_models selects
 * another model for a material. */
int fsgh_set_materials(fsgh_ctx *ctx, int32_t num_materials, const fsgh_material *mats);
/* Material model hooks: the reference's per-material eos_model and
 * strength_model pointers (K/state.h:140-170, set in K/input.cpp:115-124).
 * eos: FSGH_EOS_IDEAL_GAS (K/eos.cpp:12-46) or FSGH_EOS_USER, the device
 * function fsgh::user_eos of fierro_amd/csrc/fsgh_user_eos.h compiled into
 * the library (the counterpart of user_eos_model, K/user_mat.cpp:14-40).
 * strength: FSGH_STRENGTH_NONE, or FSGH_STRENGTH_HYPO / _HYPER, which call the
 * device function fsgh::user_strength (fierro_amd/csrc/fsgh_user_strength.h,
 * the counterpart of user_strength_model, K/user_mat.cpp:47-90) from the
 * force kernel after the corner forces (K/force_sgh.cpp:369-391) or from the
 * state update before the EOS (K/properties.cpp:46-95), with the element's
 * velocity gradient (K/momentum.cpp:60-148).  Any strength model allocates a
 * stress field (zero, K/setup.cpp:214-220) that the force kernels then use
 * (tau = stress - p I, K/force_sgh.cpp:127-137) and rk_init carries.  2D:
 * HYPO only.  The adjoint and launch tuning cover the strength-free cycle.
 * Call after fsgh_set_materials and before fsgh_setup_fills. */
#define FSGH_EOS_IDEAL_GAS 0
#define FSGH_EOS_USER 1
#define FSGH_STRENGTH_NONE 0
#define FSGH_STRENGTH_HYPO 1
#define FSGH_STRENGTH_HYPER 2
int fsgh_set_material_models(fsgh_ctx *ctx, int32_t material, int32_t eos, int32_t strength);
/* Parameters (up to FSGH_STRENGTH_PARAMS) passed to the user strength model
 * of `material` (zero until set). */
#define FSGH_STRENGTH_PARAMS 8
int fsgh_set_strength_params(fsgh_ctx *ctx, int32_t material, int32_t num_params,
                             const double *params);

/* ----------------------------------------------------- boundary sets --- */
typedef struct fsgh_boundary {
    int32_t surface;                   /* 0 x-plane, 1 y-plane, 2 z-plane, 3 cylinder, 4 sphere */
    int32_t hydro_bc;                  /* 0 fixed, 1 reflected; other values: no-op */
    double value;                      /* plane position or radius */
} fsgh_boundary;
/* Tags the boundary patches of `instance` whose 4 nodes (2 in 2D) lie
 * within 1e-7 of the surface (K/setup.cpp:487-571) and stores each set's
 * nodes.  Replaces mesh.init_bdy_sets + tag_bdys + build_boundry_node_sets.
 * 2D contexts accept surfaces 0, 1 and 3 (the reference's check_bdy reads an
 * unset third coordinate for 2 and 4). */
int fsgh_set_bcs(fsgh_ctx *ctx, int32_t instance, int32_t num_sets,
                 const fsgh_boundary *sets);
/* number of nodes in boundary set `set` of `instance`, and the nodes (caller
 * ids) in the reference's order: first seen over the boundary patches in
 * patch order -- ascending caller element, then face, then the face's node
 * order -- as build_boundry_node_sets stores bdy_nodes_in_set
 * (K/mesh.h:353-376,435-513; K/setup.cpp:591-642). */
int fsgh_bc_set_size(fsgh_ctx *ctx, int32_t instance, int32_t set, int64_t *count);
int fsgh_get_bc_set(fsgh_ctx *ctx, int32_t instance, int32_t set, int64_t *nodes);

/* ------------------------------------------------- initial conditions --- */
typedef struct fsgh_fill {
    int32_t volume;                    /* 0 global, 1 box, 2 cylinder, 3 sphere */
    int32_t mat_id;
    double x1, x2, y1, y2, z1, z2;     /* box */
    double radius1, radius2;           /* cylinder / sphere */
    int32_t velocity;                  /* 0 cartesian, 1 radial, 2 spherical, 3 and 4 no-ops,
                                          5 Taylor-Green vortex (K/setup.cpp:326-350) */
    int32_t pad_;
    double u, v, w, speed;
    double sie, den;
} fsgh_fill;
/* Paints fills in order over `instance` by element centroid, sets M = rho V,
 * e, statev, EOS, node velocity, applies the boundary velocity and the nodal
 * mass sum_{elems of n} M_e/8 (K/setup.cpp:106-424).  state_vars is
 * num_materials x num_state_vars; mat_num_state_vars[m] of them are copied.
 * 2D: V is the RZ volume per radian and the node mass is the areal mass
 * sum over the node's corners of (Barlow corner area x density)
 * (K/setup.cpp:365-420); velocity kinds 0, 1 and 5 (2: FSGH_E_UNSUPPORTED).
 * Kind 5 also sets the element's pressure 0.25 (cos 2 pi x + cos 2 pi y) + 1
 * and its energy p / (rho (statev[4] - 1)) (the material's state variable 4,
 * zero when it has fewer than five), and leaves the sound speed of the
 * fill's energy, as the reference does. */
int fsgh_setup_fills(fsgh_ctx *ctx, int32_t instance, int32_t num_fills,
                     const fsgh_fill *fills, int32_t num_materials,
                     int32_t num_state_vars, const double *state_vars,
                     const int32_t *mat_num_state_vars);

/* --------------------------------------------------------------- fields --- */
typedef enum fsgh_field {
    FSGH_NODE_COORDS = 0,  /* double [2][N][3] */
    FSGH_NODE_VEL = 1,     /* double [2][N][3] */
    FSGH_NODE_MASS = 2,    /* double [N] */
    FSGH_ELEM_DEN = 3,     /* double [E] */
    FSGH_ELEM_PRES = 4,    /* double [E] */
    FSGH_ELEM_SSPD = 5,    /* double [E] */
    FSGH_ELEM_SIE = 6,     /* double [2][E] */
    FSGH_ELEM_VOL = 7,     /* double [E] */
    FSGH_ELEM_DIV = 8,     /* double [E] */
    FSGH_ELEM_MASS = 9,    /* double [E] */
    FSGH_ELEM_MAT_ID = 10, /* int64  [E] */
    FSGH_ELEM_GAMMA = 11,  /* double [E]  = statev(e,0) */
    FSGH_ELEM_CSMIN = 12,  /* double [E]  = statev(e,1) */
    FSGH_CORNER_FORCE = 13, /* double [8E][3] */
    FSGH_ELEM_STRESS = 14   /* double [2][E][3][3] Cauchy stress; an upload allocates the
                               stress field (it is zero and absent otherwise) */
} fsgh_field;
int fsgh_upload(fsgh_ctx *ctx, int32_t field, const void *host, size_t bytes);
int fsgh_download(fsgh_ctx *ctx, int32_t field, void *host, size_t bytes);

/* ------------------------------------- per-stage operators (reference API) --- */
/* Each applies to every instance of the context, with the reference's
 * semantics and FP64 operation order.  dt and rk_alpha are scalars as in the
 * reference (K/mesh.h:1211-1373). */
int fsgh_rk_init(fsgh_ctx *ctx);                                    /* K/time_integration.cpp:9-44 */
int fsgh_get_timestep(fsgh_ctx *ctx, double time_value, double graphics_time,
                      double time_final, double dt_max, double dt_min, double dt_cfl,
                      double fuzz, double *dt_inout);               /* K/time_integration.cpp:54-154 */
int fsgh_get_divergence(fsgh_ctx *ctx);                             /* K/momentum.cpp:221-292 */
int fsgh_get_force_sgh(fsgh_ctx *ctx, double fuzz, double small, double dt,
                       double rk_alpha);                            /* K/force_sgh.cpp:12-399 */
int fsgh_update_velocity_sgh(fsgh_ctx *ctx, double rk_alpha, double dt); /* K/momentum.cpp:11-53 */
int fsgh_boundary_velocity(fsgh_ctx *ctx);                          /* K/boundary.cpp:10-55 */
int fsgh_update_energy_sgh(fsgh_ctx *ctx, double rk_alpha, double dt);   /* K/energy_sgh.cpp:6-55 */
int fsgh_update_position_sgh(fsgh_ctx *ctx, double rk_alpha, double dt); /* K/geometry.cpp:9-26 */
int fsgh_get_vol(fsgh_ctx *ctx);                                    /* K/geometry.cpp:238-310 */
int fsgh_update_state(fsgh_ctx *ctx, double dt, double rk_alpha);   /* K/properties.cpp:9-114 */
/* extensive energies IE = sum M e, KE = 1/2 sum m |v|^2 of one instance
 * (K/sgh_solve.cpp:66-104; 2D: KE = 1/2 sum m R |v|^2); synchronises. */
int fsgh_energy(fsgh_ctx *ctx, int32_t instance, double *ie, double *ke);
/* Energy tallies of caller-held DEVICE arrays, allocating nothing.  No
 * counterpart in the reference: this is a variant of fsgh_energy for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc -- and needs no context.
 * Same kernels and summation order as fsgh_energy (IE = sum M e,
 * KE = 1/2 sum m |v|^2; num_dims = 2: KE = 1/2 sum m R |v|^2 with the node
 * radius R), so equal arrays in equal order give bitwise-equal results.
 * Inputs: elem_mass, elem_sie [num_elems]; node_mass, node_u, node_v
 * [num_nodes]; node_w [num_nodes] in 3D (ignored in 2D); node_radius
 * [num_nodes] in 2D (ignored in 3D); all on hip_device.  Output: ie_ke, two
 * doubles in device memory (IE, KE).  The call is asynchronous on hip_stream
 * (NULL: the null stream) and never synchronises; it restores the caller's
 * current device.  The partial sums live in one fixed 16 KiB array per GPU
 * (1024 pairs of doubles) that the library holds until the process ends;
 * every call rewrites all of it, and calls that overlap on one GPU race on
 * it.  num_elems and num_nodes may not exceed FSGH_NOALLOC_MAX_ENTRIES
 * (FSGH_E_INVALID). */
#define FSGH_NOALLOC_MAX_ENTRIES 2147221503LL  /* 2^31 - 1 - 1024 * 256 */
int fsgh_energy_noalloc(int32_t hip_device, int32_t num_dims, int64_t num_elems,
                        const double *elem_mass, const double *elem_sie, int64_t num_nodes,
                        const double *node_mass, const double *node_u, const double *node_v,
                        const double *node_w, const double *node_radius, double *ie_ke,
                        void *hip_stream);
/* 2D RZ only.  The node extensive mass = areal node mass x radius, which
 * sgh_solve takes once at its start (K/sgh_solve.cpp:107-116), and the
 * per-stage areal mass update from it with the on-axis fix
 * (K/sgh_solve.cpp:495-581; tiny = 1e-12 in K/driver.cpp:61).  3D
 * contexts: FSGH_E_STATE. */
int fsgh_rz_extensive_mass(fsgh_ctx *ctx);
/* The velocity gradient L_ij = du_i/dx_j of every element from the current
 * coordinates, velocities and volume (get_velgrad, K/momentum.cpp:60-148;
 * 2D: get_velgrad2D, K/momentum.cpp:151-213, with L_22 = v_r / r), as
 * E x 3 x 3 in caller numbering: the input the strength models see. */
int fsgh_get_velgrad(fsgh_ctx *ctx, double *vel_grad);
int fsgh_rz_update_node_mass(fsgh_ctx *ctx, double tiny);

/* -------------------------------------------------- fused cycle loop --- */
typedef struct fsgh_run_params {
    double time_final;                 /* K/input.cpp:73 */
    double dt_start;                   /* K/input.cpp:76 */
    double dt_max, dt_min, dt_cfl;     /* K/driver.cpp:47-49 */
    double graphics_dt_ival;           /* K/driver.cpp:39 */
    int64_t graphics_cyc_ival;         /* K/driver.cpp:36 */
    int64_t cycle_stop;                /* K/driver.cpp:56 */
    double fuzz, small;                /* K/driver.cpp:60-62 */
    double tiny;                       /* K/driver.cpp:61 (2D RZ node mass) */
} fsgh_run_params;

typedef struct fsgh_instance_status {
    int64_t cycles;                    /* cycles executed */
    double time_value;
    double dt;                         /* last dt */
    double graphics_time;
    int32_t stopped;                   /* time >= time_final or cycle_stop reached */
    int32_t write_pending;             /* the last cycle met the output condition */
} fsgh_instance_status;

/* Starts the cycle loop of every instance at time 0 with dt = dt_start
 * (K/driver.cpp:269-274); params has one entry per instance. */
int fsgh_run_begin(fsgh_ctx *ctx, const fsgh_run_params *params);
/* Advances up to `ncycles` cycles of the loop of K/sgh_solve.cpp:125-634
 * with fused kernels: dt (device-resident, x1.1 growth, clamps, per instance),
 * then per RK stage divergence extremes -> force -> velocity+BC+position ->
 * energy+volume+EOS+div.  flags: FSGH_RUN_GRAPH replays one captured hipGraph
 * per cycle; FSGH_RUN_STOP_ON_OUTPUT returns after the first cycle in which an
 * instance met the output condition (K/sgh_solve.cpp:594-606; see
 * fsgh_instance_status.write_pending) so the caller can write its dump.
 * FSGH_RUN_PERSISTENT (each cycle as ONE cooperative launch, its phases
 * separated by grid-wide barriers; no counterpart in the reference) is
 * switched off in this build and returns FSGH_E_UNSUPPORTED before anything
 * runs: its first cooperative launch on the MI355X ended the process with no
 * runtime message, and the cause has not been found.
 * Returns when all instances stopped, on such an output cycle, or after
 * ncycles; synchronises. */
#define FSGH_RUN_GRAPH 1
#define FSGH_RUN_STOP_ON_OUTPUT 2
#define FSGH_RUN_PERSISTENT 4
int fsgh_run(fsgh_ctx *ctx, int64_t ncycles, int32_t flags);
int fsgh_run_status(fsgh_ctx *ctx, int32_t instance, fsgh_instance_status *out);

/* ------------------------------------------------- launch autotuning --- */
/* No counterpart in the reference (Kokkos picks its own launch shapes).  The
 * four kernels of the fused cycle run in one of several compiled variants:
 * workgroup size (elements or nodes per workgroup), an occupancy floor asked
 * of the register allocator (waves per SIMD; 1 = the compiler's choice) for
 * the two element kernels, and how work ranges are dealt to the XCDs.
 * Every variant computes bit-identical results. */
#define FSGH_KERNEL_DIV_EXTREMES 0
#define FSGH_KERNEL_FORCE 1
#define FSGH_KERNEL_VEL_BC_POS 2
#define FSGH_KERNEL_ENERGY_STATE_DIV 3
#define FSGH_NUM_TUNED_KERNELS 4

typedef struct fsgh_tuning {
    int32_t kernel;       /* FSGH_KERNEL_* */
    int32_t block;        /* 64, 128, 256 (default); 512 for the two node kernels */
    int32_t min_waves;    /* 1 (default) or 3 (force and energy kernels only) */
    int32_t xcd_order;    /* 0: launch order; 1: one contiguous range of work items
                             per XCD; k = 2..4096: chunks of k consecutive work
                             items dealt round-robin to the 8 XCDs (default 16) */
    double ms;            /* fsgh_autotune: mean time per launch of this choice */
    double ms_default;    /* fsgh_autotune: the same for the default variant */
} fsgh_tuning;

/* Times every variant of each kernel on the context's current state (reps
 * launches per variant, twice, the lower mean kept; reps <= 0 means 10) and
 * keeps the fastest.  The state is saved before and restored after, so the
 * call changes no field; it needs a set-up state (after fsgh_setup_fills or
 * uploads) and temporary device memory of about one copy of the state.
 * out: FSGH_NUM_TUNED_KERNELS entries, or NULL.  Usable before or between
 * fsgh_run calls. */
int fsgh_autotune(fsgh_ctx *ctx, int32_t reps, fsgh_tuning *out);
/* Applies saved choices (e.g. from an earlier fsgh_autotune on the same
 * mesh size); each entry names its kernel.  FSGH_E_INVALID for a variant that
 * is not compiled. */
int fsgh_set_tuning(fsgh_ctx *ctx, const fsgh_tuning *cfg, int32_t count);
/* The current choice per kernel (ms fields 0 unless set by fsgh_autotune). */
int fsgh_get_tuning(fsgh_ctx *ctx, fsgh_tuning *out, int32_t count);

/* ------------------------------------ discrete adjoint / tangent model --- */
/* No counterpart in the reference (which has no derivative of the SGH cycle).
 * The model is the fused cycle of fsgh_run with the time steps frozen to the
 * ones the forward pass used (the CFL minimum, its growth factor and the
 * clamps are not differentiated).  Controls: the initial element density
 * rho0 (element mass M = rho0 V0, node masses sum M/8, initial pressure and
 * sound speed by the EOS), the initial specific internal energy, the element
 * gamma (statev 0) and the initial node velocity.
 *
 * fsgh_forward: call right after fsgh_setup_fills (or the uploads of a full
 * initial state); runs `ncycles` cycles of every instance from time 0,
 * recording dt per cycle and instance and storing the state (x, v, e) every
 * `checkpoint_every` cycles.  FSGH_E_STATE if an instance's state was
 * advanced (fsgh_run, fsgh_forward, the state-update operators) and its
 * coordinates, velocities and energies have not all been set again since
 * (fsgh_upload of FSGH_NODE_COORDS, FSGH_NODE_VEL, FSGH_ELEM_SIE;
 * fsgh_setup_fills sets velocities and energies).  dt_replay (may be NULL; [ncycles][instances])
 * replays given time steps instead of computing them.  Every instance must
 * run all ncycles (time_final and cycle_stop not reached earlier), otherwise
 * FSGH_E_INVALID.  The state afterwards is the forward result, as after
 * fsgh_run. */
int fsgh_forward(fsgh_ctx *ctx, const fsgh_run_params *params, int64_t ncycles,
                 int64_t checkpoint_every, const double *dt_replay);
/* The recorded time steps, n = ncycles * num_instances, [cycle][instance]. */
int fsgh_dt_history(fsgh_ctx *ctx, double *out, int64_t n);
/* Reverse pass: given the adjoint seeds of the final state (node
 * coordinates and velocities [N][3], element energy [E], element density
 * M/V [E]; NULL = zero), returns the gradients of <seed, final state> with
 * respect to rho0 [E], e0 [E], gamma [E] and the initial velocity [N][3]
 * (any output may be NULL).  Recomputes each checkpoint segment; leaves the
 * forward's final state in place.  Caller numbering throughout. */
int fsgh_backward(fsgh_ctx *ctx, const double *seed_x, const double *seed_v,
                  const double *seed_e, const double *seed_den, double *grad_den0,
                  double *grad_sie0, double *grad_gamma, double *grad_vel0);
/* Tangent-linear model: the directional derivative of the final state
 * (coordinates [N][3], velocities [N][3], energy [E]) along the control
 * perturbation (d_den0 [E], d_sie0 [E], d_gamma [E], d_vel0 [N][3]; NULL =
 * zero).  Uses the forward's first checkpoint and time steps. */
int fsgh_tangent(fsgh_ctx *ctx, const double *d_den0, const double *d_sie0,
                 const double *d_gamma, const double *d_vel0, double *t_x, double *t_v,
                 double *t_e);

/* ------------------------------------------------------- diagnostics --- */
/* Per-kernel timing with HIP events on the context's stream.  enable != 0
 * starts recording (graph replays are not split per kernel). */
int fsgh_profile_enable(fsgh_ctx *ctx, int32_t enable);
/* names: '\n'-separated kernel names; ms: accumulated milliseconds; calls:
 * launch counts.  max entries in, count out. */
int fsgh_profile_read(fsgh_ctx *ctx, int32_t max, char *names, size_t names_bytes,
                      double *ms, int64_t *calls, int32_t *count);
/* Scans the state for non-finite values or non-positive volume/mass. */
int fsgh_check(fsgh_ctx *ctx);
/* HIP stream handle of the context (hipStream_t, as void*). */
void *fsgh_stream(fsgh_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* FSGH_H */
