/*
 * ppl.h -- C ABI of the MI355X collision-kinematics substrate (libppl.so).
 *
 * Plain pointers and sizes only; every entry point returns a ppl_status and
 * never exits or throws across the ABI. Reference paths below are relative to
 * code/gpu_planning/cpp/src/ of PrismPlan; each declaration names the
 * reference interface it replaces.
 *
 * Memory: "h_" arguments are host arrays, "d_" arguments device arrays
 * (HBM). Streams are hipStream_t passed as void* (NULL = default stream).
 * Calls that share one ppl_robot, ppl_edge_eval or ppl_plan must not run
 * concurrently on different streams: each handle owns its scratch buffers.
 */
#ifndef PPL_PPL_H
#define PPL_PPL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PPL_OK = 0,
    PPL_ERR_INVALID = 1,      /* bad argument, shape or table */
    PPL_ERR_HIP = 2,          /* a HIP runtime call failed */
    PPL_ERR_ALLOC = 3,        /* device or pinned allocation failed */
    PPL_ERR_UNSUPPORTED = 4,  /* valid for the reference, not handled here */
    PPL_ERR_CAPACITY = 5      /* above a documented size limit */
} ppl_status;

typedef void* ppl_stream;                 /* hipStream_t */
typedef struct ppl_robot_s* ppl_robot;
typedef struct ppl_edge_eval_s* ppl_edge_eval;
typedef struct ppl_plan_s* ppl_plan;

/* Last error message of the calling thread ("" if none). */
const char* ppl_last_error(void);
/* Library version string. */
const char* ppl_version(void);

/* ------------------------------------------------------------------------
 * Runtime robot model.
 * Replaces RobotModelDevice::create / destroy (planning/robot_model.cuh:91-122,
 * 262-293) and the RobotModelBuilder fields (python/bindings.cpp:118-160).
 * Host arrays, copied at create time. Topology arrays may be NULL for a serial
 * chain (same defaults as robot_model.cuh:205-236); approx_fixed_transforms /
 * approx_joint_types may be NULL (= full chain).
 * ------------------------------------------------------------------------ */
typedef struct {
    int n_dof;
    int n_joints;
    const float* fixed_transforms;        /* [n_joints*16] row-major 4x4 */
    const int* joint_types;               /* [n_joints] -1 fixed, 0..2 prism x/y/z, 3..5 rot x/y/z */
    int n_spheres;
    const float* spheres;                 /* [n_spheres*4] x, y, z, r (link frame) */
    const int* sphere_to_joint;           /* [n_spheres] */
    int n_self_cc_ranges;
    const int* self_cc_ranges;            /* [n*3] sphere_i, start, end (inclusive) */
    int n_approx_spheres;
    const float* approx_spheres;          /* [n_approx*4] */
    const int* approx_sphere_to_joint;    /* [n_approx] */
    int n_approx_self_cc_ranges;
    const int* approx_self_cc_ranges;     /* [n*3] */
    const float* approx_fixed_transforms; /* NULL or [n_joints*16] */
    const int* approx_joint_types;        /* NULL or [n_joints] */
    const float* joint_lower;             /* [n_dof] */
    const float* joint_upper;             /* [n_dof] */
    const int* joint_parents;             /* NULL or [n_joints] */
    const int* joint_id_to_dof;           /* NULL or [n_joints] */
    const int* t_memory_idx;              /* NULL or [n_joints] */
    const int* dfs_order;                 /* NULL or [n_joints] */
} ppl_robot_desc;

typedef struct {
    int n_dof, n_joints, n_spheres, n_approx_spheres, n_self_pairs, n_t_slots;
} ppl_robot_info;

ppl_status ppl_robot_create(const ppl_robot_desc* desc, ppl_robot* out);
ppl_status ppl_robot_destroy(ppl_robot robot);
ppl_status ppl_robot_get_info(ppl_robot robot, ppl_robot_info* info);
/* RobotModelBuilder::update_joint_limits (python/bindings.cpp:180-206). */
ppl_status ppl_robot_update_joint_limits(ppl_robot robot, const float* h_lower, const float* h_upper);

/* Device-resident robot tables for planner kernels built on ppl/ppl_device.hpp:
 * the counterpart of the RobotModel struct the reference passes by value to
 * its kernels (planning/robot_model.cuh:27-82). Tables are in the tree's DFS
 * order; the pointers stay valid until ppl_robot_destroy. */
typedef struct {
    int n_dof, n_joints, n_spheres, n_approx_spheres, n_pairs, n_approx_pairs;
    const float* ft;          /* [n_joints][12] rows 0..2 of each fixed transform */
    const int* jmeta;         /* [n_joints][4] joint type, dof (-1 none), FK slot, parent's FK slot */
    const int* sph_off;       /* [n_joints+1] spheres of DFS joint k: sph[sph_off[k] .. sph_off[k+1]) */
    const float* sph;         /* [n_spheres][4] x, y, z, r in the link frame, grouped by DFS joint */
    const int* sph_id;        /* [n_spheres] original sphere index of each sph entry */
    const int* apx_off;       /* [n_joints+1] the same for the approximate spheres */
    const float* apx;         /* [n_approx_spheres][4] */
    const int* apx_id;        /* [n_approx_spheres] */
    const float* sph_orig;    /* [n_spheres][4] by original index */
    const float* apx_orig;    /* [n_approx_spheres][4] by original index */
    const int* pairs;         /* [n_pairs][2] self pairs expanded from self_cc_ranges (original indices) */
    const int* apx_pairs;     /* [n_approx_pairs][2] */
    const float* lim;         /* [n_dof][2] joint lower, upper */
    /* The approximate spheres' own chain (RobotModel's approx_fixed_transforms /
     * approx_joint_types, runtime_kinematics.cuh:176-243), DFS order; both NULL
     * when the approximate spheres ride the full chain. */
    const float* aft;         /* [n_joints][12] rows 0..2 of each approximate fixed transform */
    const int* atype;         /* [n_joints] approximate joint types */
} ppl_robot_device_view;

ppl_status ppl_robot_get_device_view(ppl_robot robot, ppl_robot_device_view* out);

/* ------------------------------------------------------------------------
 * Scene view: non-owning DEVICE pointers with the field meanings of
 * SceneCollisionData (collision/scene_collision_data.hh:10-47).
 * ------------------------------------------------------------------------ */
typedef struct {
    const float* obb_dims;        /* [n_obbs*4] full extents dx, dy, dz, pad */
    const float* obb_pose;        /* [n_obbs*8] inverse pose tx, ty, tz, qw, qx, qy, qz, pad */
    const uint8_t* obb_enable;    /* [n_obbs] */
    int n_obbs;
    const float* voxel_params;    /* [L*4] dx, dy, dz, voxel_size (DEVICE) */
    const float* voxel_pose;      /* [L*8] inverse grid pose (DEVICE) */
    const uint8_t* voxel_enable;  /* [L] */
    const float* voxel_features;  /* [L*max_voxels_per_layer], positive inside */
    int n_voxel_layers;
    int max_voxels_per_layer;
    const uint8_t* sphere_acm_mask;             /* [S*L] or NULL */
    const uint8_t* sphere_obb_acm_mask;         /* [S*O] or NULL */
    const uint8_t* sphere_obb_acm_mask_approx;  /* [A*O] or NULL (ignored, as in the reference) */
    int esdf_interp;              /* 0: nearest voxel (reference), 1: trilinear (value + gradient) */
} ppl_scene_view;

/* ------------------------------------------------------------------------
 * Plain per-request forms (the drop-in).
 * ------------------------------------------------------------------------ */

/* pRRTC::check_collision_scene (planning/Planners.hh:141-148,
 * planning/pRRTC.cu:2388-2423): two-phase environment-only any-collision of N
 * host configurations [N*n_dof] row-major; h_results[i] = 1 on collision.
 * Synchronous. */
ppl_status ppl_check_collision_scene(ppl_robot robot, const ppl_scene_view* scene, const float* h_configs, int n,
                                     float collision_margin, uint8_t* h_results);

/* ppl::search::EdgeEvalParams (planning/edge_eval.cuh:37-51). */
typedef struct {
    int granularity;          /* states per edge when resolution <= 0 */
    float collision_margin;
    int check_self;
    int compute_clearance;    /* min ESDF clearance over the edge's states */
    int enable_mesh;          /* must be 0 */
    float resolution;         /* > 0: ceil(len / resolution) states, clamped to [2, 1024] */
} ppl_edge_params;

/* ppl::search::EdgeEvalGpu (planning/edge_batch_bridge.hh:27-111): persistent
 * device and pinned buffers for up to max_edges edges, private stream. */
ppl_status ppl_edge_eval_create(ppl_robot robot, const ppl_scene_view* scene, int dim, int max_edges,
                                ppl_edge_eval* out);
ppl_status ppl_edge_eval_destroy(ppl_edge_eval ev);
/* EdgeEvalGpu::evaluate (edge_batch_bridge.hh:60-96): states j = 1..G at
 * t = j/G of each edge; valid = no collision, cost = |to - from| if valid
 * else FLT_MAX, clearance (nullable) = min ESDF clearance if requested. */
ppl_status ppl_edge_eval_evaluate(ppl_edge_eval ev, const float* h_from, const float* h_to, int n_edges,
                                  const ppl_edge_params* params, uint8_t* h_valid, float* h_cost,
                                  float* h_clearance);

/* Host-side state counts of the reference granularity rule
 * (edge_eval.cuh:54-60) for n edges: h_G[e] states (j = 1..G), h_len[e] the
 * FP32 configuration-space length. Pure host code. */
ppl_status ppl_edge_granularities(const float* h_from, const float* h_to, int n_edges, int dim,
                                  const ppl_edge_params* params, int* h_G, float* h_len);

/* Summed-hinge cost parameters (chomp.cu:83-188, stomp.cu:206-246). */
typedef struct {
    float world_margin;       /* OBB and ESDF hinge band (raw radius) */
    float self_margin;        /* self min-clearance hinge band */
    float weight;             /* collision_cost_weight */
    int use_obb;
    int use_esdf;
    int use_self;
} ppl_cost_params;

/* Plain forms of the other two per-configuration reductions, with
 * check_collision_scene's contract (host arrays [N*n_dof] in, host array out,
 * synchronous):
 *   h_clearance[i] = scene_min_clearance_runtime of configuration i: the min
 *     ESDF clearance, collision_margin included (collision/
 *     scene_collision.cuh:339-384);
 *   h_cost[i] = weight * (summed ESDF and OBB hinge + self min-hinge), the
 *     state cost the STOMP and CHOMP kernels read (scene_collision.cuh:
 *     394-495, planning/runtime_kinematics.cuh:294-329). */
ppl_status ppl_min_clearance_scene(ppl_robot robot, const ppl_scene_view* scene, const float* h_configs, int n,
                                   float collision_margin, float* h_clearance);
ppl_status ppl_hinge_cost_scene(ppl_robot robot, const ppl_scene_view* scene, const float* h_configs, int n,
                                const ppl_cost_params* params, float* h_cost);

/* chomp_state_cost_kernel (chomp.cu:151-188) for one trajectory
 * d_theta [D][T] (dim-major): d_state_costs [D][T], D-broadcast. */
ppl_status ppl_chomp_state_cost(ppl_robot robot, const ppl_scene_view* scene, const float* d_theta, int T,
                                const ppl_cost_params* params, float* d_state_costs, ppl_stream stream);

/* chomp_obs_grad_kernel (chomp.cu:83-144) for one trajectory: central finite
 * difference of the summed hinge, d_grad [D][T], 0 at both ends. */
ppl_status ppl_chomp_fd_grad(ppl_robot robot, const ppl_scene_view* scene, const float* d_theta, int T,
                             const ppl_cost_params* params, float eps, float* d_grad, ppl_stream stream);

/* ------------------------------------------------------------------------
 * Batched forms: one persistent launch over a plan (a flattened state space
 * cut into 64-state chunks that never span two problems). No counterpart in
 * the reference, which solves one request per launch.
 * ------------------------------------------------------------------------ */
typedef enum { PPL_PLAN_CONFIGS = 0, PPL_PLAN_EDGES = 1, PPL_PLAN_TRAJ = 2 } ppl_plan_kind;

typedef struct {
    int kind;                 /* ppl_plan_kind */
    int n_units;              /* configurations, edges or trajectories */
    const int* unit_states;   /* host [n_units]: states per edge / waypoints per trajectory; NULL for configs */
    int skip_first;           /* edges: state 0 (the from-configuration) is not checked */
    int n_problems;           /* >= 1 */
    const int* problem_units; /* host [n_problems+1] unit offsets; NULL = one problem */
    const int* obb_offsets;   /* host [n_problems+1] OBB ranges; NULL = all OBBs for every problem */
    int waypoint_major;       /* trajectory plans: 1 = work is cut waypoint-major inside each problem
                                 (waypoint t of every trajectory, then t+1), which keeps the states a
                                 wave holds together in space for rollout sets such as STOMP's; the
                                 trajectories of each problem must then have equal length. Results
                                 and layouts are the same either way. */
} ppl_plan_desc;

ppl_status ppl_plan_create(ppl_robot robot, const ppl_plan_desc* desc, ppl_plan* out);
ppl_status ppl_plan_destroy(ppl_plan plan);
/* Total states of the plan (configs, checked edge states or waypoints). */
long ppl_plan_states(ppl_plan plan);

/* Forward kinematics outputs of every configuration of a configuration plan
 * (either output nullable): d_frames [n][n_joints][16], each joint's world
 * frame column-major (out[c*4 + r] = T[r][c]) in original joint order, as
 * fk_joint_transforms_runtime emits (planning/runtime_kinematics.cuh:97-171);
 * d_spheres [n][n_spheres][4] world centre and radius in original sphere
 * order (fk_runtime, planning/runtime_kinematics.cuh:16-88). */
ppl_status ppl_fk_batched(ppl_plan plan, const float* d_configs, float* d_frames, float* d_spheres, ppl_stream stream);

/* Any-collision of every configuration, d_configs [n*n_dof] row-major.
 * check_self adds the self check (two_phase_cc check_self, MIT* batch_cc). */
ppl_status ppl_check_collision_batched(ppl_plan plan, const ppl_scene_view* scene, const float* d_configs,
                                       float collision_margin, int check_self, uint8_t* d_out, ppl_stream stream);

typedef struct {
    float collision_margin;
    int check_self;
    int skip_resolved;        /* 1: states past an edge's known first hit are not checked */
} ppl_edge_batch_params;

/* Edges of a PPL_PLAN_EDGES plan: edge e has n_e = unit_states[e] (+1 if
 * skip_first) states at t = k/(n_e-1). d_first_hit[e] = smallest colliding
 * k, or -1. d_state_hits (nullable, only with skip_resolved = 0): per-state
 * verdicts in plan order. */
ppl_status ppl_edge_check_batched(ppl_plan plan, const ppl_scene_view* scene, const float* d_from, const float* d_to,
                                  const ppl_edge_batch_params* params, int* d_first_hit, uint8_t* d_state_hits,
                                  ppl_stream stream);

/* scene_min_clearance_runtime (collision/scene_collision.cuh:339-384): min
 * over spheres of ESDF clearance with radius + margin, per configuration. */
ppl_status ppl_min_clearance_batched(ppl_plan plan, const ppl_scene_view* scene, const float* d_configs,
                                     float collision_margin, float* d_out, ppl_stream stream);

/* Summed-hinge cost of every state (configs or trajectories): d_state_cost
 * [states] weighted per-state cost, d_total (nullable, trajectory plans)
 * per-trajectory sum. d_grad (nullable): d cost / d q in the input layout
 * (trajectory plans: 0 at both ends), computed by the fused adjoint. */
ppl_status ppl_hinge_cost_batched(ppl_plan plan, const ppl_scene_view* scene, const float* d_x,
                                  const ppl_cost_params* params, float* d_state_cost, float* d_total, float* d_grad,
                                  ppl_stream stream);

/* Autotuner of the cost kernel's tile shape: times the plan's summed-hinge
 * evaluation (with the fused adjoint if grad) on the caller's own inputs for
 * every lane-group size G = 1, 2, 4, ..., 16 lanes per state (64 / G states
 * per wavefront, a state's spheres split over G lanes) that fits in LDS and
 * for the grouped layout (*best_group = PPL_COST_GROUPED), and keeps the
 * fastest for this plan's later hinge / STOMP calls. reps timed runs per
 * candidate on `stream`; outputs go to plan scratch. The choice is also
 * recorded in the tuning cache (ppl_tune_cache_file) under the robot's
 * tables, the largest problem's OBB count, the state count's power of two
 * and the GPU architecture; untuned plans that match such an entry use it,
 * and a later autotuning call that finds it adopts it without timing
 * (*best_ms: the time recorded with it). No counterpart in the reference. */
ppl_status ppl_plan_autotune_cost(ppl_plan plan, const ppl_scene_view* scene, const float* d_x,
                                  const ppl_cost_params* params, int grad, int reps, int* best_group,
                                  float* best_ms, ppl_stream stream);

/* Pin the cost kernel's layout for this plan (grad: the adjoint variant):
 * lanes per state of the lane-group kernel (a power of two <= 16 that fits in
 * LDS), PPL_COST_GROUPED for the grouped kernel (16 lanes per state, the
 * state's spheres spread over them across joints, joint frames and wrenches in
 * LDS), or 0 for the built-in choice (grouped for the adjoint). */
#define PPL_COST_GROUPED (-1)
ppl_status ppl_plan_set_cost_group(ppl_plan plan, int grad, int lanes_per_state);

/* Tuning cache file: the autotuners' choices, one "key value..." line each,
 * read at the first lookup and rewritten at each new choice, so later plans
 * and processes skip the timing. Default: $PPL_TUNE_CACHE, else
 * $XDG_CACHE_HOME/ppl/tune.txt, else $HOME/.cache/ppl/tune.txt; NULL or ""
 * keeps choices in memory only. Setting it drops the entries read so far.
 * No counterpart in the reference. */
ppl_status ppl_tune_cache_file(const char* path);

/* Autotuner of the any-collision kernel's launch shape for a configuration
 * plan (d_x: configurations) or an edge plan (d_x, d_x2: the edges' start and
 * end configurations; first-hit mode with skip_resolved): times reps runs of
 * the plan's batched check on `stream` for each OBB staging tile (the largest
 * problem's OBBs in one staging, then 128, 64 and 32 records below that; the
 * stagings of a problem then pass through LDS in turn) and each grid form
 * (0: persistent at occupancy, 1: twice that, 2: one workgroup per 64-state
 * chunk), keeps the fastest for this plan's later any-collision calls and
 * records it in the tuning cache under the robot's tables, the largest
 * problem's OBB count and the GPU architecture; when the cache already holds
 * a choice for that key it is adopted without timing (*best_ms: the time
 * recorded with it). The fixed tile it replaces is the reference's
 * BATCH_SIZE (collision/scene_collision.cuh:9-11). Results do not depend on
 * the shape. No counterpart in the reference. */
ppl_status ppl_plan_autotune_collide(ppl_plan plan, const ppl_scene_view* scene, const float* d_x, const float* d_x2,
                                     int reps, int* best_tile, int* best_grid, float* best_ms, ppl_stream stream);

/* The any-collision launch shape this plan uses for problems of up to
 * max_obbs OBBs (*tile records per staging, *grid form as above) and where it
 * comes from (*source: 0 built-in, 1 tuning cache, 2 this plan's tuning or
 * ppl_plan_set_collide_shape). */
ppl_status ppl_plan_collide_shape(ppl_plan plan, int max_obbs, int* tile, int* grid, int* source);

/* Pin the any-collision launch shape (tile 0: every OBB of a problem in one
 * staging, up to 256; grid 0, 1, 2 as above); (0, 0) returns to the tuning
 * cache / built-in choice. */
ppl_status ppl_plan_set_collide_shape(ppl_plan plan, int tile, int grid);

/* Kernel timing (enable != 0: on): each later call of
 * ppl_check_collision_batched, ppl_edge_check_batched or the batched cost
 * evaluation on this plan times its main kernel (k_collide or the cost
 * kernel; not the first-hit reset or the self pass) with a HIP event pair
 * that the runtime records at the kernel's own start and end
 * (hipExtLaunchKernel), so its duration can be read without a profiler.
 * Calls under stream capture record nothing; at most 4096 launches are kept
 * until read. No counterpart in the reference. */
ppl_status ppl_plan_time_kernels(ppl_plan plan, int enable);

/* Waits for the timed launches recorded since the last read, returns their
 * summed duration (*total_ms) and count (*launches), and clears them. */
ppl_status ppl_plan_kernel_ms(ppl_plan plan, double* total_ms, int* launches);

/* Central-FD obstacle gradient over a trajectory plan (chomp_obs_grad_kernel
 * semantics), the baseline for the fused adjoint. */
ppl_status ppl_chomp_fd_grad_batched(ppl_plan plan, const ppl_scene_view* scene, const float* d_theta,
                                     const ppl_cost_params* params, float eps, float* d_grad, ppl_stream stream);

/* STOMP rollout costs (stomp.cu:206-308) over a trajectory plan whose
 * trajectories are the rollouts of each problem: per-waypoint state cost and
 * per-rollout total = sum_t state + w_u * sum_d theta_d^T R theta_d with R
 * the 5-point acceleration metric of step dt (chomp_precompute.cpp:20-38). */
ppl_status ppl_stomp_rollout_costs(ppl_plan plan, const ppl_scene_view* scene, const float* d_rollouts,
                                   const ppl_cost_params* params, float control_weight, float dt,
                                   float* d_state_cost, float* d_total, ppl_stream stream);

/* ------------------------------------------------------------------------
 * Batched CHOMP optimiser on the substrate (planning/chomp.cu:358-610
 * semantics, every trajectory of a trajectory plan at once, no host
 * round trip per iteration). Trajectories are the plan's [D][T] blocks; both
 * endpoints stay fixed. Per iteration and trajectory:
 *   obstacle gradient (fused adjoint, or central FD as chomp.cu:83-144),
 *   covariant step theta_d -= step_size * A^-1 (g_d + smoothness_weight R theta_d)
 *     with A = R + 1e-8 I (chomp_precompute.cpp:20-64; A^-1 is formed in FP64
 *     on the host, stored FP32), joint-limit clamp (chomp.cu:261-290);
 *   cost = sum_t weight (world + self hinge) + smoothness_weight theta^T R theta;
 *   validity = two-phase any-collision with self check at every waypoint and
 *     ceil(max segment / collision_check_step) sub-states per segment
 *     (chomp.cu:297-351, 459-481);
 *   best-valid tracking, convergence streak and polish count (chomp.cu:484-557).
 * ------------------------------------------------------------------------ */
typedef struct {
    float step_size;               /* 0.3 */
    float smoothness_weight;       /* lambda, 1.0 */
    float collision_cost_weight;   /* 20 */
    float world_margin;            /* OBB / ESDF hinge band, 0.02 */
    float self_margin;             /* self hinge band, 0.01 */
    float dt;                      /* 0.1 */
    float collision_check_step;    /* validity densification step (rad), 0.05; <= 0: waypoints only */
    float collision_margin;        /* validity check margin, 0 */
    float convergence_eps;         /* 1e-4 */
    int iterations_after_valid;    /* polish iterations once a valid trajectory exists, 5 */
    int use_self;                  /* self hinge in the cost */
    int fd_gradient;               /* 0: fused adjoint; 1: central finite differences */
    float fd_eps;                  /* 1e-3 */
} ppl_chomp_settings;

/* Per-trajectory record of the optimiser, 8 floats per trajectory. */
enum {
    PPL_CHOMP_BEST_COST = 0,       /* best valid total cost (+inf until valid) */
    PPL_CHOMP_PREV_COST = 1,
    PPL_CHOMP_SOLVED = 2,          /* 1: d_best holds a valid trajectory */
    PPL_CHOMP_STREAK = 3,
    PPL_CHOMP_POLISH_LEFT = 4,
    PPL_CHOMP_DONE = 5,            /* 1: converged or polished, no further updates */
    PPL_CHOMP_ITERS = 6,           /* iterations run */
    PPL_CHOMP_LAST_COST = 7        /* total cost of the current d_theta */
};

/* Runs `iterations` CHOMP iterations on every trajectory of `plan`. d_theta
 * [plan layout] is updated in place; d_best [plan layout] receives the best
 * valid trajectory of each; d_record [n_traj][8] as above. first_call = 1
 * initialises the records and evaluates d_theta before the first step (the
 * reference's initial cost / validity pass); later calls continue. */
ppl_status ppl_chomp_optimize(ppl_plan plan, const ppl_scene_view* scene, const ppl_chomp_settings* settings,
                              float* d_theta, float* d_best, float* d_record, int iterations, int first_call,
                              ppl_stream stream);

/* Single-goal CHOMP solve from host start / goal configurations, the
 * reference's CHOMP::solve_runtime_scene (planning/chomp.cu:358-610) and its
 * pybind entry chomp_solve_scene (python/bindings.cpp:872-902):
 *   theta = clamp(linear interpolation start -> goal) over num_timesteps
 *     (chomp.cu:57-69),
 *   up to num_iterations iterations of ppl_chomp_optimize, stopping when the
 *     record is done or, with time_limit_ms > 0, once the wall time since the
 *     call began exceeds it before an iteration (chomp.cu:491-495),
 *   h_path [num_timesteps][n_dof] = the best valid trajectory, else the
 *     working one, with its first and last rows set to start and goal.
 * final_state_cost / final_control_cost are those of the returned trajectory;
 * final_total_cost is the optimiser's recorded total. `bufs` (may be NULL) are
 * the reused device buffers of the reference's CHOMPBuffers (chomp.cu:371-414);
 * without them the call allocates and frees its own. Synchronous. */
typedef struct ppl_opt_buffers_s* ppl_chomp_buffers;
ppl_status ppl_chomp_buffers_create(int n_dof, int num_timesteps, ppl_chomp_buffers* out);
ppl_status ppl_chomp_buffers_destroy(ppl_chomp_buffers bufs);

typedef struct {
    int solved;
    int iterations_run;
    float final_state_cost;
    float final_control_cost;
    float final_total_cost;
    double wall_time_ms;
} ppl_opt_result;
typedef ppl_opt_result ppl_chomp_result;

ppl_status ppl_chomp_solve(ppl_robot robot, ppl_chomp_buffers bufs, const ppl_scene_view* scene, const float* h_start,
                           const float* h_goal, int num_timesteps, int num_iterations, float time_limit_ms,
                           const ppl_chomp_settings* settings, float* h_path, ppl_chomp_result* result,
                           ppl_stream stream);

/* ------------------------------------------------------------------------
 * Batched STOMP optimiser on the substrate (planning/stomp.cu:565-930
 * semantics, every trajectory of a trajectory plan at once). Per iteration
 * and trajectory, with R = dt A^T A the 5-point metric:
 *   population: num_rollouts_new fresh samples x = clamp(theta + L z) with the
 *     endpoints pinned (num_rollouts_new + num_rollouts_old on the first
 *     iteration), the num_rollouts_old cheapest rollouts of the previous
 *     population, and theta itself (stomp.cu:100-189);
 *   cost of a row per (dof, waypoint): state cost (weighted world + self
 *     hinge) + control_cost_weight x_t (R x)_t (stomp.cu:206-303);
 *   weights p = exp(-h (c - min) / max(max - min, 1e-8)) normalised over the
 *     population, per (dof, waypoint) (stomp.cu:310-350);
 *   theta += M (sum_r p_r (x_r - theta)) with both endpoints pinned and no
 *     clamp (stomp.cu:358-460);
 *   validity, best-valid tracking, streak and polish as the CHOMP optimiser.
 * L is the Cholesky factor of R^-1 rescaled to max entry 1/T, times
 * noise_scale; M is R^-1 with each column scaled to max 1/T
 * (stomp_precompute.cpp:48-101); both are formed in FP64, stored FP32.
 * z ~ N(0, 1) comes from Philox4x32-10 (Salmon et al., SC'11) with the
 * Box-Muller transform, counter (t >> 2, r | dof << 16, trajectory,
 * iteration), key = seed; the reference draws from cuRAND XORWOW instead,
 * so sampled rollouts match it in distribution, not value.
 * d_record [n_traj][8] uses the PPL_CHOMP_* layout. Limits: rollouts
 * (new + old + 1) <= 1024 and (rollouts + 1) * longest trajectory <= 16384.
 * ------------------------------------------------------------------------ */
typedef struct {
    int num_rollouts_new;                 /* 30 */
    int num_rollouts_old;                 /* 20 */
    float control_cost_weight;            /* 0.1 */
    float collision_cost_weight;          /* 20 */
    float exponentiated_cost_sensitivity; /* h, 2 */
    float world_margin;                   /* 0.02 */
    float self_margin;                    /* 0.01 */
    float dt;                             /* 0.1 */
    float noise_scale;                    /* 1 */
    float collision_check_step;           /* 0.05; <= 0: waypoints only */
    float collision_margin;               /* 0 */
    float convergence_eps;                /* 1e-4 */
    int iterations_after_valid;           /* 3 */
    int use_self;                         /* self hinge in the state cost */
    unsigned long long seed;              /* 0x5be5 */
} ppl_stomp_settings;

/* first_call = 1 initialises the records and evaluates d_theta; later calls
 * continue the same population (settings must then be unchanged). */
ppl_status ppl_stomp_optimize(ppl_plan plan, const ppl_scene_view* scene, const ppl_stomp_settings* settings,
                              float* d_theta, float* d_best, float* d_record, int iterations, int first_call,
                              ppl_stream stream);

/* Whole optimiser iterations as hipGraphs (default on). With it, a
 * ppl_chomp_optimize call of >= 2 iterations captures one iteration (update,
 * cost and adjoint, control cost, densified validity, bookkeeping) once and
 * replays it; ppl_stomp_optimize captures two iterations (reuse, sampling,
 * rollout costs, update, optimum cost, validity, bookkeeping; one per
 * keep-buffer parity) once its population has its steady shape, with the
 * sampling counter's iteration word on the device. A graph is re-captured when
 * any pointer, scene view field or setting it was captured with changes.
 * Results equal the eager launches' bit for bit. enable = 0 launches every
 * kernel eagerly. No counterpart in the reference (chomp.cu:508-555 and
 * stomp.cu:786-889 read the cost back on the host every iteration). */
ppl_status ppl_plan_set_optimizer_graphs(ppl_plan plan, int enable);

/* Single-goal STOMP solve, the reference's STOMP::solve_runtime_scene
 * (planning/stomp.cu:565-930) and pybind stomp_solve_scene
 * (python/bindings.cpp:801-831), with the same contract as ppl_chomp_solve. */
typedef struct ppl_opt_buffers_s* ppl_stomp_buffers;
ppl_status ppl_stomp_buffers_create(int n_dof, int num_timesteps, int num_rollouts_new, int num_rollouts_old,
                                    ppl_stomp_buffers* out);
ppl_status ppl_stomp_buffers_destroy(ppl_stomp_buffers bufs);
typedef ppl_opt_result ppl_stomp_result;
ppl_status ppl_stomp_solve(ppl_robot robot, ppl_stomp_buffers bufs, const ppl_scene_view* scene, const float* h_start,
                           const float* h_goal, int num_timesteps, int num_iterations, float time_limit_ms,
                           const ppl_stomp_settings* settings, float* h_path, ppl_stomp_result* result,
                           ppl_stream stream);

/* ESDF of a box world on a node grid: value = -min_o sdf_o(node), positive
 * inside. h_voxel_params [4] and h_voxel_pose [8] on the host; d_features
 * holds nx*ny*nz floats, index ix*ny*nz + iy*nz + iz. */
ppl_status ppl_esdf_from_obbs(const ppl_scene_view* obbs, const float* h_voxel_params, const float* h_voxel_pose,
                              float* d_features, ppl_stream stream);

/* ------------------------------------------------------------------------
 * hipGraph capture of a plan's cost evaluation (one optimiser iteration's
 * substrate work) for replay without per-launch host overhead.
 * ------------------------------------------------------------------------ */
typedef struct ppl_graph_s* ppl_graph;
ppl_status ppl_graph_capture_stomp(ppl_plan plan, const ppl_scene_view* scene, const float* d_rollouts,
                                   const ppl_cost_params* params, float control_weight, float dt,
                                   float* d_state_cost, float* d_total, ppl_stream stream, ppl_graph* out);
ppl_status ppl_graph_capture_hinge(ppl_plan plan, const ppl_scene_view* scene, const float* d_x,
                                   const ppl_cost_params* params, float* d_state_cost, float* d_total,
                                   float* d_grad, ppl_stream stream, ppl_graph* out);
ppl_status ppl_graph_launch(ppl_graph graph, ppl_stream stream);
ppl_status ppl_graph_destroy(ppl_graph graph);

/* ------------------------------------------------------------------------
 * MIT*'s substrate calls over device-resident data (planning/MITStar.cu). The
 * planner's own bookkeeping (adjacency, checks-done slots, tree freebies,
 * whitelists, tags) stays with the caller; these run the collision part.
 * ------------------------------------------------------------------------ */

/* batch_cc_kernel (MITStar.cu:212-276): d_cc[i] = two-phase any-collision of
 * d_samples [n][n_dof] with the self check (1 = collision). With d_clearance,
 * each FREE sample gets its min ESDF clearance (scene_min_clearance_runtime,
 * collision_margin included) when clearance_weight > 0, else 1000; colliding
 * samples keep their d_clearance entry. */
ppl_status ppl_mitstar_batch_cc(ppl_robot robot, const ppl_scene_view* scene, const float* d_samples, int n,
                                float collision_margin, uint8_t* d_cc, float* d_clearance, float clearance_weight,
                                ppl_stream stream);

/* Allocation-free form of ppl_mitstar_batch_cc. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as ppl_mitstar_batch_cc; its scratch (self
 * verdicts and clearances) is a fixed device array of the library, one per
 * GPU, held until the process ends, so n above PPL_BATCH_CC_STATIC_MAX
 * returns PPL_ERR_CAPACITY. Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (different streams or threads) race. */
#define PPL_BATCH_CC_STATIC_MAX 262144
ppl_status ppl_mitstar_batch_cc_static(ppl_robot robot, const ppl_scene_view* scene, const float* d_samples, int n,
                                       float collision_margin, uint8_t* d_cc, float* d_clearance,
                                       float clearance_weight, ppl_stream stream);

/* compute_node_clearance_kernel (MITStar.cu:283-329): for i < n_nodes, node
 * = d_node_indices ? d_node_indices[i] : node_start + i, and
 * d_clearance[node] = min ESDF clearance of d_nodes[node] (+ margin). */
ppl_status ppl_mitstar_node_clearance(ppl_robot robot, const ppl_scene_view* scene, const float* d_nodes,
                                      const int* d_node_indices, int node_start, int n_nodes, float collision_margin,
                                      float* d_clearance, ppl_stream stream);

/* sparse_cc_kernel / full_cc_kernel (MITStar.cu:832-1033, 1197-1375): edge e
 * joins d_nodes[d_src[e]] and d_nodes[d_dst[e]] ([node][dim], dim = n_dof).
 * With len the FP32 joint-space length, total = max(1, ceil(len /
 * valid_segment_length)) and target = min(num_sparse_checks, total - 1)
 * (num_sparse_checks < 0: total - 1, the full check), the breadth-first
 * midpoints k = checks_done[e] .. target - 1 (t = 1/2, 1/4, 3/4, 1/8, ...;
 * d_checks_done NULL = 0) are checked with the two-phase test and the self
 * check at q = q_src (1 - t) + q_dst t. d_cc[e] = 1 if any collides;
 * d_new_checks[e] (nullable) = the midpoints verified free (0 on collision).
 * Synchronises once on `stream` to size the launch. */
ppl_status ppl_mitstar_edge_cc(ppl_robot robot, const ppl_scene_view* scene, const float* d_nodes, int dim,
                               const int* d_src, const int* d_dst, const int* d_checks_done, int n_edges,
                               float valid_segment_length, int num_sparse_checks, float collision_margin,
                               uint8_t* d_cc, int* d_new_checks, ppl_stream stream);

/* ------------------------------------------------------------------------
 * pRRTC: GPU-parallel RRT-Connect on the substrate, the reference's
 * pRRTC::solve_runtime_scene (planning/pRRTC.cu:2016-2223, kernel
 * rrtc_runtime_scene at :1556-2010; settings planning/pRRTC_settings.hh:3-32;
 * pybind solve_scene python/bindings.cpp:405-437, 1023-1027).
 *
 * num_new_configs workgroups grow a start tree and a goal tree (one root per
 * goal). Per iteration a workgroup draws a Halton sample scaled to the joint
 * limits, steers from the nearest node of its tree by at most `range`, checks
 * the edge at `granularity` states nearest + k (new - nearest) / granularity
 * (k = 1..granularity, two-phase scene test + self check, collision_margin),
 * adds the node, then greedily extends towards the nearest node of the other
 * tree in ceil(dist / range) steps, each checked the same way. The first
 * workgroup whose extension reaches the other tree solves the problem.
 * Tree choice follows `balance` (0: halves; 1: proportional / tree_ratio;
 * 2: swap when the size gap is small); dynamic_domain keeps a per-node radius
 * (FLT_MAX until a failed extension, then dd_radius shrinking by dd_alpha to
 * dd_min_radius; grown by dd_alpha on success) and skips samples outside it.
 * The run stops at max_iters per workgroup, a full tree (max_samples nodes)
 * or time_limit_ms (> 0) of the device wall clock.
 *
 * Halton bases: workgroup 0 uses the first n_dof of the primes 3..59;
 * workgroup w > 0 a Fisher-Yates shuffle of the 16 primes drawn from
 * splitmix64 (state seed ^ (w * 0xD1B54A32D192ED03), j = next % (i + 1) for
 * i = 15..1); the reference shuffles with curand, so its per-workgroup
 * bases differ. ppl_rrtc_halton_bases returns the table. The ESDF lookup is
 * the reference's nearest voxel (esdf_interp must be 0).
 *
 * h_path receives the solution start -> goal (the pybind order), n_waypoints
 * rows of n_dof; cost sums the tree edges walked (the reference's `cost`).
 * With shortcut_path the path is shortcut by ppl_shortcut_path from the goal
 * end (margin collision_margin + shortcut_collision_margin), as the
 * reference does before its pybind layer reverses it. Synchronous on
 * `stream`. Nodes, parents and radii live in `bufs` (NULL: allocated for the
 * call); concurrent solves need separate buffers.
 * ------------------------------------------------------------------------ */
typedef struct {
    int max_samples;          /* nodes per tree */
    int max_iters;            /* iterations per workgroup */
    float time_limit_ms;      /* <= 0: none */
    int num_new_configs;      /* workgroups */
    int granularity;          /* states per edge check, 1..64 */
    float range;
    int balance;              /* 0, 1 or 2 */
    float tree_ratio;
    int dynamic_domain;
    float dd_alpha, dd_radius, dd_min_radius;
    int shortcut_path;
    float valid_segment_length;
    float collision_margin;
    float shortcut_collision_margin;
    unsigned long long seed;  /* Halton base shuffle */
} ppl_rrtc_settings;

/* pRRTC_settings defaults (planning/pRRTC_settings.hh:3-32), seed 1. */
void ppl_rrtc_default_settings(ppl_rrtc_settings* s);

typedef struct {
    int solved;
    int iters;                /* the solving workgroup's iterations, else the most any workgroup ran */
    int start_tree_size, goal_tree_size;
    int path_length;          /* tree nodes on the path (the reference's path_length) */
    int n_waypoints;          /* rows written to h_path */
    int reached_goal;         /* index of the goal root the path ends at */
    float cost;
    double wall_ms, kernel_ms;
} ppl_rrtc_result;

/* SolverBuffers (planning/solver_buffers.hh:15-54): device trees for up to
 * max_samples nodes each, num_new_configs workgroups of granularity states. */
typedef struct ppl_rrtc_buffers_s* ppl_rrtc_buffers;
ppl_status ppl_rrtc_buffers_create(int n_dof, int max_samples, int num_new_configs, int granularity,
                                   ppl_rrtc_buffers* out);
ppl_status ppl_rrtc_buffers_destroy(ppl_rrtc_buffers bufs);

/* Each tree walk of a solution holds at most 4096 nodes, so a path has at
 * most PPL_RRTC_MAX_WAYPOINTS rows; a connection whose walk is longer ends
 * the run unsolved. */
#define PPL_RRTC_MAX_WAYPOINTS 8194

/* h_goals [n_goals][n_dof]; h_path [max_waypoints][n_dof]. PPL_ERR_CAPACITY
 * when the solution has more than max_waypoints rows (result->n_waypoints
 * then holds the count needed); max_waypoints = PPL_RRTC_MAX_WAYPOINTS
 * always suffices. */
ppl_status ppl_rrtc_solve(ppl_robot robot, const ppl_scene_view* scene, const float* h_start, const float* h_goals,
                          int n_goals, const ppl_rrtc_settings* settings, ppl_rrtc_buffers bufs, float* h_path,
                          int max_waypoints, ppl_rrtc_result* result, ppl_stream stream);

/* The Halton bases [n_workgroups][n_dof] the solve uses for `seed`. */
ppl_status ppl_rrtc_halton_bases(unsigned long long seed, int n_workgroups, int n_dof, float* h_bases);

/* ppln::shortcut::shortcut_path (planning/shortcut.cuh:183-307): every skip
 * edge (i, j >= i + 2) of the n-waypoint path is checked at its interior
 * states t = s / total, s = 1 .. total - 1, total = max(1, ceil(|q_j - q_i| /
 * valid_segment_length)), q = q_i (1 - t) + q_j t (FP32, no contraction),
 * with the scene and self checks at collision_margin; then greedily from
 * waypoint 0 the farthest free j is taken (else i + 1). h_out [n][n_dof]
 * receives *n_out <= n waypoints. n <= PPL_RRTC_MAX_WAYPOINTS, the longest
 * path a solve returns (PPL_ERR_CAPACITY beyond). Synchronous on `stream`. */
ppl_status ppl_shortcut_path(ppl_robot robot, const ppl_scene_view* scene, const float* h_path, int n,
                             float valid_segment_length, float collision_margin, float* h_out, int* n_out,
                             ppl_stream stream);

/* ------------------------------------------------------------------------
 * Algorithmic flop counts (ppl/roofline.h) of one work unit for this robot:
 * the numerators of the benchmark's roofline fractions. No counterpart in
 * the reference. Returns a negative value for a bad handle or kind.
 * ------------------------------------------------------------------------ */
typedef enum {
    PPL_WORK_ANY_COLLISION = 0,   /* one state, exhaustive, n_obbs OBBs */
    PPL_WORK_HINGE_STATE = 1,     /* one state, OBB hinge (+ self min-hinge when the robot has pairs and use_self) */
    PPL_WORK_CHOMP_ADJOINT = 2,   /* one trajectory of n_waypoints: cost + fused-adjoint gradient */
    PPL_WORK_CHOMP_FD = 3,        /* one trajectory of n_waypoints: central-FD gradient */
    PPL_WORK_STOMP_STATE = 4,     /* one rollout waypoint: world + self hinge */
    PPL_WORK_ESDF_WAYPOINT = 5    /* one waypoint: trilinear ESDF hinge forward + backward */
} ppl_work_kind;
double ppl_algorithmic_flops(ppl_robot robot, int kind, int n_obbs, int n_waypoints, int use_self);

#ifdef __cplusplus
}
#endif
#endif /* PPL_PPL_H */
