/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of PrismPlan's collision-kinematics substrate, written from
 * the reference kernels' documented semantics. Two instantiations:
 *   *_f64 : the FP64 oracle every parity test checks the HIP path against;
 *   *_f32 : the FP32 CPU path timed beside the GPU numbers ("cpu_baseline").
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as a checker or a timed baseline. The product
 * (prismplan_amd/) never links or calls it.
 *
 * Inputs are always FP32 (the oracle consumes the FP32-cast values the GPU
 * sees); arithmetic runs in the instantiation's precision.
 *
 * Reference files (paths relative to code/gpu_planning/cpp/src/):
 *   planning/runtime_kinematics.cuh:16-88, 248-329, 382-416
 *   planning/utils.cuh:239-267, 549-711
 *   collision/scene_collision.cuh:24-631
 *   collision/two_phase_cc.cuh:57-138
 *   planning/edge_eval.cuh:37-165
 *   planning/pRRTC.cu:2337-2386
 *   planning/chomp.cu:83-243, planning/stomp.cu:206-308
 *   planning/chomp_precompute.cpp:20-38 (R matrix)
 */
#ifndef PRISMPLAN_ORACLE_H
#define PRISMPLAN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flat robot tables, RobotModel field meanings (planning/robot_model.cuh:27-82). */
typedef struct {
    int n_dof, n_joints;
    const float* fixed_transforms;   /* [J*16] row-major 4x4 */
    const int* joint_types;          /* [J] -1 fixed, 0..2 prism x/y/z, 3..5 rot x/y/z */
    const int* joint_parents;        /* [J] */
    const int* joint_id_to_dof;      /* [J] -1 for the base */
    const int* dfs_order;            /* [J] */
    int n_spheres; const float* spheres; const int* sphere_to_joint;           /* [S*4], [S] */
    int n_approx; const float* approx_spheres; const int* approx_sphere_to_joint;
    const float* approx_fixed_transforms; const int* approx_joint_types;     /* NULL: full */
    int n_self_ranges; const int* self_ranges;                               /* [n*3] */
    int n_approx_self_ranges; const int* approx_self_ranges;
} orc_robot;

/* SceneCollisionData field meanings (collision/scene_collision_data.hh:10-47). */
typedef struct {
    int n_obbs; const float* obb_dims; const float* obb_pose; const uint8_t* obb_enable;
    const uint8_t* sphere_obb_acm;         /* [S*O] or NULL */
    const uint8_t* sphere_obb_acm_approx;  /* [A*O] or NULL */
    int n_voxel_layers; const float* voxel_params; const float* voxel_pose;
    const uint8_t* voxel_enable; const float* voxel_features; int max_voxels_per_layer;
    const uint8_t* sphere_acm;             /* [S*L] or NULL */
    int esdf_interp;                       /* 0 nearest voxel (reference), 1 trilinear */
} orc_scene;

typedef struct {
    float world_margin;   /* hinge band for OBB and ESDF terms (raw radius) */
    float self_margin;    /* hinge band for the self min-clearance term */
    float weight;         /* collision_cost_weight */
    int use_obb, use_esdf, use_self;
} orc_cost_params;

typedef struct {
    int granularity; float collision_margin; int check_self; int compute_clearance;
    float resolution;
} orc_edge_params;

/* ---- kinematics ---- */
void orc_fk_spheres_f64(const orc_robot* r, const float* q, int n, double* out);   /* [n][S][3] */
void orc_fk_frames_f64(const orc_robot* r, const float* q, int n, double* out);    /* [n][J][12] */
void orc_fk_spheres_f32(const orc_robot* r, const float* q, int n, float* out);

/* ---- single sphere vs single OBB (scene_collision.cuh:59-112) ---- */
int    orc_sphere_obb_collide_f64(const float* dims, const float* pose, double x, double y, double z, double r);
double orc_sphere_obb_clearance_f64(const float* dims, const float* pose, double x, double y, double z, double r);
double orc_sphere_esdf_clearance_f64(const orc_scene* s, double x, double y, double z, double r, int sphere_idx);

/* ---- robot-level any-collision ----
 * mode 0: two-phase gate (reference semantics), mode 1: exhaustive exact check.
 * clr (nullable): per-config exhaustive minimum signed clearance over every
 * checked item, used to find the ambiguous band around 0. */
void orc_check_collision_f64(const orc_robot* r, const orc_scene* s, const float* q, int n,
                             float margin, int check_self, int mode, uint8_t* out, double* clr,
                             int nthreads);
void orc_check_collision_f32(const orc_robot* r, const orc_scene* s, const float* q, int n,
                             float margin, int check_self, int mode, uint8_t* out, float* clr,
                             int nthreads);

/* scene_min_clearance_runtime (ESDF only, include_obb = 0) or OBB+ESDF (include_obb = 1) */
void orc_min_clearance_f64(const orc_robot* r, const orc_scene* s, const float* q, int n,
                           float margin, int include_obb, double* out, int nthreads);
/* The same, with vamb[i] = 1 where some ESDF lookup of configuration i sits
 * within 1e-4 voxel of a face across which the looked-up value jumps (a
 * nearest lookup whose neighbour voxel holds another value or is skipped,
 * or the guarded grid boundary) and the jump can move the minimum: an FP32
 * evaluation may take either side. lo[i] (nullable) = the lowest value the
 * minimum can take with the other sides (= out[i] when none undercuts it). */
void orc_min_clearance_amb_f64(const orc_robot* r, const orc_scene* s, const float* q, int n,
                               float margin, int include_obb, double* out, uint8_t* vamb, double* lo,
                               int nthreads);

/* ---- per-configuration summed-hinge state cost and its gradient ----
 * cost[i] = weight * (obb hinge + esdf hinge + self min-hinge), chomp.cu:151-188.
 * grad[i][d] = d cost / d q_d (analytic, argmin pair); kink[i] != 0 marks
 * points within tol of a hinge kink, an argmin tie or a box-SDF kink. */
void orc_hinge_cost_f64(const orc_robot* r, const orc_scene* s, const float* q, int n,
                        const orc_cost_params* p, double* cost, int nthreads);
void orc_hinge_cost_f32(const orc_robot* r, const orc_scene* s, const float* q, int n,
                        const orc_cost_params* p, float* cost, int nthreads);
/* FP64 state cost with the value-ambiguity flag of orc_min_clearance_amb_f64
 * for the ESDF hinge terms (a sphere whose lookup sits at such a face), and
 * the number of non-zero hinge terms (OBB, ESDF, self) in each cost: the
 * parity tests' tolerance of a sum is the sum of its terms' tolerances.
 * vamb and nterm may be null. */
void orc_hinge_cost_amb_f64(const orc_robot* r, const orc_scene* s, const float* q, int n,
                            const orc_cost_params* p, double* cost, uint8_t* vamb, int* nterm,
                            int nthreads);
void orc_hinge_grad_f64(const orc_robot* r, const orc_scene* s, const float* q, int n,
                        const orc_cost_params* p, double* cost, double* grad, int* kink,
                        double tol, int nthreads);

void orc_hinge_grad_f32(const orc_robot* r, const orc_scene* s, const float* q, int n,
                        const orc_cost_params* p, float* cost, float* grad, int nthreads);

/* ---- trajectories, ragged [B] x [D][T_b] at waypoint offsets off[B+1] ---- */
/* total[b] = sum_t cost(theta_b[:, t]); grad[b][d][t] = d total / d theta, 0 at both ends */
void orc_traj_cost_grad_f64(const orc_robot* r, const orc_scene* s, const float* theta,
                            const int* off, int B, const orc_cost_params* p, double* total,
                            double* grad, int* kink, double tol, int nthreads);
/* chomp_obs_grad_kernel semantics (chomp.cu:83-144): central FD, endpoints 0 */
void orc_chomp_fd_grad_f64(const orc_robot* r, const orc_scene* s, const float* theta,
                           const int* off, int B, const orc_cost_params* p, float eps,
                           double* grad, int nthreads);
void orc_chomp_fd_grad_f32(const orc_robot* r, const orc_scene* s, const float* theta,
                           const int* off, int B, const orc_cost_params* p, float eps,
                           float* grad, int nthreads);

/* ---- edges ---- */
/* evaluate_edges_batch semantics (edge_eval.cuh:64-165) */
void orc_edge_eval_f64(const orc_robot* r, const orc_scene* s, const float* from, const float* to,
                       int E, const orc_edge_params* p, uint8_t* valid, float* cost,
                       double* clearance, int nthreads);
/* Batched edges: edge e has n_states[e] states at t = k/(n-1); states k >= k0
 * are checked (k0 = skip_first); first_hit[e] = smallest colliding k or -1. */
void orc_edge_batch_f64(const orc_robot* r, const orc_scene* s, const float* from, const float* to,
                        const int* n_states, int E, int skip_first, float margin, int check_self,
                        int* first_hit, int nthreads);
void orc_edge_batch_f32(const orc_robot* r, const orc_scene* s, const float* from, const float* to,
                        const int* n_states, int E, int skip_first, float margin, int check_self,
                        int* first_hit, int nthreads);
/* The FP32 states the batched form checks, flattened in edge order. */
void orc_edge_states_f32(const float* from, const float* to, const int* n_states, int E, int dim,
                         int skip_first, float* out);
/* Reference granularity rule (edge_eval.cuh:54-60) evaluated in FP32. */
int orc_edge_granularity(const float* from, const float* to, int dim, int granularity, float resolution,
                         float* len_out);

/* ---- smoothness metric ---- */
void orc_control_R_f64(int T, double dt, double* R);   /* chomp_precompute.cpp:20-38 */
/* per-rollout control cost: w * sum_d theta_d^T R theta_d, ragged layout */
void orc_control_cost_f64(const float* theta, const int* off, int B, int D, double dt, double w,
                          double* out);

/* ---- STOMP rollouts ----
 * Problem p: rollouts [R][D][T_p] at waypoint offset roff[p] (R*T_p waypoints),
 * OBBs [obb_off[p], obb_off[p+1]) of the concatenated scene arrays.
 * state[...] = weight*(world hinge + self min-hinge) per rollout waypoint,
 * total[p*R + r] = sum_t state + w_u * sum_d theta^T R theta.
 * nterm (may be null): non-zero hinge terms of each state cost.                   */
void orc_stomp_costs_f64(const orc_robot* r, const orc_scene* all, const int* obb_off,
                         const float* rollouts, const int* T_p, const int* roff, int P, int R,
                         const orc_cost_params* p, double w_u, double dt, double* state,
                         double* total, int nthreads, int* nterm);
void orc_stomp_costs_f32(const orc_robot* r, const orc_scene* all, const int* obb_off,
                         const float* rollouts, const int* T_p, const int* roff, int P, int R,
                         const orc_cost_params* p, double w_u, double dt, float* state,
                         float* total, int nthreads);

/* ---- algorithmic work of the gated early-exit any-collision ----
 * Per configuration, flops of the canonical order: DFS joints; per joint its
 * approximate sphere against every enabled OBB; the joint's exact spheres
 * against each OBB the approximate sphere hits; stop after the OBB at which a
 * hit is found. Costs per item: revolute joint 83, prismatic 69, sphere
 * transform 18, sphere-OBB test 30 (OBBs only; no ESDF). */
void orc_gated_flops_f64(const orc_robot* r, const orc_scene* s, const float* q, int n, float margin,
                         double* flops, int nthreads);

/* ---- ESDF from OBBs: node (i,j,k) value = -min_o SDF_o(node), positive inside ---- */
void orc_esdf_from_obbs_f64(const float* obb_dims, const float* obb_pose, int n_obbs,
                            const float* voxel_params, const float* voxel_pose, double* out,
                            int nthreads);

#ifdef __cplusplus
}
#endif
#endif
