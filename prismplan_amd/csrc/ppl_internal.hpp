// ppl_internal.hpp -- shared host/device declarations of the substrate library.
//
// Device data layout (all read-only during a launch):
//  * robot tables are re-ordered into the kinematic tree's DFS visiting order,
//    so the kernels walk joints with a plain counter and read every per-joint
//    value with a wave-uniform (scalar) load;
//  * spheres are stored in emission order (grouped by the DFS position of their
//    joint) as float4 (x, y, z, r), with their original index kept beside them
//    for ACM masks and outputs;
//  * OBBs are re-packed per launch into LDS as 16-float records
//    [R row0 | tx, R row1 | ty, R row2 | tz, hx, hy, hz, orig_index] with R the
//    rotation of the stored inverse pose, so a world point maps to box-local
//    coordinates with 9 FMAs;
//  * per-state inputs (configurations, edge endpoints, trajectories) stay in
//    the caller's layout in HBM; lanes of a wave take consecutive states.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ppl/ppl.h"

namespace ppl {

constexpr int kMaxJoints = 16;      // per-lane unrolled joint state (Baxter: 15)
constexpr int kMaxSlots = 2;        // FK transform slots (FK_T_SLOTS in the reference)
constexpr int kMaxDof = 15;
constexpr int kWave = 64;
constexpr int kMaxObbLds = 256;     // OBBs staged per problem
constexpr int kSelfGroup = 16;      // lanes per state in the self-collision pass
constexpr int kMaxGroup = 16;       // lanes per state of the cost kernel (power of two)
constexpr int kGrp = 16;            // lanes per state of the grouped cost kernel (>= kMaxJoints)
constexpr int kGrpStates = kWave / kGrp;
constexpr int kGroupKinds = 6;      // chunk tables for 64 / G states, G = 1 .. 32
constexpr int kMaxClusters = 64;    // sphere clusters per robot (cost-kernel masks in LDS)

enum JointType { kFixed = -1, kXPrism = 0, kYPrism = 1, kZPrism = 2, kXRot = 3, kYRot = 4, kZRot = 5 };

struct DRobot {
    int n_dof, n_joints, n_sph, n_apx, n_slots;
    int n_pairs, n_apx_pairs;
    const float* ft;            // [J][12] rows 0..2 of the fixed transform, DFS order
    const int4* jmeta;          // [J] (type, dof, slot, parent_slot), DFS order
    const int* sph_off;         // [J+1] sphere CSR by DFS position
    const float4* sph;          // [S] emission order
    const int* sph_id;          // [S] original index of emission slot k
    const int* apx_off;         // [J+1]
    const float4* apx;          // [A] emission order
    const int* apx_id;          // [A]
    const int* parent;          // [J] DFS position of the parent of DFS position k (root: 0)
    const int* jorig;           // [J] original joint index of DFS position k
    const float2* lim;          // [D] joint limits (lower, upper)
    const float4* sph_orig;     // [S] spheres by original index
    const int2* pairs;          // [n_pairs] exact self pairs (original indices), grouped by joint pair
    const int4* cpb;            // [n_cpb] self-pair blocks: (cluster A, cluster B, first pair, pairs), A <= B
    const unsigned* cpl;        // [n_pairs] block pair lists: slot a | slot b << 16 (emission slots)
    int n_cpb;
    const int2* apx_pairs;      // [n_apx_pairs] approx self pairs (approx original indices)
    const float4* apx_orig;     // [A] approx spheres by original index
    const int* sph_jnt;         // [S] DFS joint of emission slot k
    const int* jend;            // [J] end of DFS position k's subtree: its DFS positions are [k, jend[k])
    // The approximate spheres' own chain (DFS order), NULL when they ride the
    // full chain: rows 0..2 of each fixed transform and the joint types.
    const float* aft;
    const int* atype;
    const int* apx_jnt;         // [A] DFS joint of approximate sphere a (original index)
    // 1: every exact sphere lies inside an approximate sphere of its joint on
    // the same chain, so an exact (sphere, OBB) hit implies the approximate
    // hit of that OBB and the any-collision kernel gates each (joint, OBB)
    // pair by itself; 0: the gated form (joint flags from the approximate
    // pass over the whole scene, then the flagged joints' exact spheres).
    // Sphere clusters (contiguous emission runs inside a joint): the cost
    // kernels build one candidate-box mask per cluster around its enclosing
    // sphere (link frame, 1e-5 m padded).
    int n_clu;
    const int* clu_off;         // [J+1] clusters of DFS joint k: [clu_off[k], clu_off[k+1])
    const int* clu_beg;         // [C+1] emission slots of cluster c: [clu_beg[c], clu_beg[c+1])
    const float4* clu_enc;      // [C] enclosing sphere of cluster c
    const int* sph_clu;         // [S] cluster of emission slot k
    const int* clu_jnt;         // [C] DFS joint of cluster c
    int scene_gate_fused;
    // 1: an exact self hit implies an approximate one (every exact pair lies in
    // an approximate pair); 0: the self verdict needs the approximate pass too.
    int self_gate_implied;
};

// Non-owning scene pointers, SceneCollisionData semantics.
struct DScene {
    const float* obb_dims;      // [O][4] full extents
    const float* obb_pose;      // [O][8] inverse pose (t, qw, qx, qy, qz, pad)
    const uint8_t* obb_enable;  // [O]
    int n_obbs;
    const float* voxel_params;  // [L][4]
    const float* voxel_pose;    // [L][8]
    const uint8_t* voxel_enable;
    const float* voxel_features;
    int n_voxel_layers;
    int max_voxels_per_layer;
    const uint8_t* sphere_acm;           // [S][L]
    const uint8_t* sphere_obb_acm;       // [S][O]
    const uint8_t* sphere_obb_acm_approx;
    int esdf_interp;            // 0 nearest, 1 trilinear
};

// Where a flattened state index comes from.
enum StateSource { kSrcConfig = 0, kSrcEdge = 1, kSrcTraj = 2, kSrcDense = 3, kSrcBfs = 4 };

struct DStates {
    int source;
    long n_states;
    // kSrcConfig: q(g, d) = cfg[g * cfg_stride + d]
    const float* cfg; int cfg_stride;
    // kSrcEdge: state g in edge e with edge_off[e] <= g < edge_off[e+1];
    // k = g - edge_off[e] + k0, t = k / (n_e - 1) with n_e = edge_off[e+1]-edge_off[e]+k0
    const float* from; const float* to; const int* edge_off; int n_edges; int k0;
    // kSrcTraj: waypoint g of trajectory b (traj_off[b] <= g < traj_off[b+1]),
    // theta(b, d, t) = theta[traj_off[b] * D + d * T_b + t]
    const float* theta; const int* traj_off; int n_traj;
    // Waypoint-major trajectory plans: chunk index v of problem p maps to
    // trajectory prob[p].x + v % prob[p].y, waypoint v / prob[p].y, where
    // prob[p] = (first trajectory, trajectories, length, first state).
    const int4* prob; int wmajor;
    // kSrcDense: densified trajectories (theta, traj_off, n_traj as kSrcTraj):
    // state g of trajectory b (dense_off[b] <= g < dense_off[b+1]) is
    // f = g - dense_off[b]; f < (T_b - 1) nsub[b] lies at segment f / nsub[b],
    // fraction (f % nsub[b]) / nsub[b]; the last one is waypoint T_b - 1.
    const int* dense_off; const int* nsub;
    // kSrcBfs: graph edges node_src[e] -> node_dst[e] over node configurations
    // cfg[node * cfg_stride + d]; state g of edge e (dense_off[e] <= g <
    // dense_off[e+1]) is breadth-first midpoint k = bfs_k0[e] + g - dense_off[e]
    // (t = 1/2, 1/4, 3/4, 1/8, ...), n_traj holds the edge count.
    const int* node_src; const int* node_dst; const int* bfs_k0;
    // when set, the state count is read from device memory (*n_dyn) at launch
    const int* n_dyn;
};

// Work chunks for persistent launches: a chunk never spans two problems.
struct DChunks {
    const int4* tab;            // (problem, begin, end, first unit or -1); end - begin <= chunk
    int n_chunks;
    int chunk;                  // states per chunk (64 / lanes per state)
    const int* obb_off;         // [P+1] OBB ranges per problem; NULL = single problem
    int obb_cap;                // OBB records one LDS staging holds (a tile; 1 .. kMaxObbLds)
    int tiled;                  // some problem has more OBBs than obb_cap: they pass through LDS in tiles
                                // (k_collide, the grouped cost kernel; k_cost refuses such launches)
    const int* unit_hint;       // implicit chunks (tab NULL), nullable: unit holding state 64 k, for each k
};

struct DCostParams {
    float world_margin, self_margin, weight;
    int use_obb, use_esdf, use_self;
};

// ---- launchers (kernels.hip) ----
struct CollideArgs {
    DRobot R; DScene S; DStates X; DChunks C;
    float margin;
    int mode;                   // 0 per-state verdict, 1 edge first-hit
    int skip_resolved;          // first-hit mode: skip states past a known hit
    uint8_t* out;               // [n_states] verdicts (nullable in mode 1)
    int* first_hit;             // [n_edges] (mode 1), -1 = free
    const uint8_t* self_hit;    // [n_states] precomputed self verdicts (nullable)
};

struct CostArgs {
    DRobot R; DScene S; DStates X; DChunks C;
    DCostParams P;
    int group;                  // lanes per state (power of two <= kMaxGroup): a state's spheres are split over them
    int grad;                   // also compute d cost / d q
    int clearance_only;         // write min ESDF clearance (+margin) instead of cost
    float margin;               // clearance_only: radius margin
    const float* self_in;       // [n_states][8] self pass output (nullable)
    float* cost_out;            // [n_states] weighted cost (or clearance)
    float* grad_out;            // layout of X (kSrcConfig: [n][D]; kSrcTraj: theta layout)
};

struct SelfArgs {
    DRobot R; DStates X;
    int mode;                   // 0: any-collision bit, 1: min clearance + argmin pair + normal
    float band;                 // mode 1: pairs whose joints' enclosing spheres are >= band apart are
                                // skipped (the minimum is exact whenever it is below band)
    uint8_t* hit_out;           // [n_states]
    int* unit_first_hit;        // mode 0, nullable: per-unit first hit (-1 free); states of units with a
                                // hit are skipped and a self collision lowers the unit's entry instead of hit_out
    float* out;                 // [n_states][8]: clr, i, j (as int bits), nx, ny, nz, -, -
    DChunks C;                  // the plan's 4-state chunks (tab NULL: states in flat order); outputs are
                                // indexed by plan order either way
};

// t0 / t1 (optional): events the runtime records at the kernel's start and end
hipError_t launch_collide(const CollideArgs& a, int grid, hipStream_t s, hipEvent_t t0 = nullptr,
                          hipEvent_t t1 = nullptr);
hipError_t launch_cost(const CostArgs& a, int grid, hipStream_t s, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
hipError_t launch_cost_grp(const CostArgs& a, int grid, hipStream_t s, hipEvent_t t0 = nullptr,
                           hipEvent_t t1 = nullptr);
hipError_t launch_self(const SelfArgs& a, hipStream_t s);
hipError_t launch_fill_i32(int* p, int v, long n, hipStream_t s);
hipError_t launch_add_u32(unsigned* p, unsigned v, hipStream_t s);   // *p += v (one thread)
// CHOMP optimiser pieces (trajectory plans, dim-major [D][T] blocks).
struct ChompArgs {
    DRobot R; DStates X;        // X: trajectory source over theta
    float step_size, smoothness_weight, check_step, margin, convergence_eps, dt;
    int polish0;                // iterations after valid
    int nsub_max;               // densification cap (joint-limit box diagonal / step)
    const float* const* ainv;   // [n_traj] A^-1 (T x T) of each trajectory's length
    float* theta; float* best; float* record; const float* grad; const float* total;
    const int* first_hit;       // [n_traj] validity pass result, -1 = collision-free
    int* nsub; int* dense_off;  // [n_traj], [n_traj + 1]
};
hipError_t launch_chomp_nsub(const ChompArgs& a, hipStream_t s);   // nsub + dense_off (scan)
hipError_t launch_chomp_book(const ChompArgs& a, int initial, hipStream_t s);
hipError_t launch_chomp_update(const ChompArgs& a, int max_t, hipStream_t s);

// STOMP optimiser population of trajectory b (waypoint offset off_b, length
// T_b): slot r < nf is a fresh sample, nf <= r < n_roll - 1 a rollout kept
// from the previous population, n_roll - 1 the optimum it was drawn around.
struct StompPop {
    const float* fresh; const float* fresh_cost; int nf;   // [B][nf][D][T_b] and [B][nf][T_b], b-major
    const float* keep; const float* keep_cost; int nkeep;  // [B][nkeep][D][T_b] and [B][nkeep][T_b]
    const float* slot; const float* slot_cost;             // optimum snapshot, plan layout
};
struct StompArgs {
    DRobot R;
    int B; const int* traj_off;       // trajectory plan: [B + 1] waypoint offsets
    int n_new, n_old, n_roll;         // n_roll = n_new + n_old + 1
    float w_u, h, dt;                 // control weight, cost sensitivity, metric step
    unsigned key0, key1, iter;        // Philox key and the iteration word of the counter
    const unsigned* iter_dev;         // non-null: the iteration word is *iter_dev + iter (graph replays)
    const float* const* Lt;           // [B] sampling factor L^T (T x T, entry k * T + t = L[t][k])
    const float* const* Mt;           // [B] projection M^T (entry u * T + t = M[t][u])
    float* theta; const float* theta_cost; const float* record;
    StompPop cur, prev;
    float* fresh; float* keep; float* keep_cost; float* slot; float* slot_cost;   // writable views of cur
    float* ctrl_sum;                  // [B][n_roll][D]: sum_t of each row's control cost
    int* sel;                         // [B][n_old]: previous-population slots kept, cheapest first
};
hipError_t launch_stomp_reuse(const StompArgs& a, int max_t, hipStream_t s);    // select + gather kept rollouts
hipError_t launch_stomp_sample(const StompArgs& a, int max_t, hipStream_t s);   // optimum snapshot + fresh samples
hipError_t launch_stomp_update(const StompArgs& a, int max_t, hipStream_t s);   // weights, projection, step
size_t stomp_update_lds(int n_roll, int max_t);

// MIT* consumers: per-edge breadth-first midpoint counts + their scan (*total
// is the 64-bit sum), the edge verdicts, sample clearances, gather / scatter.
hipError_t launch_bfs_count(const float* nodes, int dim, const int* src, const int* dst, const int* done, int n,
                            float seg_len, int sparse, int* k0, int* cnt, int* off, long long* total, hipStream_t s);
// hint[k] = the unit holding state 64 k, for the units' offsets off[0..n]; k < n_hint
hipError_t launch_unit_hint(const int* off, int n, int* hint, int n_hint, hipStream_t s);
hipError_t launch_bfs_finish(const int* first_hit, const int* cnt, int n, uint8_t* cc, int* new_checks, hipStream_t s);
hipError_t launch_sample_clearance(const uint8_t* cc, const float* clr, int n, float weight, float* out, hipStream_t s);
hipError_t launch_gather_rows(const float* src, int dim, const int* idx, int base, int n, float* out, hipStream_t s);
hipError_t launch_scatter(const float* v, const int* idx, int base, int n, float* out, hipStream_t s);

// Joint frames [n][J][16] (column-major 4x4, original joint order) and / or
// sphere centres [n][S][4] (x, y, z, r, original order) of configurations.
hipError_t launch_fk_out(const DRobot& R, const DStates& X, float* frames, float* spheres, hipStream_t s);
hipError_t launch_segment_sum(const float* x, const int* off, int n_seg, float scale_unused, float* out,
                              hipStream_t s);
// t_max: the longest trajectory (0: unknown, no LDS staging); state
// (nullable): per-waypoint state costs whose per-trajectory sum joins out.
hipError_t launch_control_cost(const float* theta, const int* traj_off, int n_traj, int D, float dt,
                               float w, float* out, int accumulate, hipStream_t s, int t_max = 0,
                               const float* state = nullptr);
hipError_t launch_esdf_build(const float* obb_dims, const float* obb_pose, int n_obbs,
                             const float* voxel_params, const float* voxel_pose, float* out,
                             hipStream_t s);
size_t lds_bytes(int obb_cap, int n_dof);
// Cost kernel: OBB records + q + (grad) per-state adjoint records.
size_t lds_bytes_cost(int obb_cap, int n_dof, int n_joints, int G, bool grad);
// Floats per state of the adjoint records: 12 per joint (odd stride), d cost / d q per dof.
__host__ __device__ inline int rec_stride(int n_joints) { return 12 * n_joints + 1; }
__host__ __device__ inline int dq_stride(int n_dof) { return n_dof | 1; }
int device_cus();
int grid_collide(size_t lds);
int grid_cost(int n_joints, bool grad, size_t lds);
size_t lds_bytes_grp(int obb_cap, int n_dof, int n_joints, int n_sph, int n_clu, bool grad, bool tiled);
int grid_cost_grp(bool grad, int world, bool tiled, size_t lds);
int grp_world(const CostArgs& a);

}  // namespace ppl
