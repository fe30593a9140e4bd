// rrtc.hip -- pRRTC (GPU-parallel RRT-Connect) as a planner kernel on the
// device-inline substrate API (ppl/ppl_device.hpp) and the public C ABI,
// plus the path shortcut pass.
//
// Reference: planning/pRRTC.cu:1556-2010 (kernel rrtc_runtime_scene),
// :2016-2223 (solve_runtime_scene), :955-995 (Halton), planning/shortcut.cuh:
// 16-307 (shortcut_path), planning/pRRTC_settings.hh:3-32 (settings).
//
// Layout in HBM (one ppl_rrtc_buffers):
//   nodes   [2][max_samples][n_dof] FP32; rows not yet written hold -9999, so a
//           row that is not (fully) visible to a reader is never its nearest
//   parents [2][max_samples] int32; a root is its own parent, every other
//           node's parent has a smaller index
//   radii   [2][max_samples] FP32 dynamic-domain radius, FLT_MAX = unbounded
//   ctl     RrtcCtl: run status, per-tree counters, the solving workgroup,
//           the path walk's results
//   bases   [workgroups][16] FP32 Halton bases
//   path    [2][kPathCap][n_dof] FP32 the two tree walks of the solution
// One workgroup of 256 lanes (four waves) owns one new configuration per
// iteration; each aligned group of G = 256 / granularity lanes (a power of
// two, at most 64) checks one state of an edge with two_phase_cc<G>. Every
// access one workgroup makes to another's data (counters, node rows, radii,
// the status word) is an agent-scope atomic, coherent across the XCDs' L2s
// without fences. Each row element goes from -9999 to its value once, so a
// reader sees a row whole or with a sentinel element that keeps it from
// being nearest; parents are read only after the kernel ends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "ppl/ppl.h"
#include "ppl/ppl_device.hpp"

namespace ppl {   // capi.hip
ppl_status fail_msg(ppl_status s, const char* msg);          // sets ppl_last_error
ppl_status scene_view_check(const ppl_scene_view* v);        // pointers and alignment
}

namespace {

using namespace ppl::device;

constexpr int kMaxDim = 16;
constexpr int kMaxGran = 64;         // states per edge check
constexpr int kThreads = 256;        // four waves per workgroup
constexpr int kMaxThreads = kThreads;
constexpr int kMaxWaves = kMaxThreads / 64;
constexpr int kPathCap = 4096;       // nodes per tree walk
constexpr float kUnwritten = -9999.0f;
constexpr size_t kMaxDynLds = 56 * 1024;
constexpr long long kMaxShortcutStates = 1ll << 26;
constexpr int kHeadRows = 64;        // path rows read back with the control block
static_assert(PPL_RRTC_MAX_WAYPOINTS == 2 * kPathCap + 2, "two tree walks plus both roots");

struct RrtcCtl {
    int status;            // 0 running, 1 solved, -1 stopped (limit reached)
    int free_idx[2];       // next node index per tree (may run past max_samples)
    int completed[2];      // rows written per tree
    int win_tree, win_t, win_o, win_iter;
    int max_iter;
    int path_n[2];
    int reached_goal;
    int path_ok;
    float cost;
    int n_waypoints;       // rows of the assembled start -> goal path
};
static_assert(sizeof(RrtcCtl) <= 256, "control block must fit its 256-byte slot");

// Word offsets (into the dynamic LDS, after the sphere positions) of the LDS
// copies of the robot tables and the scene's OBBs. The edge checks read
// every table once per sphere or pair, so a copy in LDS turns dependent
// global-memory latency into LDS latency. A group that does not fit the
// budget stays in global memory (flag 0).
struct StageLayout {
    int robot, pairs, obbs;
    int ft, jmeta, sph_off, sph, sph_id, apx_off, apx, apx_id, sph_orig, apx_orig, apx_pairs, pairs_at;
    int obb_dims, obb_pose, obb_enable;
};

struct RrtcArgs {
    ppl_robot_device_view R;
    ppl_scene_view S;
    float* nodes[2];
    int* parents[2];
    float* radii[2];
    RrtcCtl* ctl;
    const float* bases;    // [workgroups][16]
    float* path;           // [2][kPathCap][dim] tree walks
    float* out;            // [2 * kPathCap + 2][dim] assembled path, start -> goal
    const float* roots;    // [1 + n_goals][dim] start, then the goals
    int n_goals;
    int dim, max_samples, max_iters, granularity, balance, dynamic_domain, n_blocks;
    float range, tree_ratio, dd_alpha, dd_radius, dd_min_radius, margin;
    long long time_ticks;  // device wall-clock ticks, 0 = no limit
    StageLayout L;         // LDS copies of the robot tables and the OBBs
};

__device__ __forceinline__ int ld_agent(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float ld_agent_f(const float* p) {
    return __int_as_float(ld_agent(reinterpret_cast<const int*>(p)));
}

__device__ __forceinline__ void st_agent_f(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<int*>(p), __float_as_int(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Dynamic-domain radius update (pRRTC.cu:1788-1799 grow on success,
// :1991-2004 shrink on failure) as one compare-and-swap on the value just
// read: an update that loses a race is dropped instead of retried. Early in a
// run hundreds of workgroups fail against the same root, and a retry loop
// serialises them on one address; a dropped update moves a radius by
// dd_alpha (1e-4 by default) at most.
__device__ void radius_update(float* r, bool grow, float alpha, float dd_radius, float dd_min) {
    int* p = reinterpret_cast<int*>(r);
    const int cur = ld_agent(p);
    const float old = __int_as_float(cur);
    float nv;
    if (grow) {
        if (old == FLT_MAX) return;
        nv = old * (1.0f + alpha);
    } else {
        nv = old == FLT_MAX ? dd_radius : fmaxf(old * (1.0f - alpha), dd_min);
    }
    if (nv != old) atomicCAS(p, cur, __float_as_int(nv));
}

// Workgroup (distance, index) minimum: smaller distance, then smaller index.
// Leaves the result in red_d[0], red_i[0] for every lane.
__device__ void block_argmin(float d, int i, float* red_d, int* red_i) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float od = __shfl_xor(d, o);
        const int oi = __shfl_xor(i, o);
        if (od < d || (od == d && oi < i)) {
            d = od;
            i = oi;
        }
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red_d[w] = d;
        red_i[w] = i;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
            if (red_d[k] < red_d[0] || (red_d[k] == red_d[0] && red_i[k] < red_i[0])) {
                red_d[0] = red_d[k];
                red_i[0] = red_i[k];
            }
    }
    __syncthreads();
}

// Nearest of the first `size` rows of a tree to q (squared L2, brute force
// over the workgroup's lanes: pRRTC.cu:1637-1666).
__device__ void nearest(const float* nodes, int size, const float* q, int dim, float* red_d, int* red_i) {
    float best = FLT_MAX;
    int bi = 0;
    for (int i = threadIdx.x; i < size; i += blockDim.x) {
        const float* x = nodes + (size_t)i * dim;
        float s = 0.0f;
        for (int k = 0; k < dim; ++k) {
            const float t = ld_agent_f(x + k) - q[k];
            s = fmaf(t, t, s);
        }
        if (s < best) {
            best = s;
            bi = i;
        }
    }
    block_argmin(best, bi, red_d, red_i);
}

// Whether any of the states base + (k + 1) * step, k = 0 .. granularity - 1
// (one per 4-lane group), collides: the two-phase scene test and the self
// check (pRRTC.cu:1680-1781, 1867-1948).
template <int G>
__device__ bool edge_hit(const RrtcArgs& A, const ppl_robot_device_view& R, const ppl_scene_view& S, const float* base,
                         const float* step, float* qbuf, float* pos, float* apos, float* frames) {
    const int g = threadIdx.x / G;
    bool hit = false;
    if (g < A.granularity) {
        float* q = qbuf + g * kMaxDim;
        for (int k = threadIdx.x & (G - 1); k < A.dim; k += G) q[k] = fmaf((float)(g + 1), step[k], base[k]);
        group_sync();
        const TwoPhaseResult r = two_phase_cc<G>(R, S, q, pos + (size_t)g * 3 * R.n_spheres,
                                                  apos + (size_t)g * 3 * R.n_approx_spheres,
                                                  frames + (size_t)g * 12 * R.n_joints, threadIdx.x, A.margin, true);
        hit = r.collision;
    }
    return __syncthreads_or(hit ? 1 : 0) != 0;
}

template <class T>
__device__ const T* stage_words(uint32_t* lds_words, int off, const T* src, int n_words) {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    for (int i = threadIdx.x; i < n_words; i += blockDim.x) lds_words[off + i] = s[i];
    return reinterpret_cast<const T*>(lds_words + off);
}

// Copy the robot tables and OBBs the layout selects into LDS and point the
// views at the copies.
__device__ void stage_tables(const RrtcArgs& A, uint32_t* W, ppl_robot_device_view& R, ppl_scene_view& S) {
    const StageLayout& L = A.L;
    const int J = R.n_joints, Sn = R.n_spheres, An = R.n_approx_spheres;
    if (L.robot) {
        R.ft = stage_words(W, L.ft, R.ft, J * 12);
        R.jmeta = stage_words(W, L.jmeta, R.jmeta, J * 4);
        R.sph_off = stage_words(W, L.sph_off, R.sph_off, J + 1);
        R.sph = stage_words(W, L.sph, R.sph, Sn * 4);
        R.sph_id = stage_words(W, L.sph_id, R.sph_id, Sn);
        R.apx_off = stage_words(W, L.apx_off, R.apx_off, J + 1);
        R.apx = stage_words(W, L.apx, R.apx, An * 4);
        R.apx_id = stage_words(W, L.apx_id, R.apx_id, An);
        R.sph_orig = stage_words(W, L.sph_orig, R.sph_orig, Sn * 4);
        R.apx_orig = stage_words(W, L.apx_orig, R.apx_orig, An * 4);
        R.apx_pairs = stage_words(W, L.apx_pairs, R.apx_pairs, R.n_approx_pairs * 2);
    }
    if (L.pairs) R.pairs = stage_words(W, L.pairs_at, R.pairs, R.n_pairs * 2);
    if (L.obbs && S.n_obbs > 0) {
        S.obb_dims = stage_words(W, L.obb_dims, S.obb_dims, S.n_obbs * 4);
        S.obb_pose = stage_words(W, L.obb_pose, S.obb_pose, S.n_obbs * 8);
        uint8_t* en = reinterpret_cast<uint8_t*>(W + L.obb_enable);
        for (int i = threadIdx.x; i < S.n_obbs; i += blockDim.x) en[i] = S.obb_enable[i];
        S.obb_enable = en;
    }
    __syncthreads();
}

// Count a row of tree t once its lanes have stored it. The count only bounds
// the readers' scans: rows are read and written with agent-scope atomics, so
// no fence orders them, and a row that is not yet (fully) visible reads as
// the -9999 sentinel.
__device__ void publish(const RrtcArgs& A, int t) {
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&A.ctl->completed[t], 1);
}

// Tree a workgroup grows this iteration (pRRTC.cu:1602-1627).
__device__ int choose_tree(const RrtcArgs& A, int iter, int prev_t) {
    const int bid = blockIdx.x;
    if (A.balance == 0 || iter == 1) return bid < A.n_blocks / 2 ? 0 : 1;
    const int f0 = ld_agent(&A.ctl->free_idx[0]), f1 = ld_agent(&A.ctl->free_idx[1]);
    if (A.balance == 1) {
        const float ratio = (float)f0 / (float)(f0 + f1);
        if ((float)abs(f0 - f1) < 1.5f * (float)A.n_blocks) return bid < (float)A.n_blocks * (1.0f - ratio) ? 0 : 1;
        return ratio < A.tree_ratio ? 0 : 1;
    }
    if (A.balance == 2) {
        const int ft = prev_t == 0 ? f0 : f1, fo = prev_t == 0 ? f1 : f0;
        const float ratio = (float)abs(ft - fo) / (float)ft;
        return ratio < A.tree_ratio ? 1 - prev_t : prev_t;
    }
    return prev_t;
}

template <int G>
__global__ __launch_bounds__(kThreads) void k_rrtc(RrtcArgs A) {
    extern __shared__ float lds[];   // per-state sphere positions and frames, then the staged tables (StageLayout)
    __shared__ float s_cfg[kMaxDim], s_near[kMaxDim], s_step[kMaxDim], s_vec[kMaxDim];
    __shared__ float s_hn[kMaxDim], s_hd[kMaxDim];
    __shared__ float s_q[kMaxGran * kMaxDim];
    __shared__ float red_d[kMaxWaves];
    __shared__ int red_i[kMaxWaves];
    __shared__ int s_t, s_stop, s_skip, s_size, s_index, s_next;
    __shared__ float s_dist;
    const int tid = threadIdx.x, dim = A.dim;
    float* pos = lds;   // [granularity][3 * n_spheres], then the approximate spheres, then the joint frames
    float* apos = lds + (size_t)A.granularity * 3 * A.R.n_spheres;
    float* frames = apos + (size_t)A.granularity * 3 * A.R.n_approx_spheres;
    ppl_robot_device_view R = A.R;
    ppl_scene_view S = A.S;
    stage_tables(A, reinterpret_cast<uint32_t*>(lds), R, S);
    const float* hb = A.bases + (size_t)blockIdx.x * kMaxDim;
    if (tid < dim) {
        s_hn[tid] = 0.0f;
        s_hd[tid] = 1.0f;
    }
    if (tid == 0) s_t = 0;
    const long long t0 = wall_clock64();
    int iter = 0;
    for (;;) {
        ++iter;
        __syncthreads();
        if (tid == 0) {
            int stop = ld_agent(&A.ctl->status) != 0;
            if (!stop && (iter > A.max_iters || (A.time_ticks > 0 && wall_clock64() - t0 > A.time_ticks))) {
                atomicCAS(&A.ctl->status, 0, -1);
                stop = 1;
            }
            s_stop = stop;
            if (!stop) s_t = choose_tree(A, iter, s_t);
        }
        if (tid < dim) {   // Halton sample scaled to the joint limits (pRRTC.cu:1634-1635)
            const float h = halton_step(hb[tid], s_hn[tid], s_hd[tid]);
            const float lo = A.R.lim[2 * tid], hi = A.R.lim[2 * tid + 1];
            s_cfg[tid] = fmaf(h, hi - lo, lo);
        }
        __syncthreads();
        if (s_stop) break;
        const int t = s_t, o = 1 - t;
        if (tid == 0) s_size = min(ld_agent(&A.ctl->free_idx[t]), ld_agent(&A.ctl->completed[t]));
        __syncthreads();
        nearest(A.nodes[t], s_size, s_cfg, dim, red_d, red_i);
        const int near = red_i[0];
        if (tid == 0) {
            const float dist = sqrtf(red_d[0]);
            s_dist = dist;
            s_skip = A.dynamic_domain && ld_agent_f(&A.radii[t][near]) < dist;
        }
        if (tid < dim) s_near[tid] = ld_agent_f(A.nodes[t] + (size_t)near * dim + tid);
        __syncthreads();
        if (s_skip) continue;
        if (tid < dim) {   // steer by at most `range` (pRRTC.cu:1668-1677)
            const float scale = fminf(1.0f, A.range / s_dist);
            const float nq = s_near[tid];
            const float c = fmaf(s_cfg[tid] - nq, scale, nq);
            s_cfg[tid] = c;
            s_step[tid] = (c - nq) / (float)A.granularity;
        }
        __syncthreads();
        if (edge_hit<G>(A, R, S, s_near, s_step, s_q, pos, apos, frames)) {
            if (A.dynamic_domain && tid == 0) radius_update(&A.radii[t][near], false, A.dd_alpha, A.dd_radius, A.dd_min_radius);
            continue;
        }
        // add the node (pRRTC.cu:1783-1812)
        if (tid == 0) {
            int idx = atomicAdd(&A.ctl->free_idx[t], 1);
            if (idx >= A.max_samples) {
                atomicCAS(&A.ctl->status, 0, -1);
                idx = -1;
            } else {
                A.parents[t][idx] = near;
                if (A.dynamic_domain) radius_update(&A.radii[t][near], true, A.dd_alpha, A.dd_radius, A.dd_min_radius);
            }
            s_index = idx;
        }
        __syncthreads();
        if (s_index < 0) break;
        int last = s_index;
        if (tid < dim) st_agent_f(A.nodes[t] + (size_t)last * dim + tid, s_cfg[tid]);
        publish(A, t);
        // connect: greedy extension towards the nearest node of the other tree (pRRTC.cu:1814-1960)
        if (tid == 0) s_size = min(ld_agent(&A.ctl->free_idx[o]), ld_agent(&A.ctl->completed[o]));
        __syncthreads();
        nearest(A.nodes[o], s_size, s_cfg, dim, red_d, red_i);
        const int onear = red_i[0];
        if (tid == 0) s_next = (int)ceilf(sqrtf(red_d[0]) / A.range);
        if (tid < dim) s_near[tid] = ld_agent_f(A.nodes[o] + (size_t)onear * dim + tid);
        __syncthreads();
        const int n_ext = s_next;
        if (tid < dim) {
            const float v = (s_near[tid] - s_cfg[tid]) / (float)n_ext;
            s_vec[tid] = v;
            s_step[tid] = v / (float)A.granularity;
        }
        __syncthreads();
        int e = 0;
        bool full = false;
        for (; e < n_ext; ++e) {
            if (edge_hit<G>(A, R, S, s_cfg, s_step, s_q, pos, apos, frames)) break;
            if (tid == 0) {
                int idx = -1;
                if (A.time_ticks > 0 && wall_clock64() - t0 > A.time_ticks) atomicCAS(&A.ctl->status, 0, -1);
                if (ld_agent(&A.ctl->status) == 0) {   // another workgroup may have finished the run
                    idx = atomicAdd(&A.ctl->free_idx[t], 1);
                    if (idx >= A.max_samples) {
                        atomicCAS(&A.ctl->status, 0, -1);
                        idx = -1;
                    } else {
                        A.parents[t][idx] = last;
                    }
                }
                s_index = idx;
            }
            __syncthreads();
            if (s_index < 0) {
                full = true;
                break;
            }
            last = s_index;
            if (tid < dim) {
                const float c = s_cfg[tid] + s_vec[tid];
                s_cfg[tid] = c;
                st_agent_f(A.nodes[t] + (size_t)last * dim + tid, c);
            }
            publish(A, t);
        }
        if (full) break;
        if (e == n_ext && tid == 0 && atomicCAS(&A.ctl->status, 0, 1) == 0) {
            A.ctl->win_tree = t;
            A.ctl->win_t = last;
            A.ctl->win_o = onear;
            A.ctl->win_iter = iter;
        }
        __syncthreads();
        if (tid == 0) s_stop = ld_agent(&A.ctl->status) != 0;
        __syncthreads();
        if (s_stop) break;
    }
    if (tid == 0) atomicMax(&A.ctl->max_iter, iter);
}

// Walk both trees from the connection to their roots (pRRTC.cu:1962-1988):
// path[tree] lists the walked nodes (roots excluded), cost sums the walked
// edges' L2 lengths, the solving tree first. A walk that leaves [0, free) or
// does not descend to smaller indices leaves path_ok = 0. Lane 0 walks; the
// whole wave then writes the assembled start -> goal rows.
__global__ void k_rrtc_path(RrtcArgs A) {
    __shared__ int s_ok, s_n0, s_n1, s_goal;
    RrtcCtl* c = A.ctl;
    const int dim = A.dim;
    if (threadIdx.x == 0) {
        s_ok = 0;
        c->path_ok = 0;
        c->path_n[0] = c->path_n[1] = 0;
        c->reached_goal = 0;
        c->cost = 0.0f;
        c->n_waypoints = 0;
        float cost = 0.0f;
        bool ok = c->status == 1;
        for (int side = 0; side < 2 && ok; ++side) {
            const int tree = side == 0 ? c->win_tree : 1 - c->win_tree;
            int cur = side == 0 ? c->win_t : c->win_o;
            const float* N = A.nodes[tree];
            const int* P = A.parents[tree];
            const int lim = min(c->free_idx[tree], A.max_samples);
            float* out = A.path + (size_t)tree * kPathCap * dim;
            int n = 0;
            for (;;) {
                if (cur < 0 || cur >= lim) {
                    ok = false;
                    break;
                }
                const int p = P[cur];
                if (p == cur) break;
                if (p < 0 || p > cur || n >= kPathCap) {
                    ok = false;
                    break;
                }
                float s = 0.0f;
                for (int k = 0; k < dim; ++k) {
                    const float d = N[(size_t)cur * dim + k] - N[(size_t)p * dim + k];
                    s = fmaf(d, d, s);
                    out[(size_t)n * dim + k] = N[(size_t)cur * dim + k];
                }
                cost += sqrtf(s);
                ++n;
                cur = p;
            }
            c->path_n[tree] = n;
            if (tree == 1) c->reached_goal = cur;
        }
        if (ok) {
            s_ok = 1;
            s_n0 = c->path_n[0];
            s_n1 = c->path_n[1];
            s_goal = c->reached_goal;
            c->n_waypoints = s_n0 + s_n1 + 2;
            c->cost = cost;
            c->path_ok = 1;
        }
    }
    __syncthreads();
    if (!s_ok) return;
    // start, the start-tree walk reversed, the goal-tree walk, the goal reached
    // (pRRTC.cu:2167-2186 builds goal -> start; python/bindings.cpp:260-272 reverses it)
    const int n0 = s_n0, n1 = s_n1, rows = n0 + n1 + 2;
    for (int e = threadIdx.x; e < rows * dim; e += blockDim.x) {
        const int r = e / dim, k = e - r * dim;
        float v;
        if (r == 0) v = A.roots[k];
        else if (r <= n0) v = A.path[(size_t)(n0 - r) * dim + k];
        else if (r <= n0 + n1) v = A.path[(size_t)(kPathCap + r - 1 - n0) * dim + k];
        else v = A.roots[(size_t)(1 + s_goal) * dim + k];
        A.out[(size_t)r * dim + k] = v;
    }
}

// Fresh run state: the control block, the roots (start in tree 0, goals in
// tree 1, each its own parent), pRRTC.cu:2096-2128.
__global__ void k_rrtc_init(RrtcArgs A) {
    const int dim = A.dim;
    if (threadIdx.x == 0) {
        RrtcCtl c{};
        c.free_idx[0] = c.completed[0] = 1;
        c.free_idx[1] = c.completed[1] = A.n_goals;
        *A.ctl = c;
        A.parents[0][0] = 0;
    }
    for (int i = threadIdx.x; i < A.n_goals; i += blockDim.x) A.parents[1][i] = i;
    for (int k = threadIdx.x; k < dim; k += blockDim.x) A.nodes[0][k] = A.roots[k];
    for (int k = threadIdx.x; k < A.n_goals * dim; k += blockDim.x) A.nodes[1][k] = A.roots[dim + k];
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Shortcut skip-edge states (shortcut.cuh:85-92): state k of pair e (pairs
// sorted by `first`) is s = k - first + 1 of total, t = s / total,
// q = q_i (1 - t) + q_j t with separately rounded products and sum.
__global__ void k_seg_expand(const float* path, int dim, const int4* pairs, int n_pairs, int n_states, float* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_states) return;
    int lo = 0, hi = n_pairs - 1;
    while (lo < hi) {   // last pair with first <= k
        const int mid = (lo + hi + 1) >> 1;
        if (pairs[mid].w <= k) lo = mid;
        else hi = mid - 1;
    }
    const int4 e = pairs[lo];
    const float t = (float)(k - e.w + 1) / (float)e.z;
    const float* a = path + (size_t)e.x * dim;
    const float* b = path + (size_t)e.y * dim;
    for (int d = 0; d < dim; ++d) out[(size_t)k * dim + d] = __fadd_rn(__fmul_rn(a[d], 1.0f - t), __fmul_rn(b[d], t));
}

__global__ void k_seg_reduce(const int4* pairs, int n_pairs, const uint8_t* hit, uint8_t* blocked) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_pairs) return;
    const int4 p = pairs[e];
    uint8_t b = 0;
    for (int k = p.w; k < p.w + p.z - 1 && !b; ++k) b = hit[k];
    blocked[e] = b;
}

// Growable device allocation.
struct Grow {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, std::max(bytes, (size_t)256));
        if (e == hipSuccess) cap = std::max(bytes, (size_t)256);
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

// Where the LDS copies go: the robot tables (without the exact pairs) first,
// then the OBBs, then the exact pairs, each only while the whole dynamic LDS
// stays within kMaxDynLds. Offsets in 4-byte words, 4-word aligned.
StageLayout stage_layout(const ppl_robot_device_view& R, int n_obbs, size_t pos_words) {
    StageLayout L{};
    const size_t cap = kMaxDynLds / 4;
    size_t at = (pos_words + 3) & ~(size_t)3;
    auto take = [&at](int n) {
        const int o = (int)at;
        at += ((size_t)n + 3) & ~(size_t)3;
        return o;
    };
    const int J = R.n_joints, Sn = R.n_spheres, An = R.n_approx_spheres;
    {
        const size_t save = at;
        StageLayout t = L;
        t.ft = take(J * 12);
        t.jmeta = take(J * 4);
        t.sph_off = take(J + 1);
        t.sph = take(Sn * 4);
        t.sph_id = take(Sn);
        t.apx_off = take(J + 1);
        t.apx = take(An * 4);
        t.apx_id = take(An);
        t.sph_orig = take(Sn * 4);
        t.apx_orig = take(An * 4);
        t.apx_pairs = take(R.n_approx_pairs * 2);
        if (at <= cap) {
            L = t;
            L.robot = 1;
        } else {
            at = save;
        }
    }
    if (n_obbs > 0) {
        const size_t save = at;
        const int d = take(n_obbs * 4), p = take(n_obbs * 8), e = take((n_obbs + 3) / 4);
        if (at <= cap) {
            L.obbs = 1;
            L.obb_dims = d;
            L.obb_pose = p;
            L.obb_enable = e;
        } else {
            at = save;
        }
    }
    if (L.robot && R.n_pairs > 0) {
        const size_t save = at;
        const int o = take(R.n_pairs * 2);
        if (at <= cap) {
            L.pairs = 1;
            L.pairs_at = o;
        } else {
            at = save;
        }
    }
    return L;
}

// Dynamic LDS words the layout uses (sphere positions included).
size_t lay_words(const StageLayout& L, const ppl_robot_device_view& R, int n_obbs, size_t pos_words) {
    size_t end = pos_words;
    auto upto = [&end](int off, int n) { end = std::max(end, (size_t)off + (size_t)n); };
    if (L.robot) upto(L.apx_pairs, R.n_approx_pairs * 2), upto(L.apx_orig, R.n_approx_spheres * 4);
    if (L.obbs) upto(L.obb_enable, (n_obbs + 3) / 4);
    if (L.pairs) upto(L.pairs_at, R.n_pairs * 2);
    return end;
}

ppl_status hip_fail(hipError_t e, const char* what) {
    static thread_local char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return ppl::fail_msg(PPL_ERR_HIP, buf);
}

#define RR_TRY(expr)                                       \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);  \
    } while (0)

#define RR_ARG(cond, msg)                                          \
    do {                                                           \
        if (!(cond)) return ppl::fail_msg(PPL_ERR_INVALID, msg);  \
    } while (0)

uint64_t splitmix64_next(uint64_t& x) {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The 16 Halton bases of workgroup w (pRRTC.cu:955-966: workgroup 0 keeps
// the primes in order, the others shuffle them).
void halton_bases(uint64_t seed, int w, float* out16) {
    static const float primes[16] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59};
    std::memcpy(out16, primes, sizeof primes);
    if (w == 0) return;
    uint64_t x = seed ^ ((uint64_t)w * 0xD1B54A32D192ED03ull);
    for (int i = 15; i > 0; --i) {
        const int j = (int)(splitmix64_next(x) % (uint64_t)(i + 1));
        std::swap(out16[i], out16[j]);
    }
}

struct Shortcut {
    Grow d_path, d_pairs, d_states, d_hit, d_blocked;
    void release() {
        d_path.release();
        d_pairs.release();
        d_states.release();
        d_hit.release();
        d_blocked.release();
    }
};

// Shortcut a path given on the host (shortcut.cuh:190-274). out may alias nothing.
ppl_status shortcut_run(ppl_robot robot, const ppl_scene_view* scene, const float* path, int n, int dim, float vsl,
                        float margin, std::vector<float>& out, Shortcut& ws, hipStream_t s) {
    out.assign(path, path + (size_t)n * dim);
    if (n <= 2) return PPL_OK;
    std::vector<int4> pairs;
    pairs.reserve((size_t)(n - 2) * (n - 1) / 2);
    long long n_states = 0;
    for (int i = 0; i < n - 2; ++i)
        for (int j = i + 2; j < n; ++j) {
            float len2 = 0.0f;
            for (int d = 0; d < dim; ++d) {
                const float df = path[(size_t)i * dim + d] - path[(size_t)j * dim + d];
                len2 += df * df;
            }
            const float total_f = std::ceil(std::sqrt(len2) / vsl);
            if (!(total_f < 2.0e9f)) return ppl::fail_msg(PPL_ERR_CAPACITY, "shortcut edge has too many states");
            const int total = std::max(1, (int)total_f);
            pairs.push_back(make_int4(i, j, total, (int)n_states));
            n_states += total - 1;
            if (n_states > kMaxShortcutStates) return ppl::fail_msg(PPL_ERR_CAPACITY, "shortcut needs more than 2^26 states");
        }
    const int n_pairs = (int)pairs.size();
    std::vector<uint8_t> blocked(n_pairs, 0);
    if (n_states > 0) {
        RR_TRY(ws.d_path.ensure((size_t)n * dim * sizeof(float)));
        RR_TRY(ws.d_pairs.ensure((size_t)n_pairs * sizeof(int4)));
        RR_TRY(ws.d_states.ensure((size_t)n_states * dim * sizeof(float)));
        RR_TRY(ws.d_hit.ensure((size_t)n_states));
        RR_TRY(ws.d_blocked.ensure((size_t)n_pairs));
        RR_TRY(hipMemcpyAsync(ws.d_path.p, path, (size_t)n * dim * sizeof(float), hipMemcpyHostToDevice, s));
        RR_TRY(hipMemcpyAsync(ws.d_pairs.p, pairs.data(), (size_t)n_pairs * sizeof(int4), hipMemcpyHostToDevice, s));
        const int ns = (int)n_states;
        hipLaunchKernelGGL(k_seg_expand, dim3((ns + 255) / 256), dim3(256), 0, s, ws.d_path.as<float>(), dim,
                           ws.d_pairs.as<int4>(), n_pairs, ns, ws.d_states.as<float>());
        RR_TRY(hipGetLastError());
        ppl_status st = ppl_mitstar_batch_cc(robot, scene, ws.d_states.as<float>(), ns, margin, ws.d_hit.as<uint8_t>(),
                                             nullptr, 0.0f, (ppl_stream)s);
        if (st != PPL_OK) return st;
        hipLaunchKernelGGL(k_seg_reduce, dim3((n_pairs + 255) / 256), dim3(256), 0, s, ws.d_pairs.as<int4>(), n_pairs,
                           ws.d_hit.as<uint8_t>(), ws.d_blocked.as<uint8_t>());
        RR_TRY(hipGetLastError());
        RR_TRY(hipMemcpyAsync(blocked.data(), ws.d_blocked.p, (size_t)n_pairs, hipMemcpyDeviceToHost, s));
        RR_TRY(hipStreamSynchronize(s));
    }
    // greedy: from waypoint i jump to the farthest j > i + 1 with a free skip edge, else i + 1
    auto pair_index = [n](int i, int j) { return i * (2 * n - i - 3) / 2 + (j - i - 2); };
    out.clear();
    out.insert(out.end(), path, path + dim);
    int i = 0;
    while (i < n - 1) {
        int far = i + 1;
        for (int j = n - 1; j > i + 1; --j)
            if (!blocked[pair_index(i, j)]) {
                far = j;
                break;
            }
        out.insert(out.end(), path + (size_t)far * dim, path + (size_t)(far + 1) * dim);
        i = far;
    }
    return PPL_OK;
}

}  // namespace

struct ppl_rrtc_buffers_s {
    int dim = 0, max_samples = 0, blocks = 0, gran = 0;
    Grow nodes, parents, radii, bases, path;
    Grow io;                    // [ctl 256 B][assembled path][roots]
    char* h_io = nullptr;       // pinned: [ctl 256 B][kHeadRows rows] then the roots staging
    size_t h_io_cap = 0;
    int hwm[2] = {0, 0};        // rows of each tree that may hold values from an earlier run
    unsigned long long bases_seed = 0;
    int bases_n = 0;            // workgroups the device table holds for bases_seed
    Shortcut sc;
    hipEvent_t ev[2] = {nullptr, nullptr};
    void release() {
        nodes.release();
        parents.release();
        radii.release();
        bases.release();
        path.release();
        io.release();
        sc.release();
        if (h_io) (void)hipHostFree(h_io);
        h_io = nullptr;
        h_io_cap = 0;
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        ev[0] = ev[1] = nullptr;
    }
};

namespace {

ppl_status rrtc_alloc(ppl_rrtc_buffers_s* b, int dim, int max_samples, int blocks, int gran) {
    b->dim = dim;
    b->max_samples = max_samples;
    b->blocks = blocks;
    b->gran = gran;
    const size_t ms = (size_t)max_samples;
    RR_TRY(b->nodes.ensure(2 * ms * dim * sizeof(float)));
    RR_TRY(b->parents.ensure(2 * ms * sizeof(int)));
    RR_TRY(b->radii.ensure(2 * ms * sizeof(float)));
    RR_TRY(b->bases.ensure((size_t)blocks * kMaxDim * sizeof(float)));
    RR_TRY(b->path.ensure((size_t)2 * kPathCap * dim * sizeof(float)));
    b->hwm[0] = b->hwm[1] = max_samples;   // fresh memory: every row is unknown
    b->bases_n = 0;
    RR_TRY(hipEventCreate(&b->ev[0]));
    RR_TRY(hipEventCreate(&b->ev[1]));
    return PPL_OK;
}

}  // namespace

extern "C" {

void ppl_rrtc_default_settings(ppl_rrtc_settings* s) {
    if (!s) return;
    s->max_samples = 1000000;
    s->max_iters = 1000000;
    s->time_limit_ms = 4000.0f;
    s->num_new_configs = 600;
    s->granularity = 16;
    s->range = 0.5f;
    s->balance = 1;
    s->tree_ratio = 0.5f;
    s->dynamic_domain = 1;
    s->dd_alpha = 0.0001f;
    s->dd_radius = 4.0f;
    s->dd_min_radius = 1.0f;
    s->shortcut_path = 0;
    s->valid_segment_length = 0.025f;
    s->collision_margin = 0.0f;
    s->shortcut_collision_margin = 0.0005f;
    s->seed = 1;
}

ppl_status ppl_rrtc_halton_bases(unsigned long long seed, int n_workgroups, int n_dof, float* h_bases) {
    RR_ARG(h_bases && n_workgroups >= 0 && n_dof >= 1 && n_dof <= kMaxDim, "bad Halton base request");
    float b[kMaxDim];
    for (int w = 0; w < n_workgroups; ++w) {
        halton_bases(seed, w, b);
        std::memcpy(h_bases + (size_t)w * n_dof, b, (size_t)n_dof * sizeof(float));
    }
    return PPL_OK;
}

ppl_status ppl_rrtc_buffers_create(int n_dof, int max_samples, int num_new_configs, int granularity,
                                   ppl_rrtc_buffers* out) {
    RR_ARG(out, "NULL output");
    *out = nullptr;
    RR_ARG(n_dof >= 1 && n_dof <= kMaxDim, "n_dof must be 1..16");
    RR_ARG(max_samples >= 2 && num_new_configs >= 1 && num_new_configs <= 65535, "bad buffer sizes");
    RR_ARG(granularity >= 1 && granularity <= kMaxGran, "granularity must be 1..64");
    auto* b = new (std::nothrow) ppl_rrtc_buffers_s;
    if (!b) return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
    ppl_status st = rrtc_alloc(b, n_dof, max_samples, num_new_configs, granularity);
    if (st != PPL_OK) {
        b->release();
        delete b;
        return st;
    }
    *out = b;
    return PPL_OK;
}

ppl_status ppl_rrtc_buffers_destroy(ppl_rrtc_buffers b) {
    if (!b) return PPL_OK;
    (void)hipDeviceSynchronize();
    b->release();
    delete b;
    return PPL_OK;
}

ppl_status ppl_shortcut_path(ppl_robot robot, const ppl_scene_view* scene, const float* h_path, int n,
                             float valid_segment_length, float collision_margin, float* h_out, int* n_out,
                             ppl_stream stream) {
    RR_ARG(robot && scene && h_out && n_out && (h_path || n == 0) && n >= 0, "NULL argument");
    RR_ARG(valid_segment_length > 0.0f, "valid_segment_length must be > 0");
    if (n > PPL_RRTC_MAX_WAYPOINTS)
        return ppl::fail_msg(PPL_ERR_CAPACITY, "shortcut paths hold at most PPL_RRTC_MAX_WAYPOINTS (8194) waypoints");
    ppl_robot_info info{};
    ppl_status st = ppl_robot_get_info(robot, &info);
    if (st != PPL_OK) return st;
    Shortcut ws;
    try {
        std::vector<float> out;
        st = shortcut_run(robot, scene, h_path, n, info.n_dof, valid_segment_length, collision_margin, out, ws,
                          (hipStream_t)stream);
        if (st == PPL_OK) {
            std::memcpy(h_out, out.data(), out.size() * sizeof(float));
            *n_out = (int)(out.size() / std::max(1, info.n_dof));
        }
    } catch (const std::bad_alloc&) {
        st = ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
    }
    ws.release();
    return st;
}

}  // extern "C"

namespace {

ppl_status rrtc_solve(ppl_robot robot, const ppl_scene_view* scene, const float* h_start, const float* h_goals,
                      int n_goals, const ppl_rrtc_settings* cfg, ppl_rrtc_buffers bufs, float* h_path,
                      int max_waypoints, ppl_rrtc_result* res, ppl_stream stream) {
    const auto wall0 = std::chrono::steady_clock::now();
    RR_ARG(robot && scene && h_start && h_goals && cfg && res, "NULL argument");
    RR_ARG(n_goals >= 1, "need at least one goal");
    RR_ARG(max_waypoints >= 0 && (h_path || max_waypoints == 0), "NULL path buffer");
    std::memset(res, 0, sizeof *res);
    ppl_robot_device_view R{};
    ppl_status st = ppl_robot_get_device_view(robot, &R);
    if (st != PPL_OK) return st;
    const int dim = R.n_dof;
    RR_ARG(dim >= 1 && dim <= kMaxDim, "pRRTC needs 1..16 degrees of freedom (one Halton base each)");
    RR_ARG(cfg->granularity >= 1 && cfg->granularity <= kMaxGran, "granularity must be 1..64");
    RR_ARG(cfg->num_new_configs >= 1 && cfg->num_new_configs <= 65535, "num_new_configs must be 1..65535");
    RR_ARG(cfg->range > 0.0f && std::isfinite(cfg->range), "range must be a positive number");
    RR_ARG(cfg->max_samples > n_goals, "max_samples must exceed the number of goals");
    RR_ARG(cfg->max_iters >= 1, "max_iters must be >= 1");
    RR_ARG(cfg->balance >= 0 && cfg->balance <= 2, "balance must be 0, 1 or 2");
    RR_ARG(!cfg->shortcut_path || cfg->valid_segment_length > 0.0f, "valid_segment_length must be > 0");
    for (int i = 0; i < dim; ++i) RR_ARG(std::isfinite(h_start[i]), "start has a non-finite coordinate");
    for (int i = 0; i < n_goals * dim; ++i) RR_ARG(std::isfinite(h_goals[i]), "a goal has a non-finite coordinate");
    if ((st = ppl::scene_view_check(scene)) != PPL_OK) return st;
    if (scene->esdf_interp != 0)
        return ppl::fail_msg(PPL_ERR_UNSUPPORTED, "pRRTC checks use the nearest-voxel ESDF lookup (esdf_interp = 0)");
    // lanes per edge state: the widest power of two that gives every state its own group
    int G = 64;
    while (G > 1 && G * cfg->granularity > kThreads) G >>= 1;
    const size_t pos_words = (size_t)cfg->granularity * (3 * (R.n_spheres + R.n_approx_spheres) + 12 * R.n_joints);
    if (pos_words * 4 > kMaxDynLds) return ppl::fail_msg(PPL_ERR_CAPACITY, "too many spheres x granularity for the LDS staging");
    StageLayout lay = stage_layout(R, scene->n_obbs, pos_words);
    const size_t lds = (size_t)lay_words(lay, R, scene->n_obbs, pos_words) * 4;

    ppl_rrtc_buffers_s local;
    ppl_rrtc_buffers_s* b = bufs;
    struct Guard {
        ppl_rrtc_buffers_s* p;
        ~Guard() {
            if (p) p->release();
        }
    } guard{nullptr};
    if (!b) {
        b = &local;
        guard.p = &local;
        if ((st = rrtc_alloc(b, dim, cfg->max_samples, cfg->num_new_configs, cfg->granularity)) != PPL_OK) return st;
    } else {
        RR_ARG(b->dim == dim, "buffers were created for another n_dof");
        RR_ARG(b->max_samples >= cfg->max_samples && b->blocks >= cfg->num_new_configs && b->gran >= cfg->granularity,
               "buffers are smaller than the settings ask for");
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t ms = (size_t)b->max_samples;
    RrtcArgs A{};
    A.R = R;
    A.S = *scene;
    for (int t = 0; t < 2; ++t) {
        A.nodes[t] = b->nodes.as<float>() + t * ms * dim;
        A.parents[t] = b->parents.as<int>() + t * ms;
        A.radii[t] = b->radii.as<float>() + t * ms;
    }
    A.bases = b->bases.as<float>();
    A.path = b->path.as<float>();
    A.dim = dim;
    A.max_samples = cfg->max_samples;
    A.max_iters = cfg->max_iters;
    A.granularity = cfg->granularity;
    A.balance = cfg->balance;
    A.dynamic_domain = cfg->dynamic_domain ? 1 : 0;
    A.n_blocks = cfg->num_new_configs;
    A.range = cfg->range;
    A.tree_ratio = cfg->tree_ratio;
    A.dd_alpha = cfg->dd_alpha;
    A.dd_radius = cfg->dd_radius;
    A.dd_min_radius = cfg->dd_min_radius;
    A.margin = cfg->collision_margin;
    A.time_ticks = 0;
    A.L = lay;
    if (cfg->time_limit_ms > 0.0f) {
        int dev = 0, khz = 0;
        RR_TRY(hipGetDevice(&dev));
        RR_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        A.time_ticks = std::max(1ll, (long long)((double)cfg->time_limit_ms * (double)khz));
    }
    // I/O: [ctl 256 B][assembled path][roots]; pinned mirror of the first two parts plus the roots staging
    const size_t out_rows = PPL_RRTC_MAX_WAYPOINTS;
    const size_t out_off = 256, roots_off = out_off + a256(out_rows * dim * sizeof(float));
    const size_t roots_bytes = (size_t)(1 + n_goals) * dim * sizeof(float);
    RR_TRY(b->io.ensure(roots_off + roots_bytes));
    const size_t head_bytes = 256 + (size_t)kHeadRows * dim * sizeof(float);
    const size_t h_need = a256(head_bytes) + roots_bytes;
    if (b->h_io_cap < h_need) {
        if (b->h_io) (void)hipHostFree(b->h_io);
        b->h_io = nullptr;
        b->h_io_cap = 0;
        RR_TRY(hipHostMalloc((void**)&b->h_io, h_need, hipHostMallocDefault));
        b->h_io_cap = h_need;
    }
    char* io = b->io.as<char>();
    A.ctl = reinterpret_cast<RrtcCtl*>(io);
    A.out = reinterpret_cast<float*>(io + out_off);
    A.roots = reinterpret_cast<const float*>(io + roots_off);
    A.n_goals = n_goals;
    float* h_roots = reinterpret_cast<float*>(b->h_io + a256(head_bytes));
    std::memcpy(h_roots, h_start, (size_t)dim * sizeof(float));
    std::memcpy(h_roots + dim, h_goals, (size_t)n_goals * dim * sizeof(float));
    // rows an earlier run may have written go back to -9999 / FLT_MAX (pRRTC.cu:2070-2083 refills everything)
    for (int t = 0; t < 2; ++t) {
        const size_t rows = (size_t)std::min(b->hwm[t], cfg->max_samples);
        if (rows == 0) continue;
        const int grid = (int)std::min<size_t>(1024, (rows * dim + 255) / 256);
        hipLaunchKernelGGL(k_fill_f32, dim3(grid), dim3(256), 0, s, A.nodes[t], rows * dim, kUnwritten);
        hipLaunchKernelGGL(k_fill_f32, dim3((int)std::min<size_t>(256, (rows + 255) / 256)), dim3(256), 0, s,
                           A.radii[t], rows, FLT_MAX);
    }
    RR_TRY(hipGetLastError());
    for (int t = 0; t < 2; ++t)   // until the run's counts come back, every row it may reach counts as written
        b->hwm[t] = std::max(b->hwm[t], cfg->max_samples);
    if (b->bases_n < cfg->num_new_configs || b->bases_seed != cfg->seed) {
        std::vector<float> bases((size_t)cfg->num_new_configs * kMaxDim);
        for (int w = 0; w < cfg->num_new_configs; ++w) halton_bases(cfg->seed, w, bases.data() + (size_t)w * kMaxDim);
        RR_TRY(hipMemcpyAsync(b->bases.p, bases.data(), bases.size() * sizeof(float), hipMemcpyHostToDevice, s));
        RR_TRY(hipStreamSynchronize(s));   // the staging vector goes out of scope
        b->bases_seed = cfg->seed;
        b->bases_n = cfg->num_new_configs;
    }
    RR_TRY(hipMemcpyAsync(io + roots_off, h_roots, roots_bytes, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_rrtc_init, dim3(1), dim3(256), 0, s, A);
    RR_TRY(hipGetLastError());
    RR_TRY(hipEventRecord(b->ev[0], s));
    switch (G) {
        case 64: hipLaunchKernelGGL(k_rrtc<64>, dim3(cfg->num_new_configs), dim3(kThreads), lds, s, A); break;
        case 32: hipLaunchKernelGGL(k_rrtc<32>, dim3(cfg->num_new_configs), dim3(kThreads), lds, s, A); break;
        case 16: hipLaunchKernelGGL(k_rrtc<16>, dim3(cfg->num_new_configs), dim3(kThreads), lds, s, A); break;
        case 8: hipLaunchKernelGGL(k_rrtc<8>, dim3(cfg->num_new_configs), dim3(kThreads), lds, s, A); break;
        default: hipLaunchKernelGGL(k_rrtc<4>, dim3(cfg->num_new_configs), dim3(kThreads), lds, s, A); break;
    }
    RR_TRY(hipGetLastError());
    RR_TRY(hipEventRecord(b->ev[1], s));
    hipLaunchKernelGGL(k_rrtc_path, dim3(1), dim3(64), 0, s, A);
    RR_TRY(hipGetLastError());
    RR_TRY(hipMemcpyAsync(b->h_io, io, head_bytes, hipMemcpyDeviceToHost, s));
    RR_TRY(hipStreamSynchronize(s));
    const RrtcCtl* hc = reinterpret_cast<const RrtcCtl*>(b->h_io);
    float kms = 0.0f;
    RR_TRY(hipEventElapsedTime(&kms, b->ev[0], b->ev[1]));
    for (int t = 0; t < 2; ++t)
        if (b->hwm[t] <= cfg->max_samples) b->hwm[t] = std::min(std::max(hc->free_idx[t], 1), cfg->max_samples);
    res->kernel_ms = kms;
    res->start_tree_size = std::min(hc->free_idx[0], cfg->max_samples);
    res->goal_tree_size = std::min(hc->free_idx[1], cfg->max_samples);
    res->solved = hc->status == 1 && hc->path_ok;
    res->iters = res->solved ? hc->win_iter : hc->max_iter;
    if (res->solved) {
        int nw = hc->n_waypoints;
        std::vector<float> path((size_t)nw * dim);
        const float* head = reinterpret_cast<const float*>(b->h_io + 256);
        std::memcpy(path.data(), head, (size_t)std::min(nw, kHeadRows) * dim * sizeof(float));
        if (nw > kHeadRows) {
            RR_TRY(hipMemcpyAsync(path.data() + (size_t)kHeadRows * dim, A.out + (size_t)kHeadRows * dim,
                                  (size_t)(nw - kHeadRows) * dim * sizeof(float), hipMemcpyDeviceToHost, s));
            RR_TRY(hipStreamSynchronize(s));
        }
        res->cost = hc->cost;
        res->path_length = nw - 2;
        res->reached_goal = hc->reached_goal;
        if (cfg->shortcut_path && nw > 2) {
            // from the goal end, as the reference shortcuts its goal -> start path (pRRTC.cu:2190-2200)
            std::vector<float> g2s((size_t)nw * dim), sc;
            for (int i = 0; i < nw; ++i)
                std::memcpy(g2s.data() + (size_t)i * dim, path.data() + (size_t)(nw - 1 - i) * dim, (size_t)dim * sizeof(float));
            st = shortcut_run(robot, scene, g2s.data(), nw, dim, cfg->valid_segment_length,
                              cfg->collision_margin + cfg->shortcut_collision_margin, sc, b->sc, s);
            if (st != PPL_OK) return st;
            nw = (int)(sc.size() / dim);
            path.resize((size_t)nw * dim);
            for (int i = 0; i < nw; ++i)
                std::memcpy(path.data() + (size_t)i * dim, sc.data() + (size_t)(nw - 1 - i) * dim, (size_t)dim * sizeof(float));
        }
        res->n_waypoints = nw;
        if (nw > max_waypoints) {
            res->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
            return ppl::fail_msg(PPL_ERR_CAPACITY, "the solution has more waypoints than max_waypoints");
        }
        std::memcpy(h_path, path.data(), path.size() * sizeof(float));
    }
    res->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    return PPL_OK;
}

}  // namespace

extern "C" ppl_status ppl_rrtc_solve(ppl_robot robot, const ppl_scene_view* scene, const float* h_start,
                                     const float* h_goals, int n_goals, const ppl_rrtc_settings* cfg,
                                     ppl_rrtc_buffers bufs, float* h_path, int max_waypoints, ppl_rrtc_result* res,
                                     ppl_stream stream) {
    try {
        return rrtc_solve(robot, scene, h_start, h_goals, n_goals, cfg, bufs, h_path, max_waypoints, res, stream);
    } catch (const std::bad_alloc&) {
        return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
    }
}
