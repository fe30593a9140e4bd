// capi.hip -- C ABI of the substrate (include/ppl/ppl.h): validation, device
// table construction, plans, and the plain / batched / differentiable entry
// points. No exceptions or exits cross the ABI; every failure returns a
// ppl_status and leaves a message for ppl_last_error().
#include "ppl/ppl.h"
#include "ppl/roofline.h"
#include "ppl_internal.hpp"

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

using namespace ppl;

namespace {

thread_local std::string g_err;

ppl_status fail(ppl_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(PPL_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define CHECK_ARG(cond, ...)                                     \
    do {                                                         \
        if (!(cond)) return fail(PPL_ERR_INVALID, __VA_ARGS__); \
    } while (0)

// Growable device buffer.
struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = std::max(bytes, (size_t)256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Growable pinned host buffer.
struct HBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = std::max(bytes, (size_t)256);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace

struct ppl_robot_s {
    DRobot R{};
    void* blob = nullptr;
    int n_slots = 1;
    int n_pairs = 0;
    int n_revolute = 0, n_prismatic = 0;   // joints that move with a dof
    std::vector<float> lower, upper;
    std::vector<std::pair<int, ppl_plan>> traj_plans;   // single-trajectory plans of the plain CHOMP forms, by T
    unsigned long long digest = 0;   // FNV-1a of the device tables: the robot's key in the tuning cache
    hipStream_t stream = nullptr;
    DBuf d_cfg, d_out, d_off, d_aux, d_self, d_mit, d_hint;
    HBuf h_stage, h_mit;
};

struct ppl_edge_eval_s {
    ppl_robot robot = nullptr;
    DScene S{};
    int dim = 0, max_edges = 0;
    hipStream_t stream = nullptr;
    // one call's inputs and first hits, packed [offsets][first hits][from][to]
    // (256-byte aligned parts) in one pinned block and one device block
    DBuf d_io, d_clr_state, d_clr_edge, d_self;
    HBuf h_io, h_clr;
};

// Byte offsets of the packed parts for n edges of dim joints.
struct EdgeIoLayout {
    size_t fh, from, to, end;
    EdgeIoLayout(int n, int dim) {
        auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t cfg = (size_t)n * dim * sizeof(float);
        fh = up((size_t)(n + 1) * 4);
        from = fh + up((size_t)n * 4);
        to = from + up(cfg);
        end = to + cfg;
    }
};

// STOMP optimiser state kept with a trajectory plan between calls: rollout
// plans over the fresh samples, per-length sampling / projection matrices,
// the population buffers, and where the run is.
struct StompState {
    int n_new = 0, n_old = 0;
    float dt = 0.0f, noise_scale = 0.0f;
    ppl_plan first_plan = nullptr;   // n_new + n_old fresh rows per trajectory (first iteration)
    ppl_plan fresh_plan = nullptr;   // n_new fresh rows per trajectory
    std::vector<std::pair<int, float*>> mats;   // per length T: [L^T | M^T]
    const float** d_lt = nullptr;   // [2 B]: L^T of each trajectory, then M^T
    const float** d_mt = nullptr;
    DBuf buf, hint;
    float *fresh = nullptr, *fresh_cost = nullptr, *keep[2] = {}, *keep_cost[2] = {};
    float *slot = nullptr, *slot_cost = nullptr, *theta_cost = nullptr, *ctrl_sum = nullptr, *totals = nullptr;
    int *sel = nullptr, *first_hit = nullptr, *nsub = nullptr, *dense_off = nullptr;
    unsigned iter = 0;
    int parity = 0, nf_prev = 0, started = 0;
};

struct ppl_plan_s {
    ppl_robot robot = nullptr;
    int kind = 0;
    int n_units = 0;
    long n_states = 0;
    int k0 = 0;
    int n_problems = 1;
    std::vector<int> h_obb_off;      // empty: single scene
    std::vector<int> h_off;          // edges / trajectories: state offsets [n_units+1]
    int* d_off = nullptr;            // [n_units+1]
    std::vector<long> pb, pe;        // state range of each problem
    int4* d_tab[kGroupKinds] = {};   // chunk tables for 64, 32, 16, 8, 4, 2 states
    int n_chunks[kGroupKinds] = {};
    int* d_obb_off = nullptr;
    int wmajor = 0;                  // waypoint-major trajectory plan
    int t_max = 0;                   // states of the longest unit (trajectory length)
    int tuned_group[2] = {0, 0};     // cost kernel lanes per state chosen by ppl_plan_autotune_cost (0: heuristic)
    // any-collision launch shape (OBB records per LDS staging, grid form) for
    // problems of up to shape_obbs OBBs; shape_src 0 built-in, 1 tuning cache,
    // 2 this plan's tuning or setting (kept whatever the OBB count)
    int ctile = 0, cgrid = 0, shape_obbs = -1, shape_src = 0;
    int cost_obbs[2] = {-1, -1}, cost_cached[2] = {0, 0};   // cost layout from the tuning cache, per OBB count
    int4* d_prob = nullptr;          // wmajor: per problem (first trajectory, trajectories, length, first state)
    DBuf d_self, d_cost, d_fdcfg, d_fdcost, d_tune;
    // CHOMP optimiser: A^-1 (FP32, T x T) per trajectory length, the
    // per-trajectory pointers to them, and scratch for gradient, costs, flags
    std::vector<std::pair<int, float*>> ainv;
    const float** d_ainv_ptr = nullptr;
    float ainv_dt = 0.0f;
    DBuf d_opt;
    StompState* stomp = nullptr;
    // Optimiser iterations captured as hipGraphs (one CHOMP iteration; two
    // STOMP iterations, one per keep-buffer parity), keyed by their inputs.
    int opt_graphs = 1;
    ppl_graph chomp_graph = nullptr, stomp_graph = nullptr;
    unsigned long long chomp_key = 0, stomp_key = 0;
    unsigned* d_iter = nullptr;      // STOMP graphs: the iteration word of the sampling counter
    // kernel timing (ppl_plan_time_kernels): an event pair per timed launch
    bool ktime = false;
    int kused = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> kev;
};

struct ppl_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

ppl_status scene_from_view(const ppl_scene_view* v, DScene& S) {
    CHECK_ARG(v != nullptr, "scene is NULL");
    CHECK_ARG(v->n_obbs >= 0 && v->n_voxel_layers >= 0, "negative scene sizes");
    if (v->n_obbs > 0) {
        CHECK_ARG(v->obb_dims && v->obb_pose && v->obb_enable, "scene has %d OBBs but NULL OBB arrays", v->n_obbs);
        CHECK_ARG(((uintptr_t)v->obb_dims & 15) == 0 && ((uintptr_t)v->obb_pose & 15) == 0,
                  "obb_dims / obb_pose must be 16-byte aligned (rows are read as float4)");
    }
    if (v->n_voxel_layers > 0)
        CHECK_ARG(v->voxel_params && v->voxel_pose && v->voxel_enable && v->voxel_features && v->max_voxels_per_layer > 0,
                  "scene has voxel layers but NULL voxel arrays");
    S.obb_dims = v->obb_dims;
    S.obb_pose = v->obb_pose;
    S.obb_enable = v->obb_enable;
    S.n_obbs = v->n_obbs;
    S.voxel_params = v->voxel_params;
    S.voxel_pose = v->voxel_pose;
    S.voxel_enable = v->voxel_enable;
    S.voxel_features = v->voxel_features;
    S.n_voxel_layers = v->n_voxel_layers;
    S.max_voxels_per_layer = v->max_voxels_per_layer;
    S.sphere_acm = v->sphere_acm_mask;
    S.sphere_obb_acm = v->sphere_obb_acm_mask;
    S.sphere_obb_acm_approx = v->sphere_obb_acm_mask_approx;
    S.esdf_interp = v->esdf_interp ? 1 : 0;
    return PPL_OK;
}

enum KernelKind { kKCollide, kKCost, kKCostGrad, kKCostGrp, kKCostGrpGrad };

// LDS OBB records of a launch whose largest problem has max_obbs OBBs: all of
// them, or one tile of kMaxObbLds when they do not fit (tiled launch).
int obb_tile(int max_obbs) { return std::max(1, std::min(max_obbs, kMaxObbLds)); }

// Persistent grid for a kernel family at the LDS footprint of this launch.
int grid_for(KernelKind k, const DRobot& R, int obb_cap, int G) {
    if (k == kKCollide) return grid_collide(lds_bytes(obb_cap, R.n_dof));
    if (k == kKCostGrp || k == kKCostGrpGrad) return INT_MAX;   // launch_cost_grp caps it per kernel variant
    const bool grad = k == kKCostGrad;
    return grid_cost(R.n_joints, grad, lds_bytes_cost(obb_cap, R.n_dof, R.n_joints, G, grad));
}

// max_obbs: OBBs of the largest problem of the launch.
DChunks implicit_chunks(long n_states, KernelKind k, const DRobot& R, int max_obbs, int* grid_out, int G = 1) {
    const int grid_cap = grid_for(k, R, obb_tile(max_obbs), G);
    DChunks C{};
    C.obb_cap = obb_tile(max_obbs);
    C.tiled = max_obbs > kMaxObbLds;
    C.tab = nullptr;
    C.chunk = kWave / G;
    C.n_chunks = (int)((n_states + C.chunk - 1) / C.chunk);
    C.obb_off = nullptr;
    *grid_out = std::max(1, std::min(grid_cap, C.n_chunks));
    return C;
}

// The cost kernel over n_states implicit chunks: k_cost with one lane per
// state when the scene's OBBs fit one LDS staging, else the grouped kernel,
// which streams them through LDS in tiles.
// Kernel timing: with the plan's timing on, the batched checks and the cost
// evaluation time their main kernel with an event pair the runtime records
// at the kernel's start and end (not under stream capture; at most
// kMaxTimedLaunches pairs until read).
constexpr int kMaxTimedLaunches = 4096;

bool timing_now(ppl_plan P, hipStream_t s) {
    if (!P->ktime || P->kused >= kMaxTimedLaunches) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone;
}

// The plan's next event pair for a timed launch (created on first use); the
// launch passes it to the runtime, which records it at the kernel's own
// start and end.
hipError_t timing_pair(ppl_plan P, hipEvent_t* e0, hipEvent_t* e1) {
    if ((int)P->kev.size() == P->kused) {
        hipEvent_t a = nullptr, b = nullptr;
        hipError_t e = hipEventCreate(&a);
        if (e != hipSuccess) return e;
        if ((e = hipEventCreate(&b)) != hipSuccess) {
            (void)hipEventDestroy(a);
            return e;
        }
        P->kev.emplace_back(a, b);
    }
    *e0 = P->kev[P->kused].first;
    *e1 = P->kev[P->kused].second;
    return hipSuccess;
}

hipError_t launch_cost_flat(CostArgs& a, long n_states, hipStream_t s) {
    int grid = 0;
    const bool tiled = a.S.n_obbs > kMaxObbLds;
    const KernelKind k = tiled ? (a.grad ? kKCostGrpGrad : kKCostGrp) : (a.grad ? kKCostGrad : kKCost);
    a.C = implicit_chunks(n_states, k, a.R, a.S.n_obbs, &grid, tiled ? kGrp : 1);
    a.group = tiled ? kGrp : 1;
    return tiled ? launch_cost_grp(a, grid, s) : launch_cost(a, grid, s);
}

// ---------------------------------------------------------------------------
// Tuning cache: the launch shapes the autotuners choose, kept in a text file
// ("key value..." per line) so that later plans and processes skip the
// timing. A key names the kernel, the robot tables' digest, the largest
// problem's OBB count, (cost kernel) the state count's power of two, and the
// GPU architecture. The file is $PPL_TUNE_CACHE, else
// $XDG_CACHE_HOME/ppl/tune.txt, else $HOME/.cache/ppl/tune.txt; an empty
// path (or ppl_tune_cache_file(NULL)) keeps the cache in memory only.
// ---------------------------------------------------------------------------
std::mutex g_tune_mu;
bool g_tune_loaded = false, g_tune_path_set = false;
std::string g_tune_path;
std::map<std::string, std::string> g_tune;

std::string default_tune_path() {
    if (const char* e = getenv("PPL_TUNE_CACHE")) return e;
    const char* xdg = getenv("XDG_CACHE_HOME");
    if (xdg && *xdg) return std::string(xdg) + "/ppl/tune.txt";
    const char* home = getenv("HOME");
    return home && *home ? std::string(home) + "/.cache/ppl/tune.txt" : std::string();
}

void tune_load_locked() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    if (!g_tune_path_set) g_tune_path = default_tune_path();
    if (g_tune_path.empty()) return;
    FILE* f = fopen(g_tune_path.c_str(), "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        std::string l(line);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
        const size_t sp = l.find(' ');
        if (l.empty() || l[0] == '#' || sp == std::string::npos || sp == 0) continue;
        g_tune[l.substr(0, sp)] = l.substr(sp + 1);
    }
    fclose(f);
}

bool tune_get(const std::string& key, std::string& val) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    const auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    val = it->second;
    return true;
}

// Records key -> val and rewrites the file (parent directories created, a
// sibling temporary renamed over it). The cache only saves timing, so a file
// that cannot be written leaves the in-memory entry alone.
void tune_put(const std::string& key, const std::string& val) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    g_tune[key] = val;
    if (g_tune_path.empty()) return;
    for (size_t k = 1; (k = g_tune_path.find('/', k)) != std::string::npos; ++k)
        (void)mkdir(g_tune_path.substr(0, k).c_str(), 0755);
    const std::string tmp = g_tune_path + ".tmp." + std::to_string((long)getpid());
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) return;
    fprintf(f, "# ppl tuning cache: key value...\n");
    for (const auto& e : g_tune) fprintf(f, "%s %s\n", e.first.c_str(), e.second.c_str());
    const bool ok = fclose(f) == 0;
    if (!ok || rename(tmp.c_str(), g_tune_path.c_str()) != 0) (void)remove(tmp.c_str());
}

// Architecture of the current device ("gfx950"), without feature suffixes.
std::string gpu_arch() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    static std::mutex mu;
    static std::map<int, std::string> names;
    std::lock_guard<std::mutex> lk(mu);
    auto it = names.find(dev);
    if (it != names.end()) return it->second;
    hipDeviceProp_t prop;
    std::string a = "unknown";
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        a = prop.gcnArchName;
        const size_t c = a.find(':');
        if (c != std::string::npos) a.resize(c);
    }
    names[dev] = a;
    return a;
}

std::string hex64(unsigned long long v) {
    char b[24];
    snprintf(b, sizeof b, "%016llx", v);
    return b;
}

std::string collide_key(const ppl_robot_s* r, int max_obbs) {
    return "collide/" + hex64(r->digest) + "/" + std::to_string(max_obbs) + "/" + gpu_arch();
}

std::string cost_key(const ppl_robot_s* r, bool grad, int max_obbs, long n_states) {
    int lg = 0;
    while ((1L << (lg + 1)) <= n_states) ++lg;
    return std::string(grad ? "adjoint/" : "cost/") + hex64(r->digest) + "/" + std::to_string(max_obbs) + "/" +
           std::to_string(lg) + "/" + gpu_arch();
}

int group_index(int G) { return G >= 32 ? 5 : G >= 16 ? 4 : G >= 8 ? 3 : G >= 4 ? 2 : G >= 2 ? 1 : 0; }

// Chunk table of a plan for `G` lanes per state (64 / G states per chunk),
// built on first use; chunks never cross a problem boundary.
hipError_t ensure_table(ppl_plan_s* P, int G) {
    const int gi = group_index(G);
    if (P->d_tab[gi] || P->n_states == 0) return hipSuccess;
    const long chunk = kWave / G;
    std::vector<int4> tab;
    // w = the unit (edge / trajectory) holding the chunk's first state
    for (size_t p = 0; p < P->pb.size(); ++p)
        for (long s = P->pb[p]; s < P->pe[p]; s += chunk) {
            int u = (int)s;
            if (P->wmajor) u = -1;   // the kernel maps waypoint-major indices arithmetically
            else if (P->kind != PPL_PLAN_CONFIGS)
                u = (int)(std::upper_bound(P->h_off.begin(), P->h_off.end(), (int)s) - P->h_off.begin()) - 1;
            tab.push_back(make_int4((int)p, (int)s, (int)std::min(s + chunk, P->pe[p]), u));
        }
    if (tab.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void**)&P->d_tab[gi], tab.size() * sizeof(int4));
    if (e == hipSuccess) e = hipMemcpy(P->d_tab[gi], tab.data(), tab.size() * sizeof(int4), hipMemcpyHostToDevice);
    if (e == hipSuccess) P->n_chunks[gi] = (int)tab.size();
    return e;
}

DChunks plan_chunks(const ppl_plan_s* P, KernelKind k, int max_obbs, int G, int* grid_out) {
    const int gi = group_index(G);
    const int grid_cap = grid_for(k, P->robot->R, obb_tile(max_obbs), G);
    DChunks C{};
    C.obb_cap = obb_tile(max_obbs);
    C.tiled = max_obbs > kMaxObbLds;
    C.tab = P->d_tab[gi];
    C.n_chunks = P->n_chunks[gi];
    C.chunk = kWave / G;
    C.obb_off = P->d_obb_off;
    *grid_out = std::max(1, std::min(grid_cap, C.n_chunks));
    return C;
}

// The any-collision launch shape of plan P for problems of up to max_obbs
// OBBs: this plan's tuned or set shape, else the tuning cache's entry for the
// robot, OBB count and GPU, else the built-in one (every OBB of a problem in
// one staging, persistent grid at occupancy). tile: OBB records per LDS
// staging; gmode 0 persistent grid at occupancy, 1 twice that, 2 one
// workgroup per chunk.
void collide_shape(ppl_plan_s* P, int max_obbs, int* tile, int* gmode) {
    if (P->shape_src != 2 && P->shape_obbs != max_obbs) {
        P->shape_obbs = max_obbs;
        P->ctile = P->cgrid = P->shape_src = 0;
        std::string v;
        int t = 0, g = 0;
        if (tune_get(collide_key(P->robot, max_obbs), v) && sscanf(v.c_str(), "%d %d", &t, &g) == 2 && t >= 0 &&
            t <= kMaxObbLds && g >= 0 && g <= 2) {
            P->ctile = t;
            P->cgrid = g;
            P->shape_src = 1;
        }
    }
    *tile = P->ctile > 0 ? std::min(P->ctile, obb_tile(max_obbs)) : obb_tile(max_obbs);
    *gmode = P->cgrid;
}

DChunks plan_collide_chunks(ppl_plan_s* P, int max_obbs, int* grid_out) {
    int tile = 0, gm = 0;
    collide_shape(P, max_obbs, &tile, &gm);
    DChunks C{};
    C.obb_cap = tile;
    C.tiled = max_obbs > tile;
    C.tab = P->d_tab[0];
    C.n_chunks = P->n_chunks[0];
    C.chunk = kWave;
    C.obb_off = P->d_obb_off;
    const int occ = grid_collide(lds_bytes(tile, P->robot->R.n_dof));
    const int g = gm == 2 ? C.n_chunks : (gm == 1 ? 2 * occ : occ);
    *grid_out = std::max(1, std::min(g, C.n_chunks));
    return C;
}

// Lanes per state for the cost kernel: enough lanes for about four waves per
// SIMD across the chip, at most kMaxGroup. The adjoint also keeps its per-state
// records within 16 KiB of LDS per wave.
int pick_group(long n_states, const DRobot& R, bool grad) {
    const long want = (long)device_cus() * 4 * 4 * kWave;
    int G = 1;
    while (G < kMaxGroup && n_states * G < want) G <<= 1;
    if (grad)
        while (G < kMaxGroup && (size_t)(kWave / G) * (rec_stride(R.n_joints) + dq_stride(R.n_dof)) * sizeof(float) > 16384)
            G <<= 1;
    return G;
}

DStates plan_states(const ppl_plan_s* P, const float* x) {
    DStates X{};
    X.n_states = P->n_states;
    const int D = P->robot->R.n_dof;
    if (P->kind == PPL_PLAN_CONFIGS) {
        X.source = kSrcConfig;
        X.cfg = x;
        X.cfg_stride = D;
    } else if (P->kind == PPL_PLAN_EDGES) {
        X.source = kSrcEdge;
        X.edge_off = P->d_off;
        X.n_edges = P->n_units;
        X.k0 = P->k0;
    } else {
        X.source = kSrcTraj;
        X.theta = x;
        X.traj_off = P->d_off;
        X.n_traj = P->n_units;
        X.prob = P->d_prob;
        X.wmajor = P->wmajor;
    }
    return X;
}

// Reference granularity rule (edge_eval.cuh:54-60) in FP32 with fused
// accumulation, identical to the device interpolation's inputs.
int edge_g(const float* a, const float* b, int dim, const ppl_edge_params& p, float* len_out) {
    float len2 = 0.0f;
    for (int i = 0; i < dim; ++i) {
        const float d = b[i] - a[i];
        len2 = std::fma(d, d, len2);
    }
    const float len = std::sqrt(len2);
    *len_out = len;
    if (!(p.resolution > 0.0f)) return p.granularity;
    const float gf = std::ceil(len / p.resolution);   // clamped before the cast: huge or NaN -> 1024
    const int g = gf < 1024.0f ? (int)gf : 1024;
    return std::max(g, 2);
}

// Self pass. mode 0: any-collision; mode 1: min clearance, exact below `band`
// (hinge callers pass their self margin).
ppl_status run_self(const ppl_robot_s* r, const DStates& X, int mode, uint8_t* hit_out, float* out, hipStream_t s,
                    float band = FLT_MAX, int* unit_first_hit = nullptr, ppl_plan_s* P = nullptr,
                    const int* unit_hint = nullptr) {
    SelfArgs a{};
    a.R = r->R;
    a.X = X;
    if (P && X.source != kSrcDense) {
        // the plan's 4-state chunks: unit hints, and waypoint-major order where the plan has it
        HIP_TRY(ensure_table(P, kSelfGroup));
        a.C.tab = P->d_tab[group_index(kSelfGroup)];
        a.C.n_chunks = P->n_chunks[group_index(kSelfGroup)];
        a.C.chunk = kWave / kSelfGroup;
    } else {
        a.X.wmajor = 0;  // flat plan order
        a.C.unit_hint = unit_hint;
    }
    a.mode = mode;
    a.band = band;
    a.unit_first_hit = unit_first_hit;
    a.hit_out = hit_out;
    a.out = out;
    HIP_TRY(launch_self(a, s));
    return PPL_OK;
}

}  // namespace

namespace ppl {
ppl_status fail_msg(ppl_status s, const char* msg) {
    g_err = msg;
    return s;
}
ppl_status scene_view_check(const ppl_scene_view* v) {
    DScene S{};
    ppl_status st = scene_from_view(v, S);
    return st;
}
}  // namespace ppl

extern "C" {

const char* ppl_last_error(void) { return g_err.c_str(); }
const char* ppl_version(void) { return "ppl 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// Robot
// ---------------------------------------------------------------------------
ppl_status ppl_robot_create(const ppl_robot_desc* d, ppl_robot* out) try {
    CHECK_ARG(d && out, "NULL argument");
    *out = nullptr;
    const int J = d->n_joints, D = d->n_dof, S = d->n_spheres, A = d->n_approx_spheres;
    CHECK_ARG(D >= 1 && D <= kMaxDof, "n_dof %d outside [1, %d]", D, kMaxDof);
    CHECK_ARG(J >= 1 && J <= kMaxJoints, "n_joints %d outside [1, %d]", J, kMaxJoints);
    CHECK_ARG(S >= 0 && A >= 0, "negative sphere counts");
    CHECK_ARG(d->fixed_transforms && d->joint_types && d->joint_lower && d->joint_upper, "NULL robot tables");
    CHECK_ARG(S == 0 || (d->spheres && d->sphere_to_joint), "NULL sphere tables");
    CHECK_ARG(A == 0 || (d->approx_spheres && d->approx_sphere_to_joint), "NULL approx sphere tables");
    CHECK_ARG(d->n_self_cc_ranges >= 0 && (d->n_self_cc_ranges == 0 || d->self_cc_ranges), "bad self ranges");
    CHECK_ARG(d->n_approx_self_cc_ranges >= 0 && (d->n_approx_self_cc_ranges == 0 || d->approx_self_cc_ranges),
              "bad approx self ranges");
    // Serial-chain defaults, robot_model.cuh:205-236.
    std::vector<int> parents(J), id2dof(J), tmem(J), dfs(J);
    for (int i = 0; i < J; ++i) {
        parents[i] = d->joint_parents ? d->joint_parents[i] : (i > 0 ? i - 1 : 0);
        id2dof[i] = d->joint_id_to_dof ? d->joint_id_to_dof[i] : i - 1;
        tmem[i] = d->t_memory_idx ? d->t_memory_idx[i] : (i == 0 ? 0 : 1);
        dfs[i] = d->dfs_order ? d->dfs_order[i] : i;
    }
    std::vector<int> pos(J, -1);
    for (int k = 0; k < J; ++k) {
        CHECK_ARG(dfs[k] >= 0 && dfs[k] < J && pos[dfs[k]] < 0, "dfs_order is not a permutation");
        pos[dfs[k]] = k;
    }
    const int root = dfs[0];
    CHECK_ARG(parents[root] == root, "dfs_order[0] must be the root (its own parent)");
    int n_slots = 1;
    for (int i = 0; i < J; ++i) {
        CHECK_ARG(parents[i] >= 0 && parents[i] < J, "joint_parents[%d] out of range", i);
        if (i != root) CHECK_ARG(pos[parents[i]] < pos[i], "joint %d visited before its parent", i);
        CHECK_ARG(id2dof[i] >= -1 && id2dof[i] < D, "joint_id_to_dof[%d] out of range", i);
        CHECK_ARG(tmem[i] >= 0, "negative t_memory_idx");
        n_slots = std::max(n_slots, tmem[i] + 1);
        const int t = d->joint_types[i];
        CHECK_ARG(t >= -1 && t <= 5, "joint_types[%d] = %d unknown", i, t);
        const float* ft = d->fixed_transforms + 16 * i;
        CHECK_ARG(std::fabs(ft[12]) < 1e-6f && std::fabs(ft[13]) < 1e-6f && std::fabs(ft[14]) < 1e-6f &&
                      std::fabs(ft[15] - 1.0f) < 1e-6f,
                  "fixed_transforms[%d] last row is not [0 0 0 1]", i);
        if (d->approx_fixed_transforms) {
            const float* af = d->approx_fixed_transforms + 16 * i;
            CHECK_ARG(std::fabs(af[12]) < 1e-6f && std::fabs(af[13]) < 1e-6f && std::fabs(af[14]) < 1e-6f &&
                          std::fabs(af[15] - 1.0f) < 1e-6f,
                      "approx_fixed_transforms[%d] last row is not [0 0 0 1]", i);
        }
        if (d->approx_joint_types)
            CHECK_ARG(d->approx_joint_types[i] >= -1 && d->approx_joint_types[i] <= 5, "approx_joint_types[%d] = %d unknown",
                      i, d->approx_joint_types[i]);
    }
    // The approximate spheres ride a chain of their own when the model gives
    // one that differs from the full chain (bindings.cpp:161 copies the full
    // chain when none is set; runtime_kinematics.cuh:176-243 walks it).
    // Without approximate transforms the full chain serves, types included.
    bool own_chain = false;
    for (int i = 0; i < J && !own_chain && d->approx_fixed_transforms; ++i) {
        if (d->approx_joint_types && d->approx_joint_types[i] != d->joint_types[i]) own_chain = true;
        for (int k = 0; k < 16; ++k)
            if (d->approx_fixed_transforms[16 * i + k] != d->fixed_transforms[16 * i + k]) own_chain = true;
    }
    if (n_slots > kMaxSlots) return fail(PPL_ERR_UNSUPPORTED, "%d FK slots requested, %d supported", n_slots, kMaxSlots);
    // The slot each joint reads must still hold its parent's frame.
    {
        std::vector<int> owner(n_slots, -1);
        for (int k = 0; k < J; ++k) {
            const int i = dfs[k];
            if (k > 0) CHECK_ARG(owner[tmem[parents[i]]] == parents[i], "slot of joint %d's parent was overwritten", i);
            owner[tmem[i]] = i;
        }
    }
    for (int s = 0; s < S; ++s) CHECK_ARG(d->sphere_to_joint[s] >= 0 && d->sphere_to_joint[s] < J, "sphere_to_joint[%d]", s);
    for (int a = 0; a < A; ++a)
        CHECK_ARG(d->approx_sphere_to_joint[a] >= 0 && d->approx_sphere_to_joint[a] < J, "approx_sphere_to_joint[%d]", a);
    std::vector<int2> pairs, apairs;
    for (int e = 0; e < d->n_self_cc_ranges; ++e) {
        const int i = d->self_cc_ranges[3 * e], a = d->self_cc_ranges[3 * e + 1], b = d->self_cc_ranges[3 * e + 2];
        CHECK_ARG(i >= 0 && i < S && a >= 0 && b < S, "self_cc_ranges[%d] out of range", e);
        for (int j = a; j <= b; ++j) pairs.push_back(make_int2(i, j));
    }
    for (int e = 0; e < d->n_approx_self_cc_ranges; ++e) {
        const int i = d->approx_self_cc_ranges[3 * e], a = d->approx_self_cc_ranges[3 * e + 1],
                  b = d->approx_self_cc_ranges[3 * e + 2];
        CHECK_ARG(i >= 0 && i < A && a >= 0 && b < A, "approx_self_cc_ranges[%d] out of range", e);
        for (int j = a; j <= b; ++j) apairs.push_back(make_int2(i, j));
    }
    // The reference checks self collision in two phases, exact pairs only
    // after an approximate pair hits (two_phase_cc.cuh:114-131). When every
    // exact pair's spheres lie inside approximate spheres of their joints that
    // form an approximate pair (the loader's models), an exact hit implies an
    // approximate one and the self pass checks exact pairs alone; otherwise
    // the self pass runs the approximate pairs first.
    bool self_gate_implied = !own_chain || pairs.empty();
    if (!own_chain) {
        std::vector<char> apair((size_t)A * A, 0);
        for (const int2& p : apairs) apair[(size_t)p.x * A + p.y] = apair[(size_t)p.y * A + p.x] = 1;
        auto inside = [&](int s, int a) {
            const float* f = d->spheres + 4 * s;
            const float* g = d->approx_spheres + 4 * a;
            const double dx = (double)f[0] - g[0], dy = (double)f[1] - g[1], dz = (double)f[2] - g[2];
            return d->approx_sphere_to_joint[a] == d->sphere_to_joint[s] &&
                   std::sqrt(dx * dx + dy * dy + dz * dz) + f[3] <= (double)g[3] + 1e-6;
        };
        for (const int2& p : pairs) {
            bool ok = false;
            for (int a = 0; a < A && !ok; ++a) {
                if (!inside(p.x, a)) continue;
                for (int b = 0; b < A && !ok; ++b) ok = apair[(size_t)a * A + b] && inside(p.y, b);
            }
            if (!ok) {
                self_gate_implied = false;
                break;
            }
        }
    }
    // The any-collision kernel may gate each (joint, OBB) pair on its own when
    // every exact sphere lies inside an approximate sphere of its joint on the
    // same chain: an exact hit of a box then implies the approximate hit of
    // that box, so flagging joints first (scene_collision.cuh:519-628) adds
    // nothing. Otherwise the kernel flags joints over the whole scene first.
    bool scene_gate_fused = !own_chain;
    for (int si = 0; si < S && scene_gate_fused; ++si) {
        const float* f = d->spheres + 4 * si;
        bool in = false;
        for (int a = 0; a < A && !in; ++a) {
            if (d->approx_sphere_to_joint[a] != d->sphere_to_joint[si]) continue;
            const float* g = d->approx_spheres + 4 * a;
            const double dx = (double)f[0] - g[0], dy = (double)f[1] - g[1], dz = (double)f[2] - g[2];
            in = std::sqrt(dx * dx + dy * dy + dz * dz) + f[3] <= (double)g[3] + 1e-6;
        }
        scene_gate_fused = in;
    }
    // DFS-ordered tables.
    std::vector<float> ft(12 * J);
    std::vector<int4> meta(J);
    for (int k = 0; k < J; ++k) {
        const int i = dfs[k];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) ft[12 * k + 4 * r + c] = d->fixed_transforms[16 * i + 4 * r + c];
        meta[k] = make_int4(d->joint_types[i], id2dof[i], tmem[i], tmem[parents[i]]);
    }
    std::vector<float> aft;
    std::vector<int> atype;
    if (own_chain) {
        aft.resize(12 * J);
        atype.resize(J);
        const float* src = d->approx_fixed_transforms;
        for (int k = 0; k < J; ++k) {
            const int i = dfs[k];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 4; ++c) aft[12 * k + 4 * r + c] = src[16 * i + 4 * r + c];
            atype[k] = d->approx_joint_types ? d->approx_joint_types[i] : d->joint_types[i];
        }
    }
    std::vector<int> apx_jnt(A);
    for (int a = 0; a < A; ++a) apx_jnt[a] = pos[d->approx_sphere_to_joint[a]];
    auto emit = [&](int n, const float* sp, const int* s2j, std::vector<int>& off, std::vector<float4>& v,
                    std::vector<int>& id) {
        off.assign(J + 1, 0);
        for (int k = 0; k < J; ++k) {
            off[k] = (int)v.size();
            for (int s = 0; s < n; ++s)
                if (s2j[s] == dfs[k]) {
                    v.push_back(make_float4(sp[4 * s], sp[4 * s + 1], sp[4 * s + 2], sp[4 * s + 3]));
                    id.push_back(s);
                }
        }
        off[J] = (int)v.size();
    };
    std::vector<int> sph_off, apx_off, sph_id, apx_id;
    std::vector<float4> sph, apx;
    emit(S, d->spheres, d->sphere_to_joint, sph_off, sph, sph_id);
    emit(A, d->approx_spheres, d->approx_sphere_to_joint, apx_off, apx, apx_id);
    // Sphere clusters for the cost kernels' candidate masks: each joint's
    // spheres are bisected along their principal axis (FP64, median split)
    // while a part holds more than 16 spheres, or more than 4 with an
    // enclosing radius above 0.15 m; the leaves become contiguous runs of
    // emission slots (the joint's spheres are re-emitted in leaf order).
    // A mask built around a cluster's enclosing sphere holds far fewer
    // boxes than one around a large link's.
    std::vector<int> clu_off(J + 1, 0), clu_beg;
    std::vector<float4> clu_enc;
    {
        auto enclose = [&](const std::vector<int>& ix, double* c) {
            c[0] = c[1] = c[2] = 0.0;
            for (int t : ix) { c[0] += sph[t].x; c[1] += sph[t].y; c[2] += sph[t].z; }
            for (int k = 0; k < 3; ++k) c[k] /= (double)ix.size();
            double m = 0.0;
            for (int t : ix) {
                const double dx = sph[t].x - c[0], dy = sph[t].y - c[1], dz = sph[t].z - c[2];
                m = std::max(m, std::sqrt(dx * dx + dy * dy + dz * dz) + sph[t].w);
            }
            return m;
        };
        std::vector<float4> sph2(sph);
        std::vector<int> id2(sph_id);
        std::vector<std::vector<int>> leaves;
        std::function<void(std::vector<int>)> split = [&](std::vector<int> ix) {
            double c[3];
            const double rad = enclose(ix, c);
            if (ix.size() <= 1 || (ix.size() <= 16 && (ix.size() <= 4 || rad <= 0.15))) {
                leaves.push_back(ix);
                return;
            }
            // principal axis by power iteration on the 3x3 covariance
            double C[9] = {0};
            for (int t : ix) {
                const double v[3] = {sph[t].x - c[0], sph[t].y - c[1], sph[t].z - c[2]};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) C[3 * a + b] += v[a] * v[b];
            }
            double u[3] = {1.0, 0.7, 0.4};
            for (int it = 0; it < 64; ++it) {
                double w[3] = {0, 0, 0};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) w[a] += C[3 * a + b] * u[b];
                const double n = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                if (!(n > 0.0)) break;
                for (int a = 0; a < 3; ++a) u[a] = w[a] / n;
            }
            std::vector<std::pair<double, int>> key;
            for (int t : ix) key.push_back({(sph[t].x - c[0]) * u[0] + (sph[t].y - c[1]) * u[1] + (sph[t].z - c[2]) * u[2], t});
            std::stable_sort(key.begin(), key.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            const size_t h = ix.size() / 2;
            std::vector<int> lo, hi;
            for (size_t k = 0; k < key.size(); ++k) (k < h ? lo : hi).push_back(key[k].second);
            split(lo);
            split(hi);
        };
        for (int k = 0; k < J; ++k) {
            clu_off[k] = (int)clu_beg.size();
            if (sph_off[k] == sph_off[k + 1]) continue;
            std::vector<int> ix;
            for (int t = sph_off[k]; t < sph_off[k + 1]; ++t) ix.push_back(t);
            leaves.clear();
            split(ix);
            int at = sph_off[k];
            for (const auto& lf : leaves) {
                clu_beg.push_back(at);
                double c[3];
                const double rad = enclose(lf, c);
                // 1e-5 m larger, so FP32 rounding cannot make a sphere poke out
                clu_enc.push_back(make_float4((float)c[0], (float)c[1], (float)c[2], (float)(rad + 1e-5)));
                for (int t : lf) {
                    sph2[at] = sph[t];
                    id2[at] = sph_id[t];
                    ++at;
                }
            }
        }
        clu_off[J] = (int)clu_beg.size();
        clu_beg.push_back(S);
        sph.swap(sph2);
        sph_id.swap(id2);
    }
    const int n_clu = (int)clu_enc.size();
    CHECK_ARG(n_clu <= kMaxClusters, "%d sphere clusters, at most %d supported", n_clu, kMaxClusters);
    // Self-pair blocks for the self pass: the exact pairs grouped by the
    // sphere clusters (A, B), A <= B, of their two spheres, each block a list
    // of (slot a | slot b << 16) in slot order. A state tests each block's two
    // cluster enclosing spheres and walks the pair lists of the blocks that
    // pass, 16 pairs per step.
    CHECK_ARG(S <= 65535, "%d spheres, at most 65535 supported", S);
    std::vector<int4> cpb;
    std::vector<unsigned> cpl;
    {
        std::vector<int> slot_of(S), clu_of_slot(S);
        for (int k = 0; k < S; ++k) slot_of[sph_id[k]] = k;
        for (int c = 0; c + 1 < (int)clu_beg.size(); ++c)
            for (int t = clu_beg[c]; t < clu_beg[c + 1]; ++t) clu_of_slot[t] = c;
        std::vector<std::array<int, 4>> ent;   // (cluster A, cluster B, slot a, slot b)
        ent.reserve(pairs.size());
        for (const int2& p : pairs) {
            int u = slot_of[p.x], v = slot_of[p.y];
            if (clu_of_slot[u] > clu_of_slot[v] || (clu_of_slot[u] == clu_of_slot[v] && u > v)) std::swap(u, v);
            ent.push_back({clu_of_slot[u], clu_of_slot[v], u, v});
        }
        std::sort(ent.begin(), ent.end());
        for (size_t k = 0; k < ent.size();) {
            size_t e = k;
            while (e < ent.size() && ent[e][0] == ent[k][0] && ent[e][1] == ent[k][1]) ++e;
            cpb.push_back(make_int4(ent[k][0], ent[k][1], (int)cpl.size(), (int)(e - k)));
            for (size_t t = k; t < e; ++t) cpl.push_back((unsigned)ent[t][2] | ((unsigned)ent[t][3] << 16));
            k = e;
        }
    }
    std::vector<float4> sph_orig(S), apx_orig(A);
    for (int s = 0; s < S; ++s) sph_orig[s] = make_float4(d->spheres[4 * s], d->spheres[4 * s + 1], d->spheres[4 * s + 2], d->spheres[4 * s + 3]);
    for (int a = 0; a < A; ++a)
        apx_orig[a] = make_float4(d->approx_spheres[4 * a], d->approx_spheres[4 * a + 1], d->approx_spheres[4 * a + 2],
                                  d->approx_spheres[4 * a + 3]);
    // Parent of each DFS position, as a DFS position (root: itself).
    std::vector<int> close(J, 0);
    for (int k = 0; k < J; ++k) close[k] = k == 0 ? 0 : pos[parents[dfs[k]]];
    // Subtree ends in DFS order (a subtree is a contiguous run of DFS
    // positions) and the DFS joint of every emitted sphere.
    std::vector<int> jend(J);
    for (int k = 0; k < J; ++k) jend[k] = k + 1;
    for (int k = J - 1; k >= 1; --k) jend[close[k]] = std::max(jend[close[k]], jend[k]);
    std::vector<int> sph_jnt(S), sph_clu(S), clu_jnt(n_clu);
    for (int k = 0; k < J; ++k)
        for (int t = sph_off[k]; t < sph_off[k + 1]; ++t) sph_jnt[t] = k;
    for (int k = 0; k < J; ++k)
        for (int c = clu_off[k]; c < clu_off[k + 1]; ++c) {
            clu_jnt[c] = k;
            for (int t = clu_beg[c]; t < clu_beg[c + 1]; ++t) sph_clu[t] = c;
        }
    // One device blob for every table.
    size_t off = 0;
    auto place = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 15) & ~(size_t)15;
        return o;
    };
    const size_t o_ft = place(ft.size() * 4), o_meta = place(meta.size() * 16), o_soff = place(sph_off.size() * 4),
                 o_sph = place(sph.size() * 16), o_sid = place(sph_id.size() * 4), o_aoff = place(apx_off.size() * 4),
                 o_apx = place(apx.size() * 16), o_aid = place(apx_id.size() * 4), o_close = place(close.size() * 4),
                 o_sorig = place(sph_orig.size() * 16), o_pairs = place(pairs.size() * 8),
                 o_apairs = place(apairs.size() * 8), o_aorig = place(apx_orig.size() * 16),
                 o_cpb = place(cpb.size() * 16), o_cpl = place(cpl.size() * 4), o_jorig = place(J * 4),
                 o_lim = place((size_t)D * 8),
                 o_sjnt = place((size_t)S * 4), o_jend = place((size_t)J * 4),
                 o_aft = place(aft.size() * 4), o_atype = place(atype.size() * 4), o_ajnt = place(apx_jnt.size() * 4),
                 o_coff = place(clu_off.size() * 4), o_cbeg = place(clu_beg.size() * 4), o_cenc = place(clu_enc.size() * 16),
                 o_sclu = place(sph_clu.size() * 4), o_cjnt = place(clu_jnt.size() * 4);
    std::vector<char> host(off + 16, 0);
    auto put = [&](size_t o, const void* src, size_t bytes) { if (bytes) memcpy(host.data() + o, src, bytes); };
    put(o_ft, ft.data(), ft.size() * 4);
    put(o_meta, meta.data(), meta.size() * 16);
    put(o_soff, sph_off.data(), sph_off.size() * 4);
    put(o_sph, sph.data(), sph.size() * 16);
    put(o_sid, sph_id.data(), sph_id.size() * 4);
    put(o_aoff, apx_off.data(), apx_off.size() * 4);
    put(o_apx, apx.data(), apx.size() * 16);
    put(o_aid, apx_id.data(), apx_id.size() * 4);
    put(o_close, close.data(), close.size() * 4);
    put(o_sorig, sph_orig.data(), sph_orig.size() * 16);
    put(o_pairs, pairs.data(), pairs.size() * 8);
    put(o_apairs, apairs.data(), apairs.size() * 8);
    put(o_aorig, apx_orig.data(), apx_orig.size() * 16);
    put(o_cpb, cpb.data(), cpb.size() * 16);
    put(o_cpl, cpl.data(), cpl.size() * 4);
    put(o_jorig, dfs.data(), (size_t)J * 4);
    std::vector<float2> lim(D);
    for (int k = 0; k < D; ++k) lim[k] = make_float2(d->joint_lower[k], d->joint_upper[k]);
    put(o_lim, lim.data(), (size_t)D * 8);
    put(o_sjnt, sph_jnt.data(), (size_t)S * 4);
    put(o_jend, jend.data(), (size_t)J * 4);
    put(o_aft, aft.data(), aft.size() * 4);
    put(o_atype, atype.data(), atype.size() * 4);
    put(o_ajnt, apx_jnt.data(), apx_jnt.size() * 4);
    put(o_coff, clu_off.data(), clu_off.size() * 4);
    put(o_cbeg, clu_beg.data(), clu_beg.size() * 4);
    put(o_cenc, clu_enc.data(), clu_enc.size() * 16);
    put(o_sclu, sph_clu.data(), sph_clu.size() * 4);
    put(o_cjnt, clu_jnt.data(), clu_jnt.size() * 4);

    ppl_robot_s* r = new (std::nothrow) ppl_robot_s();
    if (!r) return fail(PPL_ERR_ALLOC, "out of host memory");
    r->digest = 1469598103934665603ull;
    for (char c : host) r->digest = (r->digest ^ (unsigned char)c) * 1099511628211ull;
    if (hipMalloc(&r->blob, host.size()) != hipSuccess) {
        delete r;
        return fail(PPL_ERR_ALLOC, "hipMalloc of %zu bytes for robot tables failed", host.size());
    }
    hipError_t e = hipMemcpy(r->blob, host.data(), host.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        ppl_robot_destroy(r);
        return fail(PPL_ERR_HIP, "robot upload failed: %s", hipGetErrorString(e));
    }
    char* b = static_cast<char*>(r->blob);
    DRobot& R = r->R;
    R.n_dof = D; R.n_joints = J; R.n_sph = S; R.n_apx = A; R.n_slots = n_slots;
    R.n_pairs = (int)pairs.size(); R.n_apx_pairs = (int)apairs.size();
    R.ft = reinterpret_cast<const float*>(b + o_ft);
    R.jmeta = reinterpret_cast<const int4*>(b + o_meta);
    R.sph_off = reinterpret_cast<const int*>(b + o_soff);
    R.sph = reinterpret_cast<const float4*>(b + o_sph);
    R.sph_id = reinterpret_cast<const int*>(b + o_sid);
    R.apx_off = reinterpret_cast<const int*>(b + o_aoff);
    R.apx = reinterpret_cast<const float4*>(b + o_apx);
    R.apx_id = reinterpret_cast<const int*>(b + o_aid);
    R.parent = reinterpret_cast<const int*>(b + o_close);
    R.sph_orig = reinterpret_cast<const float4*>(b + o_sorig);
    R.pairs = reinterpret_cast<const int2*>(b + o_pairs);
    R.apx_pairs = reinterpret_cast<const int2*>(b + o_apairs);
    R.apx_orig = reinterpret_cast<const float4*>(b + o_aorig);
    R.cpb = reinterpret_cast<const int4*>(b + o_cpb);
    R.cpl = reinterpret_cast<const unsigned*>(b + o_cpl);
    R.jorig = reinterpret_cast<const int*>(b + o_jorig);
    R.lim = reinterpret_cast<const float2*>(b + o_lim);
    R.sph_jnt = reinterpret_cast<const int*>(b + o_sjnt);
    R.jend = reinterpret_cast<const int*>(b + o_jend);
    R.aft = own_chain ? reinterpret_cast<const float*>(b + o_aft) : nullptr;
    R.atype = own_chain ? reinterpret_cast<const int*>(b + o_atype) : nullptr;
    R.apx_jnt = reinterpret_cast<const int*>(b + o_ajnt);
    R.n_clu = n_clu;
    R.clu_off = reinterpret_cast<const int*>(b + o_coff);
    R.clu_beg = reinterpret_cast<const int*>(b + o_cbeg);
    R.clu_enc = reinterpret_cast<const float4*>(b + o_cenc);
    R.sph_clu = reinterpret_cast<const int*>(b + o_sclu);
    R.clu_jnt = reinterpret_cast<const int*>(b + o_cjnt);
    R.scene_gate_fused = scene_gate_fused ? 1 : 0;
    R.self_gate_implied = self_gate_implied ? 1 : 0;
    R.n_cpb = (int)cpb.size();
    r->n_slots = n_slots;
    r->n_pairs = (int)pairs.size();
    for (int i = 0; i < J; ++i) {
        if (i == root || id2dof[i] < 0) continue;
        if (d->joint_types[i] >= kXRot) ++r->n_revolute;
        else if (d->joint_types[i] >= kXPrism) ++r->n_prismatic;
    }
    r->lower.assign(d->joint_lower, d->joint_lower + D);
    r->upper.assign(d->joint_upper, d->joint_upper + D);
    *out = r;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_robot_get_device_view(ppl_robot r, ppl_robot_device_view* v) try {
    CHECK_ARG(r && v, "NULL argument");
    const DRobot& R = r->R;
    v->n_dof = R.n_dof;
    v->n_joints = R.n_joints;
    v->n_spheres = R.n_sph;
    v->n_approx_spheres = R.n_apx;
    v->n_pairs = R.n_pairs;
    v->n_approx_pairs = R.n_apx_pairs;
    v->ft = R.ft;
    v->jmeta = reinterpret_cast<const int*>(R.jmeta);
    v->sph_off = R.sph_off;
    v->sph = reinterpret_cast<const float*>(R.sph);
    v->sph_id = R.sph_id;
    v->apx_off = R.apx_off;
    v->apx = reinterpret_cast<const float*>(R.apx);
    v->apx_id = R.apx_id;
    v->sph_orig = reinterpret_cast<const float*>(R.sph_orig);
    v->apx_orig = reinterpret_cast<const float*>(R.apx_orig);
    v->pairs = reinterpret_cast<const int*>(R.pairs);
    v->apx_pairs = reinterpret_cast<const int*>(R.apx_pairs);
    v->lim = reinterpret_cast<const float*>(R.lim);
    v->aft = R.aft;
    v->atype = R.atype;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_robot_destroy(ppl_robot r) try {
    if (!r) return PPL_OK;
    for (auto& e : r->traj_plans) ppl_plan_destroy(e.second);
    r->traj_plans.clear();
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    if (r->blob) (void)hipFree(r->blob);
    r->d_cfg.release(); r->d_out.release(); r->d_off.release(); r->d_aux.release(); r->d_self.release();
    r->d_mit.release(); r->d_hint.release(); r->h_mit.release();
    r->h_stage.release();
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_robot_get_info(ppl_robot r, ppl_robot_info* info) try {
    CHECK_ARG(r && info, "NULL argument");
    info->n_dof = r->R.n_dof;
    info->n_joints = r->R.n_joints;
    info->n_spheres = r->R.n_sph;
    info->n_approx_spheres = r->R.n_apx;
    info->n_self_pairs = r->n_pairs;
    info->n_t_slots = r->n_slots;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_robot_update_joint_limits(ppl_robot r, const float* lo, const float* hi) try {
    CHECK_ARG(r && lo && hi, "NULL argument");
    const int D = r->R.n_dof;
    r->lower.assign(lo, lo + D);
    r->upper.assign(hi, hi + D);
    std::vector<float2> lim(D);
    for (int k = 0; k < D; ++k) lim[k] = make_float2(lo[k], hi[k]);
    HIP_TRY(hipMemcpy(const_cast<float2*>(r->R.lim), lim.data(), (size_t)D * sizeof(float2), hipMemcpyHostToDevice));
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

// ---------------------------------------------------------------------------
// Plain any-collision (pRRTC::check_collision_scene)
// ---------------------------------------------------------------------------
ppl_status ppl_check_collision_scene(ppl_robot r, const ppl_scene_view* view, const float* h_cfg, int n, float margin,
                                     uint8_t* h_out) try {
    CHECK_ARG(r, "robot is NULL");
    CHECK_ARG(n >= 0, "negative N");
    if (n == 0) return PPL_OK;
    CHECK_ARG(h_cfg && h_out, "NULL host arrays");
    DScene S{};
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    const int D = r->R.n_dof;
    const size_t cfg_bytes = (size_t)n * D * sizeof(float);
    HIP_TRY(r->d_cfg.ensure(cfg_bytes));
    HIP_TRY(r->d_out.ensure((size_t)n));
    HIP_TRY(r->h_stage.ensure(std::max(cfg_bytes, (size_t)n)));
    memcpy(r->h_stage.p, h_cfg, cfg_bytes);
    HIP_TRY(hipMemcpyAsync(r->d_cfg.p, r->h_stage.p, cfg_bytes, hipMemcpyHostToDevice, r->stream));
    CollideArgs a{};
    a.R = r->R;
    a.S = S;
    a.X.source = kSrcConfig;
    a.X.n_states = n;
    a.X.cfg = r->d_cfg.as<float>();
    a.X.cfg_stride = D;
    int grid = 0;
    a.C = implicit_chunks(n, kKCollide, r->R, S.n_obbs, &grid);
    a.margin = margin;
    a.mode = 0;
    a.out = r->d_out.as<uint8_t>();
    HIP_TRY(launch_collide(a, grid, r->stream));
    HIP_TRY(hipMemcpyAsync(r->h_stage.p, r->d_out.p, (size_t)n, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    memcpy(h_out, r->h_stage.p, (size_t)n);
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

// ---------------------------------------------------------------------------
// Plain edge evaluator (EdgeEvalGpu)
// ---------------------------------------------------------------------------
ppl_status ppl_edge_eval_create(ppl_robot r, const ppl_scene_view* view, int dim, int max_edges, ppl_edge_eval* out) try {
    CHECK_ARG(r && out, "NULL argument");
    *out = nullptr;
    CHECK_ARG(dim == r->R.n_dof, "dim %d != robot n_dof %d", dim, r->R.n_dof);
    CHECK_ARG(max_edges > 0, "max_edges must be positive");
    DScene S{};
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    ppl_edge_eval_s* ev = new (std::nothrow) ppl_edge_eval_s();
    if (!ev) return fail(PPL_ERR_ALLOC, "out of host memory");
    ev->robot = r;
    ev->S = S;
    ev->dim = dim;
    ev->max_edges = max_edges;
    const EdgeIoLayout io(max_edges, dim);
    hipError_t e = hipStreamCreateWithFlags(&ev->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = ev->d_io.ensure(io.end);
    if (e == hipSuccess) e = ev->h_io.ensure(io.end);
    if (e == hipSuccess) e = ev->h_clr.ensure((size_t)max_edges * 4);
    if (e != hipSuccess) {
        ppl_edge_eval_destroy(ev);
        return fail(PPL_ERR_ALLOC, "edge evaluator buffers: %s", hipGetErrorString(e));
    }
    *out = ev;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_edge_eval_destroy(ppl_edge_eval ev) try {
    if (!ev) return PPL_OK;
    if (ev->stream) (void)hipStreamSynchronize(ev->stream);
    ev->d_io.release(); ev->d_clr_state.release(); ev->d_clr_edge.release(); ev->d_self.release();
    ev->h_io.release(); ev->h_clr.release();
    if (ev->stream) (void)hipStreamDestroy(ev->stream);
    delete ev;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

}  // extern "C"

namespace {
__global__ void k_segment_min(const float* x, const int* off, int n_seg, float* out) {
    const int b = blockIdx.x;
    if (b >= n_seg) return;
    const int lane = threadIdx.x;
    float m = FLT_MAX;
    for (int i = off[b] + lane; i < off[b + 1]; i += 64) m = fminf(m, x[i]);
    for (int o = 32; o >= 1; o >>= 1) m = fminf(m, __shfl_xor(m, o));
    if (lane == 0) out[b] = m;
}
}  // namespace

extern "C" {

ppl_status ppl_edge_granularities(const float* h_from, const float* h_to, int n, int dim, const ppl_edge_params* p,
                                  int* h_G, float* h_len) try {
    CHECK_ARG(h_from && h_to && p && h_G, "NULL argument");
    CHECK_ARG(n >= 0 && dim >= 1, "bad sizes");
    for (int e = 0; e < n; ++e) {
        float len = 0.0f;
        h_G[e] = edge_g(h_from + (size_t)e * dim, h_to + (size_t)e * dim, dim, *p, &len);
        if (h_len) h_len[e] = len;
    }
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_edge_eval_evaluate(ppl_edge_eval ev, const float* h_from, const float* h_to, int n, const ppl_edge_params* p,
                                  uint8_t* h_valid, float* h_cost, float* h_clr) try {
    CHECK_ARG(ev && p, "NULL argument");
    CHECK_ARG(n >= 0, "negative edge count");
    if (n == 0) return PPL_OK;
    if (n > ev->max_edges) return fail(PPL_ERR_CAPACITY, "num_edges %d exceeds max_edges %d", n, ev->max_edges);
    CHECK_ARG(h_from && h_to && h_valid && h_cost, "NULL host arrays");
    if (p->enable_mesh) return fail(PPL_ERR_UNSUPPORTED, "mesh edge checks are not supported");
    ppl_robot r = ev->robot;
    const int D = ev->dim;
    const EdgeIoLayout io(n, D);
    char* hio = ev->h_io.as<char>();
    char* dio = ev->d_io.as<char>();
    int* off = reinterpret_cast<int*>(hio);
    std::vector<float> len(n);
    off[0] = 0;
    for (int e = 0; e < n; ++e) {
        const int G = edge_g(h_from + (size_t)e * D, h_to + (size_t)e * D, D, *p, &len[e]);
        const long next = (long)off[e] + std::max(G, 0);
        if (next > INT_MAX) return fail(PPL_ERR_CAPACITY, "more than 2^31 - 1 edge states in one call");
        off[e + 1] = (int)next;
    }
    const long total = off[n];
    const size_t cfg = (size_t)n * D * sizeof(float);
    int* h_fh = reinterpret_cast<int*>(hio + io.fh);
    for (int e = 0; e < n; ++e) h_fh[e] = -1;   // first hits start at "none"
    memcpy(hio + io.from, h_from, cfg);
    memcpy(hio + io.to, h_to, cfg);
    HIP_TRY(hipMemcpyAsync(dio, hio, io.end, hipMemcpyHostToDevice, ev->stream));
    int* d_off = reinterpret_cast<int*>(dio);
    int* d_fh = reinterpret_cast<int*>(dio + io.fh);
    DStates X{};
    X.source = kSrcEdge;
    X.n_states = total;
    X.from = reinterpret_cast<const float*>(dio + io.from);
    X.to = reinterpret_cast<const float*>(dio + io.to);
    X.edge_off = d_off;
    X.n_edges = n;
    X.k0 = 1;
    // With G = granularity <= 1 the +1 state count gives t = k/G as the reference.
    const uint8_t* self_hit = nullptr;
    if (total > 0 && p->check_self && r->n_pairs > 0) {
        HIP_TRY(ev->d_self.ensure((size_t)total));
        ppl_status st = run_self(r, X, 0, ev->d_self.as<uint8_t>(), nullptr, ev->stream);
        if (st != PPL_OK) return st;
        self_hit = ev->d_self.as<uint8_t>();
    }
    if (total > 0) {
        CollideArgs a{};
        a.R = r->R;
        a.S = ev->S;
        a.X = X;
        int grid = 0;
        a.C = implicit_chunks(total, kKCollide, r->R, ev->S.n_obbs, &grid);
        a.margin = p->collision_margin;
        a.mode = 1;
        a.skip_resolved = 1;
        a.first_hit = d_fh;
        a.self_hit = self_hit;
        HIP_TRY(launch_collide(a, grid, ev->stream));
    }
    const bool want_clr = p->compute_clearance && h_clr;
    if (want_clr && total > 0) {
        HIP_TRY(ev->d_clr_state.ensure((size_t)total * 4));
        HIP_TRY(ev->d_clr_edge.ensure((size_t)n * 4));
        CostArgs c{};
        c.R = r->R;
        c.S = ev->S;
        c.X = X;
        c.P.weight = 1.0f;
        c.clearance_only = 1;
        c.margin = p->collision_margin;
        c.cost_out = ev->d_clr_state.as<float>();
        HIP_TRY(launch_cost_flat(c, total, ev->stream));
        hipLaunchKernelGGL(k_segment_min, dim3(n), dim3(64), 0, ev->stream, ev->d_clr_state.as<float>(), d_off, n,
                           ev->d_clr_edge.as<float>());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(ev->h_clr.p, ev->d_clr_edge.p, (size_t)n * 4, hipMemcpyDeviceToHost, ev->stream));
    }
    HIP_TRY(hipMemcpyAsync(h_fh, d_fh, (size_t)n * 4, hipMemcpyDeviceToHost, ev->stream));
    HIP_TRY(hipStreamSynchronize(ev->stream));
    const int* fh = h_fh;
    for (int e = 0; e < n; ++e) {
        const bool ok = fh[e] < 0;
        h_valid[e] = ok ? 1 : 0;
        h_cost[e] = ok ? len[e] : FLT_MAX;
    }
    if (h_clr) {
        for (int e = 0; e < n; ++e) {
            const bool has = want_clr && off[e + 1] > off[e];
            h_clr[e] = has ? ev->h_clr.as<float>()[e] : FLT_MAX;
        }
    }
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
ppl_status ppl_plan_create(ppl_robot r, const ppl_plan_desc* d, ppl_plan* out) try {
    CHECK_ARG(r && d && out, "NULL argument");
    *out = nullptr;
    CHECK_ARG(d->kind >= PPL_PLAN_CONFIGS && d->kind <= PPL_PLAN_TRAJ, "unknown plan kind %d", d->kind);
    CHECK_ARG(d->n_units >= 0, "negative n_units");
    CHECK_ARG(d->n_problems >= 1, "n_problems must be >= 1");
    CHECK_ARG(d->kind == PPL_PLAN_CONFIGS || d->unit_states, "unit_states required for edge/trajectory plans");
    ppl_plan_s* P = new (std::nothrow) ppl_plan_s();
    if (!P) return fail(PPL_ERR_ALLOC, "out of host memory");
    P->robot = r;
    P->kind = d->kind;
    P->n_units = d->n_units;
    P->k0 = (d->kind == PPL_PLAN_EDGES && d->skip_first) ? 1 : 0;
    P->n_problems = d->n_problems;
    // unit -> state offsets
    std::vector<long> uoff(d->n_units + 1, 0);
    for (int u = 0; u < d->n_units; ++u) {
        long c = 1;
        if (d->kind == PPL_PLAN_EDGES) c = std::max(0, d->unit_states[u] - P->k0);
        if (d->kind == PPL_PLAN_TRAJ) c = d->unit_states[u];
        if (c < 0 || (d->kind == PPL_PLAN_TRAJ && c < 1)) {
            delete P;
            return fail(PPL_ERR_INVALID, "unit %d has invalid state count", u);
        }
        uoff[u + 1] = uoff[u] + c;
    }
    if (uoff[d->n_units] > (long)INT32_MAX - 64) {
        delete P;
        return fail(PPL_ERR_CAPACITY, "more than 2^31 states in one plan");
    }
    P->n_states = uoff[d->n_units];
    std::vector<int> pu(d->n_problems + 1);
    for (int p = 0; p <= d->n_problems; ++p) pu[p] = d->problem_units ? d->problem_units[p] : (p == 0 ? 0 : d->n_units);
    if (!d->problem_units && d->n_problems != 1) {
        delete P;
        return fail(PPL_ERR_INVALID, "problem_units required when n_problems > 1");
    }
    for (int p = 0; p < d->n_problems; ++p)
        if (pu[p] < 0 || pu[p] > pu[p + 1] || pu[p + 1] > d->n_units) {
            delete P;
            return fail(PPL_ERR_INVALID, "problem_units not monotone in [0, n_units]");
        }
    if (pu[0] != 0 || pu[d->n_problems] != d->n_units) {
        delete P;
        return fail(PPL_ERR_INVALID, "problem_units must cover [0, n_units]");
    }
    if (d->obb_offsets) {
        P->h_obb_off.assign(d->obb_offsets, d->obb_offsets + d->n_problems + 1);
        for (int p = 0; p < d->n_problems; ++p) {
            const int n = P->h_obb_off[p + 1] - P->h_obb_off[p];
            if (n < 0) {
                delete P;
                return fail(PPL_ERR_INVALID, "problem %d has a negative OBB range", p);
            }
        }
    }
    for (int p = 0; p < d->n_problems; ++p) {
        P->pb.push_back(uoff[pu[p]]);
        P->pe.push_back(uoff[pu[p + 1]]);
    }
    std::vector<int4> prob;
    if (d->waypoint_major) {
        if (d->kind != PPL_PLAN_TRAJ) {
            delete P;
            return fail(PPL_ERR_INVALID, "waypoint_major needs a trajectory plan");
        }
        for (int p = 0; p < d->n_problems; ++p) {
            const int u0 = pu[p], nu = pu[p + 1] - pu[p];
            const int T = nu > 0 ? d->unit_states[u0] : 1;
            for (int u = u0; u < pu[p + 1]; ++u)
                if (d->unit_states[u] != T) {
                    delete P;
                    return fail(PPL_ERR_INVALID, "waypoint_major: trajectories of problem %d differ in length", p);
                }
            prob.push_back(make_int4(u0, std::max(nu, 1), T, (int)uoff[u0]));
        }
        P->wmajor = 1;
    }
    P->h_off.assign(uoff.size(), 0);
    for (size_t i = 0; i < uoff.size(); ++i) P->h_off[i] = (int)uoff[i];
    for (size_t i = 0; i + 1 < uoff.size(); ++i) P->t_max = std::max(P->t_max, (int)(uoff[i + 1] - uoff[i]));
    hipError_t e = ensure_table(P, 1);
    if (e == hipSuccess && !prob.empty()) {
        e = hipMalloc((void**)&P->d_prob, prob.size() * sizeof(int4));
        if (e == hipSuccess) e = hipMemcpy(P->d_prob, prob.data(), prob.size() * sizeof(int4), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMalloc((void**)&P->d_off, P->h_off.size() * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(P->d_off, P->h_off.data(), P->h_off.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess && !P->h_obb_off.empty()) {
        e = hipMalloc((void**)&P->d_obb_off, P->h_obb_off.size() * sizeof(int));
        if (e == hipSuccess)
            e = hipMemcpy(P->d_obb_off, P->h_obb_off.data(), P->h_obb_off.size() * sizeof(int), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        ppl_plan_destroy(P);
        return fail(PPL_ERR_HIP, "plan upload: %s", hipGetErrorString(e));
    }
    *out = P;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_plan_destroy(ppl_plan P) try {
    if (!P) return PPL_OK;
    (void)hipDeviceSynchronize();
    for (int i = 0; i < kGroupKinds; ++i)
        if (P->d_tab[i]) (void)hipFree(P->d_tab[i]);
    if (P->d_off) (void)hipFree(P->d_off);
    if (P->d_obb_off) (void)hipFree(P->d_obb_off);
    if (P->d_prob) (void)hipFree(P->d_prob);
    for (auto& e : P->ainv) (void)hipFree(e.second);
    if (P->d_ainv_ptr) (void)hipFree(P->d_ainv_ptr);
    ppl_graph_destroy(P->chomp_graph);
    ppl_graph_destroy(P->stomp_graph);
    if (P->d_iter) (void)hipFree(P->d_iter);
    if (StompState* z = P->stomp) {
        ppl_plan_destroy(z->first_plan);
        ppl_plan_destroy(z->fresh_plan);
        for (auto& e : z->mats) (void)hipFree(e.second);
        if (z->d_lt) (void)hipFree((void*)z->d_lt);
        z->buf.release();
        z->hint.release();
        delete z;
    }
    P->d_opt.release();
    P->d_self.release(); P->d_cost.release(); P->d_fdcfg.release(); P->d_fdcost.release();
    for (auto& e : P->kev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    delete P;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

long ppl_plan_states(ppl_plan P) { return P ? P->n_states : -1; }

ppl_status ppl_plan_time_kernels(ppl_plan P, int enable) try {
    CHECK_ARG(P, "NULL plan");
    P->ktime = enable != 0;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_plan_kernel_ms(ppl_plan P, double* total_ms, int* launches) try {
    CHECK_ARG(P && total_ms && launches, "NULL argument");
    double sum = 0.0;
    hipError_t e = hipSuccess;
    for (int i = 0; i < P->kused && e == hipSuccess; ++i) {
        float ms = 0.0f;
        e = hipEventSynchronize(P->kev[i].second);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, P->kev[i].first, P->kev[i].second);
        sum += ms;
    }
    const int n = P->kused;
    P->kused = 0;
    if (e != hipSuccess) return fail(PPL_ERR_HIP, "kernel timing: %s", hipGetErrorString(e));
    *total_ms = sum;
    *launches = n;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

double ppl_algorithmic_flops(ppl_robot r, int kind, int n_obbs, int n_waypoints, int use_self) {
    if (!r || n_obbs < 0) return -1.0;
    const int nr = r->n_revolute, np = r->n_prismatic, ns = r->R.n_sph, D = r->R.n_dof;
    const int pairs = use_self ? r->n_pairs : 0;
    switch (kind) {
        case PPL_WORK_ANY_COLLISION: return ppl_flops_any_collision(nr, np, ns, n_obbs);
        case PPL_WORK_HINGE_STATE: return ppl_flops_hinge_state(nr, np, ns, n_obbs, pairs);
        case PPL_WORK_CHOMP_ADJOINT:
            return n_waypoints >= 2 ? ppl_flops_chomp_adjoint_traj(nr, np, ns, n_obbs, n_waypoints) : -1.0;
        case PPL_WORK_CHOMP_FD: return n_waypoints >= 2 ? ppl_flops_chomp_fd_traj(nr, np, ns, n_obbs, n_waypoints, D) : -1.0;
        case PPL_WORK_STOMP_STATE: return ppl_flops_stomp_state(nr, np, ns, n_obbs, pairs);
        case PPL_WORK_ESDF_WAYPOINT: return ppl_flops_esdf_waypoint(nr, np, ns);
        default: return -1.0;
    }
}

}  // extern "C"

namespace {

// LDS records needed: the largest per-problem OBB range.
int plan_obb_cap(const ppl_plan_s* P, const DScene& S) {
    if (P->h_obb_off.empty()) return S.n_obbs;
    int m = 0;
    for (size_t p = 0; p + 1 < P->h_obb_off.size(); ++p) m = std::max(m, P->h_obb_off[p + 1] - P->h_obb_off[p]);
    return m;
}

ppl_status plan_scene(const ppl_plan_s* P, const ppl_scene_view* view, DScene& S) {
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    if (P->h_obb_off.empty()) return PPL_OK;
    if (P->h_obb_off.back() > S.n_obbs)
        return fail(PPL_ERR_INVALID, "plan OBB offsets reach %d but the scene has %d OBBs", P->h_obb_off.back(), S.n_obbs);
    return PPL_OK;
}

// Self pass (min clearance + argmin) for cost plans, into the plan's scratch.
ppl_status cost_self(ppl_plan_s* P, const DStates& X, const ppl_cost_params* prm, hipStream_t s, const float** self_in) {
    *self_in = nullptr;
    if (!prm->use_self || P->robot->n_pairs == 0 || X.n_states == 0) return PPL_OK;
    HIP_TRY(P->d_self.ensure((size_t)X.n_states * 8 * sizeof(float)));
    ppl_status st = run_self(P->robot, X, 1, nullptr, P->d_self.as<float>(), s, prm->self_margin, nullptr, P);
    if (st != PPL_OK) return st;
    *self_in = P->d_self.as<float>();
    return PPL_OK;
}

// Whether lane group G (or the grouped layout) fits LDS for problems of up to
// max_obbs OBBs; only the grouped layout takes problems beyond one staging.
bool group_fits(const ppl_plan_s* P, int G, bool grad, int max_obbs) {
    const DRobot& R = P->robot->R;
    const bool tiled = max_obbs > kMaxObbLds;
    if (G == PPL_COST_GROUPED)
        return lds_bytes_grp(obb_tile(max_obbs), R.n_dof, R.n_joints, R.n_sph, R.n_clu, grad, tiled) <= 65536;
    return !tiled && lds_bytes_cost(max_obbs, R.n_dof, R.n_joints, G, grad) <= 65536;
}

// force_G > 0 overrides the plan's lane-group choice (autotuning).
ppl_status hinge_impl(ppl_plan_s* P, const ppl_scene_view* view, const float* d_x, const ppl_cost_params* prm,
                      float* d_state, float* d_total, float* d_grad, hipStream_t s, int force_G = 0) {
    CHECK_ARG(P && prm && d_x, "NULL argument");
    CHECK_ARG(P->kind != PPL_PLAN_EDGES, "hinge cost needs a configuration or trajectory plan");
    DScene S{};
    ppl_status st = plan_scene(P, view, S);
    if (st != PPL_OK) return st;
    if (P->n_states == 0) return PPL_OK;
    DStates X = plan_states(P, d_x);
    const float* self_in = nullptr;
    if ((st = cost_self(P, X, prm, s, &self_in)) != PPL_OK) return st;
    float* state = d_state;
    if (!state) {
        HIP_TRY(P->d_cost.ensure((size_t)P->n_states * 4));
        state = P->d_cost.as<float>();
    }
    CostArgs a{};
    a.R = P->robot->R;
    a.S = S;
    a.X = X;
    int grid = 0;
    // problems with more OBBs than one LDS staging take the grouped kernel (OBB tiles)
    const int max_obbs = plan_obb_cap(P, S);
    const int gi = d_grad ? 1 : 0;
    int tuned = P->tuned_group[gi];
    if (tuned == 0 && force_G == 0) {
        // the tuning cache's layout for this robot, OBB count, state count and GPU
        if (P->cost_obbs[gi] != max_obbs) {
            P->cost_obbs[gi] = max_obbs;
            P->cost_cached[gi] = 0;
            std::string v;
            int g = 0;
            if (tune_get(cost_key(P->robot, gi != 0, max_obbs, P->n_states), v) && sscanf(v.c_str(), "%d", &g) == 1 &&
                (g == PPL_COST_GROUPED || (g >= 1 && g <= kMaxGroup && (g & (g - 1)) == 0)) &&
                group_fits(P, g, gi != 0, max_obbs))
                P->cost_cached[gi] = g;
        }
        tuned = P->cost_cached[gi];
    }
    // default: the grouped layout for the adjoint and for ESDF-only costs
    // (its ESDF variant keeps five waves per SIMD on the grid lookups), the
    // lane-group heuristic for the cost alone over boxes
    const bool esdf_only = !(prm->use_obb && S.n_obbs > 0) && prm->use_esdf && S.n_voxel_layers > 0;
    const int G = max_obbs > kMaxObbLds ? PPL_COST_GROUPED
                : force_G != 0          ? force_G
                : tuned != 0            ? tuned
                : d_grad || esdf_only   ? PPL_COST_GROUPED
                                        : pick_group(P->n_states, P->robot->R, false);
    const bool grouped = G == PPL_COST_GROUPED;
    HIP_TRY(ensure_table(P, grouped ? kGrp : G));
    a.C = plan_chunks(P, grouped ? (d_grad ? kKCostGrpGrad : kKCostGrp) : (d_grad ? kKCostGrad : kKCost),
                      max_obbs, grouped ? kGrp : G, &grid);
    a.group = grouped ? kGrp : G;
    a.P = {prm->world_margin, prm->self_margin, prm->weight, prm->use_obb, prm->use_esdf, prm->use_self};
    a.grad = d_grad != nullptr;
    a.self_in = self_in;
    a.cost_out = state;
    a.grad_out = d_grad;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing_now(P, s)) HIP_TRY(timing_pair(P, &t0, &t1));
    HIP_TRY(grouped ? launch_cost_grp(a, grid, s, t0, t1) : launch_cost(a, grid, s, t0, t1));
    if (t0) ++P->kused;
    if (d_total && P->kind == PPL_PLAN_TRAJ) HIP_TRY(launch_segment_sum(state, P->d_off, P->n_units, 0.0f, d_total, s));
    return PPL_OK;
}

__global__ void k_fd_expand(const float* theta, const int* traj_off, int n_traj, long n_wp, int D, float eps, float* cfg) {
    // one thread per (waypoint, dim, sign): cfg[((w*D + d)*2 + sgn)*D + k]
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_wp * D * 2) return;
    const long w = i / (2 * D);
    const int d = (int)((i / 2) % D);
    const int sgn = (int)(i & 1);
    int lo = 0, hi = n_traj;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (traj_off[mid] <= w) lo = mid; else hi = mid;
    }
    const int t0 = traj_off[lo], Tn = traj_off[lo + 1] - t0, t = (int)(w - t0);
    const float* th = theta + (size_t)t0 * D;
    float* q = cfg + (size_t)i * D;
    for (int k = 0; k < D; ++k) q[k] = th[(size_t)k * Tn + t];
    // chomp.cu:108-133: q[d] += eps; then q[d] -= 2 eps (FP32).
    const float plus = q[d] + eps;
    q[d] = sgn == 0 ? plus : plus - 2.0f * eps;
}

__global__ void k_fd_combine(const float* cost, const int* traj_off, int n_traj, long n_wp, int D, float eps, float w,
                             float* grad) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_wp * D) return;
    const long wp = i / D;
    const int d = (int)(i % D);
    int lo = 0, hi = n_traj;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (traj_off[mid] <= wp) lo = mid; else hi = mid;
    }
    const int t0 = traj_off[lo], Tn = traj_off[lo + 1] - t0, t = (int)(wp - t0);
    const float cp = cost[2 * i], cm = cost[2 * i + 1];
    const bool end = t == 0 || t == Tn - 1;
    grad[(size_t)t0 * D + (size_t)d * Tn + t] = end ? 0.0f : w * (cp - cm) / (2.0f * eps);
}

ppl_status fd_impl(ppl_plan_s* P, const ppl_scene_view* view, const float* d_theta, const ppl_cost_params* prm, float eps,
                   float* d_grad, hipStream_t s) {
    CHECK_ARG(P && prm && d_theta && d_grad, "NULL argument");
    CHECK_ARG(P->kind == PPL_PLAN_TRAJ, "finite-difference gradient needs a trajectory plan");
    CHECK_ARG(P->h_obb_off.empty(), "finite-difference gradient supports a single scene");
    DScene S{};
    ppl_status st = plan_scene(P, view, S);
    if (st != PPL_OK) return st;
    const long n_wp = P->n_states;
    if (n_wp == 0) return PPL_OK;
    const int D = P->robot->R.n_dof;
    const long n_cfg = n_wp * D * 2;
    HIP_TRY(P->d_fdcfg.ensure((size_t)n_cfg * D * 4));
    HIP_TRY(P->d_fdcost.ensure((size_t)n_cfg * 4));
    {
        const long th = n_cfg;
        hipLaunchKernelGGL(k_fd_expand, dim3((unsigned)((th + 255) / 256)), dim3(256), 0, s, d_theta, P->d_off, P->n_units,
                           n_wp, D, eps, P->d_fdcfg.as<float>());
        HIP_TRY(hipGetLastError());
    }
    DStates X{};
    X.source = kSrcConfig;
    X.n_states = n_cfg;
    X.cfg = P->d_fdcfg.as<float>();
    X.cfg_stride = D;
    const float* self_in = nullptr;
    if (prm->use_self && P->robot->n_pairs > 0) {
        HIP_TRY(P->d_self.ensure((size_t)n_cfg * 8 * 4));
        if ((st = run_self(P->robot, X, 1, nullptr, P->d_self.as<float>(), s, prm->self_margin)) != PPL_OK) return st;
        self_in = P->d_self.as<float>();
    }
    CostArgs a{};
    a.R = P->robot->R;
    a.S = S;
    a.X = X;
    a.P = {prm->world_margin, prm->self_margin, 1.0f, prm->use_obb, prm->use_esdf, prm->use_self};
    a.self_in = self_in;
    a.cost_out = P->d_fdcost.as<float>();
    HIP_TRY(launch_cost_flat(a, n_cfg, s));
    {
        const long th = n_wp * D;
        hipLaunchKernelGGL(k_fd_combine, dim3((unsigned)((th + 255) / 256)), dim3(256), 0, s, P->d_fdcost.as<float>(),
                           P->d_off, P->n_units, n_wp, D, eps, prm->weight, d_grad);
        HIP_TRY(hipGetLastError());
    }
    return PPL_OK;
}

ppl_status stomp_impl(ppl_plan_s* P, const ppl_scene_view* view, const float* d_roll, const ppl_cost_params* prm, float w_u,
                      float dt, float* d_state, float* d_total, hipStream_t s) {
    CHECK_ARG(P && prm && d_roll && d_total, "NULL argument");
    CHECK_ARG(P->kind == PPL_PLAN_TRAJ, "rollout costs need a trajectory plan");
    float* state = d_state;
    if (!state) {
        HIP_TRY(P->d_cost.ensure((size_t)P->n_states * 4));
        state = P->d_cost.as<float>();
    }
    ppl_status st = hinge_impl(P, view, d_roll, prm, state, nullptr, nullptr, s);
    if (st != PPL_OK) return st;
    // rollout totals: the state costs' sum and the control cost in one launch
    HIP_TRY(launch_control_cost(d_roll, P->d_off, P->n_units, P->robot->R.n_dof, dt, w_u, d_total, 0, s, P->t_max, state));
    return PPL_OK;
}

// R + 1e-8 I in FP64 for the 5-point acceleration metric R = dt (A^T A)
// restricted to T x T from a grid padded by 6 (chomp_precompute.cpp:20-64,
// stomp_precompute.cpp:24-44), row-major.
void smooth_metric64(int T, float dt, std::vector<double>& R) {
    const int pad = 6, N = T + 2 * pad;
    const double c[5] = {-1.0 / 12.0, 16.0 / 12.0, -30.0 / 12.0, 16.0 / 12.0, -1.0 / 12.0};
    const double inv_dt2 = 1.0 / ((double)dt * (double)dt);
    // R[i][j] = dt * sum_r A[r][i + pad] A[r][j + pad], A[r][k] = inv_dt2 c[k - r + 2]
    R.assign((size_t)T * T, 0.0);
    for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) {
            double acc = 0.0;
            for (int r = 0; r < N; ++r) {
                const int a = i + pad - r + 2, b = j + pad - r + 2;
                if (a < 0 || a > 4 || b < 0 || b > 4) continue;
                acc += c[a] * c[b];
            }
            R[(size_t)i * T + j] = (double)dt * inv_dt2 * inv_dt2 * acc + (i == j ? 1e-8 : 0.0);
        }
}

// Lower Cholesky factor of a symmetric positive definite T x T matrix (FP64).
bool cholesky64(const std::vector<double>& A, int T, std::vector<double>& L) {
    L.assign((size_t)T * T, 0.0);
    for (int j = 0; j < T; ++j) {
        double d = A[(size_t)j * T + j];
        for (int k = 0; k < j; ++k) d -= L[(size_t)j * T + k] * L[(size_t)j * T + k];
        if (!(d > 0.0)) return false;
        const double ljj = std::sqrt(d);
        L[(size_t)j * T + j] = ljj;
        for (int i = j + 1; i < T; ++i) {
            double v = A[(size_t)i * T + j];
            for (int k = 0; k < j; ++k) v -= L[(size_t)i * T + k] * L[(size_t)j * T + k];
            L[(size_t)i * T + j] = v / ljj;
        }
    }
    return true;
}

// Inverse of a symmetric positive definite matrix: Cholesky, then two
// triangular solves per column (FP64).
bool spd_inverse64(const std::vector<double>& A, int T, std::vector<double>& inv) {
    std::vector<double> L;
    if (!cholesky64(A, T, L)) return false;
    inv.assign((size_t)T * T, 0.0);
    std::vector<double> y(T), x(T);
    for (int col = 0; col < T; ++col) {
        for (int i = 0; i < T; ++i) {
            double v = i == col ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) v -= L[(size_t)i * T + k] * y[k];
            y[i] = v / L[(size_t)i * T + i];
        }
        for (int i = T - 1; i >= 0; --i) {
            double v = y[i];
            for (int k = i + 1; k < T; ++k) v -= L[(size_t)k * T + i] * x[k];
            x[i] = v / L[(size_t)i * T + i];
        }
        for (int i = 0; i < T; ++i) inv[(size_t)i * T + col] = x[i];
    }
    return true;
}

// (R + 1e-8 I)^-1, stored FP32 row-major (chomp_precompute.cpp:40-64).
bool chomp_ainv(int T, float dt, std::vector<float>& out) {
    std::vector<double> R, inv;
    smooth_metric64(T, dt, R);
    if (!spd_inverse64(R, T, inv)) return false;
    out.assign(inv.begin(), inv.end());
    return true;
}

// STOMP's sampling factor and projection (stomp_precompute.cpp:48-101), out =
// [L^T | M^T] (2 T^2 floats, entry k * T + t of each = L[t][k] / M[t][k]):
// L = chol(S) * noise_scale, S = R^-1 scaled so its largest entry is 1/T and
// symmetrised; M = R^-1 with column j scaled so its largest entry is 1/T.
bool stomp_mats(int T, float dt, float noise_scale, std::vector<float>& out) {
    std::vector<double> R, inv, S, L;
    smooth_metric64(T, dt, R);
    if (!spd_inverse64(R, T, inv)) return false;
    double mx = inv[0];
    for (double v : inv) mx = std::max(mx, v);
    const double sc = (1.0 / T) / std::max(mx, 1e-12);
    S.resize(inv.size());
    for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j)
            S[(size_t)i * T + j] = 0.5 * (inv[(size_t)i * T + j] * sc + inv[(size_t)j * T + i] * sc);
    if (!cholesky64(S, T, L)) return false;
    const double ns = noise_scale > 0.0f ? (double)noise_scale : 1.0;
    out.assign(2 * (size_t)T * T, 0.0f);
    float* Lt = out.data();
    float* Mt = Lt + (size_t)T * T;
    for (int t = 0; t < T; ++t)
        for (int k = 0; k <= t; ++k) Lt[(size_t)k * T + t] = (float)(L[(size_t)t * T + k] * ns);
    for (int j = 0; j < T; ++j) {
        double cm = inv[j];
        for (int i = 1; i < T; ++i) cm = std::max(cm, inv[(size_t)i * T + j]);
        const double cs = (1.0 / T) / std::max(cm, 1e-12);
        for (int i = 0; i < T; ++i) Mt[(size_t)j * T + i] = (float)(inv[(size_t)i * T + j] * cs);
    }
    return true;
}

// A^-1 of every trajectory length of the plan for step dt, and the
// per-trajectory pointer table the update kernel reads.
ppl_status ensure_ainv(ppl_plan_s* P, float dt) {
    if (P->d_ainv_ptr && P->ainv_dt == dt) return PPL_OK;
    for (auto& e : P->ainv) (void)hipFree(e.second);
    P->ainv.clear();
    std::vector<const float*> ptr(P->n_units);
    for (int u = 0; u < P->n_units; ++u) {
        const int T = P->h_off[u + 1] - P->h_off[u];
        float* d = nullptr;
        for (auto& e : P->ainv)
            if (e.first == T) d = e.second;
        if (!d) {
            std::vector<float> h;
            if (!chomp_ainv(T, dt, h))
                return fail(PPL_ERR_INVALID, "smoothness metric of length %d is not positive definite", T);
            HIP_TRY(hipMalloc((void**)&d, h.size() * sizeof(float)));
            P->ainv.emplace_back(T, d);
            HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        ptr[u] = d;
    }
    if (!P->d_ainv_ptr) HIP_TRY(hipMalloc((void**)&P->d_ainv_ptr, (size_t)P->n_units * sizeof(float*)));
    HIP_TRY(hipMemcpy(P->d_ainv_ptr, ptr.data(), (size_t)P->n_units * sizeof(float*), hipMemcpyHostToDevice));
    P->ainv_dt = dt;
    return PPL_OK;
}

// Every trajectory has >= 2 waypoints and at most 4096; *max_t the longest.
ppl_status traj_lengths(const ppl_plan_s* P, int* max_t) {
    *max_t = 0;
    for (int u = 0; u < P->n_units; ++u) {
        const int T = P->h_off[u + 1] - P->h_off[u];
        CHECK_ARG(T >= 2, "trajectory %d has fewer than 2 waypoints", u);
        *max_t = std::max(*max_t, T);
    }
    CHECK_ARG(*max_t <= 4096, "trajectories longer than 4096 waypoints are not supported");
    return PPL_OK;
}

// Densification cap: no segment of a trajectory inside the joint limits is
// longer than the limit box's diagonal (times `slack`).
int nsub_cap(const ppl_robot_s* r, float step, double slack) {
    if (!(step > 0.0f)) return 1;
    double diag2 = 0.0;
    for (int k = 0; k < r->R.n_dof; ++k) diag2 += (double)(r->upper[k] - r->lower[k]) * (r->upper[k] - r->lower[k]);
    return std::max(1, (int)std::ceil(slack * std::sqrt(diag2) / step));
}

// Densified validity of every trajectory of a plan (chomp.cu:297-351,
// 459-481; stomp.cu:462-560, 662-686): every waypoint plus nsub - 1
// interpolants per segment, nsub = ceil(max segment / step) from the current
// theta on the device; environment first, then the self check on the
// trajectories still free (the order of two_phase_cc,
// collision/two_phase_cc.cuh:89-131). first_hit[b] = -1 when valid.
struct DenseCheck {
    CollideArgs va{};
    DStates Xd{};
    int vgrid = 0;
    int* first_hit = nullptr;
    int* hint = nullptr;        // unit holding each 64-state block (built on the device per evaluation)
    int n_hint = 0;
    long n_traj = 0;
};

// Upper bound of the densified state count, and the hint entries it needs.
size_t dense_bound(const ppl_plan_s* P, int nsub_max) {
    size_t n_dense = 0;
    for (int u = 0; u < P->n_units; ++u) n_dense += (size_t)(P->h_off[u + 1] - P->h_off[u] - 1) * nsub_max + 1;
    return n_dense;
}

ppl_status dense_check_setup(const ppl_plan_s* P, const DScene& S, float step, float margin, int nsub_max,
                             int* first_hit, int* nsub, int* dense_off, int* hint, ChompArgs& a, DenseCheck& dc) {
    const size_t n_dense = dense_bound(P, nsub_max);
    CHECK_ARG(n_dense < (size_t)INT32_MAX, "densified validity states exceed 2^31");
    a.check_step = step;
    a.margin = margin;
    a.nsub_max = nsub_max;
    a.first_hit = first_hit;
    a.nsub = nsub;
    a.dense_off = dense_off;
    dc.Xd = a.X;
    dc.Xd.source = kSrcDense;
    dc.Xd.dense_off = dense_off;
    dc.Xd.nsub = nsub;
    dc.Xd.n_dyn = dense_off + P->n_units;
    dc.Xd.n_states = (long)n_dense;
    dc.va.R = P->robot->R;
    dc.va.S = S;
    dc.va.X = dc.Xd;
    dc.va.C = implicit_chunks((long)n_dense, kKCollide, P->robot->R, S.n_obbs, &dc.vgrid);
    dc.va.margin = margin;
    dc.va.mode = 1;
    dc.va.skip_resolved = 1;
    dc.va.first_hit = first_hit;
    dc.va.self_hit = nullptr;
    dc.va.C.unit_hint = hint;
    dc.first_hit = first_hit;
    dc.hint = hint;
    dc.n_hint = (int)(n_dense / 64 + 1);
    dc.n_traj = P->n_units;
    return PPL_OK;
}

ppl_status dense_check_run(const ppl_robot_s* r, const ChompArgs& a, DenseCheck& dc, hipStream_t s) {
    HIP_TRY(launch_chomp_nsub(a, s));
    HIP_TRY(launch_unit_hint(a.dense_off, (int)dc.n_traj, dc.hint, dc.n_hint, s));
    HIP_TRY(launch_fill_i32(dc.first_hit, -1, dc.n_traj, s));
    HIP_TRY(launch_collide(dc.va, dc.vgrid, s));
    if (r->n_pairs > 0) return run_self(r, dc.Xd, 0, nullptr, nullptr, s, FLT_MAX, dc.first_hit, nullptr, dc.hint);
    return PPL_OK;
}

// Rollout plan over the fresh rows: nf rollouts of each trajectory, b-major,
// cut waypoint-major inside each trajectory's rollout set.
ppl_status stomp_rollout_plan(ppl_plan_s* P, int nf, ppl_plan* out) {
    const int B = P->n_units;
    std::vector<int> len((size_t)B * nf), pu(B + 1);
    for (int b = 0; b < B; ++b) {
        pu[b] = b * nf;
        for (int k = 0; k < nf; ++k) len[(size_t)b * nf + k] = P->h_off[b + 1] - P->h_off[b];
    }
    pu[B] = B * nf;
    ppl_plan_desc d{};
    d.kind = PPL_PLAN_TRAJ;
    d.n_units = B * nf;
    d.unit_states = len.data();
    d.n_problems = B;
    d.problem_units = pu.data();
    d.waypoint_major = 1;
    return ppl_plan_create(P->robot, &d, out);
}

// STOMP state of a plan for these settings: rollout plans and buffers for
// the rollout counts, matrices for (dt, noise_scale), reused when unchanged.
ppl_status ensure_stomp(ppl_plan_s* P, const ppl_stomp_settings* ss) {
    if (!P->stomp) {
        P->stomp = new (std::nothrow) StompState;
        CHECK_ARG(P->stomp, "out of host memory");
    }
    StompState* z = P->stomp;
    const int B = P->n_units, D = P->robot->R.n_dof;
    const size_t n = (size_t)P->n_states;
    const int nn = ss->num_rollouts_new, no = ss->num_rollouts_old, n_roll = nn + no + 1;
    if (!z->fresh_plan || z->n_new != nn || z->n_old != no) {
        ppl_plan_destroy(z->first_plan);
        ppl_plan_destroy(z->fresh_plan);
        z->first_plan = z->fresh_plan = nullptr;
        z->started = 0;
        ppl_status st = stomp_rollout_plan(P, nn + no, &z->first_plan);
        if (st != PPL_OK) return st;
        if ((st = stomp_rollout_plan(P, nn, &z->fresh_plan)) != PPL_OK) return st;
        const size_t nf = (size_t)(nn + no);
        const size_t floats = nf * n * D + nf * n + 2 * ((size_t)no * n * D + (size_t)no * n) + n * D + 2 * n +
                              (size_t)B * n_roll * D + B;
        const size_t ints = (size_t)B * no + 3 * (size_t)B + 1;
        HIP_TRY(z->buf.ensure((floats + ints) * sizeof(float) + 64));
        float* f = z->buf.as<float>();
        z->fresh = f; f += nf * n * D;
        z->fresh_cost = f; f += nf * n;
        for (int k = 0; k < 2; ++k) {
            z->keep[k] = f; f += (size_t)no * n * D;
            z->keep_cost[k] = f; f += (size_t)no * n;
        }
        z->slot = f; f += n * D;
        z->slot_cost = f; f += n;
        z->theta_cost = f; f += n;
        z->ctrl_sum = f; f += (size_t)B * n_roll * D;
        z->totals = f; f += B;
        int* i = reinterpret_cast<int*>(f);
        z->sel = i; i += (size_t)B * no;
        z->first_hit = i; i += B;
        z->nsub = i; i += B;
        z->dense_off = i;
        z->n_new = nn;
        z->n_old = no;
    }
    if (!z->d_lt || z->dt != ss->dt || z->noise_scale != ss->noise_scale) {
        for (auto& e : z->mats) (void)hipFree(e.second);
        z->mats.clear();
        std::vector<const float*> ptr(2 * (size_t)B);
        for (int b = 0; b < B; ++b) {
            const int T = P->h_off[b + 1] - P->h_off[b];
            float* d = nullptr;
            for (auto& e : z->mats)
                if (e.first == T) d = e.second;
            if (!d) {
                std::vector<float> h;
                if (!stomp_mats(T, ss->dt, ss->noise_scale, h))
                    return fail(PPL_ERR_INVALID, "smoothness metric of length %d is not positive definite", T);
                HIP_TRY(hipMalloc((void**)&d, h.size() * sizeof(float)));
                z->mats.emplace_back(T, d);
                HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
            }
            ptr[b] = d;
            ptr[(size_t)B + b] = d + (size_t)T * T;
        }
        if (!z->d_lt) HIP_TRY(hipMalloc((void**)&z->d_lt, 2 * (size_t)B * sizeof(float*)));
        z->d_mt = z->d_lt + B;
        HIP_TRY(hipMemcpy((void*)z->d_lt, ptr.data(), 2 * (size_t)B * sizeof(float*), hipMemcpyHostToDevice));
        z->dt = ss->dt;
        z->noise_scale = ss->noise_scale;
        z->started = 0;
    }
    return PPL_OK;
}

// Capture the launches `body` issues on a private stream as one graph.
template <typename F>
ppl_status capture(F&& body, ppl_graph* out) {
    CHECK_ARG(out, "NULL argument");
    *out = nullptr;
    hipStream_t cs = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    ppl_graph_s* g = new (std::nothrow) ppl_graph_s();
    if (!g) {
        (void)hipStreamDestroy(cs);
        return fail(PPL_ERR_ALLOC, "out of host memory");
    }
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    ppl_status st = PPL_OK;
    if (e == hipSuccess) {
        st = body(cs);
        hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
        if (st == PPL_OK && e2 != hipSuccess) st = fail(PPL_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
    } else {
        st = fail(PPL_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    if (st == PPL_OK) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) st = fail(PPL_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    (void)hipStreamDestroy(cs);
    if (st != PPL_OK) {
        ppl_graph_destroy(g);
        return st;
    }
    *out = g;
    return PPL_OK;
}

// Key of a captured optimiser iteration: every pointer and value its
// launches take (FNV-1a over the bytes).
struct KeyHash {
    unsigned long long h = 1469598103934665603ull;
    template <typename T> KeyHash& add(const T& v) {
        const unsigned char* b = reinterpret_cast<const unsigned char*>(&v);
        for (size_t i = 0; i < sizeof(T); ++i) h = (h ^ b[i]) * 1099511628211ull;
        return *this;
    }
};

unsigned long long plan_buffers_key(const ppl_plan_s* P, KeyHash k) {
    k.add(P->d_opt.p).add(P->d_self.p).add(P->d_cost.p).add(P->d_fdcfg.p).add(P->d_fdcost.p).add(P->tuned_group);
    for (int i = 0; i < kGroupKinds; ++i) k.add(P->d_tab[i]);
    return k.h;
}

}  // namespace

extern "C" {

ppl_status ppl_chomp_optimize(ppl_plan P, const ppl_scene_view* view, const ppl_chomp_settings* cs, float* d_theta,
                              float* d_best, float* d_record, int iterations, int first_call, ppl_stream stream) try {
    CHECK_ARG(P && view && cs && d_theta && d_best && d_record, "NULL argument");
    CHECK_ARG(P->kind == PPL_PLAN_TRAJ && !P->wmajor, "CHOMP needs a (unit-major) trajectory plan");
    CHECK_ARG(P->h_obb_off.empty(), "CHOMP optimises trajectories of one scene (no per-problem OBB ranges)");
    CHECK_ARG(iterations >= 0, "negative iteration count");
    CHECK_ARG(cs->dt > 0.0f, "dt must be positive");
    if (P->n_units == 0) return PPL_OK;
    int max_t = 0;
    ppl_status st = traj_lengths(P, &max_t);
    if (st != PPL_OK) return st;
    if ((st = ensure_ainv(P, cs->dt)) != PPL_OK) return st;
    DScene S{};
    if ((st = plan_scene(P, view, S)) != PPL_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const ppl_robot_s* r = P->robot;
    const int D = r->R.n_dof;
    const size_t n = (size_t)P->n_states, B = (size_t)P->n_units;
    const int nsub_max = nsub_cap(r, cs->collision_check_step, 1.0);
    // scratch: gradient [n D], state cost [n], totals [B], first hit [B], nsub [B], dense offsets [B + 1],
    // block hints [n_dense / 64 + 1]
    const size_t n_hint = dense_bound(P, nsub_max) / 64 + 1;
    HIP_TRY(P->d_opt.ensure((n * D + n + B) * sizeof(float) + (3 * B + 1 + n_hint) * sizeof(int)));
    float* grad = P->d_opt.as<float>();
    float* state = grad + n * D;
    float* total = state + n;
    int* first_hit = reinterpret_cast<int*>(total + B);
    ppl_cost_params cp{cs->world_margin, cs->self_margin, cs->collision_cost_weight, 1, 1, cs->use_self};
    ChompArgs a{};
    a.R = P->robot->R;
    a.X = plan_states(P, d_theta);
    a.step_size = cs->step_size;
    a.smoothness_weight = cs->smoothness_weight;
    a.convergence_eps = cs->convergence_eps;
    a.dt = cs->dt;
    a.polish0 = cs->iterations_after_valid;
    a.ainv = P->d_ainv_ptr;
    a.theta = d_theta;
    a.best = d_best;
    a.record = d_record;
    a.grad = grad;
    a.total = total;
    DenseCheck dc{};
    if ((st = dense_check_setup(P, S, cs->collision_check_step, cs->collision_margin, nsub_max, first_hit,
                                first_hit + B, first_hit + 2 * B, first_hit + 3 * B + 1, a, dc)) != PPL_OK)
        return st;
    // cost (+ gradient) of the current theta, its totals and its validity
    auto evaluate = [&](hipStream_t q) -> ppl_status {
        ppl_status e = cs->fd_gradient ? fd_impl(P, view, d_theta, &cp, cs->fd_eps, grad, q) : PPL_OK;
        if (e != PPL_OK) return e;
        e = hinge_impl(P, view, d_theta, &cp, state, total, cs->fd_gradient ? nullptr : grad, q);
        if (e != PPL_OK) return e;
        HIP_TRY(launch_control_cost(d_theta, P->d_off, P->n_units, D, cs->dt, cs->smoothness_weight, total, 1, q,
                                    P->t_max));
        return dense_check_run(r, a, dc, q);
    };
    auto iteration = [&](hipStream_t q) -> ppl_status {
        HIP_TRY(launch_chomp_update(a, max_t, q));
        ppl_status e = evaluate(q);
        if (e != PPL_OK) return e;
        HIP_TRY(launch_chomp_book(a, 0, q));
        return PPL_OK;
    };
    // the gradient at the current theta feeds the first step of this call
    // (this evaluation also sizes every buffer an iteration touches)
    if (first_call || iterations > 0) {
        if ((st = evaluate(s)) != PPL_OK) return st;
    }
    if (first_call) HIP_TRY(launch_chomp_book(a, 1, s));
    if (P->opt_graphs && iterations >= 2) {
        // one iteration captured once, replayed: no host work between iterations
        KeyHash k;
        k.add(d_theta).add(d_best).add(d_record).add(*view).add(*cs).add(max_t);
        const unsigned long long key = plan_buffers_key(P, k);
        if (!P->chomp_graph || P->chomp_key != key) {
            ppl_graph_destroy(P->chomp_graph);
            P->chomp_graph = nullptr;
            if ((st = capture(iteration, &P->chomp_graph)) != PPL_OK) return st;
            P->chomp_key = key;
        }
        for (int it = 0; it < iterations; ++it) HIP_TRY(hipGraphLaunch(P->chomp_graph->exec, s));
        return PPL_OK;
    }
    for (int it = 0; it < iterations; ++it)
        if ((st = iteration(s)) != PPL_OK) return st;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_stomp_optimize(ppl_plan P, const ppl_scene_view* view, const ppl_stomp_settings* ss, float* d_theta,
                              float* d_best, float* d_record, int iterations, int first_call, ppl_stream stream) try {
    CHECK_ARG(P && view && ss && d_theta && d_best && d_record, "NULL argument");
    CHECK_ARG(P->kind == PPL_PLAN_TRAJ && !P->wmajor, "STOMP needs a (unit-major) trajectory plan");
    CHECK_ARG(P->h_obb_off.empty(), "STOMP optimises trajectories of one scene (no per-problem OBB ranges)");
    CHECK_ARG(iterations >= 0, "negative iteration count");
    CHECK_ARG(ss->dt > 0.0f, "dt must be positive");
    CHECK_ARG(ss->num_rollouts_new >= 1 && ss->num_rollouts_old >= 0, "need >= 1 new and >= 0 kept rollouts");
    const int n_roll = ss->num_rollouts_new + ss->num_rollouts_old + 1;
    CHECK_ARG(n_roll <= 1024, "at most 1024 rollouts per trajectory (new + old + 1 = %d)", n_roll);
    CHECK_ARG(P->n_units <= 65535, "at most 65535 trajectories per STOMP plan");
    if (P->n_units == 0) return PPL_OK;
    int max_t = 0;
    ppl_status st = traj_lengths(P, &max_t);
    if (st != PPL_OK) return st;
    if (stomp_update_lds(n_roll, max_t) > 65536)
        return fail(PPL_ERR_CAPACITY, "(rollouts + 1) x longest trajectory = %d x %d exceeds 16384", n_roll + 1, max_t);
    StompState* z = P->stomp;
    if (!first_call) {
        CHECK_ARG(z && z->started, "the first STOMP call on a plan must have first_call = 1");
        CHECK_ARG(z->n_new == ss->num_rollouts_new && z->n_old == ss->num_rollouts_old && z->dt == ss->dt &&
                      z->noise_scale == ss->noise_scale,
                  "rollout counts, dt and noise_scale cannot change between calls of one run");
    }
    if ((st = ensure_stomp(P, ss)) != PPL_OK) return st;
    z = P->stomp;
    DScene S{};
    if ((st = plan_scene(P, view, S)) != PPL_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const ppl_robot_s* r = P->robot;
    const int D = r->R.n_dof;
    const int B = P->n_units;
    ppl_cost_params cp{ss->world_margin, ss->self_margin, ss->collision_cost_weight, 1, 1, ss->use_self};
    // theta may leave the joint-limit box (the update is not clamped): twice the diagonal caps a segment
    const int nsub_max = nsub_cap(r, ss->collision_check_step, 2.0);
    ChompArgs a{};
    a.R = r->R;
    a.X = plan_states(P, d_theta);
    a.convergence_eps = ss->convergence_eps;
    a.dt = ss->dt;
    a.polish0 = ss->iterations_after_valid;
    a.theta = d_theta;
    a.best = d_best;
    a.record = d_record;
    a.total = z->totals;
    DenseCheck dc{};
    HIP_TRY(z->hint.ensure((dense_bound(P, nsub_max) / 64 + 1) * sizeof(int)));
    if ((st = dense_check_setup(P, S, ss->collision_check_step, ss->collision_margin, nsub_max, z->first_hit, z->nsub,
                                z->dense_off, z->hint.as<int>(), a, dc)) != PPL_OK)
        return st;
    auto evaluate = [&](hipStream_t q) -> ppl_status {
        ppl_status e = hinge_impl(P, view, d_theta, &cp, z->theta_cost, z->totals, nullptr, q);
        if (e != PPL_OK) return e;
        HIP_TRY(launch_control_cost(d_theta, P->d_off, B, D, ss->dt, ss->control_cost_weight, z->totals, 1, q,
                                    P->t_max));
        return dense_check_run(r, a, dc, q);
    };
    if (first_call) {
        z->iter = 0;
        z->parity = 0;
        z->nf_prev = 0;
        z->started = 1;
        if ((st = evaluate(s)) != PPL_OK) return st;
        HIP_TRY(launch_chomp_book(a, 1, s));
    }
    StompArgs g{};
    g.R = r->R;
    g.B = B;
    g.traj_off = P->d_off;
    g.n_new = ss->num_rollouts_new;
    g.n_old = ss->num_rollouts_old;
    g.n_roll = n_roll;
    g.w_u = ss->control_cost_weight;
    g.h = ss->exponentiated_cost_sensitivity;
    g.dt = ss->dt;
    g.key0 = (unsigned)(ss->seed & 0xffffffffull);
    g.key1 = (unsigned)(ss->seed >> 32);
    g.Lt = z->d_lt;
    g.Mt = z->d_mt;
    g.theta = d_theta;
    g.theta_cost = z->theta_cost;
    g.record = d_record;
    g.fresh = z->fresh;
    g.slot = z->slot;
    g.slot_cost = z->slot_cost;
    g.ctrl_sum = z->ctrl_sum;
    g.sel = z->sel;
    // One iteration on stream q. iter_dev: the sampling counter's iteration
    // word comes from device memory plus g.iter (graph replays), else g.iter.
    auto iteration = [&](hipStream_t q, unsigned iter_word, int parity, int nf_prev, const unsigned* iter_dev) -> ppl_status {
        const bool first = iter_dev == nullptr && z->iter == 0;
        g.iter = iter_word;
        g.iter_dev = iter_dev;
        g.keep = z->keep[parity];
        g.keep_cost = z->keep_cost[parity];
        g.cur = StompPop{z->fresh, z->fresh_cost, first ? g.n_new + g.n_old : g.n_new, g.keep, g.keep_cost,
                         first ? 0 : g.n_old, z->slot, z->slot_cost};
        g.prev = StompPop{z->fresh, z->fresh_cost, nf_prev, z->keep[parity ^ 1], z->keep_cost[parity ^ 1],
                          nf_prev == g.n_new + g.n_old ? 0 : g.n_old, z->slot, z->slot_cost};
        if (!first) HIP_TRY(launch_stomp_reuse(g, max_t, q));
        HIP_TRY(launch_stomp_sample(g, max_t, q));
        ppl_status e = hinge_impl(first ? z->first_plan : z->fresh_plan, view, z->fresh, &cp, z->fresh_cost, nullptr,
                                  nullptr, q);
        if (e != PPL_OK) return e;
        HIP_TRY(launch_stomp_update(g, max_t, q));
        if ((e = evaluate(q)) != PPL_OK) return e;
        HIP_TRY(launch_chomp_book(a, 0, q));
        return PPL_OK;
    };
    auto eager = [&]() -> ppl_status {
        ppl_status e = iteration(s, (unsigned)z->iter, z->parity, z->nf_prev, nullptr);
        if (e != PPL_OK) return e;
        z->nf_prev = g.cur.nf;
        z->parity ^= 1;
        ++z->iter;
        return PPL_OK;
    };
    int it = 0;
    // eagerly until the population is in its steady shape: the first
    // iteration samples new + old rows, the second reuses none of them
    while (it < iterations && (z->iter == 0 || z->nf_prev != g.n_new)) {
        if ((st = eager()) != PPL_OK) return st;
        ++it;
    }
    if (P->opt_graphs && iterations - it >= 4) {
        // two iterations (one per keep-buffer parity) captured once and
        // replayed; the sampling counter's iteration word lives on the device
        if (!P->d_iter) HIP_TRY(hipMalloc((void**)&P->d_iter, sizeof(unsigned)));
        KeyHash k;
        k.add(d_theta).add(d_best).add(d_record).add(*view).add(*ss).add(max_t).add(z->parity).add(z->nf_prev)
            .add(z->buf.p).add(z->hint.p).add(z->first_plan).add(z->fresh_plan).add(z->d_lt);
        const unsigned long long key = plan_buffers_key(z->fresh_plan, KeyHash{}.add(plan_buffers_key(P, k)));
        if (!P->stomp_graph || P->stomp_key != key) {
            ppl_graph_destroy(P->stomp_graph);
            P->stomp_graph = nullptr;
            const int p0 = z->parity, nf = z->nf_prev;
            st = capture([&](hipStream_t q) -> ppl_status {
                ppl_status e = iteration(q, 0u, p0, nf, P->d_iter);
                if (e != PPL_OK) return e;
                if ((e = iteration(q, 1u, p0 ^ 1, g.n_new, P->d_iter)) != PPL_OK) return e;
                HIP_TRY(launch_add_u32(P->d_iter, 2u, q));
                return PPL_OK;
            }, &P->stomp_graph);
            if (st != PPL_OK) return st;
            P->stomp_key = key;
        }
        HIP_TRY(launch_fill_i32(reinterpret_cast<int*>(P->d_iter), z->iter, 1, s));
        for (; it + 2 <= iterations; it += 2) {
            HIP_TRY(hipGraphLaunch(P->stomp_graph->exec, s));
            z->nf_prev = g.n_new;
            z->iter += 2;
        }
    }
    for (; it < iterations; ++it)
        if ((st = eager()) != PPL_OK) return st;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_fk_batched(ppl_plan P, const float* d_cfg, float* d_frames, float* d_spheres, ppl_stream stream) try {
    CHECK_ARG(P && d_cfg, "NULL argument");
    CHECK_ARG(P->kind == PPL_PLAN_CONFIGS, "configuration plan required");
    if (P->n_states == 0 || (!d_frames && !d_spheres)) return PPL_OK;
    HIP_TRY(launch_fk_out(P->robot->R, plan_states(P, d_cfg), d_frames, d_spheres, (hipStream_t)stream));
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_check_collision_batched(ppl_plan P, const ppl_scene_view* view, const float* d_cfg, float margin,
                                       int check_self, uint8_t* d_out, ppl_stream stream) try {
    CHECK_ARG(P && d_cfg && d_out, "NULL argument");
    CHECK_ARG(P->kind == PPL_PLAN_CONFIGS, "configuration plan required");
    DScene S{};
    ppl_status st = plan_scene(P, view, S);
    if (st != PPL_OK) return st;
    if (P->n_states == 0) return PPL_OK;
    hipStream_t s = (hipStream_t)stream;
    DStates X = plan_states(P, d_cfg);
    const uint8_t* self_hit = nullptr;
    if (check_self && P->robot->n_pairs > 0) {
        HIP_TRY(P->d_self.ensure((size_t)P->n_states));
        if ((st = run_self(P->robot, X, 0, P->d_self.as<uint8_t>(), nullptr, s, FLT_MAX, nullptr, P)) != PPL_OK)
            return st;
        self_hit = P->d_self.as<uint8_t>();
    }
    CollideArgs a{};
    a.R = P->robot->R;
    a.S = S;
    a.X = X;
    int grid = 0;
    a.C = plan_collide_chunks(P, plan_obb_cap(P, S), &grid);
    a.margin = margin;
    a.mode = 0;
    a.out = d_out;
    a.self_hit = self_hit;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing_now(P, s)) HIP_TRY(timing_pair(P, &t0, &t1));
    HIP_TRY(launch_collide(a, grid, s, t0, t1));
    if (t0) ++P->kused;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_edge_check_batched(ppl_plan P, const ppl_scene_view* view, const float* d_from, const float* d_to,
                                  const ppl_edge_batch_params* prm, int* d_first_hit, uint8_t* d_state_hits,
                                  ppl_stream stream) try {
    CHECK_ARG(P && prm && d_from && d_to && d_first_hit, "NULL argument");
    CHECK_ARG(P->kind == PPL_PLAN_EDGES, "edge plan required");
    CHECK_ARG(!(d_state_hits && prm->skip_resolved), "per-state verdicts need skip_resolved = 0");
    DScene S{};
    ppl_status st = plan_scene(P, view, S);
    if (st != PPL_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(launch_fill_i32(d_first_hit, -1, P->n_units, s));
    if (P->n_states == 0) return PPL_OK;
    DStates X = plan_states(P, nullptr);
    X.from = d_from;
    X.to = d_to;
    const uint8_t* self_hit = nullptr;
    if (prm->check_self && P->robot->n_pairs > 0) {
        HIP_TRY(P->d_self.ensure((size_t)P->n_states));
        if ((st = run_self(P->robot, X, 0, P->d_self.as<uint8_t>(), nullptr, s, FLT_MAX, nullptr, P)) != PPL_OK)
            return st;
        self_hit = P->d_self.as<uint8_t>();
    }
    CollideArgs a{};
    a.R = P->robot->R;
    a.S = S;
    a.X = X;
    int grid = 0;
    a.C = plan_collide_chunks(P, plan_obb_cap(P, S), &grid);
    a.margin = prm->collision_margin;
    a.mode = 1;
    a.skip_resolved = prm->skip_resolved;
    a.out = d_state_hits;
    a.first_hit = d_first_hit;
    a.self_hit = self_hit;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing_now(P, s)) HIP_TRY(timing_pair(P, &t0, &t1));
    HIP_TRY(launch_collide(a, grid, s, t0, t1));
    if (t0) ++P->kused;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_min_clearance_batched(ppl_plan P, const ppl_scene_view* view, const float* d_cfg, float margin, float* d_out,
                                     ppl_stream stream) try {
    CHECK_ARG(P && d_cfg && d_out, "NULL argument");
    CHECK_ARG(P->kind == PPL_PLAN_CONFIGS, "configuration plan required");
    DScene S{};
    ppl_status st = plan_scene(P, view, S);
    if (st != PPL_OK) return st;
    if (P->n_states == 0) return PPL_OK;
    CostArgs a{};
    a.R = P->robot->R;
    a.S = S;
    a.X = plan_states(P, d_cfg);
    int grid = 0;
    const int max_obbs = plan_obb_cap(P, S);
    const bool tiled = max_obbs > kMaxObbLds;   // the grouped kernel streams the OBBs in tiles
    HIP_TRY(ensure_table(P, tiled ? kGrp : 1));
    a.C = plan_chunks(P, tiled ? kKCostGrp : kKCost, max_obbs, tiled ? kGrp : 1, &grid);
    a.group = tiled ? kGrp : 1;
    a.P.weight = 1.0f;
    a.clearance_only = 1;
    a.margin = margin;
    a.cost_out = d_out;
    HIP_TRY(tiled ? launch_cost_grp(a, grid, (hipStream_t)stream) : launch_cost(a, grid, (hipStream_t)stream));
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

}  // extern "C"

namespace {
// Min ESDF clearance (+ margin) of n configurations d_cfg [n][D] into d_out,
// scene_min_clearance_runtime semantics (the clearance mode of the cost kernel).
ppl_status configs_clearance(const ppl_robot_s* r, const DScene& S, const float* d_cfg, int n, float margin,
                             float* d_out, hipStream_t s) {
    if (n <= 0) return PPL_OK;
    CostArgs c{};
    c.R = r->R;
    c.S = S;
    c.X.source = kSrcConfig;
    c.X.n_states = n;
    c.X.cfg = d_cfg;
    c.X.cfg_stride = r->R.n_dof;
    c.P.weight = 1.0f;
    c.clearance_only = 1;
    c.margin = margin;
    c.cost_out = d_out;
    HIP_TRY(launch_cost_flat(c, n, s));
    return PPL_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

extern "C" {

// Stage n host configurations into the robot's device buffer (synchronous
// plain forms); returns the device copy.
namespace {
ppl_status stage_configs(ppl_robot_s* r, const float* h_cfg, int n, float** d_cfg) {
    const size_t bytes = (size_t)n * r->R.n_dof * sizeof(float);
    HIP_TRY(r->d_cfg.ensure(bytes));
    HIP_TRY(r->h_stage.ensure(std::max(bytes, (size_t)n * sizeof(float))));
    memcpy(r->h_stage.p, h_cfg, bytes);
    HIP_TRY(hipMemcpyAsync(r->d_cfg.p, r->h_stage.p, bytes, hipMemcpyHostToDevice, r->stream));
    *d_cfg = r->d_cfg.as<float>();
    return PPL_OK;
}

ppl_status fetch_floats(ppl_robot_s* r, const float* d_out, int n, float* h_out) {
    HIP_TRY(hipMemcpyAsync(r->h_stage.p, d_out, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    memcpy(h_out, r->h_stage.p, (size_t)n * sizeof(float));
    return PPL_OK;
}
}  // namespace

ppl_status ppl_min_clearance_scene(ppl_robot r, const ppl_scene_view* view, const float* h_configs, int n,
                                   float margin, float* h_clearance) try {
    CHECK_ARG(r, "robot is NULL");
    CHECK_ARG(n >= 0, "negative N");
    if (n == 0) return PPL_OK;
    CHECK_ARG(h_configs && h_clearance, "NULL host arrays");
    DScene S{};
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    float* d_cfg = nullptr;
    if ((st = stage_configs(r, h_configs, n, &d_cfg)) != PPL_OK) return st;
    HIP_TRY(r->d_out.ensure((size_t)n * sizeof(float)));
    if ((st = configs_clearance(r, S, d_cfg, n, margin, r->d_out.as<float>(), r->stream)) != PPL_OK) return st;
    return fetch_floats(r, r->d_out.as<float>(), n, h_clearance);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_hinge_cost_scene(ppl_robot r, const ppl_scene_view* view, const float* h_configs, int n,
                                const ppl_cost_params* prm, float* h_cost) try {
    CHECK_ARG(r && prm, "NULL argument");
    CHECK_ARG(n >= 0, "negative N");
    if (n == 0) return PPL_OK;
    CHECK_ARG(h_configs && h_cost, "NULL host arrays");
    DScene S{};
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    float* d_cfg = nullptr;
    if ((st = stage_configs(r, h_configs, n, &d_cfg)) != PPL_OK) return st;
    CostArgs c{};
    c.R = r->R;
    c.S = S;
    c.X.source = kSrcConfig;
    c.X.n_states = n;
    c.X.cfg = d_cfg;
    c.X.cfg_stride = r->R.n_dof;
    if (prm->use_self && r->n_pairs > 0) {
        HIP_TRY(r->d_self.ensure((size_t)n * 8 * sizeof(float)));
        if ((st = run_self(r, c.X, 1, nullptr, r->d_self.as<float>(), r->stream, prm->self_margin)) != PPL_OK) return st;
        c.self_in = r->d_self.as<float>();
    }
    HIP_TRY(r->d_out.ensure((size_t)n * sizeof(float)));
    c.P = {prm->world_margin, prm->self_margin, prm->weight, prm->use_obb, prm->use_esdf, prm->use_self};
    c.cost_out = r->d_out.as<float>();
    HIP_TRY(launch_cost_flat(c, n, r->stream));
    return fetch_floats(r, r->d_out.as<float>(), n, h_cost);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

}  // extern "C"

// The allocation-free form's scratch: self verdicts then clearances for up to
// PPL_BATCH_CC_STATIC_MAX samples, one array per GPU, held until the process
// ends (1.25 MiB). Host code never passes `scratch` itself to a kernel (on the
// host the name is a stand-in, not device memory); it uses the device address
// hipGetSymbolAddress returns on each device, kept in g_scratch_dev below.
constexpr size_t kStaticSelfBytes = ((size_t)PPL_BATCH_CC_STATIC_MAX + 255) & ~(size_t)255;
constexpr size_t kStaticScratchBytes = kStaticSelfBytes + (size_t)PPL_BATCH_CC_STATIC_MAX * sizeof(float);
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {
// batch_cc over caller-provided scratch: self verdicts [n] bytes at `self`,
// clearances [n] floats at `clr`. Every entry below n is written before it is
// read, so the scratch needs no clearing. Allocates nothing on the device.
ppl_status batch_cc_impl(ppl_robot r, const DScene& S, const float* d_samples, int n, float margin, uint8_t* d_cc,
                         float* d_clearance, float clearance_weight, uint8_t* self, float* clr, hipStream_t s) {
    ppl_status st = PPL_OK;
    CollideArgs a{};
    a.R = r->R;
    a.S = S;
    a.X.source = kSrcConfig;
    a.X.n_states = n;
    a.X.cfg = d_samples;
    a.X.cfg_stride = r->R.n_dof;
    if (r->n_pairs > 0) {
        if ((st = run_self(r, a.X, 0, self, nullptr, s)) != PPL_OK) return st;
        a.self_hit = self;
    }
    int grid = 0;
    a.C = implicit_chunks(n, kKCollide, r->R, S.n_obbs, &grid);
    a.margin = margin;
    a.mode = 0;
    a.out = d_cc;
    HIP_TRY(launch_collide(a, grid, s));
    if (d_clearance) {
        if (clearance_weight > 0.0f && (st = configs_clearance(r, S, d_samples, n, margin, clr, s)) != PPL_OK) return st;
        HIP_TRY(launch_sample_clearance(d_cc, clr, n, clearance_weight, d_clearance, s));
    }
    return PPL_OK;
}

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};

ppl_status static_scratch(uint8_t** out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) return fail(PPL_ERR_UNSUPPORTED, "device index %d >= %d", dev, kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PPL_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = static_cast<uint8_t*>(g_scratch_dev[dev]);
    return PPL_OK;
}
}  // namespace

extern "C" {

ppl_status ppl_mitstar_batch_cc(ppl_robot r, const ppl_scene_view* view, const float* d_samples, int n, float margin,
                                uint8_t* d_cc, float* d_clearance, float clearance_weight, ppl_stream stream) try {
    CHECK_ARG(r && d_samples && d_cc, "NULL argument");
    CHECK_ARG(n >= 0, "negative sample count");
    if (n == 0) return PPL_OK;
    DScene S{};
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    // scratch: self verdicts [n] bytes, clearances [n] floats
    HIP_TRY(r->d_mit.ensure(align256((size_t)n) + (size_t)n * sizeof(float)));
    uint8_t* self = r->d_mit.as<uint8_t>();
    float* clr = reinterpret_cast<float*>(self + align256((size_t)n));
    return batch_cc_impl(r, S, d_samples, n, margin, d_cc, d_clearance, clearance_weight, self, clr,
                         (hipStream_t)stream);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_mitstar_batch_cc_static(ppl_robot r, const ppl_scene_view* view, const float* d_samples, int n,
                                       float margin, uint8_t* d_cc, float* d_clearance, float clearance_weight,
                                       ppl_stream stream) try {
    CHECK_ARG(r && d_samples && d_cc, "NULL argument");
    CHECK_ARG(n >= 0, "negative sample count");
    if (n > PPL_BATCH_CC_STATIC_MAX)
        return fail(PPL_ERR_CAPACITY, "batch_cc_static: %d samples exceed the fixed scratch of %d", n,
                    PPL_BATCH_CC_STATIC_MAX);
    if (n == 0) return PPL_OK;
    DScene S{};
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    uint8_t* base = nullptr;
    if ((st = static_scratch(&base)) != PPL_OK) return st;
    return batch_cc_impl(r, S, d_samples, n, margin, d_cc, d_clearance, clearance_weight, base,
                         reinterpret_cast<float*>(base + kStaticSelfBytes), (hipStream_t)stream);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_mitstar_node_clearance(ppl_robot r, const ppl_scene_view* view, const float* d_nodes,
                                      const int* d_node_indices, int node_start, int n_nodes, float margin,
                                      float* d_clearance, ppl_stream stream) try {
    CHECK_ARG(r && d_nodes && d_clearance, "NULL argument");
    CHECK_ARG(n_nodes >= 0 && (d_node_indices || node_start >= 0), "bad node range");
    if (n_nodes == 0) return PPL_OK;
    DScene S{};
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const int D = r->R.n_dof;
    const size_t cfg_bytes = align256((size_t)n_nodes * D * sizeof(float));
    HIP_TRY(r->d_mit.ensure(cfg_bytes + (size_t)n_nodes * sizeof(float)));
    float* cfg = r->d_mit.as<float>();
    float* clr = reinterpret_cast<float*>(r->d_mit.as<char>() + cfg_bytes);
    HIP_TRY(launch_gather_rows(d_nodes, D, d_node_indices, node_start, n_nodes, cfg, s));
    if ((st = configs_clearance(r, S, cfg, n_nodes, margin, clr, s)) != PPL_OK) return st;
    HIP_TRY(launch_scatter(clr, d_node_indices, node_start, n_nodes, d_clearance, s));
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_mitstar_edge_cc(ppl_robot r, const ppl_scene_view* view, const float* d_nodes, int dim, const int* d_src,
                               const int* d_dst, const int* d_checks_done, int n_edges, float valid_segment_length,
                               int num_sparse_checks, float margin, uint8_t* d_cc, int* d_new_checks,
                               ppl_stream stream) try {
    CHECK_ARG(r && d_nodes && d_src && d_dst && d_cc, "NULL argument");
    CHECK_ARG(dim == r->R.n_dof, "dim %d != robot n_dof %d", dim, r->R.n_dof);
    CHECK_ARG(n_edges >= 0, "negative edge count");
    CHECK_ARG(valid_segment_length > 0.0f, "valid_segment_length must be positive");
    if (n_edges == 0) return PPL_OK;
    DScene S{};
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const size_t E = (size_t)n_edges;
    // scratch: first midpoint [E], count [E], offsets [E + 1], first hit [E], 64-bit total
    HIP_TRY(r->d_mit.ensure(align256((4 * E + 1) * sizeof(int)) + sizeof(long long)));
    int* k0 = r->d_mit.as<int>();
    int* cnt = k0 + E;
    int* off = cnt + E;
    int* first_hit = off + E + 1;
    long long* d_total = reinterpret_cast<long long*>(r->d_mit.as<char>() + align256((4 * E + 1) * sizeof(int)));
    HIP_TRY(r->h_mit.ensure(sizeof(long long)));
    HIP_TRY(launch_bfs_count(d_nodes, dim, d_src, d_dst, d_checks_done, n_edges, valid_segment_length,
                             num_sparse_checks, k0, cnt, off, d_total, s));
    HIP_TRY(hipMemcpyAsync(r->h_mit.p, d_total, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(launch_fill_i32(first_hit, -1, n_edges, s));
    HIP_TRY(hipStreamSynchronize(s));
    const long long total = *r->h_mit.as<long long>();
    if (total >= (long long)INT32_MAX)
        return fail(PPL_ERR_CAPACITY, "%lld edge midpoints exceed 2^31 in one call", total);
    if (total > 0) {
        HIP_TRY(r->d_hint.ensure((size_t)(total / 64 + 1) * sizeof(int)));
        int* hint = r->d_hint.as<int>();
        HIP_TRY(launch_unit_hint(off, n_edges, hint, (int)(total / 64 + 1), s));
        DStates X{};
        X.source = kSrcBfs;
        X.n_states = (long)total;
        X.cfg = d_nodes;
        X.cfg_stride = dim;
        X.node_src = d_src;
        X.node_dst = d_dst;
        X.bfs_k0 = k0;
        X.dense_off = off;
        X.n_traj = n_edges;
        CollideArgs a{};
        a.R = r->R;
        a.S = S;
        a.X = X;
        int grid = 0;
        a.C = implicit_chunks((long)total, kKCollide, r->R, S.n_obbs, &grid);
        a.C.unit_hint = hint;
        a.margin = margin;
        a.mode = 1;
        a.skip_resolved = 1;
        a.first_hit = first_hit;
        HIP_TRY(launch_collide(a, grid, s));
        // self after the environment, on the edges still free (two_phase_cc's order)
        if (r->n_pairs > 0 &&
            (st = run_self(r, X, 0, nullptr, nullptr, s, FLT_MAX, first_hit, nullptr, hint)) != PPL_OK)
            return st;
    }
    HIP_TRY(launch_bfs_finish(first_hit, cnt, n_edges, d_cc, d_new_checks, s));
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_plan_autotune_cost(ppl_plan P, const ppl_scene_view* view, const float* d_x, const ppl_cost_params* prm,
                                  int grad, int reps, int* best_group, float* best_ms, ppl_stream stream) try {
    CHECK_ARG(P && view && d_x && prm, "NULL argument");
    CHECK_ARG(P->kind != PPL_PLAN_EDGES, "cost autotuning needs a configuration or trajectory plan");
    CHECK_ARG(reps >= 1 && reps <= 1000, "reps must be in [1, 1000]");
    if (P->n_states == 0) return PPL_OK;
    DScene S{};
    ppl_status st = plan_scene(P, view, S);
    if (st != PPL_OK) return st;
    {
        // a choice already in the tuning cache is adopted without timing
        std::string v;
        int g = 0;
        float ms = 0.0f;
        const int cap0 = plan_obb_cap(P, S);
        if (tune_get(cost_key(P->robot, grad != 0, cap0, P->n_states), v) && sscanf(v.c_str(), "%d %f", &g, &ms) == 2 &&
            (g == PPL_COST_GROUPED || (g >= 1 && g <= kMaxGroup && (g & (g - 1)) == 0)) &&
            group_fits(P, g, grad != 0, cap0)) {
            P->tuned_group[grad ? 1 : 0] = g;
            if (best_group) *best_group = g;
            if (best_ms) *best_ms = ms;
            return PPL_OK;
        }
    }
    const int D = P->robot->R.n_dof;
    const size_t n = (size_t)P->n_states;
    HIP_TRY(P->d_tune.ensure((n * (grad ? D + 1 : 1) + (size_t)P->n_units) * sizeof(float)));
    float* state = P->d_tune.as<float>();
    float* total = state + n;
    float* g = grad ? total + P->n_units : nullptr;
    if (P->kind != PPL_PLAN_TRAJ) total = nullptr;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) { (void)hipEventDestroy(e0); return fail(PPL_ERR_HIP, "hipEventCreate failed"); }
    int best = 0;
    float best_t = 0.0f;
    const int cap = plan_obb_cap(P, S);
    // candidates: lane groups G = 1 .. 16 of k_cost, then the grouped layout
    for (int G = 1; G <= 2 * kMaxGroup && st == PPL_OK; G <<= 1) {
        if (G > kMaxGroup) G = PPL_COST_GROUPED;
        if (!group_fits(P, G, grad != 0, cap)) {
            if (G == PPL_COST_GROUPED) break;
            continue;
        }
        st = hinge_impl(P, view, d_x, prm, state, total, g, s, G);   // warm-up (builds the chunk table)
        he = hipEventRecord(e0, s);
        for (int k = 0; k < reps && st == PPL_OK && he == hipSuccess; ++k)
            st = hinge_impl(P, view, d_x, prm, state, total, g, s, G);
        if (he == hipSuccess) he = hipEventRecord(e1, s);
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess) st = fail(PPL_ERR_HIP, "autotune timing: %s", hipGetErrorString(he));
        if (st == PPL_OK && (best == 0 || ms < best_t)) { best = G; best_t = ms; }
        if (G == PPL_COST_GROUPED) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != PPL_OK) return st;
    P->tuned_group[grad ? 1 : 0] = best;
    {
        char v[64];
        snprintf(v, sizeof v, "%d %.6f", best, best_t / reps);
        tune_put(cost_key(P->robot, grad != 0, cap, P->n_states), v);
    }
    if (best_group) *best_group = best;
    if (best_ms) *best_ms = best_t / reps;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_tune_cache_file(const char* path) try {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_path = path ? path : "";
    g_tune_path_set = true;
    g_tune.clear();
    g_tune_loaded = false;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_plan_autotune_collide(ppl_plan P, const ppl_scene_view* view, const float* d_x, const float* d_x2,
                                     int reps, int* best_tile, int* best_grid, float* best_ms, ppl_stream stream) try {
    CHECK_ARG(P && view && d_x, "NULL argument");
    CHECK_ARG(P->kind == PPL_PLAN_CONFIGS || P->kind == PPL_PLAN_EDGES, "any-collision tuning needs a configuration or edge plan");
    CHECK_ARG(P->kind != PPL_PLAN_EDGES || d_x2, "edge plans take the edges' start (d_x) and end (d_x2) configurations");
    CHECK_ARG(reps >= 1 && reps <= 1000, "reps must be in [1, 1000]");
    if (P->n_states == 0) return PPL_OK;
    DScene S{};
    ppl_status st = plan_scene(P, view, S);
    if (st != PPL_OK) return st;
    const int max_obbs = plan_obb_cap(P, S);
    const bool edges = P->kind == PPL_PLAN_EDGES;
    {
        // a choice already in the tuning cache is adopted without timing
        std::string v;
        int t = 0, g = 0;
        float ms = 0.0f;
        if (tune_get(collide_key(P->robot, max_obbs), v) && sscanf(v.c_str(), "%d %d %f", &t, &g, &ms) == 3 && t >= 0 &&
            t <= kMaxObbLds && g >= 0 && g <= 2) {
            P->ctile = t;
            P->cgrid = g;
            P->shape_src = 2;
            P->shape_obbs = max_obbs;
            if (best_tile) *best_tile = t > 0 ? std::min(t, obb_tile(max_obbs)) : obb_tile(max_obbs);
            if (best_grid) *best_grid = g;
            if (best_ms) *best_ms = ms;
            return PPL_OK;
        }
    }
    HIP_TRY(P->d_tune.ensure(edges ? (size_t)P->n_units * 4 : (size_t)P->n_states));
    hipStream_t s = (hipStream_t)stream;
    ppl_edge_batch_params ep{};
    ep.skip_resolved = 1;
    auto run = [&]() {
        return edges ? ppl_edge_check_batched(P, view, d_x, d_x2, &ep, P->d_tune.as<int>(), nullptr, stream)
                     : ppl_check_collision_batched(P, view, d_x, 0.0f, 0, P->d_tune.as<uint8_t>(), stream);
    };
    // candidates: the whole problem in one staging, then tiles of 128, 64 and
    // 32 records below it; each with the three grid forms
    std::vector<int> tiles{obb_tile(max_obbs)};
    for (int t : {128, 64, 32})
        if (t < tiles[0]) tiles.push_back(t);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) { (void)hipEventDestroy(e0); return fail(PPL_ERR_HIP, "hipEventCreate failed"); }
    const int keep_tile = P->ctile, keep_grid = P->cgrid, keep_src = P->shape_src, keep_obbs = P->shape_obbs;
    int bt = -1, bg = 0;
    float bms = 0.0f;
    for (int t : tiles)
        for (int g = 0; g <= 2 && st == PPL_OK && he == hipSuccess; ++g) {
            P->ctile = t;
            P->cgrid = g;
            P->shape_src = 2;
            st = run();   // warm-up
            if (st == PPL_OK) he = hipEventRecord(e0, s);
            for (int k = 0; k < reps && st == PPL_OK && he == hipSuccess; ++k) st = run();
            if (st == PPL_OK && he == hipSuccess) he = hipEventRecord(e1, s);
            if (st == PPL_OK && he == hipSuccess) he = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (st == PPL_OK && he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
            if (st == PPL_OK && he == hipSuccess && (bt < 0 || ms < bms)) { bt = t; bg = g; bms = ms; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st == PPL_OK && he != hipSuccess) st = fail(PPL_ERR_HIP, "collide autotune timing: %s", hipGetErrorString(he));
    if (st != PPL_OK) {
        P->ctile = keep_tile; P->cgrid = keep_grid; P->shape_src = keep_src; P->shape_obbs = keep_obbs;
        return st;
    }
    // the whole-problem tile is stored as 0 (the built-in choice for any OBB count)
    P->ctile = bt == tiles[0] ? 0 : bt;
    P->cgrid = bg;
    P->shape_src = 2;
    P->shape_obbs = max_obbs;
    char v[64];
    snprintf(v, sizeof v, "%d %d %.6f", P->ctile, bg, bms / reps);
    tune_put(collide_key(P->robot, max_obbs), v);
    if (best_tile) *best_tile = bt;
    if (best_grid) *best_grid = bg;
    if (best_ms) *best_ms = bms / reps;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_plan_collide_shape(ppl_plan P, int max_obbs, int* tile, int* grid, int* source) try {
    CHECK_ARG(P, "NULL argument");
    CHECK_ARG(max_obbs >= 0, "negative OBB count");
    int t = 0, g = 0;
    collide_shape(P, max_obbs, &t, &g);
    if (tile) *tile = t;
    if (grid) *grid = g;
    if (source) *source = P->shape_src;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_plan_set_collide_shape(ppl_plan P, int tile, int grid) try {
    CHECK_ARG(P, "NULL argument");
    CHECK_ARG(tile >= 0 && tile <= kMaxObbLds, "tile must be in [0, %d]", kMaxObbLds);
    CHECK_ARG(grid >= 0 && grid <= 2, "grid must be 0, 1 or 2");
    P->ctile = tile;
    P->cgrid = grid;
    P->shape_src = tile == 0 && grid == 0 ? 0 : 2;
    P->shape_obbs = -1;   // built-in again: the tuning cache is consulted at the next launch
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_plan_set_optimizer_graphs(ppl_plan P, int enable) try {
    CHECK_ARG(P, "NULL argument");
    P->opt_graphs = enable ? 1 : 0;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_plan_set_cost_group(ppl_plan P, int grad, int lanes_per_state) try {
    CHECK_ARG(P, "NULL argument");
    CHECK_ARG(lanes_per_state == 0 || lanes_per_state == PPL_COST_GROUPED ||
                  (lanes_per_state >= 1 && lanes_per_state <= kMaxGroup && (lanes_per_state & (lanes_per_state - 1)) == 0),
              "lanes_per_state must be 0, PPL_COST_GROUPED or a power of two in [1, %d]", kMaxGroup);
    if (lanes_per_state != 0) {
        const int cap = P->h_obb_off.empty() ? kMaxObbLds : 0;
        int need = cap;
        for (size_t p = 0; p + 1 < P->h_obb_off.size(); ++p) need = std::max(need, P->h_obb_off[p + 1] - P->h_obb_off[p]);
        // problems beyond one OBB staging take the grouped layout whatever is set here
        CHECK_ARG(group_fits(P, lanes_per_state, grad != 0, std::min(need, kMaxObbLds)), "tile does not fit in LDS");
    }
    P->tuned_group[grad ? 1 : 0] = lanes_per_state;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_hinge_cost_batched(ppl_plan P, const ppl_scene_view* view, const float* d_x, const ppl_cost_params* prm,
                                  float* d_state, float* d_total, float* d_grad, ppl_stream stream) try {
    return hinge_impl(P, view, d_x, prm, d_state, d_total, d_grad, (hipStream_t)stream);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_chomp_fd_grad_batched(ppl_plan P, const ppl_scene_view* view, const float* d_theta,
                                     const ppl_cost_params* prm, float eps, float* d_grad, ppl_stream stream) try {
    return fd_impl(P, view, d_theta, prm, eps, d_grad, (hipStream_t)stream);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_stomp_rollout_costs(ppl_plan P, const ppl_scene_view* view, const float* d_roll, const ppl_cost_params* prm,
                                   float w_u, float dt, float* d_state, float* d_total, ppl_stream stream) try {
    return stomp_impl(P, view, d_roll, prm, w_u, dt, d_state, d_total, (hipStream_t)stream);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

// Single-trajectory plain forms share one cached plan per (robot, T).
}  // extern "C"

namespace {
// Single-trajectory plan of length T, created once per robot and T and kept
// with the robot (the plain per-trajectory forms then allocate nothing).
ppl_status single_traj_plan(ppl_robot r, int T, ppl_plan* out) {
    for (auto& e : r->traj_plans)
        if (e.first == T) { *out = e.second; return PPL_OK; }
    ppl_plan_desc d{};
    d.kind = PPL_PLAN_TRAJ;
    d.n_units = 1;
    d.unit_states = &T;
    d.n_problems = 1;
    ppl_status st = ppl_plan_create(r, &d, out);
    if (st == PPL_OK) r->traj_plans.emplace_back(T, *out);
    return st;
}

__global__ void k_broadcast_rows(float* x, int D, int T) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const float v = x[t];
    for (int d = 1; d < D; ++d) x[(size_t)d * T + t] = v;
}
}  // namespace

extern "C" {

ppl_status ppl_chomp_state_cost(ppl_robot r, const ppl_scene_view* view, const float* d_theta, int T,
                                const ppl_cost_params* prm, float* d_state_costs, ppl_stream stream) try {
    CHECK_ARG(r && prm && d_theta && d_state_costs, "NULL argument");
    CHECK_ARG(T >= 1, "T must be >= 1");
    ppl_plan P = nullptr;
    ppl_status st = single_traj_plan(r, T, &P);
    if (st != PPL_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    st = hinge_impl(P, view, d_theta, prm, d_state_costs, nullptr, nullptr, s);
    if (st == PPL_OK) {
        hipLaunchKernelGGL(k_broadcast_rows, dim3((T + 63) / 64), dim3(64), 0, s, d_state_costs, r->R.n_dof, T);
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(PPL_ERR_HIP, "chomp_state_cost: %s", hipGetErrorString(e));
    }
    return st;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_chomp_fd_grad(ppl_robot r, const ppl_scene_view* view, const float* d_theta, int T,
                             const ppl_cost_params* prm, float eps, float* d_grad, ppl_stream stream) try {
    CHECK_ARG(r && prm && d_theta && d_grad, "NULL argument");
    CHECK_ARG(T >= 1, "T must be >= 1");
    ppl_plan P = nullptr;
    ppl_status st = single_traj_plan(r, T, &P);
    if (st != PPL_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    st = fd_impl(P, view, d_theta, prm, eps, d_grad, s);
    if (st == PPL_OK) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(PPL_ERR_HIP, "chomp_fd_grad: %s", hipGetErrorString(e));
    }
    return st;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

}  // extern "C"

struct ppl_opt_buffers_s {
    int n_dof = 0, T = 0;
    DBuf dev;    // theta [D T], best [D T], record [8], state / control totals [2]
    HBuf host;   // record [8], totals [2], path [D T]
    float* theta() const { return dev.as<float>(); }
    float* best() const { return theta() + (size_t)n_dof * T; }
    float* record() const { return best() + (size_t)n_dof * T; }
    float* totals() const { return record() + 8; }
    float* h_record() const { return host.as<float>(); }
    float* h_totals() const { return h_record() + 8; }
    float* h_path() const { return h_totals() + 2; }
};

namespace {
ppl_status opt_buffers_alloc(int n_dof, int T, ppl_opt_buffers_s* b) {
    b->n_dof = n_dof;
    b->T = T;
    const size_t floats = 2 * (size_t)n_dof * T + 10;
    HIP_TRY(b->dev.ensure(floats * sizeof(float)));
    HIP_TRY(b->host.ensure((10 + (size_t)n_dof * T) * sizeof(float)));
    return PPL_OK;
}

ppl_status opt_buffers_create(int n_dof, int T, ppl_opt_buffers_s** out) {
    CHECK_ARG(out, "NULL argument");
    *out = nullptr;
    CHECK_ARG(n_dof >= 1 && n_dof <= kMaxDof, "n_dof must be in [1, %d]", kMaxDof);
    CHECK_ARG(T >= 2 && T <= 4096, "num_timesteps must be in [2, 4096]");
    auto* b = new (std::nothrow) ppl_opt_buffers_s;
    CHECK_ARG(b, "out of host memory");
    const ppl_status st = opt_buffers_alloc(n_dof, T, b);
    if (st != PPL_OK) {
        b->dev.release();
        b->host.release();
        delete b;
        return st;
    }
    *out = b;
    return PPL_OK;
}

// Single-goal solve shared by CHOMP and STOMP (chomp.cu:358-610,
// stomp.cu:565-930): clamped straight line, iterations of `optimize` in
// chunks until the record is done (one by one under a time limit, checked
// before each iteration), then the best valid trajectory or the working one
// with its costs, endpoints set to start and goal.
template <typename Optimize>
ppl_status solve_single(ppl_robot r, ppl_opt_buffers_s* bufs, const ppl_scene_view* view, const float* h_start,
                        const float* h_goal, int T, int num_iterations, float time_limit_ms,
                        const ppl_cost_params& cp, float ctrl_weight, float dt, Optimize&& optimize, float* h_path,
                        ppl_opt_result* res, hipStream_t s) {
    const auto wall0 = std::chrono::steady_clock::now();
    CHECK_ARG(r && view && h_start && h_goal && h_path && res, "NULL argument");
    const int D = r->R.n_dof;
    CHECK_ARG(T >= 2 && T <= 4096, "num_timesteps must be in [2, 4096]");
    CHECK_ARG(num_iterations >= 0, "negative iteration count");
    CHECK_ARG(!bufs || (bufs->n_dof == D && bufs->T == T),
              "buffers were created for n_dof %d, T %d; the solve needs n_dof %d, T %d", bufs ? bufs->n_dof : 0,
              bufs ? bufs->T : 0, D, T);
    ppl_opt_buffers_s local;
    ppl_opt_buffers_s* b = bufs;
    ppl_status st = PPL_OK;
    if (!b) {
        if ((st = opt_buffers_alloc(D, T, &local)) != PPL_OK) {
            local.dev.release();
            local.host.release();
            return st;
        }
        b = &local;
    }
    // the temporary buffers of a call without `bufs` live until the call returns
    struct Release {
        ppl_opt_buffers_s* b;
        ~Release() {
            if (b) {
                b->dev.release();
                b->host.release();
            }
        }
    } release{bufs ? nullptr : &local};
    ppl_plan P = nullptr;
    if ((st = single_traj_plan(r, T, &P)) != PPL_OK) return st;
    // chomp.cu:57-69 / stomp.cu:58-70: joint-space line start -> goal, clamped
    float* h_theta = b->h_path();
    for (int d = 0; d < D; ++d)
        for (int t = 0; t < T; ++t) {
            const float frac = (float)t / (float)(T - 1);
            const float v = fmaf(h_goal[d], frac, h_start[d] * (1.0f - frac));
            h_theta[(size_t)d * T + t] = std::max(r->lower[d], std::min(r->upper[d], v));
        }
    HIP_TRY(hipMemcpyAsync(b->theta(), h_theta, (size_t)D * T * sizeof(float), hipMemcpyHostToDevice, s));
    // The device stops updating a finished trajectory by itself; the chunks
    // only bound the launches queued after it finished.
    const int chunk = time_limit_ms > 0.0f ? 1 : 8;
    int done_iters = 0;
    for (bool first = true;; first = false) {
        const int k = std::min(chunk, num_iterations - done_iters);
        if (!first && k <= 0) break;
        if (!first && time_limit_ms > 0.0f) {
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
            if (ms > time_limit_ms) break;
        }
        if ((st = optimize(P, b->theta(), b->best(), b->record(), k, first ? 1 : 0, s)) != PPL_OK) return st;
        done_iters += k;
        HIP_TRY(hipMemcpyAsync(b->h_record(), b->record(), 8 * sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (b->h_record()[PPL_CHOMP_DONE] != 0.0f || done_iters >= num_iterations) break;
    }
    // chomp.cu:560-600, stomp.cu:889-925
    const bool solved = b->h_record()[PPL_CHOMP_SOLVED] != 0.0f;
    const float* d_res = solved ? b->best() : b->theta();
    if ((st = hinge_impl(P, view, d_res, &cp, nullptr, b->totals(), nullptr, s)) != PPL_OK) return st;
    HIP_TRY(launch_control_cost(d_res, P->d_off, 1, D, dt, ctrl_weight, b->totals() + 1, 0, s));
    HIP_TRY(hipMemcpyAsync(b->h_totals(), b->totals(), 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(b->h_path(), d_res, (size_t)D * T * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int t = 0; t < T; ++t)
        for (int d = 0; d < D; ++d) h_path[(size_t)t * D + d] = b->h_path()[(size_t)d * T + t];
    for (int d = 0; d < D; ++d) {
        h_path[d] = h_start[d];
        h_path[(size_t)(T - 1) * D + d] = h_goal[d];
    }
    res->solved = solved ? 1 : 0;
    res->iterations_run = (int)b->h_record()[PPL_CHOMP_ITERS];
    res->final_state_cost = b->h_totals()[0];
    res->final_control_cost = b->h_totals()[1];
    res->final_total_cost = solved ? b->h_record()[PPL_CHOMP_BEST_COST] : b->h_record()[PPL_CHOMP_LAST_COST];
    res->wall_time_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    return PPL_OK;
}
}  // namespace

extern "C" {

ppl_status ppl_chomp_buffers_create(int n_dof, int num_timesteps, ppl_chomp_buffers* out) try {
    return opt_buffers_create(n_dof, num_timesteps, out);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_stomp_buffers_create(int n_dof, int num_timesteps, int num_rollouts_new, int num_rollouts_old,
                                    ppl_stomp_buffers* out) try {
    CHECK_ARG(num_rollouts_new >= 1 && num_rollouts_old >= 0 && num_rollouts_new + num_rollouts_old + 1 <= 1024,
              "need 1 <= new, 0 <= old, new + old + 1 <= 1024 rollouts");
    return opt_buffers_create(n_dof, num_timesteps, out);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_chomp_buffers_destroy(ppl_chomp_buffers b) try {
    if (!b) return PPL_OK;
    b->dev.release();
    b->host.release();
    delete b;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_stomp_buffers_destroy(ppl_stomp_buffers b) { return ppl_chomp_buffers_destroy(b); }

ppl_status ppl_chomp_solve(ppl_robot r, ppl_chomp_buffers bufs, const ppl_scene_view* view, const float* h_start,
                           const float* h_goal, int num_timesteps, int num_iterations, float time_limit_ms,
                           const ppl_chomp_settings* cs, float* h_path, ppl_chomp_result* res, ppl_stream stream) try {
    CHECK_ARG(cs, "NULL argument");
    const ppl_cost_params cp{cs->world_margin, cs->self_margin, cs->collision_cost_weight, 1, 1, cs->use_self};
    auto opt = [&](ppl_plan P, float* th, float* best, float* rec, int k, int first, hipStream_t s) {
        return ppl_chomp_optimize(P, view, cs, th, best, rec, k, first, s);
    };
    return solve_single(r, bufs, view, h_start, h_goal, num_timesteps, num_iterations, time_limit_ms, cp,
                        cs->smoothness_weight, cs->dt, opt, h_path, res, (hipStream_t)stream);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_stomp_solve(ppl_robot r, ppl_stomp_buffers bufs, const ppl_scene_view* view, const float* h_start,
                           const float* h_goal, int num_timesteps, int num_iterations, float time_limit_ms,
                           const ppl_stomp_settings* ss, float* h_path, ppl_stomp_result* res, ppl_stream stream) try {
    CHECK_ARG(ss, "NULL argument");
    const ppl_cost_params cp{ss->world_margin, ss->self_margin, ss->collision_cost_weight, 1, 1, ss->use_self};
    auto opt = [&](ppl_plan P, float* th, float* best, float* rec, int k, int first, hipStream_t s) {
        return ppl_stomp_optimize(P, view, ss, th, best, rec, k, first, s);
    };
    return solve_single(r, bufs, view, h_start, h_goal, num_timesteps, num_iterations, time_limit_ms, cp,
                        ss->control_cost_weight, ss->dt, opt, h_path, res, (hipStream_t)stream);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_esdf_from_obbs(const ppl_scene_view* view, const float* h_params, const float* h_pose, float* d_features,
                              ppl_stream stream) try {
    CHECK_ARG(view && h_params && h_pose && d_features, "NULL argument");
    DScene S{};
    ppl_status st = scene_from_view(view, S);
    if (st != PPL_OK) return st;
    CHECK_ARG(h_params[3] > 0.0f, "voxel size must be positive");
    HIP_TRY(launch_esdf_build(S.obb_dims, S.obb_pose, S.n_obbs, h_params, h_pose, d_features, (hipStream_t)stream));
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

// ---------------------------------------------------------------------------
// hipGraph capture (on a private stream; replay on any stream)
// ---------------------------------------------------------------------------
}  // extern "C"


extern "C" {

ppl_status ppl_graph_capture_stomp(ppl_plan P, const ppl_scene_view* view, const float* d_roll, const ppl_cost_params* prm,
                                   float w_u, float dt, float* d_state, float* d_total, ppl_stream stream,
                                   ppl_graph* out) try {
    CHECK_ARG(P && prm, "NULL argument");
    // Allocate every scratch buffer before capture (no allocation inside a graph).
    ppl_status st = stomp_impl(P, view, d_roll, prm, w_u, dt, d_state, d_total, (hipStream_t)stream);
    if (st != PPL_OK) return st;
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return capture([&](hipStream_t cs) { return stomp_impl(P, view, d_roll, prm, w_u, dt, d_state, d_total, cs); }, out);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_graph_capture_hinge(ppl_plan P, const ppl_scene_view* view, const float* d_x, const ppl_cost_params* prm,
                                   float* d_state, float* d_total, float* d_grad, ppl_stream stream, ppl_graph* out) try {
    CHECK_ARG(P && prm, "NULL argument");
    ppl_status st = hinge_impl(P, view, d_x, prm, d_state, d_total, d_grad, (hipStream_t)stream);
    if (st != PPL_OK) return st;
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return capture([&](hipStream_t cs) { return hinge_impl(P, view, d_x, prm, d_state, d_total, d_grad, cs); }, out);
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_graph_launch(ppl_graph g, ppl_stream stream) try {
    CHECK_ARG(g && g->exec, "NULL graph");
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

ppl_status ppl_graph_destroy(ppl_graph g) try {
    if (!g) return PPL_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return PPL_OK;
} catch (const std::bad_alloc&) {
    return ppl::fail_msg(PPL_ERR_ALLOC, "out of host memory");
}

}  // extern "C"
