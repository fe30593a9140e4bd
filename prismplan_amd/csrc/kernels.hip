// kernels.hip -- gfx950 kernels of the collision-kinematics substrate.
//
// Execution model (wave64, CDNA4):
//  * one lane = one robot state (configuration, edge state or waypoint);
//    a 64-lane workgroup is one wavefront and runs persistently over a
//    static, XCD-aware sweep of 64-state chunks (chunks never span two
//    problems, so each wave sees one OBB set at a time);
//  * the problem's enabled OBBs are re-packed into LDS once per problem
//    switch; every lane then reads the same record (LDS broadcast);
//  * FK walks the DFS-ordered joint table with wave-uniform (scalar) loads,
//    keeping the live 3x4 transforms in VGPRs;
//  * the any-collision predicate gates each (joint, OBB) pair on the joint's
//    approximate bounding sphere with a wave ballot, tests the joint's exact
//    spheres in box-local coordinates through one composed 3x4 matrix, and
//    retires the wave as soon as every lane has a verdict;
//  * the summed-hinge cost keeps per-joint adjoint accumulators in registers
//    (joint loop unrolled to a compile-time bound) and differentiates through
//    the argmin OBB of each active sphere in the same pass.
#include "ppl_internal.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <mutex>

namespace ppl {
namespace {

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// ---------------------------------------------------------------------------
// 3x4 affine helpers (row-major, implicit last row 0 0 0 1)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void aff_identity(float* T) {
#pragma unroll
    for (int i = 0; i < 12; ++i) T[i] = 0.0f;
    T[0] = T[5] = T[10] = 1.0f;
}

// S = FT * M(q), restating xrot_fn / yrot_fn / zrot_fn / prism_fn
// (planning/utils.cuh:560-702). FIXED or unknown types move nothing.
__device__ __forceinline__ void joint_step(const float* __restrict__ ft, int type, float q, float* S) {
#pragma unroll
    for (int i = 0; i < 12; ++i) S[i] = ft[i];
    if (type >= kXRot) {
        float s, c;
        sincosf(q, &s, &c);
        const int a = (type == kXRot) ? 1 : (type == kYRot ? 2 : 0);
        const int b = (type == kXRot) ? 2 : (type == kYRot ? 0 : 1);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float fa = ft[r * 4 + a], fb = ft[r * 4 + b];
            S[r * 4 + a] = fmaf(fa, c, fb * s);
            S[r * 4 + b] = fmaf(fb, c, -fa * s);
        }
    } else if (type >= kXPrism) {
#pragma unroll
        for (int r = 0; r < 3; ++r) S[r * 4 + 3] = fmaf(ft[r * 4 + type], q, ft[r * 4 + 3]);
    }
}

// T = P * S
__device__ __forceinline__ void aff_compose(const float* P, const float* S, float* T) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float acc = fmaf(P[r * 4 + 2], S[8 + c], fmaf(P[r * 4 + 1], S[4 + c], P[r * 4 + 0] * S[c]));
            if (c == 3) acc += P[r * 4 + 3];
            T[r * 4 + c] = acc;
        }
    }
}

__device__ __forceinline__ void aff_point(const float* T, float x, float y, float z, float& px, float& py, float& pz) {
    px = fmaf(T[0], x, fmaf(T[1], y, fmaf(T[2], z, T[3])));
    py = fmaf(T[4], x, fmaf(T[5], y, fmaf(T[6], z, T[7])));
    pz = fmaf(T[8], x, fmaf(T[9], y, fmaf(T[10], z, T[11])));
}

// Transform slots (the reference's t_memory_idx scratch), kept in VGPRs. The
// slot index is wave-uniform, so the switch compiles to scalar branches.
struct Slots {
    float s[kMaxSlots][12];
    __device__ __forceinline__ void load(int i, float* T) const {
        if (i == 0) {
#pragma unroll
            for (int k = 0; k < 12; ++k) T[k] = s[0][k];
        } else {
#pragma unroll
            for (int k = 0; k < 12; ++k) T[k] = s[1][k];
        }
    }
    __device__ __forceinline__ void store(int i, const float* T) {
        if (i == 0) {
#pragma unroll
            for (int k = 0; k < 12; ++k) s[0][k] = T[k];
        } else {
#pragma unroll
            for (int k = 0; k < 12; ++k) s[1][k] = T[k];
        }
    }
};

// Transform of DFS position jj: identity at the root, parent * step otherwise.
// APPROX walks the approximate spheres' own chain (R.aft / R.atype, set when
// the model has one): same tree, dofs and slots, its own fixed transforms and
// joint types.
template <bool APPROX = false>
__device__ __forceinline__ void fk_joint(const DRobot& R, int jj, const float* s_q, Slots& sl, float* T, int4& meta) {
    meta = R.jmeta[jj];
    if (jj == 0) {
        aff_identity(T);
    } else {
        float P[12], S[12];
        sl.load(meta.w, P);
        const float q = meta.y >= 0 ? s_q[meta.y * kWave] : 0.0f;
        joint_step(APPROX ? R.aft + 12 * jj : R.ft + 12 * jj, APPROX ? R.atype[jj] : meta.x, q, S);
        aff_compose(P, S, T);
    }
    sl.store(meta.z, T);
}

// ---------------------------------------------------------------------------
// OBB staging: enabled OBBs of [o0, o1) -> LDS records (see ppl_internal.hpp)
// ---------------------------------------------------------------------------
// Record of OBB o from its stored inverse pose and full extents.
__device__ __forceinline__ void obb_record(const float4& d4, const float4& p0, const float4& p1, int o, float4* rec) {
    const float w = p0.w, x = p1.x, y = p1.y, z = p1.z;
    // Matrix of v -> v + w t + u x t, t = 2 u x v (scene_collision.cuh:36-45).
    const float r00 = 1.0f - 2.0f * (y * y + z * z), r01 = 2.0f * (x * y - w * z), r02 = 2.0f * (x * z + w * y);
    const float r10 = 2.0f * (x * y + w * z), r11 = 1.0f - 2.0f * (x * x + z * z), r12 = 2.0f * (y * z - w * x);
    const float r20 = 2.0f * (x * z - w * y), r21 = 2.0f * (y * z + w * x), r22 = 1.0f - 2.0f * (x * x + y * y);
    rec[0] = make_float4(r00, r01, r02, p0.x);
    rec[1] = make_float4(r10, r11, r12, p0.y);
    rec[2] = make_float4(r20, r21, r22, p0.z);
    rec[3] = make_float4(0.5f * d4.x, 0.5f * d4.y, 0.5f * d4.z, __int_as_float(o));
}

__device__ __forceinline__ void obb_record(const DScene& S, int o, float4* rec) {
    obb_record(*reinterpret_cast<const float4*>(S.obb_dims + 4 * (size_t)o),
               *reinterpret_cast<const float4*>(S.obb_pose + 8 * (size_t)o),
               *reinterpret_cast<const float4*>(S.obb_pose + 8 * (size_t)o + 4), o, rec);
}

// Stages the enabled OBBs of [o0, o1) (at most cap of them: callers pass
// ranges of at most cap OBBs) and returns their count.
__device__ int stage_obbs(const DScene& S, int o0, int o1, float4* s_obb, int cap) {
    const int lane = lane_id();
    int n = 0;
    for (int base = o0; base < o1; base += kWave) {
        const int o = base + lane;
        // enable flag, extents and pose in one round trip, then compaction
        float4 d4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), p0 = d4, p1 = d4;
        bool ok = false;
        if (o < o1) {
            ok = S.obb_enable[o] != 0;
            d4 = *reinterpret_cast<const float4*>(S.obb_dims + 4 * (size_t)o);
            p0 = *reinterpret_cast<const float4*>(S.obb_pose + 8 * (size_t)o);
            p1 = *reinterpret_cast<const float4*>(S.obb_pose + 8 * (size_t)o + 4);
        }
        const unsigned long long m = __ballot(ok);
        if (ok) {
            const int pos = n + __popcll(m & ((1ull << lane) - 1ull));
            if (pos < cap) obb_record(d4, p0, p1, o, s_obb + 4 * pos);
        }
        n += __popcll(m);
    }
    __syncthreads();
    return n < cap ? n : cap;
}

// OBB tiles of a problem's range [o0, o1): ranges of at most `cap` OBBs
// (cap >= 1), one LDS staging each.
__device__ __forceinline__ int obb_tiles(int o0, int o1, int cap) {
    return o1 - o0 <= cap ? 1 : (o1 - o0 + cap - 1) / cap;
}

__device__ __forceinline__ void to_box(const float4& r0, const float4& r1, const float4& r2, float px, float py, float pz,
                                       float& lx, float& ly, float& lz) {
    lx = fmaf(r0.x, px, fmaf(r0.y, py, fmaf(r0.z, pz, r0.w)));
    ly = fmaf(r1.x, px, fmaf(r1.y, py, fmaf(r1.z, pz, r1.w)));
    lz = fmaf(r2.x, px, fmaf(r2.y, py, fmaf(r2.z, pz, r2.w)));
}

// Collision iff squared distance to the box < r^2 (scene_collision.cuh:59-78).
__device__ __forceinline__ bool box_hit(float lx, float ly, float lz, const float4& h, float r2) {
    const float cx = fmaxf(fabsf(lx) - h.x, 0.0f);
    const float cy = fmaxf(fabsf(ly) - h.y, 0.0f);
    const float cz = fmaxf(fabsf(lz) - h.z, 0.0f);
    return fmaf(cz, cz, fmaf(cy, cy, cx * cx)) < r2;
}

// Order-preserving key of the box signed distance (Quilez form,
// scene_collision.cuh:88-112): key = |max(q,0)|^2 - min(max q, 0)^2 is
// sdf*|sdf|, so the minimum needs no square root per pair.
__device__ __forceinline__ float box_key(float lx, float ly, float lz, const float4& h) {
    const float qx = fabsf(lx) - h.x, qy = fabsf(ly) - h.y, qz = fabsf(lz) - h.z;
    const float ox = fmaxf(qx, 0.0f), oy = fmaxf(qy, 0.0f), oz = fmaxf(qz, 0.0f);
    const float in = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
    return fmaf(-in, in, fmaf(oz, oz, fmaf(oy, oy, ox * ox)));
}

__device__ __forceinline__ float key_to_sdf(float k) { return k >= 0.0f ? sqrtf(k) : -sqrtf(-k); }

// World-frame gradient of the box signed distance at world point p.
__device__ __forceinline__ void box_sdf_grad(const float4& r0, const float4& r1, const float4& r2, const float4& h,
                                             float px, float py, float pz, float& gx, float& gy, float& gz) {
    float lx, ly, lz;
    to_box(r0, r1, r2, px, py, pz, lx, ly, lz);
    const float sx = lx >= 0.0f ? 1.0f : -1.0f, sy = ly >= 0.0f ? 1.0f : -1.0f, sz = lz >= 0.0f ? 1.0f : -1.0f;
    const float qx = fabsf(lx) - h.x, qy = fabsf(ly) - h.y, qz = fabsf(lz) - h.z;
    float ax, ay, az;
    if (fmaxf(qx, fmaxf(qy, qz)) > 0.0f) {
        const float ox = fmaxf(qx, 0.0f), oy = fmaxf(qy, 0.0f), oz = fmaxf(qz, 0.0f);
        const float inv = 1.0f / sqrtf(fmaf(oz, oz, fmaf(oy, oy, ox * ox)));
        ax = ox * inv * sx; ay = oy * inv * sy; az = oz * inv * sz;
    } else {
        const bool xb = qx >= qy && qx >= qz;
        const bool yb = !xb && qy >= qz;
        ax = xb ? sx : 0.0f; ay = yb ? sy : 0.0f; az = (!xb && !yb) ? sz : 0.0f;
    }
    // d/dp = R^T d/dl
    gx = fmaf(r0.x, ax, fmaf(r1.x, ay, r2.x * az));
    gy = fmaf(r0.y, ax, fmaf(r1.y, ay, r2.y * az));
    gz = fmaf(r0.z, ax, fmaf(r1.z, ay, r2.z * az));
}

// ---------------------------------------------------------------------------
// ESDF (scene_collision.cuh:186-331); trilinear mode adds value + gradient
// ---------------------------------------------------------------------------
__device__ __forceinline__ int robust_floor(float x) {
    const float nearest = roundf(x);
    if (x - nearest >= 1e-4f) return (int)floorf(x);
    return (int)nearest;
}

__device__ __forceinline__ void quat_local(const float* pose, float sx, float sy, float sz, float& lx, float& ly, float& lz) {
    const float qw = pose[3], qx = pose[4], qy = pose[5], qz = pose[6];
    const float tx = 2.0f * (qy * sz - qz * sy);
    const float ty = 2.0f * (qz * sx - qx * sz);
    const float tz = 2.0f * (qx * sy - qy * sx);
    lx = sx + qw * tx + (qy * tz - qz * ty) + pose[0];
    ly = sy + qw * ty + (qz * tx - qx * tz) + pose[1];
    lz = sz + qw * tz + (qx * ty - qy * tx) + pose[2];
}

// Minimum ESDF clearance (-dist - r) over enabled layers, 1000 when none
// applies. With `want_grad` (trilinear only) returns d dist / d p (world)
// of the winning layer in g*.
__device__ float esdf_clearance(const DScene& S, float px, float py, float pz, float r, int sph_orig, bool want_grad,
                                float& gx, float& gy, float& gz) {
    float best = 1000.0f;
    gx = gy = gz = 0.0f;
    for (int layer = 0; layer < S.n_voxel_layers; ++layer) {
        if (S.voxel_enable[layer] == 0) continue;
        if (sph_orig >= 0 && S.sphere_acm && S.sphere_acm[sph_orig * S.n_voxel_layers + layer] == 0) continue;
        const float* prm = S.voxel_params + 4 * layer;
        const float* pose = S.voxel_pose + 8 * layer;
        const float vs = prm[3];
        if (vs <= 0.0f) continue;
        const float inv_vs = 1.0f / vs;
        const int nx = robust_floor(prm[0] * inv_vs) + 1;
        const int ny = robust_floor(prm[1] * inv_vs) + 1;
        const int nz = robust_floor(prm[2] * inv_vs) + 1;
        float lx, ly, lz;
        quat_local(pose, px, py, pz, lx, ly, lz);
        const float ux = (lx + prm[0] * 0.5f) * inv_vs, uy = (ly + prm[1] * 0.5f) * inv_vs, uz = (lz + prm[2] * 0.5f) * inv_vs;
        const float* feat = S.voxel_features + (size_t)layer * S.max_voxels_per_layer;
        if (S.esdf_interp == 0) {
            const int ix = (int)ux, iy = (int)uy, iz = (int)uz;
            if (ix >= nx - 2 || iy >= ny - 2 || iz >= nz - 2 || ix <= 2 || iy <= 2 || iz <= 2) continue;
            const long lin = (long)ix * ny * nz + (long)iy * nz + iz;
            if (lin >= S.max_voxels_per_layer) continue;
            const float d = feat[lin];
            if (d <= -999.0f) continue;
            const float c = -d - r;
            if (c < best) { best = c; gx = gy = gz = 0.0f; }
        } else {
            const float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
            const int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
            if (ix >= nx - 2 || iy >= ny - 2 || iz >= nz - 2 || ix <= 2 || iy <= 2 || iz <= 2) continue;
            const long base = (long)ix * ny * nz + (long)iy * nz + iz;
            if (base + (long)ny * nz + nz + 1 >= S.max_voxels_per_layer) continue;
            const long sx = (long)ny * nz, sy = nz;
            const float c000 = feat[base], c001 = feat[base + 1], c010 = feat[base + sy], c011 = feat[base + sy + 1];
            const float c100 = feat[base + sx], c101 = feat[base + sx + 1], c110 = feat[base + sx + sy], c111 = feat[base + sx + sy + 1];
            const float mn = fminf(fminf(fminf(c000, c001), fminf(c010, c011)), fminf(fminf(c100, c101), fminf(c110, c111)));
            if (mn <= -999.0f) continue;
            const float fx = ux - fx0, fy = uy - fy0, fz = uz - fz0;
            const float c00 = fmaf(c100 - c000, fx, c000), c01 = fmaf(c101 - c001, fx, c001);
            const float c10 = fmaf(c110 - c010, fx, c010), c11 = fmaf(c111 - c011, fx, c011);
            const float c0 = fmaf(c10 - c00, fy, c00), c1 = fmaf(c11 - c01, fy, c01);
            const float d = fmaf(c1 - c0, fz, c0);
            const float c = -d - r;
            if (c < best) {
                best = c;
                if (want_grad) {
                    const float ofy = 1.0f - fy, ofz = 1.0f - fz;
                    const float dx = ((c100 - c000) * ofy * ofz + (c101 - c001) * ofy * fz + (c110 - c010) * fy * ofz +
                                      (c111 - c011) * fy * fz) * inv_vs;
                    const float dy = ((c10 - c00) * ofz + (c11 - c01) * fz) * inv_vs;
                    const float dz = (c1 - c0) * inv_vs;
                    // grid-local -> world: inverse quaternion rotation
                    const float qw = pose[3], qx = -pose[4], qy = -pose[5], qz = -pose[6];
                    const float tx = 2.0f * (qy * dz - qz * dy), ty = 2.0f * (qz * dx - qx * dz), tz = 2.0f * (qx * dy - qy * dx);
                    gx = dx + qw * tx + (qy * tz - qz * ty);
                    gy = dy + qw * ty + (qz * tx - qx * tz);
                    gz = dz + qw * tz + (qx * ty - qy * tx);
                }
            }
        }
    }
    return best;
}

__device__ __forceinline__ bool esdf_hit(const DScene& S, float px, float py, float pz, float r, int sph_orig) {
    float gx, gy, gz;
    // min clearance < 0  <=>  some layer has dist > -r
    return esdf_clearance(S, px, py, pz, r, sph_orig, false, gx, gy, gz) < 0.0f;
}

// ---------------------------------------------------------------------------
// State resolution: flattened state index -> per-lane q in LDS
// ---------------------------------------------------------------------------
__device__ __forceinline__ int upper_bound_idx(const int* off, int n, long g) {
    // largest e in [0, n) with off[e] <= g
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

struct StateRef {
    int unit;      // config / edge / trajectory index
    int k;         // edge state index or waypoint index
    int n;         // edge states (incl. skipped) or trajectory length
    long u0;       // plan-order index of the unit's first state
    long gs;       // plan-order index of the state (differs from the chunk index when waypoint-major)
};

// State ranges of units u0 .. u0+63 into LDS: s_end[i] = off[u0+i+1] and
// s_end[64+i] = off[u0+i] (whole wave; no-op for configuration sources or
// when the chunk's first unit is unknown, u0 < 0).
__device__ __forceinline__ void stage_unit_ends(const DStates& X, int u0, int* s_end) {
    if (X.source == kSrcConfig || u0 < 0) return;
    const bool dense = X.source == kSrcDense || X.source == kSrcBfs;
    const int* off = X.source == kSrcEdge ? X.edge_off : dense ? X.dense_off : X.traj_off;
    const int n = X.source == kSrcEdge ? X.n_edges : X.n_traj;
    const int u = u0 + lane_id();
    const bool in = u < n;
    s_end[lane_id()] = in ? off[u + 1] : INT_MAX;
    s_end[kWave + lane_id()] = in ? off[u] : INT_MAX;
}

// The same two loads into registers, for a chunk ahead.
__device__ __forceinline__ void prefetch_unit_ends(const DStates& X, int u0, int& e_end, int& e_beg) {
    if (X.source == kSrcConfig || u0 < 0) return;
    const bool dense = X.source == kSrcDense || X.source == kSrcBfs;
    const int* off = X.source == kSrcEdge ? X.edge_off : dense ? X.dense_off : X.traj_off;
    const int n = X.source == kSrcEdge ? X.n_edges : X.n_traj;
    const int u = u0 + lane_id();
    const bool in = u < n;
    e_end = in ? off[u + 1] : INT_MAX;
    e_beg = in ? off[u] : INT_MAX;
}

// Unit holding state g and its state range [*u_begin, *u_end): the first of
// units u0 .. u0+63 whose end exceeds g (bisection over the staged ends),
// else a binary search over the offset table.
__device__ __forceinline__ int find_unit(const int* off, int n, long g, int u0, const int* s_end, int* u_begin,
                                         int* u_end) {
    if (u0 >= 0 && s_end[kWave - 1] > g) {
        int lo = 0, hi = kWave - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s_end[mid] > g) hi = mid; else lo = mid + 1;
        }
        *u_begin = s_end[kWave + lo];
        *u_end = s_end[lo];
        return u0 + lo;
    }
    const int u = upper_bound_idx(off, n, g);
    *u_begin = off[u];
    *u_end = off[u + 1];
    return u;
}

// Writes q_d into s_q[d * 64 + lane] for d < D (inactive lanes write nothing).
// g is the chunk-order index; prob the chunk's problem (waypoint-major plans).
// qcol / qstride: where q_d goes, s_q[d * qstride + qcol] (default: the
// lane's own column of a [D][64] array).
// store = false resolves the state without writing q (lanes that share a
// column with a storing lane).
__device__ __forceinline__ StateRef load_state(const DStates& X, int D, long g, bool active, float* s_q,
                                               int u0 = -1, const int* s_end = nullptr, int prob = 0,
                                               int qcol = -1, int qstride = kWave, bool store = true) {
    StateRef ref{0, 0, 0, 0, g};
    const int lane = qcol >= 0 ? qcol : lane_id();
    const int nd = store ? D : 0;   // dimensions to write
    if (!active) return ref;
    if (X.source == kSrcDense) {
        int d0, d1;
        const int b = find_unit(X.dense_off, X.n_traj, g, u0, s_end, &d0, &d1);
        const int t0 = X.traj_off[b], T = X.traj_off[b + 1] - t0;
        const int ns = X.nsub[b];
        const int f = (int)(g - d0);
        ref.unit = b;
        ref.k = f;
        ref.n = (T - 1) * ns + 1;
        ref.u0 = d0;
        const float* th = X.theta + (size_t)t0 * D;
        if (f >= (T - 1) * ns) {
            for (int d = 0; d < nd; ++d) s_q[d * qstride + lane] = th[(size_t)d * T + T - 1];
        } else {
            const int seg = f / ns, sub = f - seg * ns;
            const float frac = __fdiv_rn((float)sub, (float)ns);
            for (int d = 0; d < nd; ++d) {
                const float x0 = th[(size_t)d * T + seg], x1 = th[(size_t)d * T + seg + 1];
                s_q[d * qstride + lane] = __fadd_rn(x0, __fmul_rn(__fsub_rn(x1, x0), frac));
            }
        }
        return ref;
    }
    if (X.source == kSrcBfs) {
        // MIT*'s edge midpoints (MITStar.cu:938-965, 1316-1330): level
        // L = floor(log2(k + 1)), t = (2 (k + 1 - 2^L) + 1) / 2^(L + 1), and
        // q = q_src (1 - t) + q_dst t as written (no contraction)
        int e0, e1;
        const int e = find_unit(X.dense_off, X.n_traj, g, u0, s_end, &e0, &e1);
        const int f = (int)(g - e0);
        const int k = X.bfs_k0[e] + f;
        ref.unit = e;
        ref.k = f;
        ref.n = e1 - e0;
        ref.u0 = e0;
        const int kk = k + 1;
        const int lvl = 31 - __clz(kk);
        const float t = __fdiv_rn((float)(2 * (kk - (1 << lvl)) + 1), (float)(1 << (lvl + 1)));
        const float* a = X.cfg + (size_t)X.node_src[e] * X.cfg_stride;
        const float* b = X.cfg + (size_t)X.node_dst[e] * X.cfg_stride;
        for (int d = 0; d < nd; ++d)
            s_q[d * qstride + lane] = __fadd_rn(__fmul_rn(a[d], 1.0f - t), __fmul_rn(b[d], t));
        return ref;
    }
    if (X.source == kSrcTraj && X.wmajor) {
        const int4 pr = X.prob[prob];
        const int v = (int)(g - pr.w);
        const int t = v / pr.y, r = v - t * pr.y;
        ref.unit = pr.x + r;
        ref.k = t;
        ref.n = pr.z;
        ref.u0 = (long)pr.w + (long)r * pr.z;
        ref.gs = ref.u0 + t;
        const float* th = X.theta + (size_t)ref.u0 * D;
        for (int d = 0; d < nd; ++d) s_q[d * qstride + lane] = th[(size_t)d * pr.z + t];
        return ref;
    }
    if (X.source == kSrcConfig) {
        ref.unit = (int)g;
        ref.u0 = g;
        const float* q = X.cfg + (size_t)g * X.cfg_stride;
        for (int d = 0; d < nd; ++d) s_q[d * qstride + lane] = q[d];
    } else if (X.source == kSrcEdge) {
        int b0, b1;
        const int e = find_unit(X.edge_off, X.n_edges, g, u0, s_end, &b0, &b1);
        const int cnt = b1 - b0;
        ref.unit = e;
        ref.u0 = b0;
        ref.k = (int)(g - b0) + X.k0;
        ref.n = cnt + X.k0;
        const float t = ref.n > 1 ? (float)ref.k / (float)(ref.n - 1) : 0.0f;
        const float* a = X.from + (size_t)e * D;
        const float* b = X.to + (size_t)e * D;
        for (int d = 0; d < nd; ++d) s_q[d * qstride + lane] = fmaf(t, b[d] - a[d], a[d]);
    } else {
        int t0, t1;
        const int b = find_unit(X.traj_off, X.n_traj, g, u0, s_end, &t0, &t1);
        const int Tn = t1 - t0;
        ref.unit = b;
        ref.u0 = t0;
        ref.k = (int)(g - t0);
        ref.n = Tn;
        const float* th = X.theta + (size_t)t0 * D;
        for (int d = 0; d < nd; ++d) s_q[d * qstride + lane] = th[(size_t)d * Tn + ref.k];
    }
    return ref;
}

// Static, XCD-aware chunk schedule (no atomics). Workgroups are dispatched
// round-robin over the 8 XCDs, so workgroup w runs on XCD w % 8; its logical
// rank puts every XCD on a contiguous block of chunks in each sweep, which
// keeps neighbouring states (shared edge endpoints, trajectories, OBB sets)
// in one L2. Chunk c = rank + k * grid for k = 0, 1, ...: every wave reaches
// the end of the table, and chunks are visited roughly in table order.
struct ChunkCursor {
    int next;
    int stride;
};

__device__ __forceinline__ ChunkCursor chunk_cursor() {
    const int W = gridDim.x, w = blockIdx.x;
    constexpr int kXcds = 8;
    const int rank = (W % kXcds == 0) ? (w % kXcds) * (W / kXcds) + w / kXcds : w;
    return ChunkCursor{rank, W};
}

// Chunk count of a launch: the table's, or ceil(states / chunk) for implicit
// chunks (the state count may come from device memory).
__device__ __forceinline__ int chunk_count(const DChunks& C, long n_states) {
    return C.tab ? C.n_chunks : (int)((n_states + C.chunk - 1) / C.chunk);
}

__device__ __forceinline__ long state_count(const DStates& X) { return X.n_dyn ? (long)uniform(*X.n_dyn) : X.n_states; }

__device__ __forceinline__ bool next_chunk(const DChunks& C, long n_states, int n_chunks, ChunkCursor& cur, int4& ch) {
    const int c = cur.next;
    if (c >= n_chunks) return false;
    cur.next += cur.stride;
    if (C.tab) {
        ch = C.tab[c];
        ch.x = uniform(ch.x); ch.y = uniform(ch.y); ch.z = uniform(ch.z); ch.w = uniform(ch.w);
    } else {
        const long b = (long)c * C.chunk;
        const int u = C.unit_hint ? uniform(C.unit_hint[b >> 6]) : -1;
        ch = make_int4(0, (int)b, (int)(b + C.chunk < n_states ? b + C.chunk : n_states), u);
    }
    return true;
}

__device__ __forceinline__ void problem_obbs(const DChunks& C, const DScene& S, int prob, int& o0, int& o1) {
    if (C.obb_off) { o0 = C.obb_off[prob]; o1 = C.obb_off[prob + 1]; }
    else { o0 = 0; o1 = S.n_obbs; }
}

// Wave-uniform OBB candidate set: bit o set for OBB o (up to kMaxObbLds).
constexpr int kMaskWords = kMaxObbLds / 64;

struct ObbMask {
    unsigned long long w[kMaskWords];
};

__device__ __forceinline__ float read_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ===========================================================================
// Any-collision: two-phase gate semantics (two_phase_cc.cuh:57-138 with the
// check kernel's environment pass, pRRTC.cu:2337-2386).
// ===========================================================================
// Five waves per SIMD (95 VGPRs, no spills): the chunk loads and the joint
// loop's LDS reads leave half the wave time waiting, which a fifth resident
// wave hides better than the register room four waves would give.
__global__ __launch_bounds__(64, 5) void k_collide(CollideArgs a) {
    extern __shared__ float4 s_lds[];
    float4* s_obb = s_lds;
    float* s_q = reinterpret_cast<float*>(s_lds + 4 * a.C.obb_cap);
    const DRobot& R = a.R;
    const int lane = lane_id();
    const int J = R.n_joints, D = R.n_dof;
    int* s_end = reinterpret_cast<int*>(s_q + D * kWave);
    const bool has_esdf = a.S.n_voxel_layers > 0;
    const uint8_t* acm = a.S.sphere_obb_acm;
    const int cap = a.C.obb_cap;
    int cur_prob = -1, n_obb = 0, o0 = 0, o1 = 0, n_tiles = 1;
    int4 ch;
    ChunkCursor cur = chunk_cursor();
    const long n_states = state_count(a.X);
    const int n_chunks = chunk_count(a.C, n_states);
    // the next chunk's table entry is loaded at the top of each round and its
    // unit ends once this chunk's states are in, so both arrive during the
    // joint loop instead of heading the next round's chain of loads
    bool have = next_chunk(a.C, n_states, n_chunks, cur, ch);
    int pf_end = 0, pf_beg = 0;
    if (have) prefetch_unit_ends(a.X, ch.w, pf_end, pf_beg);
    while (have) {
        int4 chn;
        const bool have_n = next_chunk(a.C, n_states, n_chunks, cur, chn);
        if (ch.x != cur_prob) {
            problem_obbs(a.C, a.S, ch.x, o0, o1);
            n_tiles = uniform(obb_tiles(o0, o1, cap));
            if (n_tiles == 1) {
                __syncthreads();
                n_obb = uniform(stage_obbs(a.S, o0, o1, s_obb, cap));
            }
            cur_prob = ch.x;
        }
        const long g = (long)ch.y + lane;
        bool active = g < ch.z;
        if (a.X.source != kSrcConfig && ch.w >= 0) {
            s_end[lane] = pf_end;
            s_end[kWave + lane] = pf_beg;
        }
        __syncthreads();
        StateRef ref = load_state(a.X, D, g, active, s_q, ch.w, s_end, ch.x);
        bool hit = false;
        if (active && a.self_hit && a.self_hit[ref.gs]) { hit = true; active = false; }
        if (active && a.mode == 1 && a.skip_resolved) {
            const unsigned fh = (unsigned)a.first_hit[ref.unit];
            if (fh < (unsigned)ref.k) active = false;
        }
        if (have_n) prefetch_unit_ends(a.X, chn.w, pf_end, pf_beg);
        __syncthreads();
        // A problem with more OBBs than one staging holds is checked tile by
        // tile (FK redone per tile); bit jj of `flagged`: joint jj's
        // approximate spheres hit an OBB of an earlier tile, which the ESDF
        // pass of the last tile counts as the joint's flag.
        unsigned flagged = 0u;
        for (int t = 0; t < n_tiles && wave_any(active); ++t) {
            const bool last = t == n_tiles - 1;
            if (n_tiles > 1) {
                const int t0 = o0 + t * cap;
                __syncthreads();
                n_obb = uniform(stage_obbs(a.S, t0, min(o1, t0 + cap), s_obb, cap));
            }
            Slots sl;
            for (int jj = 0; jj < J && wave_any(active); ++jj) {
                float T[12];
                int4 meta;
                fk_joint(R, jj, s_q + lane, sl, T, meta);
                const int s0 = R.sph_off[jj], s1 = R.sph_off[jj + 1];
                const int a0 = R.apx_off[jj], a1 = R.apx_off[jj + 1];
                if (s1 == s0 || a1 == a0) continue;
                // Approximate spheres of this joint in world coordinates (usually one).
                float apx[3], apr2;
                const float4 as0 = R.apx[a0];
                aff_point(T, as0.x, as0.y, as0.z, apx[0], apx[1], apx[2]);
                {
                    const float rr = as0.w + a.margin;
                    apr2 = rr * rr;
                }
                // Broad phase: the OBBs some active lane's approximate sphere can
                // reach. The active lanes' centres lie within rw of the first
                // active lane's centre c0, so every OBB an approximate sphere
                // overlaps lies within rw + r of c0 (slack covers rounding).
                const bool broad = a1 - a0 == 1;
                float cx = 0.0f, cy = 0.0f, cz = 0.0f, reach2 = 0.0f;
                if (broad) {
                    const int l0 = __ffsll(__ballot(active)) - 1;
                    cx = read_lane(apx[0], l0); cy = read_lane(apx[1], l0); cz = read_lane(apx[2], l0);
                    const float dx = apx[0] - cx, dy = apx[1] - cy, dz = apx[2] - cz;
                    const float d2 = wave_max(active ? fmaf(dz, dz, fmaf(dy, dy, dx * dx)) : 0.0f);
                    const float reach = sqrtf(d2) * 1.0001f + fabsf(as0.w + a.margin) + 1e-5f;
                    reach2 = uniform(reach * reach);
                }
                bool flag = false;
                bool done = false;
                for (int base = 0; base < n_obb && !done; base += kWave) {
                    // candidates of this block of 64 OBBs: lane i tests OBB base + i
                    unsigned long long bits;
                    {
                        const int o = base + lane;
                        bool c = o < n_obb;
                        if (broad && c) {
                            float lx, ly, lz;
                            to_box(s_obb[4 * o], s_obb[4 * o + 1], s_obb[4 * o + 2], cx, cy, cz, lx, ly, lz);
                            c = box_hit(lx, ly, lz, s_obb[4 * o + 3], reach2);
                        }
                        bits = __ballot(c);
                    }
                    while (bits && !done) {
                        const int o = base + (__ffsll(bits) - 1);
                        bits &= bits - 1ull;
                        const float4 r0 = s_obb[4 * o], r1 = s_obb[4 * o + 1], r2 = s_obb[4 * o + 2], h = s_obb[4 * o + 3];
                        float lx, ly, lz;
                        to_box(r0, r1, r2, apx[0], apx[1], apx[2], lx, ly, lz);
                        bool ah = box_hit(lx, ly, lz, h, apr2);
                        for (int ai = a0 + 1; ai < a1; ++ai) {
                            const float4 as = R.apx[ai];
                            float x, y, z;
                            aff_point(T, as.x, as.y, as.z, x, y, z);
                            to_box(r0, r1, r2, x, y, z, lx, ly, lz);
                            const float rr = as.w + a.margin;
                            ah = ah || box_hit(lx, ly, lz, h, rr * rr);
                        }
                        flag = flag || ah;
                        if (!wave_any(ah && active)) continue;
                        // M = R_o * T_j: sphere-local -> box-local in one affine map.
                        float M[12];
    #pragma unroll
                        for (int rr = 0; rr < 3; ++rr) {
                            const float4 row = rr == 0 ? r0 : (rr == 1 ? r1 : r2);
    #pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                float v = fmaf(row.z, T[8 + c], fmaf(row.y, T[4 + c], row.x * T[c]));
                                if (c == 3) v += row.w;
                                M[rr * 4 + c] = v;
                            }
                        }
                        const int o_orig = __float_as_int(h.w);
                        bool h_any = false;
                        for (int s = s0; s < s1; ++s) {
                            const float4 sp = R.sph[s];
                            float x, y, z;
                            aff_point(M, sp.x, sp.y, sp.z, x, y, z);
                            const float rr = sp.w + a.margin;
                            bool hs = box_hit(x, y, z, h, rr * rr);
                            if (acm) hs = hs && acm[(size_t)R.sph_id[s] * a.S.n_obbs + o_orig] != 0;
                            h_any = h_any || hs;
                        }
                        hit = hit || (active && ah && h_any);
                        active = active && !hit;
                        done = !wave_any(active);
                    }
                }
                if (!last) {
                    if (flag) flagged |= 1u << jj;
                    continue;
                }
                flag = flag || ((flagged >> jj) & 1u) != 0u;
                if (has_esdf) {
                    // Approx ESDF pass (ACM disabled, scene_collision.cuh:574-582),
                    // then the joint's exact spheres when it is flagged.
                    for (int ai = a0; ai < a1; ++ai) {
                        const float4 as = R.apx[ai];
                        float x, y, z;
                        aff_point(T, as.x, as.y, as.z, x, y, z);
                        flag = flag || (active && esdf_hit(a.S, x, y, z, as.w + a.margin, -1));
                    }
                    if (wave_any(flag && active)) {
                        for (int s = s0; s < s1; ++s) {
                            const float4 sp = R.sph[s];
                            float x, y, z;
                            aff_point(T, sp.x, sp.y, sp.z, x, y, z);
                            const bool hs = flag && active && esdf_hit(a.S, x, y, z, sp.w + a.margin, R.sph_id[s]);
                            hit = hit || hs;
                        }
                        active = active && !hit;
                    }
                }
            }
        }
        if (g < ch.z && a.out) a.out[ref.gs] = hit ? 1 : 0;
        if (a.mode == 1) {
            // One atomic per (chunk, edge): a hit lane reports only when no
            // lower lane of the same edge hit (lanes hold consecutive states).
            const bool rep = hit && g < ch.z;
            const unsigned long long hm = __ballot(rep);
            if (rep) {
                const int l_e0 = ref.u0 > ch.y ? (int)(ref.u0 - ch.y) : 0;
                const unsigned long long same = ((1ull << lane) - 1ull) & ~((1ull << l_e0) - 1ull);
                if ((hm & same) == 0ull) atomicMin((unsigned*)&a.first_hit[ref.unit], (unsigned)ref.k);
            }
        }
        ch = chn;
        have = have_n;
    }
}

// ===========================================================================
// Any-collision in the gated two-phase form (two_phase_cc.cuh:57-104,
// scene_collision.cuh:519-628), for models whose exact (sphere, box) hits
// need not imply the approximate hit of the same box: approximate spheres
// that do not enclose the exact ones, or that ride a chain of their own.
// The approximate pass flags joints over the whole scene (every enabled box,
// no ACM, then the ESDF layers), then the exact spheres of flagged joints are
// checked against the whole scene (box and layer ACMs). One lane per state,
// OBBs in LDS tiles as in k_collide.
// ===========================================================================
__global__ __launch_bounds__(64) void k_collide_gated(CollideArgs a) {
    extern __shared__ float4 s_lds[];
    float4* s_obb = s_lds;
    float* s_q = reinterpret_cast<float*>(s_lds + 4 * a.C.obb_cap);
    const DRobot& R = a.R;
    const int lane = lane_id();
    const int J = R.n_joints, D = R.n_dof;
    int* s_end = reinterpret_cast<int*>(s_q + D * kWave);
    const bool has_esdf = a.S.n_voxel_layers > 0;
    const bool own_chain = R.aft != nullptr;
    const uint8_t* acm = a.S.sphere_obb_acm;
    const int cap = a.C.obb_cap;
    int cur_prob = -1, n_obb = 0, o0 = 0, o1 = 0, n_tiles = 1;
    int4 ch;
    ChunkCursor cur = chunk_cursor();
    const long n_states = state_count(a.X);
    const int n_chunks = chunk_count(a.C, n_states);
    while (next_chunk(a.C, n_states, n_chunks, cur, ch)) {
        if (ch.x != cur_prob) {
            problem_obbs(a.C, a.S, ch.x, o0, o1);
            n_tiles = uniform(obb_tiles(o0, o1, cap));
            if (n_tiles == 1) {
                __syncthreads();
                n_obb = uniform(stage_obbs(a.S, o0, o1, s_obb, cap));
            }
            cur_prob = ch.x;
        }
        const long g = (long)ch.y + lane;
        bool active = g < ch.z;
        stage_unit_ends(a.X, ch.w, s_end);
        __syncthreads();
        StateRef ref = load_state(a.X, D, g, active, s_q, ch.w, s_end, ch.x);
        bool hit = false;
        if (active && a.self_hit && a.self_hit[ref.gs]) { hit = true; active = false; }
        if (active && a.mode == 1 && a.skip_resolved) {
            const unsigned fh = (unsigned)a.first_hit[ref.unit];
            if (fh < (unsigned)ref.k) active = false;
        }
        __syncthreads();
        // 1. approximate pass: bit jj of `flagged` once an approximate sphere
        // of joint jj (its own chain's frame) hits an enabled box or a layer
        unsigned flagged = 0u;
        for (int t = 0; t < n_tiles && wave_any(active); ++t) {
            if (n_tiles > 1) {
                const int t0 = o0 + t * cap;
                __syncthreads();
                n_obb = uniform(stage_obbs(a.S, t0, min(o1, t0 + cap), s_obb, cap));
            }
            const bool last = t == n_tiles - 1;
            Slots sl;
            for (int jj = 0; jj < J; ++jj) {
                float T[12];
                int4 meta;
                if (own_chain) fk_joint<true>(R, jj, s_q + lane, sl, T, meta);
                else fk_joint(R, jj, s_q + lane, sl, T, meta);
                for (int ai = R.apx_off[jj]; ai < R.apx_off[jj + 1]; ++ai) {
                    const float4 as = R.apx[ai];
                    float x, y, z;
                    aff_point(T, as.x, as.y, as.z, x, y, z);
                    const float rr = as.w + a.margin;
                    bool f = !active || ((flagged >> jj) & 1u) != 0u;   // nothing left to find
                    for (int o = 0; o < n_obb && wave_any(!f); ++o) {
                        float lx, ly, lz;
                        to_box(s_obb[4 * o], s_obb[4 * o + 1], s_obb[4 * o + 2], x, y, z, lx, ly, lz);
                        f = f || box_hit(lx, ly, lz, s_obb[4 * o + 3], rr * rr);
                    }
                    if (last && has_esdf && !f) f = esdf_hit(a.S, x, y, z, rr, -1);
                    if (active && f) flagged |= 1u << jj;
                }
            }
        }
        // 2. exact pass over the flagged joints' spheres
        for (int t = 0; t < n_tiles && wave_any(active && flagged != 0u); ++t) {
            if (n_tiles > 1) {
                const int t0 = o0 + t * cap;
                __syncthreads();
                n_obb = uniform(stage_obbs(a.S, t0, min(o1, t0 + cap), s_obb, cap));
            }
            const bool last = t == n_tiles - 1;
            Slots sl;
            for (int jj = 0; jj < J && wave_any(active); ++jj) {
                float T[12];
                int4 meta;
                fk_joint(R, jj, s_q + lane, sl, T, meta);
                const bool fj = active && ((flagged >> jj) & 1u) != 0u;
                if (!wave_any(fj)) continue;
                for (int s = R.sph_off[jj]; s < R.sph_off[jj + 1]; ++s) {
                    const float4 sp = R.sph[s];
                    float x, y, z;
                    aff_point(T, sp.x, sp.y, sp.z, x, y, z);
                    const float rr = sp.w + a.margin;
                    const uint8_t* acm_row = acm ? acm + (size_t)R.sph_id[s] * a.S.n_obbs : nullptr;
                    bool hs = false;
                    for (int o = 0; o < n_obb && wave_any(fj && !hs); ++o) {
                        const float4 h = s_obb[4 * o + 3];
                        if (acm_row && acm_row[__float_as_int(h.w)] == 0) continue;
                        float lx, ly, lz;
                        to_box(s_obb[4 * o], s_obb[4 * o + 1], s_obb[4 * o + 2], x, y, z, lx, ly, lz);
                        hs = hs || (fj && box_hit(lx, ly, lz, h, rr * rr));
                    }
                    if (last && has_esdf && fj && !hs) hs = esdf_hit(a.S, x, y, z, rr, R.sph_id[s]);
                    hit = hit || hs;
                }
                active = active && !hit;
            }
        }
        if (g < ch.z && a.out) a.out[ref.gs] = hit ? 1 : 0;
        if (a.mode == 1) {
            const bool rep = hit && g < ch.z;
            const unsigned long long hm = __ballot(rep);
            if (rep) {
                const int l_e0 = ref.u0 > ch.y ? (int)(ref.u0 - ch.y) : 0;
                const unsigned long long same = ((1ull << lane) - 1ull) & ~((1ull << l_e0) - 1ull);
                if ((hm & same) == 0ull) atomicMin((unsigned*)&a.first_hit[ref.unit], (unsigned)ref.k);
            }
        }
    }
}

// ===========================================================================
// Summed-hinge state cost (+ adjoint). chomp.cu:151-188 / stomp.cu:206-246
// cost = w * (sum_s relu(m - min_o clr(s,o)) + sum_s relu(m - esdf clr(s))
//             + relu(m_self - min_pairs clr)), raw radii.
// Gradient: every sphere s with a non-zero point gradient g_s adds, for each
// ancestor joint j of its link, g_s . (z_j x (p_s - o_j)) (revolute) or
// g_s . z_j (prismatic), with z_j, o_j the joint's world axis and origin
// recorded during the forward walk. Lever arms are taken relative to each
// joint, so no large world-origin moments are summed and subtracted.
// ===========================================================================
// Candidate OBBs of one sphere cluster for the hinge: a wave-uniform bitmask
// of the OBBs some active lane's spheres of the cluster can come within the
// hinge band of. Every sphere s of the cluster lies inside its enclosing
// sphere `as` (centre a, radius R; DRobot::clu_enc), so clr(s, o) >=
// sdf_o(a) - R; the active lanes' centres lie within rw of the first active
// lane's centre c0, so sdf_o(a) >= sdf_o(c0) - rw. An OBB with sdf_o(c0) >=
// band + R + rw therefore has clearance >= band from every sphere: it cannot
// change a hinge, its argmin or its gradient. One pass per 64 OBBs (lane i
// tests OBB base + i).
__device__ __forceinline__ void near_obbs(const float4* __restrict__ s_obb, int n_obb, const float* T, const float4& as,
                                          float band, bool active, ObbMask& mk) {
    float ax, ay, az;
    aff_point(T, as.x, as.y, as.z, ax, ay, az);
    const int l0 = __ffsll(__ballot(active)) - 1;
    if (l0 < 0) {
#pragma unroll
        for (int k = 0; k < kMaskWords; ++k) mk.w[k] = 0ull;
        return;
    }
    const float cx = read_lane(ax, l0), cy = read_lane(ay, l0), cz = read_lane(az, l0);
    const float dx = ax - cx, dy = ay - cy, dz = az - cz;
    const float d2 = wave_max(active ? fmaf(dz, dz, fmaf(dy, dy, dx * dx)) : 0.0f);
    // sdf(c0) < reach  <=>  key(c0) < reach * |reach|; slack covers rounding
    const float reach = band + as.w + sqrtf(d2) * 1.0001f + 1e-5f;
    const float thr = uniform(reach * fabsf(reach));
#pragma unroll
    for (int k = 0; k < kMaskWords; ++k) {
        bool c = false;
        if (k * 64 < n_obb) {
            const int o = k * 64 + lane_id();
            if (o < n_obb) {
                float lx, ly, lz;
                to_box(s_obb[4 * o], s_obb[4 * o + 1], s_obb[4 * o + 2], cx, cy, cz, lx, ly, lz);
                c = box_key(lx, ly, lz, s_obb[4 * o + 3]) < thr;
            }
        }
        mk.w[k] = __ballot(c);
    }
}

// Smallest box key over the candidate OBBs of `mk` at world point p, and its
// OBB (LDS record index; -1 when there is none). Ties keep the lowest index.
// A (sphere, OBB) pair the scene's ACM masks out is skipped
// (collision/scene_collision.cuh:140-167: sphere_obb_clearance by sphere index).
// PAIR: two candidates per step, both records read before either test, so
// one record's LDS latency hides the other's (the cost-only kernel has the
// registers for it; the adjoint instantiation would spill).
template <bool PAIR>
__device__ __forceinline__ void obb_argmin(const float4* __restrict__ s_obb, const ObbMask& mk, float px, float py,
                                           float pz, const uint8_t* __restrict__ acm_row, float& best, int& arg) {
    best = FLT_MAX;
    arg = -1;
#pragma unroll
    for (int k = 0; k < kMaskWords; ++k) {
        unsigned long long bits = mk.w[k];
        if (PAIR) {
            while (bits) {
                const int oa = k * 64 + (__ffsll(bits) - 1);
                bits &= bits - 1ull;
                const bool two = bits != 0ull;
                const int ob = two ? k * 64 + (__ffsll(bits) - 1) : oa;
                if (two) bits &= bits - 1ull;
                const float4 a0 = s_obb[4 * oa], a1 = s_obb[4 * oa + 1], a2 = s_obb[4 * oa + 2], ah = s_obb[4 * oa + 3];
                const float4 b0 = s_obb[4 * ob], b1 = s_obb[4 * ob + 1], b2 = s_obb[4 * ob + 2], bh = s_obb[4 * ob + 3];
                float lx, ly, lz;
                to_box(a0, a1, a2, px, py, pz, lx, ly, lz);
                float ka = box_key(lx, ly, lz, ah);
                to_box(b0, b1, b2, px, py, pz, lx, ly, lz);
                float kb = box_key(lx, ly, lz, bh);
                if (acm_row) {
                    if (acm_row[__float_as_int(ah.w)] == 0) ka = FLT_MAX;
                    if (acm_row[__float_as_int(bh.w)] == 0) kb = FLT_MAX;
                }
                if (ka < best) { best = ka; arg = oa; }   // oa < ob: ties keep the lower index
                if (two && kb < best) { best = kb; arg = ob; }
            }
        } else {
            while (bits) {
                const int o = k * 64 + (__ffsll(bits) - 1);
                bits &= bits - 1ull;
                if (acm_row && acm_row[__float_as_int(s_obb[4 * o + 3].w)] == 0) continue;
                float lx, ly, lz;
                to_box(s_obb[4 * o], s_obb[4 * o + 1], s_obb[4 * o + 2], px, py, pz, lx, ly, lz);
                const float key = box_key(lx, ly, lz, s_obb[4 * o + 3]);
                if (key < best) { best = key; arg = o; }
            }
        }
    }
}

// Hinge terms of one sphere at world position p given its OBB argmin (best
// key; has_arg: there is one; rec(r0, r1, r2, r3) fetches the argmin OBB's
// record, called only for the gradient of an active hinge): the summed hinge
// (OBB + ESDF) and its point gradient d hinge / d p in g*; in clearance-only
// mode the ESDF clearance with radius + margin.
template <bool GRAD, class Rec>
__device__ __forceinline__ float sphere_hinge_at(const CostArgs& a, bool has_arg, const Rec& rec, float best, float px,
                                                 float py, float pz, float r, int sph_id, bool use_obb, bool use_esdf,
                                                 float& gx, float& gy, float& gz) {
    const float m = a.P.world_margin;
    float world = 0.0f;
    gx = gy = gz = 0.0f;
    // clearance mode is ESDF-only; no layer gives sphere_esdf_clearance's 1000
    if (a.clearance_only && !use_esdf) return 1000.0f;
    if (use_obb) {
        const float clr = has_arg ? key_to_sdf(best) - r : 1000.0f;
        const float hh = m - clr;
        if (hh > 0.0f) {
            world += hh;
            if (GRAD) {
                float dx, dy, dz;
                float4 r0, r1, r2, r3;
                rec(r0, r1, r2, r3);
                box_sdf_grad(r0, r1, r2, r3, px, py, pz, dx, dy, dz);
                gx -= dx; gy -= dy; gz -= dz;
            }
        }
    }
    if (use_esdf) {
        float dx, dy, dz;
        const float rr = a.clearance_only ? r + a.margin : r;
        const float c = esdf_clearance(a.S, px, py, pz, rr, sph_id, GRAD, dx, dy, dz);
        if (a.clearance_only) return c;
        const float hh = m - c;
        if (hh > 0.0f) {
            world += hh;
            if (GRAD) { gx += dx; gy += dy; gz += dz; }
        }
    }
    return world;
}

// Sum / min over the G lanes of a state group (G a power of two <= 64; the
// groups are aligned, so xor offsets below G stay inside the group).
__device__ __forceinline__ float group_sum(float v, int G) {
    for (int o = 1; o < G; o <<= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float group_min(float v, int G) {
    for (int o = 1; o < G; o <<= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

template <bool GRAD>
__global__ __launch_bounds__(64, 4) void k_cost(CostArgs a) {
    extern __shared__ float4 s_lds[];
    float4* s_obb = s_lds;
    float* s_q = reinterpret_cast<float*>(s_lds + 4 * a.C.obb_cap);
    const DRobot& R = a.R;
    const int lane = lane_id();
    const int G = a.group;                    // lanes per state
    const int sub = lane & (G - 1);
    const int slot = lane / G;                // state slot of this lane group
    const int J = R.n_joints, D = R.n_dof;
    // Adjoint records (GRAD): per state slot and joint, the joint origin o,
    // axis z, and the force f / moment tau (about o) of its own spheres'
    // hinge gradients; then d cost / d q per dof.
    int* s_end = reinterpret_cast<int*>(s_q + D * kWave);
    const int rs = rec_stride(J), ds = dq_stride(D);
    float* rec = s_q + (D + 2) * kWave + slot * rs;
    float* dqs = s_q + (D + 2) * kWave + (kWave / G) * rs + slot * ds;
    const bool use_obb = a.P.use_obb != 0 && !a.clearance_only;
    const bool use_esdf = (a.P.use_esdf != 0 || a.clearance_only) && a.S.n_voxel_layers > 0;
    const float wt = a.P.weight;
    int cur_prob = -1, n_obb = 0;
    int4 ch;
    ChunkCursor cur = chunk_cursor();
    const long n_states = state_count(a.X);
    const int n_chunks = chunk_count(a.C, n_states);
    while (next_chunk(a.C, n_states, n_chunks, cur, ch)) {
        if (ch.x != cur_prob) {
            int o0, o1;
            problem_obbs(a.C, a.S, ch.x, o0, o1);
            __syncthreads();
            n_obb = uniform(stage_obbs(a.S, o0, o1, s_obb, a.C.obb_cap));
            cur_prob = ch.x;
        }
        const long g = (long)ch.y + slot;
        const bool active = g < ch.z;
        stage_unit_ends(a.X, ch.w, s_end);
        __syncthreads();
        StateRef ref = load_state(a.X, D, g, active, s_q, ch.w, s_end, ch.x);
        __syncthreads();
        // Self term from the self pass: hinge and its two point gradients.
        float self_h = 0.0f, snx = 0.0f, sny = 0.0f, snz = 0.0f;
        int si = -1, sj = -1;
        if (a.self_in && active && !a.clearance_only) {
            const float* so = a.self_in + 8 * (size_t)ref.gs;
            self_h = fmaxf(a.P.self_margin - so[0], 0.0f);
            if (self_h > 0.0f) { si = __float_as_int(so[1]); sj = __float_as_int(so[2]); snx = so[3]; sny = so[4]; snz = so[5]; }
        }
        float world = 0.0f, minclr = 1000.0f;
        Slots sl;
        for (int jj = 0; jj < J; ++jj) {
            float T[12];
            int4 meta;
            fk_joint(R, jj, s_q + lane, sl, T, meta);
            float fx = 0.0f, fy = 0.0f, fz = 0.0f, tx = 0.0f, ty = 0.0f, tz = 0.0f;
            const int s0 = R.sph_off[jj], s1 = R.sph_off[jj + 1];
            // one candidate mask per sphere cluster of the joint
            for (int c = (s1 > s0 ? R.clu_off[jj] : 0); c < (s1 > s0 ? R.clu_off[jj + 1] : 0); ++c) {
                ObbMask mk;
                if (use_obb) near_obbs(s_obb, n_obb, T, R.clu_enc[c], a.P.world_margin, active, mk);
                for (int s = R.clu_beg[c] + sub; s < R.clu_beg[c + 1]; s += G) {
                    const float4 sp = R.sph[s];
                    float px, py, pz;
                    aff_point(T, sp.x, sp.y, sp.z, px, py, pz);
                    float best = FLT_MAX;
                    int arg = -1;
                    const uint8_t* acm_row =
                        a.S.sphere_obb_acm ? a.S.sphere_obb_acm + (size_t)R.sph_id[s] * a.S.n_obbs : nullptr;
                    if (use_obb) obb_argmin<!GRAD>(s_obb, mk, px, py, pz, acm_row, best, arg);
                    float gx, gy, gz;
                    auto rec = [&](float4& r0, float4& r1, float4& r2, float4& r3) {
                        r0 = s_obb[4 * arg]; r1 = s_obb[4 * arg + 1]; r2 = s_obb[4 * arg + 2]; r3 = s_obb[4 * arg + 3];
                    };
                    const float hs = sphere_hinge_at<GRAD>(a, arg >= 0, rec, best, px, py, pz, sp.w, R.sph_id[s],
                                                           use_obb, use_esdf, gx, gy, gz);
                    if (a.clearance_only) minclr = fminf(minclr, hs);
                    else world += hs;
                    if (GRAD) {
                        if (si >= 0) {
                            const int id = R.sph_id[s];
                            if (id == si) { gx -= snx; gy -= sny; gz -= snz; }
                            if (id == sj) { gx += snx; gy += sny; gz += snz; }
                        }
                        // wrench about the joint origin: f += g, tau += (p - o) x g
                        const float rx = px - T[3], ry = py - T[7], rz = pz - T[11];
                        fx += gx; fy += gy; fz += gz;
                        tx = fmaf(ry, gz, fmaf(-rz, gy, tx));
                        ty = fmaf(rz, gx, fmaf(-rx, gz, ty));
                        tz = fmaf(rx, gy, fmaf(-ry, gx, tz));
                    }
                }
            }
            if (GRAD && !a.clearance_only) {
                const bool nz = fx != 0.0f || fy != 0.0f || fz != 0.0f || tx != 0.0f || ty != 0.0f || tz != 0.0f;
                if (G > 1 && wave_any(nz)) {
                    fx = group_sum(fx, G); fy = group_sum(fy, G); fz = group_sum(fz, G);
                    tx = group_sum(tx, G); ty = group_sum(ty, G); tz = group_sum(tz, G);
                }
                if (sub == 0) {
                    const int type = meta.x;
                    const int ax = type >= kXRot ? type - kXRot : (type >= 0 ? type : 0);
                    float* r = rec + 12 * jj;
                    r[0] = T[3]; r[1] = T[7]; r[2] = T[11];
                    r[3] = T[ax]; r[4] = T[4 + ax]; r[5] = T[8 + ax];
                    r[6] = fx; r[7] = fy; r[8] = fz;
                    r[9] = tx; r[10] = ty; r[11] = tz;
                }
            }
        }
        if (GRAD && !a.clearance_only && sub == 0 && active) {
            // Reverse DFS sweep: every joint's subtree wrench is complete once
            // its descendants (higher DFS positions) have been folded in.
            for (int d = 0; d < D; ++d) dqs[d] = 0.0f;
            for (int jj = J - 1; jj >= 1; --jj) {
                const float* r = rec + 12 * jj;
                const float ox = r[0], oy = r[1], oz = r[2];
                const float fx = r[6], fy = r[7], fz = r[8];
                const float tx = r[9], ty = r[10], tz = r[11];
                const int4 meta = R.jmeta[jj];
                if (meta.y >= 0 && meta.x >= 0) {
                    const float dq = meta.x >= kXRot ? fmaf(r[3], tx, fmaf(r[4], ty, r[5] * tz))
                                                     : fmaf(r[3], fx, fmaf(r[4], fy, r[5] * fz));
                    dqs[meta.y] += dq;
                }
                float* q = rec + 12 * R.parent[jj];
                const float lx = ox - q[0], ly = oy - q[1], lz = oz - q[2];
                q[6] += fx; q[7] += fy; q[8] += fz;
                q[9] += tx + (ly * fz - lz * fy);
                q[10] += ty + (lz * fx - lx * fz);
                q[11] += tz + (lx * fy - ly * fx);
            }
            bool endpoint = false;
            size_t base = (size_t)ref.gs * D, stride = 1;
            if (a.X.source == kSrcTraj) {
                endpoint = (ref.k == 0 || ref.k == ref.n - 1);
                base = (size_t)a.X.traj_off[ref.unit] * D + ref.k;
                stride = ref.n;
            }
            for (int d = 0; d < D; ++d) a.grad_out[base + (size_t)d * stride] = endpoint ? 0.0f : wt * dqs[d];
        }
        if (a.clearance_only) {
            minclr = group_min(minclr, G);
            if (active && sub == 0) a.cost_out[ref.gs] = minclr;
        } else {
            world = group_sum(world, G);
            if (active && sub == 0 && a.cost_out) a.cost_out[ref.gs] = wt * (world + self_h);
        }
        __syncthreads();   // records and q of this chunk are dead before the next load_state
    }
}

// ===========================================================================
// Grouped summed-hinge cost (+ adjoint): one state per aligned group of
// kGrp = 16 lanes, four states per wave, with the state's spheres spread over
// the group's lanes across joints instead of joint by joint.
//  1. Forward kinematics in parallel: lane jj of the group computes joint
//     jj's step (one sin/cos per lane) and the frames are composed by
//     pointer jumping over the DFS tree in ceil(log2(depth)) rounds through
//     LDS. Then every sphere cluster's enclosing sphere is placed at once
//     and one wave pass per cluster over the OBBs gives its candidate mask
//     (the broad phase of near_obbs over the wave's four states).
//  2. Sphere pass: emission slot k (spheres are emitted joint by joint,
//     cluster by cluster) goes to lane k mod 16, which reads its joint's
//     frame and its cluster's mask from LDS,
//     finds the argmin OBB and the hinge terms. Panda's 59 spheres take four
//     rounds instead of one round per joint with most lanes idle.
//  3. Adjoint: every sphere with a non-zero point gradient g adds its force g
//     and moment (p - o_j) x g about its joint's origin o_j into the joint's
//     wrench in LDS. A prefix scan over the group's lanes (one joint per lane)
//     sums each joint's DFS subtree -- a contiguous run of positions -- into a
//     total wrench about o_jj, and lane jj reads d cost / d q = z . tau
//     (revolute) or z . f (prismatic).
// Results equal k_cost's up to FP32 summation order.
// ===========================================================================
// Joint frames of the state of a 16-lane group (one state per group, a
// 1-wave workgroup), DFS positions 0..J-1, into F (12 floats each, row-major
// 3x4), by pointer jumping over the DFS tree: lane jj of the group starts
// from its joint step A = S_jj (the root's is the identity) and its parent
// link P; each round composes the frame of link P onto A and jumps P to P's
// link, until P is the root. Then A is the product of the steps from the root
// down to jj: the frame. ceil(log2(depth)) rounds through LDS. qs: the
// state's q (q_d = qs[d * qstride]); par: J ints of LDS for the links.
// APPROX: the approximate spheres' own chain (R.aft / R.atype).
// Contains block barriers: every lane of the workgroup must call it.
template <bool APPROX = false>
__device__ __forceinline__ void group_frames(const DRobot& R, const float* qs, int qstride, float* F, int* par,
                                             int sub) {
    const int J = R.n_joints;
    const bool own = sub < J;
    int P = 0;
    if (own) {
        float A[12];
        if (sub == 0) {
            aff_identity(A);
        } else {
            const int4 meta = R.jmeta[sub];
            const float q = meta.y >= 0 ? qs[meta.y * qstride] : 0.0f;
            joint_step(APPROX ? R.aft + 12 * sub : R.ft + 12 * sub, APPROX ? R.atype[sub] : meta.x, q, A);
            P = R.parent[sub];
        }
        float4* d = reinterpret_cast<float4*>(F + 12 * sub);
        d[0] = make_float4(A[0], A[1], A[2], A[3]);
        d[1] = make_float4(A[4], A[5], A[6], A[7]);
        d[2] = make_float4(A[8], A[9], A[10], A[11]);
        par[sub] = P;
    }
    __syncthreads();
    while (wave_any(own && P != 0)) {
        const bool jump = own && P != 0;
        float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0, c2 = c0;
        int PP = 0;
        if (jump) {
            const float4* b = reinterpret_cast<const float4*>(F + 12 * P);
            const float4* m = reinterpret_cast<const float4*>(F + 12 * sub);
            const float4 m0 = m[0], m1 = m[1], m2 = m[2];
            auto row = [&](const float4 br) {
                return make_float4(fmaf(br.z, m2.x, fmaf(br.y, m1.x, br.x * m0.x)),
                                   fmaf(br.z, m2.y, fmaf(br.y, m1.y, br.x * m0.y)),
                                   fmaf(br.z, m2.z, fmaf(br.y, m1.z, br.x * m0.z)),
                                   fmaf(br.z, m2.w, fmaf(br.y, m1.w, br.x * m0.w)) + br.w);
            };
            c0 = row(b[0]);
            c1 = row(b[1]);
            c2 = row(b[2]);
            PP = par[P];
        }
        __syncthreads();
        if (jump) {
            float4* d = reinterpret_cast<float4*>(F + 12 * sub);
            d[0] = c0;
            d[1] = c1;
            d[2] = c2;
            P = PP;
            par[sub] = P;
        }
        __syncthreads();
    }
}

static_assert(kGrp >= kMaxJoints, "the grouped kernel gives every joint its own lane of the group");

size_t lds_floats_grp(int obb_cap, int n_dof, int n_joints, int n_sph, int n_clu, bool grad, bool tiled) {
    size_t f = (size_t)obb_cap * 16 + (size_t)n_clu * kMaskWords * 2 +
               (((size_t)n_dof * kGrpStates + 3) & ~(size_t)3) + 2 * kWave + (size_t)kGrpStates * (n_joints * 13 + 4);
    if (grad) f += (size_t)kGrpStates * ((size_t)n_joints * 6 + kMaxDof + 1);
    if (tiled) f += (size_t)kGrpStates * n_sph * 2;
    return f;
}

// Candidate masks of every sphere cluster of the wave's four states over the
// staged OBBs (the broad phase of near_obbs, all clusters' enclosing spheres
// placed at once, 16 clusters per pass): lane k of each group places cluster
// c0 + k's enclosing sphere; the four states' spread about state slot 0's
// centre sets the cluster's reach; one pass of 64 lanes over the OBBs per
// cluster gives its mask.
__device__ __forceinline__ void grp_masks(const DRobot& R, const float4* __restrict__ s_obb, int n_obb, const float* Ts,
                                          float band, bool active, int sub, unsigned long long* s_mask) {
    const int C = R.n_clu, lane = lane_id();
    for (int c0 = 0; c0 < C; c0 += kGrp) {
        const int cl = c0 + sub;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        if (cl < C && active) {
            const float4 ce = R.clu_enc[cl];
            const float4* t = reinterpret_cast<const float4*>(Ts + 12 * R.clu_jnt[cl]);
            const float4 t0 = t[0], t1 = t[1], t2 = t[2];
            ax = fmaf(t0.x, ce.x, fmaf(t0.y, ce.y, fmaf(t0.z, ce.z, t0.w)));
            ay = fmaf(t1.x, ce.x, fmaf(t1.y, ce.y, fmaf(t1.z, ce.z, t1.w)));
            az = fmaf(t2.x, ce.x, fmaf(t2.y, ce.y, fmaf(t2.z, ce.z, t2.w)));
        }
        const float cx = __shfl(ax, sub), cy = __shfl(ay, sub), cz = __shfl(az, sub);   // slot 0's centre
        const float dx = ax - cx, dy = ay - cy, dz = az - cz;
        float d2 = (cl < C && active) ? fmaf(dz, dz, fmaf(dy, dy, dx * dx)) : 0.0f;
        d2 = fmaxf(d2, __shfl_xor(d2, 16));
        d2 = fmaxf(d2, __shfl_xor(d2, 32));
        const int nk = min(kGrp, C - c0);
        // up to 64 staged boxes: each lane reads its box's record once for
        // every cluster of the pass
        float4 q0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), q1 = q0, q2 = q0, qh = q0;
        const bool one_word = n_obb <= 64;
        if (one_word && lane < n_obb) {
            q0 = s_obb[4 * lane]; q1 = s_obb[4 * lane + 1]; q2 = s_obb[4 * lane + 2]; qh = s_obb[4 * lane + 3];
        }
        for (int k = 0; k < nk; ++k) {
            const float4 ce = R.clu_enc[c0 + k];
            ObbMask mk;
            const float c0x = read_lane(cx, k), c0y = read_lane(cy, k), c0z = read_lane(cz, k);
            const float reach = band + ce.w + sqrtf(read_lane(d2, k)) * 1.0001f + 1e-5f;
            const float thr = uniform(reach * fabsf(reach));
            if (one_word) {
                bool c = false;
                if (lane < n_obb) {
                    float lx, ly, lz;
                    to_box(q0, q1, q2, c0x, c0y, c0z, lx, ly, lz);
                    c = box_key(lx, ly, lz, qh) < thr;
                }
                mk.w[0] = __ballot(c);
#pragma unroll
                for (int w = 1; w < kMaskWords; ++w) mk.w[w] = 0ull;
            } else
#pragma unroll
            for (int w = 0; w < kMaskWords; ++w) {
                bool c = false;
                if (w * 64 < n_obb) {
                    const int o = w * 64 + lane;
                    if (o < n_obb) {
                        float lx, ly, lz;
                        to_box(s_obb[4 * o], s_obb[4 * o + 1], s_obb[4 * o + 2], c0x, c0y, c0z, lx, ly, lz);
                        c = box_key(lx, ly, lz, s_obb[4 * o + 3]) < thr;
                    }
                }
                mk.w[w] = __ballot(c);
            }
            if (lane < kMaskWords) {
                unsigned long long v = 0ull;
#pragma unroll
                for (int w = 0; w < kMaskWords; ++w) if (w == lane) v = mk.w[w];
                s_mask[(c0 + k) * kMaskWords + lane] = v;
            }
        }
    }
}

// Smallest box key at p over the candidate OBBs of `own` (a joint's mask in
// LDS) and its LDS record index (-1: none). Two candidates per step: both
// records are read before either test, so the LDS latency of one hides the
// other's. Ties keep the lower index.
__device__ __forceinline__ void grp_argmin(const float4* __restrict__ s_obb, const unsigned long long* own, bool valid,
                                           float px, float py, float pz, const uint8_t* __restrict__ acm_row,
                                           float& best, int& arg) {
    best = FLT_MAX;
    arg = -1;
#pragma unroll
    for (int k = 0; k < kMaskWords; ++k) {
        unsigned long long bits = valid ? own[k] : 0ull;
        while (bits) {
            const int oa = k * 64 + (__ffsll(bits) - 1);
            bits &= bits - 1ull;
            const bool two = bits != 0ull;
            const int ob = two ? k * 64 + (__ffsll(bits) - 1) : oa;
            if (two) bits &= bits - 1ull;
            const float4 a0 = s_obb[4 * oa], a1 = s_obb[4 * oa + 1], a2 = s_obb[4 * oa + 2], ah = s_obb[4 * oa + 3];
            const float4 b0 = s_obb[4 * ob], b1 = s_obb[4 * ob + 1], b2 = s_obb[4 * ob + 2], bh = s_obb[4 * ob + 3];
            float lx, ly, lz;
            to_box(a0, a1, a2, px, py, pz, lx, ly, lz);
            float ka = box_key(lx, ly, lz, ah);
            to_box(b0, b1, b2, px, py, pz, lx, ly, lz);
            float kb = box_key(lx, ly, lz, bh);
            if (acm_row) {
                if (acm_row[__float_as_int(ah.w)] == 0) ka = FLT_MAX;
                if (acm_row[__float_as_int(bh.w)] == 0) kb = FLT_MAX;
            }
            if (ka < best) { best = ka; arg = oa; }        // oa < ob: ties keep the lower index
            if (two && kb < best) { best = kb; arg = ob; }
        }
    }
}

// WORLD: bit 0 the scene's boxes may enter (OBB hinge / argmin), bit 1 its
// ESDF layers may; the launch picks the smallest variant the call needs, which
// keeps the other path's registers out of the sphere pass. TILED: some
// problem's OBBs pass through LDS in tiles (a variant of its own, so the
// folding across tiles costs the one-staging kernels no registers).
template <bool GRAD, int WORLD, bool TILED>
__device__ __forceinline__ void cost_grp(const CostArgs& a) {
    extern __shared__ float4 s_lds[];
    const DRobot& R = a.R;
    const int J = R.n_joints, D = R.n_dof, S = R.n_sph;
    float4* s_obb = s_lds;
    unsigned long long* s_mask = reinterpret_cast<unsigned long long*>(s_lds + 4 * a.C.obb_cap);
    float* s_q = reinterpret_cast<float*>(s_mask + (size_t)R.n_clu * kMaskWords);
    int* s_end = reinterpret_cast<int*>(s_q + D * kGrpStates);   // s_q: [D][slot]
    // frames [slot][J][12], slot stride padded by 4 floats so the four
    // slots' frames of one joint start in different LDS banks
    const int fstride = J * 12 + 4;
    float* s_T = reinterpret_cast<float*>(s_end + 2 * kWave);
    int* s_par = reinterpret_cast<int*>(s_T + kGrpStates * fstride);   // [slot][J] pointer-jumping links
    float* s_w = reinterpret_cast<float*>(s_par + kGrpStates * J);    // GRAD: [slot][J][6] wrenches
    float* s_dq = s_w + kGrpStates * J * 6;                       // GRAD: [slot][kMaxDof + 1]
    // tiled problems: per state slot and emission slot, the smallest box key
    // over the tiles so far and its OBB (original index, -1: none)
    float* s_best = GRAD ? s_dq + kGrpStates * (kMaxDof + 1) : s_w;
    int* s_barg = reinterpret_cast<int*>(s_best + kGrpStates * S);
    const int lane = lane_id();
    const int sub = lane & (kGrp - 1);
    const int slot = lane / kGrp;
    float* Ts = s_T + slot * fstride;
    const bool use_obb = (WORLD & 1) && a.P.use_obb != 0 && !a.clearance_only;
    const bool use_esdf = (WORLD & 2) && (a.P.use_esdf != 0 || a.clearance_only) && a.S.n_voxel_layers > 0;
    const bool want_grad = GRAD && !a.clearance_only;
    const float wt = a.P.weight;
    const int cap = a.C.obb_cap;
    int cur_prob = -1, n_obb = 0, o0 = 0, o1 = 0, n_tiles = 1;
    int4 ch;
    ChunkCursor cur = chunk_cursor();
    const long n_states = state_count(a.X);
    const int n_chunks = chunk_count(a.C, n_states);
    while (next_chunk(a.C, n_states, n_chunks, cur, ch)) {
        if ((WORLD & 1) && ch.x != cur_prob) {
            problem_obbs(a.C, a.S, ch.x, o0, o1);
            n_tiles = TILED ? uniform(obb_tiles(o0, o1, cap)) : 1;
            if (n_tiles == 1) {
                __syncthreads();
                n_obb = uniform(stage_obbs(a.S, o0, o1, s_obb, cap));
            }
            cur_prob = ch.x;
        }
        const bool tiled = TILED && use_obb && n_tiles > 1;
        const long g = (long)ch.y + slot;
        const bool active = g < ch.z;
        stage_unit_ends(a.X, ch.w, s_end);
        if (want_grad)
            for (int i = lane; i < kGrpStates * (J * 6 + kMaxDof + 1); i += kWave) s_w[i] = 0.0f;
        __syncthreads();
        StateRef ref = load_state(a.X, D, g, active, s_q, ch.w, s_end, ch.x, slot, kGrpStates, sub == 0);
        __syncthreads();
        // 1a. joint frames (pointer jumping over the DFS tree)
        group_frames(R, s_q + slot, kGrpStates, Ts, s_par + slot * J, sub);
        // 1b. candidate masks per sphere cluster; a tiled problem's OBBs pass through
        // LDS tile by tile, each tile's argmins folded into s_best / s_barg
        // (tiles in index order and a strict < keep the lowest index on ties)
        if (use_obb && !tiled) grp_masks(R, s_obb, n_obb, Ts, a.P.world_margin, active, sub, s_mask);
        for (int t = 0; tiled && t < n_tiles; ++t) {
            const int t0 = o0 + t * cap;
            __syncthreads();
            n_obb = uniform(stage_obbs(a.S, t0, min(o1, t0 + cap), s_obb, cap));
            grp_masks(R, s_obb, n_obb, Ts, a.P.world_margin, active, sub, s_mask);
            __syncthreads();
            for (int r0 = 0; r0 < S; r0 += kGrp) {
                const int s = r0 + sub;
                const bool valid = active && s < S;
                const int sc = valid ? s : r0;
                const int jj = R.sph_jnt[sc];
                const float4* Tq = reinterpret_cast<const float4*>(Ts + 12 * jj);
                const float4 t0q = Tq[0], t1q = Tq[1], t2q = Tq[2];
                const float4 sp = R.sph[sc];
                const int sid = R.sph_id[sc];
                const float px = fmaf(t0q.x, sp.x, fmaf(t0q.y, sp.y, fmaf(t0q.z, sp.z, t0q.w)));
                const float py = fmaf(t1q.x, sp.x, fmaf(t1q.y, sp.y, fmaf(t1q.z, sp.z, t1q.w)));
                const float pz = fmaf(t2q.x, sp.x, fmaf(t2q.y, sp.y, fmaf(t2q.z, sp.z, t2q.w)));
                const uint8_t* acm_row = a.S.sphere_obb_acm ? a.S.sphere_obb_acm + (size_t)sid * a.S.n_obbs : nullptr;
                float best;
                int arg;
                grp_argmin(s_obb, s_mask + R.sph_clu[sc] * kMaskWords, valid, px, py, pz, acm_row, best, arg);
                if (valid) {
                    float* bk = s_best + slot * S + s;
                    int* ba = s_barg + slot * S + s;
                    if (t == 0 || best < *bk) {
                        *bk = best;
                        *ba = arg >= 0 ? __float_as_int(s_obb[4 * arg + 3].w) : -1;
                    }
                }
            }
        }
        float self_h = 0.0f, snx = 0.0f, sny = 0.0f, snz = 0.0f;
        int si = -1, sj = -1;
        if (a.self_in && active && !a.clearance_only) {
            const float* so = a.self_in + 8 * (size_t)ref.gs;
            self_h = fmaxf(a.P.self_margin - so[0], 0.0f);
            if (self_h > 0.0f) { si = __float_as_int(so[1]); sj = __float_as_int(so[2]); snx = so[3]; sny = so[4]; snz = so[5]; }
        }
        // 2. sphere pass, in rounds of 16 emission slots per state: each lane
        // places its sphere with its joint's frame and walks its cluster's
        // candidate mask for the argmin box (tiled: reads the folded argmin
        // and rebuilds its record from the scene when the gradient needs it).
        float world = 0.0f, minclr = 1000.0f;
        // In this pass lane = 4 * sphere slot + state: the four states of a
        // sphere walk its cluster's mask together, so each quarter-wave of an
        // LDS record read touches 4 records (broadcast) instead of 16 (bank
        // conflicts). Each state's spheres keep their lanes' relative order,
        // so sums and atomics run in the same order as with lane = 16 *
        // state + slot.
        const int rslot = lane & (kGrpStates - 1), rsub = lane >> 2;
        const bool ractive = (long)ch.y + rslot < ch.z;
        const float* Tr = s_T + rslot * fstride;
        const int r_si = __shfl(si, rslot * kGrp), r_sj = __shfl(sj, rslot * kGrp);
        const float r_snx = __shfl(snx, rslot * kGrp), r_sny = __shfl(sny, rslot * kGrp), r_snz = __shfl(snz, rslot * kGrp);
        // wrench of a sphere's point gradient g at p about its joint's origin o
        auto add_wrench = [&](int jj, int sid, float px, float py, float pz, float ox, float oy, float oz, float gx,
                              float gy, float gz) {
            if (r_si >= 0) {
                if (sid == r_si) { gx -= r_snx; gy -= r_sny; gz -= r_snz; }
                if (sid == r_sj) { gx += r_snx; gy += r_sny; gz += r_snz; }
            }
            if (gx != 0.0f || gy != 0.0f || gz != 0.0f) {
                const float rx = px - ox, ry = py - oy, rz = pz - oz;
                float* w = s_w + (rslot * J + jj) * 6;
                atomicAdd(w + 0, gx);
                atomicAdd(w + 1, gy);
                atomicAdd(w + 2, gz);
                atomicAdd(w + 3, ry * gz - rz * gy);
                atomicAdd(w + 4, rz * gx - rx * gz);
                atomicAdd(w + 5, rx * gy - ry * gx);
            }
        };
        for (int r0 = 0; r0 < S; r0 += kGrp) {
            const int s = r0 + rsub;
            const bool valid = ractive && s < S;
            const int sc = valid ? s : r0;
            const int jj = R.sph_jnt[sc];
            const float4* Tq = reinterpret_cast<const float4*>(Tr + 12 * jj);
            const float4 t0 = Tq[0], t1 = Tq[1], t2 = Tq[2];
            const float4 sp = R.sph[sc];
            const int sid = R.sph_id[sc];
            const float px = fmaf(t0.x, sp.x, fmaf(t0.y, sp.y, fmaf(t0.z, sp.z, t0.w)));
            const float py = fmaf(t1.x, sp.x, fmaf(t1.y, sp.y, fmaf(t1.z, sp.z, t1.w)));
            const float pz = fmaf(t2.x, sp.x, fmaf(t2.y, sp.y, fmaf(t2.z, sp.z, t2.w)));
            float best = FLT_MAX;
            int arg = -1;   // tiled: the argmin's original OBB index; else its LDS record
            if (tiled) {
                if (valid) {
                    best = s_best[rslot * S + s];
                    arg = s_barg[rslot * S + s];
                }
            } else if (use_obb) {
                const uint8_t* acm_row = a.S.sphere_obb_acm ? a.S.sphere_obb_acm + (size_t)sid * a.S.n_obbs : nullptr;
                grp_argmin(s_obb, s_mask + R.sph_clu[sc] * kMaskWords, valid, px, py, pz, acm_row, best, arg);
            }
            if (!valid) continue;
            auto rec = [&](float4& r0, float4& r1, float4& r2, float4& r3) {
                if (tiled) {
                    float4 t[4];
                    obb_record(a.S, arg, t);
                    r0 = t[0]; r1 = t[1]; r2 = t[2]; r3 = t[3];
                } else {
                    r0 = s_obb[4 * arg]; r1 = s_obb[4 * arg + 1]; r2 = s_obb[4 * arg + 2]; r3 = s_obb[4 * arg + 3];
                }
            };
            float gx, gy, gz;
            const float hs =
                sphere_hinge_at<GRAD>(a, arg >= 0, rec, best, px, py, pz, sp.w, sid, use_obb, use_esdf, gx, gy, gz);
            if (a.clearance_only) minclr = fminf(minclr, hs);
            else world += hs;
            if (want_grad) add_wrench(jj, sid, px, py, pz, t0.w, t1.w, t2.w, gx, gy, gz);
        }
        // each state's partial sums sit in the lanes with its lane & 3; the
        // pairwise tree matches a 16-lane group sum's
        for (int o = kGrpStates; o < kWave; o <<= 1) {
            world += __shfl_xor(world, o);
            minclr = fminf(minclr, __shfl_xor(minclr, o));
        }
        world = __shfl(world, slot);
        minclr = __shfl(minclr, slot);
        // 3. adjoint: subtree wrench of joint `sub` -> d cost / d q
        if (want_grad) {
            __syncthreads();
            float* dq = s_dq + slot * (kMaxDof + 1);
            // Subtree sums by one prefix scan over the group's lanes: lane
            // i holds joint i's force f_i and its moment about the world
            // origin m_i = tau_i + o_i x f_i; DFS subtrees are contiguous, so
            // joint j's subtree force and moment are P(jend[j] - 1) -
            // P(j - 1), and its moment about o_j is M - o_j x F.
            float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f, m0 = 0.0f, m1 = 0.0f, m2 = 0.0f;
            float ox = 0.0f, oy = 0.0f, oz = 0.0f;
            if (sub < J) {
                const float* w = s_w + (slot * J + sub) * 6;
                const float* Tj = Ts + 12 * sub;
                ox = Tj[3]; oy = Tj[7]; oz = Tj[11];
                f0 = w[0]; f1 = w[1]; f2 = w[2];
                m0 = w[3] + (oy * f2 - oz * f1);
                m1 = w[4] + (oz * f0 - ox * f2);
                m2 = w[5] + (ox * f1 - oy * f0);
            }
#pragma unroll
            for (int off = 1; off < kGrp; off <<= 1) {
                const float u0 = __shfl_up(f0, off, kGrp), u1 = __shfl_up(f1, off, kGrp), u2 = __shfl_up(f2, off, kGrp);
                const float u3 = __shfl_up(m0, off, kGrp), u4 = __shfl_up(m1, off, kGrp), u5 = __shfl_up(m2, off, kGrp);
                if (sub >= off) { f0 += u0; f1 += u1; f2 += u2; m0 += u3; m1 += u4; m2 += u5; }
            }
            const int gbase = lane - sub;
            const int e = sub < J ? R.jend[sub] : sub + 1;
            const int hi = gbase + e - 1, lo = gbase + (sub > 0 ? sub - 1 : 0);
            float F0 = __shfl(f0, hi), F1 = __shfl(f1, hi), F2 = __shfl(f2, hi);
            float M0 = __shfl(m0, hi), M1 = __shfl(m1, hi), M2 = __shfl(m2, hi);
            const float l0 = __shfl(f0, lo), l1 = __shfl(f1, lo), l2 = __shfl(f2, lo);
            const float l3 = __shfl(m0, lo), l4 = __shfl(m1, lo), l5 = __shfl(m2, lo);
            if (sub > 0) { F0 -= l0; F1 -= l1; F2 -= l2; M0 -= l3; M1 -= l4; M2 -= l5; }
            if (active && sub >= 1 && sub < J) {
                const int4 meta = R.jmeta[sub];
                if (meta.y >= 0 && meta.x >= 0) {
                    const float* Tj = Ts + 12 * sub;
                    const float tx = M0 - (oy * F2 - oz * F1), ty = M1 - (oz * F0 - ox * F2), tz = M2 - (ox * F1 - oy * F0);
                    const int type = meta.x;
                    const int ax = type >= kXRot ? type - kXRot : type;
                    const float zx = Tj[ax], zy = Tj[4 + ax], zz = Tj[8 + ax];
                    const float v = type >= kXRot ? fmaf(zx, tx, fmaf(zy, ty, zz * tz)) : fmaf(zx, F0, fmaf(zy, F1, zz * F2));
                    atomicAdd(dq + meta.y, v);
                }
            }
            __syncthreads();
            if (active && sub < D) {
                bool endpoint = false;
                size_t base = (size_t)ref.gs * D, stride = 1;
                if (a.X.source == kSrcTraj) {
                    endpoint = (ref.k == 0 || ref.k == ref.n - 1);
                    base = (size_t)a.X.traj_off[ref.unit] * D + ref.k;
                    stride = ref.n;
                }
                a.grad_out[base + (size_t)sub * stride] = endpoint ? 0.0f : wt * dq[sub];
            }
        }
        if (a.clearance_only) {
            if (active && sub == 0) a.cost_out[ref.gs] = minclr;
        } else {
            if (active && sub == 0 && a.cost_out) a.cost_out[ref.gs] = wt * (world + self_h);
        }
        __syncthreads();   // frames, masks and wrenches of this chunk are dead before the next one
    }
}

// Four waves per SIMD for the one-staging variants (the adjoint's 128 VGPRs);
// the tiled ones would spill there.
template <bool GRAD, int WORLD, bool TILED>
__global__ __launch_bounds__(64, TILED ? 1 : 4) void k_cost_grp(CostArgs a) {
    cost_grp<GRAD, WORLD, TILED>(a);
}

// ESDF-only launches (no OBB path): five waves per SIMD fit in the register
// budget, which the lookups' memory latency rewards.
template <bool GRAD>
__global__ __launch_bounds__(64, 5) void k_cost_grp_esdf(CostArgs a) {
    cost_grp<GRAD, 2, false>(a);
}

// ===========================================================================
// Self collision: 16 lanes per state, sphere positions in LDS.
// mode 0: any pair with |d|^2 - (ri+rj)^2 < 0 (utils.cuh:239-267,
//         runtime_kinematics.cuh:248-283);
// mode 1: min over pairs of |d| - ri - rj with argmin pair and unit normal
//         (runtime_kinematics.cuh:294-329).
// ===========================================================================
constexpr int kSelfStates = kWave / kSelfGroup;
static_assert(kSelfGroup >= kMaxJoints, "the self pass composes frames with one joint per lane of a state group");

__global__ __launch_bounds__(64) void k_self(SelfArgs a) {
    extern __shared__ float4 s_dyn4[];
    const DRobot& R = a.R;
    const int lane = lane_id();
    const int sub = lane & (kSelfGroup - 1);
    const int st = lane / kSelfGroup;
    const int S = R.n_sph, D = R.n_dof, J = R.n_joints;
    float4* s_pos = s_dyn4;                                          // [kSelfStates][S] (x, y, z, r)
    const int NC = R.n_clu;
    float4* s_cenc = s_dyn4 + kSelfStates * S;                        // [kSelfStates][NC] cluster enclosing spheres
    const int fstride = J * 12 + 4;                                   // frame block per state, padded
    float* s_T = reinterpret_cast<float*>(s_cenc + kSelfStates * NC); // [kSelfStates][J][12] joint frames
    float* s_q = s_T + kSelfStates * fstride;                         // [kMaxDof][kSelfStates]
    int* s_par = reinterpret_cast<int*>(s_q + kMaxDof * kSelfStates); // [kSelfStates][J] pointer-jumping links
    int* s_end = s_par + kSelfStates * J;                             // [2][64] unit ends / starts of a chunk
    float4* s_apos = reinterpret_cast<float4*>(s_end + 2 * kWave);    // gated mode 0: [kSelfStates][A] approx spheres
    const bool gated = a.mode == 0 && !R.self_gate_implied;
    const long n_states = state_count(a.X);
    // With the plan's chunk table a chunk names its first unit (staged unit
    // ends make the state -> unit lookup an LDS bisection) or, waypoint-major,
    // its problem; without it states go in flat order.
    const bool chunked = a.C.tab != nullptr;
    const long n_work = chunked ? (long)a.C.n_chunks : (n_states + kSelfStates - 1) / kSelfStates;
    for (long w = blockIdx.x; w < n_work; w += gridDim.x) {
    long base, g;
    bool active;
    int u0 = -1, prob = 0;
    if (chunked) {
        const int4 ch = a.C.tab[w];
        base = ch.y;
        g = base + st;
        active = g < ch.z;
        u0 = ch.w;
        prob = ch.x;
        stage_unit_ends(a.X, u0, s_end);
        __syncthreads();
    } else {
        base = w * kSelfStates;
        g = base + st;
        active = g < n_states;
        if (a.C.unit_hint) {
            u0 = uniform(a.C.unit_hint[base >> 6]);
            stage_unit_ends(a.X, u0, s_end);
            __syncthreads();
        }
    }
    // units already decided need no configuration, FK or pairs (the unit
    // first-hit mode runs over densified trajectories and edge midpoints)
    const bool by_unit = a.unit_first_hit && (a.X.source == kSrcDense || a.X.source == kSrcBfs);
    if (by_unit) {
        int b0, b1;
        const int u = find_unit(a.X.dense_off, a.X.n_traj, active ? g : base, u0, s_end, &b0, &b1);
        if (active && a.unit_first_hit[u] >= 0) active = false;
        if (!wave_any(active)) continue;
    }
    // Every lane of a state group resolves the state; lane 0 stores its q.
    const StateRef ref = load_state(a.X, D, active ? g : base, true, s_q, u0, s_end, prob, st, kSelfStates, sub == 0);
    if (a.unit_first_hit && !by_unit) {
        if (active && a.unit_first_hit[ref.unit] >= 0) active = false;
        if (!wave_any(active)) continue;
    }
    __syncthreads();
    float4* pos = s_pos + st * S;
    float4* cenc = s_cenc + st * NC;
    float* Ts = s_T + st * fstride;
    group_frames(R, s_q + st, kSelfStates, Ts, s_par + st * J, sub);
    // sphere centres (emission slot k: a joint's spheres are contiguous) and
    // each sphere cluster's enclosing sphere, from the frames
    for (int k = sub; k < S; k += kSelfGroup) {
        const float4 sp = R.sph[k];
        const float4* Tq = reinterpret_cast<const float4*>(Ts + 12 * R.sph_jnt[k]);
        const float4 t0 = Tq[0], t1 = Tq[1], t2 = Tq[2];
        pos[k] = make_float4(fmaf(t0.x, sp.x, fmaf(t0.y, sp.y, fmaf(t0.z, sp.z, t0.w))),
                             fmaf(t1.x, sp.x, fmaf(t1.y, sp.y, fmaf(t1.z, sp.z, t1.w))),
                             fmaf(t2.x, sp.x, fmaf(t2.y, sp.y, fmaf(t2.z, sp.z, t2.w))), sp.w);
    }
    for (int c = sub; c < NC; c += kSelfGroup) {
        const float4 ce = R.clu_enc[c];
        const float4* Tq = reinterpret_cast<const float4*>(Ts + 12 * R.clu_jnt[c]);
        const float4 t0 = Tq[0], t1 = Tq[1], t2 = Tq[2];
        cenc[c] = make_float4(fmaf(t0.x, ce.x, fmaf(t0.y, ce.y, fmaf(t0.z, ce.z, t0.w))),
                              fmaf(t1.x, ce.x, fmaf(t1.y, ce.y, fmaf(t1.z, ce.z, t1.w))),
                              fmaf(t2.x, ce.x, fmaf(t2.y, ce.y, fmaf(t2.z, ce.z, t2.w))), ce.w);
    }
    __syncthreads();
    bool live = active;   // states whose pairs are checked (gated: an approximate pair collides)
    if (gated) {
        // Approximate pairs first (two_phase_cc.cuh:114-131): a state none of
        // whose approximate pairs collide has no self collision. Their frames
        // come from the approximate chain when the model has its own (the
        // exact positions above no longer need the full chain's).
        if (R.aft) {
            group_frames<true>(R, s_q + st, kSelfStates, Ts, s_par + st * J, sub);
        }
        float4* apos = s_apos + st * R.n_apx;
        for (int k = sub; k < R.n_apx; k += kSelfGroup) {
            const float4 sp = R.apx_orig[k];
            const float4* Tq = reinterpret_cast<const float4*>(Ts + 12 * R.apx_jnt[k]);
            const float4 t0 = Tq[0], t1 = Tq[1], t2 = Tq[2];
            apos[k] = make_float4(fmaf(t0.x, sp.x, fmaf(t0.y, sp.y, fmaf(t0.z, sp.z, t0.w))),
                                  fmaf(t1.x, sp.x, fmaf(t1.y, sp.y, fmaf(t1.z, sp.z, t1.w))),
                                  fmaf(t2.x, sp.x, fmaf(t2.y, sp.y, fmaf(t2.z, sp.z, t2.w))), sp.w);
        }
        __syncthreads();
        bool ah = false;
        for (int p = sub; p < R.n_apx_pairs && live; p += kSelfGroup) {
            const int2 pr = R.apx_pairs[p];
            const float4 u = apos[pr.x], v = apos[pr.y];
            const float ux = u.x - v.x, uy = u.y - v.y, uz = u.z - v.z;
            const float r2 = u.w + v.w;
            ah = ah || fmaf(uz, uz, fmaf(uy, uy, ux * ux)) - r2 * r2 < 0.0f;
        }
        live = live && ((__ballot(ah) >> (st * kSelfGroup)) & 0xFFFFull) != 0ull;
    }
    // Pairs come in blocks by the sphere clusters of their two spheres; a
    // cluster's enclosing sphere contains all of its spheres, so a block
    // whose two enclosing spheres are at clearance >= band holds no pair
    // below band (band = 0 for the any-collision check, where "below" is the
    // strict d^2 < (r_i + r_j)^2 test). The 16 lanes of a state test 16
    // blocks at once; then every block that passes for some state of the
    // wave has its pair list walked, 16 pairs per step.
    const float band = a.mode == 0 ? 0.0f : a.band;
    auto near = [&](float4 e, float4 f) {
        const float dx = e.x - f.x, dy = e.y - f.y, dz = e.z - f.z;
        const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const float rs = e.w + f.w;
        return a.mode == 0 ? d2 < rs * rs : sqrtf(d2) - rs < band;
    };
    float best = 1000.0f;
    int bkey = INT_MAX;   // (block << 16 | pair) of the best pair: ties go to the first
    bool hit = false;
    // lane sub holds the entry of block b0 + sub, the next pass's loaded ahead;
    // a walked block's list comes from its lane in the first state group
    int4 nb4 = sub < R.n_cpb ? R.cpb[sub] : make_int4(0, 0, 0, 0);
    for (int b0 = 0; b0 < R.n_cpb; b0 += kSelfGroup) {
        const int bb = b0 + sub;
        const int4 hb = nb4;
        if (bb + kSelfGroup < R.n_cpb) nb4 = R.cpb[bb + kSelfGroup];
        const bool g = live && !hit && bb < R.n_cpb && near(cenc[hb.x], cenc[hb.y]);
        const unsigned long long bal = __ballot(g);
        const unsigned mine = (unsigned)(bal >> (st * kSelfGroup)) & 0xFFFFu;
        unsigned todo = (unsigned)(bal | (bal >> 16) | (bal >> 32) | (bal >> 48)) & 0xFFFFu;   // wave-uniform
        while (todo) {
            const int t = __builtin_ctz(todo);
            todo &= todo - 1u;
            const int b = b0 + t;
            const int first = __builtin_amdgcn_readlane(hb.z, t), n = __builtin_amdgcn_readlane(hb.w, t);
            const bool on = ((mine >> t) & 1u) != 0u && !hit;
            for (int k0 = 0; k0 < n; k0 += kSelfGroup) {
                const int k = k0 + sub;
                if (!on || k >= n) continue;
                const unsigned pr = R.cpl[first + k];
                const float4 u = pos[pr & 0xFFFFu], v = pos[pr >> 16];
                const float ux = u.x - v.x, uy = u.y - v.y, uz = u.z - v.z;
                const float e2 = fmaf(uz, uz, fmaf(uy, uy, ux * ux));
                if (a.mode == 0) {
                    const float r2 = u.w + v.w;
                    if (e2 - r2 * r2 < 0.0f) hit = true;
                } else {
                    const float cl = sqrtf(e2) - (u.w + v.w);
                    const int key = (b << 16) | k;
                    if (cl < best || (cl == best && key < bkey)) {
                        best = cl;
                        bkey = key;
                    }
                }
            }
            if (a.mode == 0) {
                // share the state's verdict so its lanes stop
                hit = ((__ballot(hit) >> (st * kSelfGroup)) & 0xFFFFull) != 0ull;
            }
        }
        if (a.mode == 0 && __ballot(live && !hit) == 0ull) break;
    }
    if (a.mode == 0) {
        if (active && sub == 0) {
            if (a.unit_first_hit) {
                if (hit) atomicMin((unsigned*)&a.unit_first_hit[ref.unit], (unsigned)ref.k);
            } else {
                a.hit_out[ref.gs] = hit ? 1 : 0;
            }
        }
    } else {
#pragma unroll
        for (int off = 1; off < kSelfGroup; off <<= 1) {
            const float ob = __shfl_xor(best, off, kSelfGroup);
            const int ok = __shfl_xor(bkey, off, kSelfGroup);
            if (ob < best || (ob == best && ok < bkey)) {
                best = ob;
                bkey = ok;
            }
        }
        if (active && sub == 0) {
            float* o = a.out + 8 * (size_t)ref.gs;
            o[0] = best;
            if (bkey != INT_MAX) {
                const unsigned pr = R.cpl[R.cpb[bkey >> 16].z + (bkey & 0xFFFF)];
                const int si = (int)(pr & 0xFFFFu), sj = (int)(pr >> 16);
                const float4 u = pos[si], v = pos[sj];
                const float dx = u.x - v.x, dy = u.y - v.y, dz = u.z - v.z;
                const float inv = 1.0f / sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
                o[1] = __int_as_float(R.sph_id[si]); o[2] = __int_as_float(R.sph_id[sj]);
                o[3] = dx * inv; o[4] = dy * inv; o[5] = dz * inv;
            } else {
                o[1] = __int_as_float(-1); o[2] = __int_as_float(-1);
                o[3] = o[4] = o[5] = 0.0f;
            }
            o[6] = o[7] = 0.0f;
        }
    }
    __syncthreads();   // positions of this round are dead before the next load
    }
}

// ===========================================================================
// FK outputs: every joint's world frame (fk_joint_transforms_runtime,
// planning/runtime_kinematics.cuh:97-171: column-major, out[c*4 + r] = T[r][c])
// and every sphere's world centre (fk_runtime). One lane per configuration.
// ===========================================================================
__global__ __launch_bounds__(64) void k_fk_out(DRobot R, DStates X, float* frames, float* spheres) {
    __shared__ float s_q[kMaxDof * kWave];
    const int lane = lane_id();
    const long g = (long)blockIdx.x * kWave + lane;
    const bool active = g < X.n_states;
    load_state(X, R.n_dof, g, active, s_q);
    __syncthreads();
    if (!active) return;
    const int J = R.n_joints;
    Slots sl;
    for (int jj = 0; jj < J; ++jj) {
        float T[12];
        int4 meta;
        fk_joint(R, jj, s_q + lane, sl, T, meta);
        if (frames) {
            float* o = frames + ((size_t)g * J + R.jorig[jj]) * 16;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                o[c * 4 + 0] = T[c];
                o[c * 4 + 1] = T[4 + c];
                o[c * 4 + 2] = T[8 + c];
                o[c * 4 + 3] = c == 3 ? 1.0f : 0.0f;
            }
        }
        if (spheres) {
            for (int s = R.sph_off[jj]; s < R.sph_off[jj + 1]; ++s) {
                const float4 sp = R.sph[s];
                float x, y, z;
                aff_point(T, sp.x, sp.y, sp.z, x, y, z);
                reinterpret_cast<float4*>(spheres)[(size_t)g * R.n_sph + R.sph_id[s]] = make_float4(x, y, z, sp.w);
            }
        }
    }
}

// ===========================================================================
// CHOMP optimiser (planning/chomp.cu:261-351, 459-557 semantics, batched)
// ===========================================================================
// Sub-states per segment of each trajectory: ceil(max segment length / step)
// when the longest segment exceeds the step, else 1 (chomp.cu:459-475, FP32 as
// the host code: squares summed in dof order, no contraction), capped at
// nsub_max (the joint-limit box diagonal over the step, which no clamped
// trajectory exceeds). One wave per trajectory.
__global__ __launch_bounds__(64) void k_chomp_nsub(ChompArgs a) {
    const int b = blockIdx.x;
    if (b >= a.X.n_traj) return;
    const int lane = lane_id();
    const int D = a.R.n_dof;
    const int t0 = a.X.traj_off[b], T = a.X.traj_off[b + 1] - t0;
    const float* th = a.theta + (size_t)t0 * D;
    float m = 0.0f;
    for (int t = lane; t < T - 1; t += kWave) {
        float d2 = 0.0f;
        for (int d = 0; d < D; ++d) {
            const float diff = __fsub_rn(th[(size_t)d * T + t + 1], th[(size_t)d * T + t]);
            d2 = __fadd_rn(d2, __fmul_rn(diff, diff));
        }
        m = fmaxf(m, __fsqrt_rn(d2));
    }
    m = wave_max(m);
    if (lane == 0) {
        int n = 1;
        if (a.check_step > 0.0f && m > a.check_step)
            n = (int)fminf(ceilf(__fdiv_rn(m, a.check_step)), (float)a.nsub_max);
        a.nsub[b] = n;
    }
}

// dense_off = exclusive scan of (T_b - 1) nsub[b] + 1 (dense_off[B] = total).
// One workgroup of 256 threads, each scanning a contiguous slice.
__global__ __launch_bounds__(256) void k_chomp_scan(ChompArgs a) {
    __shared__ int part[256];
    const int B = a.X.n_traj, tid = threadIdx.x;
    const int per = (B + 255) / 256, lo = tid * per, hi = min(B, lo + per);
    int acc = 0;
    for (int b = lo; b < hi; ++b) acc += (a.X.traj_off[b + 1] - a.X.traj_off[b] - 1) * a.nsub[b] + 1;
    part[tid] = acc;
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int k = 0; k < 256; ++k) {
            const int v = part[k];
            part[k] = run;
            run += v;
        }
        a.dense_off[B] = run;
    }
    __syncthreads();
    int run = part[tid];
    for (int b = lo; b < hi; ++b) {
        a.dense_off[b] = run;
        run += (a.X.traj_off[b + 1] - a.X.traj_off[b] - 1) * a.nsub[b] + 1;
    }
}

// Per-trajectory bookkeeping after an evaluation (chomp.cu:484-557): best
// valid snapshot, convergence streak on the best cost, polish count, done.
// One wave per trajectory (the snapshot copy uses every lane).
__global__ __launch_bounds__(64) void k_chomp_book(ChompArgs a, int initial) {
    const int b = blockIdx.x;
    if (b >= a.X.n_traj) return;
    const int lane = lane_id();
    const int D = a.R.n_dof;
    float* rec = a.record + 8 * (size_t)b;
    const float total = a.total[b];
    const bool valid = a.first_hit[b] < 0;
    bool copy = false;
    if (initial) {
        copy = valid;
    } else {
        copy = rec[PPL_CHOMP_DONE] == 0.0f && valid && total < rec[PPL_CHOMP_BEST_COST];
    }
    __syncthreads();
    if (lane == 0) {
        if (initial) {
            rec[PPL_CHOMP_BEST_COST] = valid ? total : INFINITY;
            rec[PPL_CHOMP_PREV_COST] = INFINITY;
            rec[PPL_CHOMP_SOLVED] = valid ? 1.0f : 0.0f;
            rec[PPL_CHOMP_STREAK] = 0.0f;
            rec[PPL_CHOMP_POLISH_LEFT] = (float)a.polish0;
            rec[PPL_CHOMP_DONE] = 0.0f;
            rec[PPL_CHOMP_ITERS] = 0.0f;
        } else if (rec[PPL_CHOMP_DONE] == 0.0f) {
            rec[PPL_CHOMP_ITERS] += 1.0f;
            if (copy) {
                rec[PPL_CHOMP_BEST_COST] = total;
                rec[PPL_CHOMP_SOLVED] = 1.0f;
            }
            if (rec[PPL_CHOMP_SOLVED] != 0.0f) {
                const float impr = rec[PPL_CHOMP_PREV_COST] - rec[PPL_CHOMP_BEST_COST];
                rec[PPL_CHOMP_STREAK] = (impr >= 0.0f && impr < a.convergence_eps) ? rec[PPL_CHOMP_STREAK] + 1.0f : 0.0f;
                rec[PPL_CHOMP_PREV_COST] = rec[PPL_CHOMP_BEST_COST];
                rec[PPL_CHOMP_POLISH_LEFT] -= 1.0f;
                if (rec[PPL_CHOMP_STREAK] >= 2.0f) rec[PPL_CHOMP_DONE] = 1.0f;
            }
        }
        // the next iteration starts only while no valid result is polished out
        if (rec[PPL_CHOMP_SOLVED] != 0.0f && rec[PPL_CHOMP_POLISH_LEFT] <= 0.0f) rec[PPL_CHOMP_DONE] = 1.0f;
        rec[PPL_CHOMP_LAST_COST] = total;
    }
    if (copy) {
        const size_t o = (size_t)a.X.traj_off[b] * D;
        const int n = (a.X.traj_off[b + 1] - a.X.traj_off[b]) * D;
        for (int i = lane; i < n; i += kWave) a.best[o + i] = a.theta[o + i];
    }
}

// Covariant step for one (dof, trajectory) (chomp.cu:261-290):
//   g = grad_d + lambda R theta_d with R theta = dt A^T (A theta) for the
//   5-point stencil A on the zero-padded grid, step = A^-1 g,
//   theta_d[t] -= step_size step[t] for 0 < t < T-1, clamped to the limits.
__global__ __launch_bounds__(256) void k_chomp_update(ChompArgs a) {
    extern __shared__ float s_u[];
    const int d = blockIdx.x, b = blockIdx.y;
    const float* rec = a.record + 8 * (size_t)b;
    if (rec[PPL_CHOMP_DONE] != 0.0f) return;
    const int D = a.R.n_dof;
    const int t0 = a.X.traj_off[b], T = a.X.traj_off[b + 1] - t0;
    float* x = s_u;             // [T + 8], theta with 4 zeros each side
    float* v = x + T + 8;       // [T + 4], A theta on [-2, T + 2)
    float* g = v + T + 4;       // [T]
    float* row = a.theta + (size_t)t0 * D + (size_t)d * T;
    const float* gr = a.grad + (size_t)t0 * D + (size_t)d * T;
    const float dt = a.dt;
    const float inv = 1.0f / (12.0f * dt * dt);
    for (int i = threadIdx.x; i < T + 8; i += blockDim.x) x[i] = (i >= 4 && i < T + 4) ? row[i - 4] : 0.0f;
    __syncthreads();
    for (int i = threadIdx.x; i < T + 4; i += blockDim.x) {
        const float* c = x + i + 2;   // centre at grid index i - 2
        v[i] = (-c[-2] + 16.0f * c[-1] - 30.0f * c[0] + 16.0f * c[1] - c[2]) * inv;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        const float* c = v + t + 2;
        const float rt = dt * ((-c[-2] + 16.0f * c[-1] - 30.0f * c[0] + 16.0f * c[1] - c[2]) * inv);
        g[t] = gr[t] + a.smoothness_weight * rt;
    }
    __syncthreads();
    const float* ai = a.ainv[b];
    const float2 lim = a.R.lim[d];
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        float step = 0.0f;
        for (int u = 0; u < T; ++u) step = fmaf(ai[(size_t)u * T + t], g[u], step);   // A^-1 symmetric
        if (t == 0 || t == T - 1) continue;
        const float nv = x[t + 4] - a.step_size * step;
        row[t] = fmaxf(lim.x, fminf(lim.y, nv));
    }
}

// ===========================================================================
// STOMP optimiser (planning/stomp.cu:565-930). Population rows are [D][T_b]
// blocks; every kernel finds a slot's row through stomp_row.
// ===========================================================================
__device__ __forceinline__ void stomp_row(const StompPop& p, int off_b, int T, int D, int r, int n_roll,
                                          const float** x, const float** cost) {
    size_t base;
    if (r < p.nf) {
        base = (size_t)p.nf * off_b + (size_t)r * T;
        *x = p.fresh + base * D;
        *cost = p.fresh_cost + base;
    } else if (r < n_roll - 1) {
        base = (size_t)p.nkeep * off_b + (size_t)(r - p.nf) * T;
        *x = p.keep + base * D;
        *cost = p.keep_cost + base;
    } else {
        *x = p.slot + (size_t)off_b * D;
        *cost = p.slot_cost + off_b;
    }
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of two
// 32 x 32 -> 64-bit products, the key advanced by the Weyl constants.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, unsigned k0, unsigned k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        if (i) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const unsigned hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const unsigned hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    }
    return c;
}

// z ~ N(0, 1) for waypoint t of row (b, r, d): Box-Muller on the Philox words
// of counter (t >> 2, r | d << 16, b, iter); t % 4 picks the pair and cos / sin.
__device__ __forceinline__ float stomp_normal(const StompArgs& a, int b, int r, int d, int t) {
    const unsigned it = a.iter_dev ? *a.iter_dev + a.iter : a.iter;
    const uint4 w = philox4x32_10(make_uint4((unsigned)t >> 2, (unsigned)r | ((unsigned)d << 16), (unsigned)b, it),
                                  a.key0, a.key1);
    const int j = t & 3;
    const unsigned w1 = j < 2 ? w.x : w.z, w2 = j < 2 ? w.y : w.w;
    const float u1 = ((float)(w1 >> 9) + 0.5f) * 1.1920928955078125e-7f;   // (0, 1), exact
    const float u2 = ((float)(w2 >> 9) + 0.5f) * 1.1920928955078125e-7f;
    const float rad = sqrtf(-2.0f * logf(u1));
    float sn, cs;
    sincospif(2.0f * u2, &sn, &cs);
    return rad * ((j & 1) ? sn : cs);
}

// Kept rollouts (stomp.cu:131-170): the n_old cheapest slots of the previous
// population by total cost (state sum + control sum), ties to the lower slot,
// copied with their state costs into slots nf.. of the current one.
__global__ __launch_bounds__(64) void k_stomp_select(StompArgs a) {
    extern __shared__ float s_tot[];
    const int b = blockIdx.x;
    if (b >= a.B) return;
    const int lane = lane_id();
    const int D = a.R.n_dof, off = a.traj_off[b], T = a.traj_off[b + 1] - off;
    for (int r = lane; r < a.n_roll; r += kWave) {
        const float *x, *c;
        stomp_row(a.prev, off, T, D, r, a.n_roll, &x, &c);
        float s = 0.0f;
        for (int t = 0; t < T; ++t) s += c[t];
        const float* cs = a.ctrl_sum + ((size_t)b * a.n_roll + r) * D;
        for (int d = 0; d < D; ++d) s += cs[d];
        s_tot[r] = s == s ? s : INFINITY;
    }
    __syncthreads();
    for (int r = lane; r < a.n_roll; r += kWave) {
        const float v = s_tot[r];
        int rank = 0;
        for (int j = 0; j < a.n_roll; ++j) {
            const float u = s_tot[j];
            rank += (u < v || (u == v && j < r)) ? 1 : 0;
        }
        if (rank < a.n_old) a.sel[(size_t)b * a.n_old + rank] = r;
    }
}

__global__ __launch_bounds__(256) void k_stomp_gather(StompArgs a) {
    const int k = blockIdx.x, b = blockIdx.y;
    const int D = a.R.n_dof, off = a.traj_off[b], T = a.traj_off[b + 1] - off;
    const int src = a.sel[(size_t)b * a.n_old + k];
    const float *x, *c;
    stomp_row(a.prev, off, T, D, src, a.n_roll, &x, &c);
    const size_t base = (size_t)a.n_old * off + (size_t)k * T;
    float* dx = a.keep + base * D;
    float* dc = a.keep_cost + base;
    for (int i = threadIdx.x; i < D * T; i += blockDim.x) dx[i] = x[i];
    for (int i = threadIdx.x; i < T; i += blockDim.x) dc[i] = c[i];
}

// The optimum as the last slot (stomp.cu:172-189).
__global__ __launch_bounds__(256) void k_stomp_seed(StompArgs a) {
    const int b = blockIdx.x;
    const int D = a.R.n_dof, off = a.traj_off[b], T = a.traj_off[b + 1] - off;
    const size_t o = (size_t)off * D;
    for (int i = threadIdx.x; i < D * T; i += blockDim.x) a.slot[o + i] = a.theta[o + i];
    for (int i = threadIdx.x; i < T; i += blockDim.x) a.slot_cost[off + i] = a.theta_cost[off + i];
}

// Fresh sample rows (stomp.cu:100-127): x = clamp(theta + L z) with the
// endpoints pinned; one wave per (dof, slot, trajectory).
__global__ __launch_bounds__(64) void k_stomp_sample(StompArgs a) {
    extern __shared__ float s_z[];
    const int d = blockIdx.x, r = blockIdx.y, b = blockIdx.z;
    const int lane = lane_id();
    const int D = a.R.n_dof, off = a.traj_off[b], T = a.traj_off[b + 1] - off;
    for (int t = lane; t < T; t += kWave) s_z[t] = stomp_normal(a, b, r, d, t);
    __syncthreads();
    const float* Lt = a.Lt[b];
    const float* th = a.theta + (size_t)off * D + (size_t)d * T;
    float* out = a.fresh + ((size_t)a.cur.nf * off + (size_t)r * T) * D + (size_t)d * T;
    const float2 lim = a.R.lim[d];
    for (int t = lane; t < T; t += kWave) {
        float e = 0.0f;
        for (int k = 0; k <= t; ++k) e += Lt[(size_t)k * T + t] * s_z[k];
        if (t == 0 || t == T - 1) e = 0.0f;
        out[t] = fmaxf(lim.x, fminf(lim.y, th[t] + e));
    }
}

// Per (dof, trajectory): control cost of every population row, the
// exponentiated-cost probabilities per waypoint, the weighted noise average,
// its M projection and the unconditional step (stomp.cu:255-403, 406-460).
// Control: w x_t (R x)_t with R = dt A^T A, (A^T A x)_t the 9-point
// autocorrelation stencil of the 5-point acceleration rule, zero beyond the
// ends (the padded-grid block of stomp_precompute.cpp:24-44).
__global__ __launch_bounds__(256) void k_stomp_update(StompArgs a) {
    extern __shared__ float s_u[];
    const int d = blockIdx.x, b = blockIdx.y;
    const int D = a.R.n_dof, off = a.traj_off[b], T = a.traj_off[b + 1] - off;
    const int NR = a.n_roll;
    float* tot = s_u;             // [NR][T]: total cost, then the unnormalised weight
    float* delta = tot + NR * T;  // [T]
    const int lane = lane_id(), wave = threadIdx.x / kWave, n_waves = blockDim.x / kWave;
    const float scale = 1.0f / (144.0f * a.dt * a.dt * a.dt);
    for (int r = wave; r < NR; r += n_waves) {
        const float *x, *c;
        stomp_row(a.cur, off, T, D, r, NR, &x, &c);
        x += (size_t)d * T;
        auto X = [&](int k) { return (k >= 0 && k < T) ? x[k] : 0.0f; };
        float part = 0.0f;
        for (int t = lane; t < T; t += kWave) {
            const float xt = x[t];
            // the weights sum to zero, so the stencil is taken over x_{t+k} - x_t:
            // FP32 error then scales with the local variation, not with |x|
            auto P2 = [&](int k) { return (X(t - k) - xt) + (X(t + k) - xt); };
            const float st = P2(4) - 32.0f * P2(3) + 316.0f * P2(2) - 992.0f * P2(1);
            const float ctrl = a.w_u * xt * (st * scale);
            tot[r * T + t] = c[t] + ctrl;
            part += ctrl;
        }
        part = wave_sum(part);
        if (lane == 0) a.ctrl_sum[((size_t)b * NR + r) * D + d] = part;
    }
    __syncthreads();
    const float* th = a.theta + (size_t)off * D + (size_t)d * T;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        float mn = tot[t], mx = tot[t];
        for (int r = 1; r < NR; ++r) {
            mn = fminf(mn, tot[r * T + t]);
            mx = fmaxf(mx, tot[r * T + t]);
        }
        const float denom = fmaxf(mx - mn, 1.0e-8f);
        float sum = 0.0f;
        for (int r = 0; r < NR; ++r) {
            const float u = expf(-a.h * (tot[r * T + t] - mn) / denom);
            tot[r * T + t] = u;
            sum += u;
        }
        sum = fmaxf(sum, 1.0e-12f);
        const float tht = th[t];
        float acc = 0.0f;
        for (int r = 0; r < NR; ++r) {
            const float *x, *c;
            stomp_row(a.cur, off, T, D, r, NR, &x, &c);
            acc += (x[(size_t)d * T + t] - tht) * (tot[r * T + t] / sum);
        }
        delta[t] = (t == 0 || t == T - 1) ? 0.0f : acc;
    }
    __syncthreads();
    if (a.record[8 * (size_t)b + PPL_CHOMP_DONE] != 0.0f) return;
    const float* Mt = a.Mt[b];
    float* row = a.theta + (size_t)off * D + (size_t)d * T;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        float acc = 0.0f;
        for (int u = 0; u < T; ++u) acc += Mt[(size_t)u * T + t] * delta[u];
        if (t != 0 && t != T - 1) row[t] = th[t] + acc;
    }
}

// ===========================================================================
// MIT* substrate consumers (planning/MITStar.cu:212-329, 832-1375)
// ===========================================================================
// Per edge: FP32 length in the kernels' order, segment count
// max(1, ceil(len / seg)), and the breadth-first midpoints still to check:
// [done, min(sparse, total - 1)), or [done, total - 1) when sparse < 0.
__global__ void k_bfs_count(const float* nodes, int dim, const int* src, const int* dst, const int* done, int n_edges,
                            float seg_len, int sparse, int* k0, int* cnt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_edges) return;
    const float* a = nodes + (size_t)src[e] * dim;
    const float* b = nodes + (size_t)dst[e] * dim;
    float len = 0.0f;
    for (int i = 0; i < dim; ++i) {
        const float diff = __fsub_rn(a[i], b[i]);
        len = fmaf(diff, diff, len);
    }
    const int total = max(1, (int)ceilf(__fdiv_rn(__fsqrt_rn(len), seg_len)));
    const int target = sparse >= 0 ? min(sparse, total - 1) : total - 1;
    const int start = done ? done[e] : 0;
    k0[e] = start;
    cnt[e] = max(0, target - start);
}

// off = exclusive scan of cnt (off[n] = total): one workgroup of 1024 threads
// takes passes of 16,384 counts, 16 consecutive counts per thread loaded
// together (one memory latency per pass), scanned in registers, then across
// the workgroup by wave scans of the thread totals. Offsets are only used when
// the 64-bit total (*total64, summed separately) is below 2^31, so every
// partial sum then fits 32 bits.
constexpr int kScanPer = 16;
__global__ __launch_bounds__(1024) void k_scan_i32(const int* cnt, int n, int* off, long long* total64) {
    __shared__ int wsum[16];
    __shared__ long long part[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int carry = 0;
    long long acc = 0;
    for (int base = 0; base < n; base += 1024 * kScanPer) {
        const int i0 = base + tid * kScanPer;
        int v[kScanPer];
        int local = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) v[k] = i0 + k < n ? cnt[i0 + k] : 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            local += v[k];
            acc += v[k];
        }
        int x = local;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int t = __shfl_up(x, o);
            if (lane >= o) x += t;
        }
        if (lane == kWave - 1) wsum[wave] = x;
        __syncthreads();
        int before = 0, tile = 0;
        for (int w = 0; w < 16; ++w) {
            const int y = wsum[w];
            before += w < wave ? y : 0;
            tile += y;
        }
        int run = carry + before + x - local;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            if (i0 + k < n) off[i0 + k] = run;
            run += v[k];
        }
        carry += tile;
        __syncthreads();
    }
    // 64-bit total, reduced from 32-bit halves
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const int lo = __shfl_xor((int)(unsigned)(acc & 0xffffffffll), o);
        const int hi = __shfl_xor((int)(acc >> 32), o);
        acc += ((long long)hi << 32) | (unsigned)lo;
    }
    if (lane == 0) part[wave] = acc;
    __syncthreads();
    if (tid == 0) {
        long long t = 0;
        for (int w = 0; w < 16; ++w) t += part[w];
        *total64 = t;
        off[n] = (int)min(t, (long long)INT_MAX);
    }
}

// hint[c] = the unit holding state 64 c, for every 64-state block below the
// total (one thread per block; c < n_hint, the buffer's capacity)
__global__ void k_unit_hint(const int* off, int n, int* hint, int n_hint) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_hint) return;
    const long g = (long)c << 6;
    if (g >= off[n]) return;
    hint[c] = upper_bound_idx(off, n, g);
}

__global__ void k_bfs_finish(const int* first_hit, const int* cnt, int n, uint8_t* cc, int* new_checks) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const bool hit = first_hit[e] >= 0;
    cc[e] = hit ? 1 : 0;
    if (new_checks) new_checks[e] = hit ? 0 : cnt[e];   // only verified-free midpoints count
}

// batch_cc: clearance of the free samples only (weight > 0: the min
// clearance, else 1000); colliding samples keep their entry.
__global__ void k_sample_clearance(const uint8_t* cc, const float* clr, int n, float weight, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || cc[i]) return;
    out[i] = weight > 0.0f ? clr[i] : 1000.0f;
}

__global__ void k_gather_rows(const float* src, int dim, const int* idx, int base, int n, float* out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * dim) return;
    const int r = (int)(i / dim), d = (int)(i % dim);
    const int node = idx ? idx[r] : base + r;
    out[i] = src[(size_t)node * dim + d];
}

__global__ void k_scatter(const float* v, const int* idx, int base, int n, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[idx ? idx[i] : base + i] = v[i];
}

// ===========================================================================
// Small helpers
// ===========================================================================
__global__ void k_add_u32(unsigned* p, unsigned v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}

__global__ void k_fill_i32(int* p, int v, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// out[b] = sum of x over segment b (one wave per segment, fixed order).
__global__ __launch_bounds__(64) void k_segment_sum(const float* x, const int* off, int n_seg, float* out) {
    const int b = blockIdx.x;
    if (b >= n_seg) return;
    const int lane = lane_id();
    float acc = 0.0f;
    for (int i = off[b] + lane; i < off[b + 1]; i += kWave) acc += x[i];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) out[b] = acc;
}

// Control cost w * sum_d theta_d^T R theta_d for the 5-point acceleration
// metric R = dt * (A^T A)[pad..pad+T] (chomp_precompute.cpp:20-38), evaluated
// as dt * sum_i acc_i^2 over the stencil rows that touch the trajectory,
// acc_i = (-x[i-2] + 16x[i-1] - 30x[i] + 16x[i+1] - x[i+2]) / (12 dt^2) with
// x = 0 outside [0, T). One wave per trajectory; with `stage` the wave first
// copies its trajectory's D x T block into LDS rows padded by four zeros on
// each side (eight loads per lane in flight at a time), then evaluates the
// stencil from LDS. state (nullable): the trajectory's per-waypoint state
// costs, whose sum joins the total (the STOMP rollout total in one launch).
__global__ __launch_bounds__(64) void k_control(const float* theta, const int* traj_off, int n_traj, int D, float dt,
                                                float w, const float* state, float* out, int accumulate, int stage) {
    extern __shared__ float s_x[];
    const int b = blockIdx.x;
    if (b >= n_traj) return;
    const int lane = lane_id();
    const int t0 = traj_off[b], Tn = traj_off[b + 1] - t0;
    const float* th = theta + (size_t)t0 * D;
    const float inv = 1.0f / (12.0f * dt * dt);
    float ssum = 0.0f;
    if (state)
        for (int i = lane; i < Tn; i += kWave) ssum += state[t0 + i];
    float acc = 0.0f;
    if (stage) {
        const int rs = Tn + 8, n = D * rs;
        for (int k0 = 0; k0 < n; k0 += 8 * kWave) {
            float v[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int k = k0 + c * kWave + lane;
                const int d = k / rs, i = k - d * rs - 4;
                v[c] = (k < n && i >= 0 && i < Tn) ? th[(size_t)d * Tn + i] : 0.0f;
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int k = k0 + c * kWave + lane;
                if (k < n) s_x[k] = v[c];
            }
        }
        __syncthreads();
        for (int d = 0; d < D; ++d) {
            const float* x = s_x + d * rs + 4;
            for (int i = lane - 2; i < Tn + 2; i += kWave) {
                const float v = (-x[i - 2] + 16.0f * x[i - 1] - 30.0f * x[i] + 16.0f * x[i + 1] - x[i + 2]) * inv;
                acc = fmaf(v, v, acc);
            }
        }
    } else {
        for (int d = 0; d < D; ++d) {
            const float* x = th + (size_t)d * Tn;
            for (int i = lane - 2; i < Tn + 2; i += kWave) {
                auto X = [&](int k) { return (k >= 0 && k < Tn) ? x[k] : 0.0f; };
                const float v = (-X(i - 2) + 16.0f * X(i - 1) - 30.0f * X(i) + 16.0f * X(i + 1) - X(i + 2)) * inv;
                acc = fmaf(v, v, acc);
            }
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        acc += __shfl_xor(acc, o);
        ssum += __shfl_xor(ssum, o);
    }
    if (lane == 0) {
        const float c = w * dt * acc;
        const float base = accumulate ? out[b] : 0.0f;
        out[b] = state ? base + (ssum + c) : base + c;
    }
}

// ESDF node values = -min_o sdf_o(node), positive inside. One thread per node.
__global__ __launch_bounds__(256) void k_esdf_build(const float* obb_dims, const float* obb_pose, int n_obbs,
                                                    float4 prm, float4 gp0, float4 gp1, int nx, int ny, int nz, float* out) {
    const float gpose[8] = {gp0.x, gp0.y, gp0.z, gp0.w, gp1.x, gp1.y, gp1.z, gp1.w};
    __shared__ float4 s_obb[4 * kMaxObbLds];
    // OBBs [t0, t1) -> LDS records (the collision kernels' format)
    auto stage = [&](int t0, int t1) {
        for (int o = t0 + (int)threadIdx.x; o < t1; o += blockDim.x) {
            const float* d = obb_dims + 4 * o;
            const float* p = obb_pose + 8 * o;
            const float w = p[3], x = p[4], y = p[5], z = p[6];
            float4* r = s_obb + 4 * (o - t0);
            r[0] = make_float4(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - w * z), 2.0f * (x * z + w * y), p[0]);
            r[1] = make_float4(2.0f * (x * y + w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - w * x), p[1]);
            r[2] = make_float4(2.0f * (x * z - w * y), 2.0f * (y * z + w * x), 1.0f - 2.0f * (x * x + y * y), p[2]);
            r[3] = make_float4(0.5f * d[0], 0.5f * d[1], 0.5f * d[2], 0.0f);
        }
    };
    // more OBBs than one staging holds: every block batch of voxels walks
    // the OBBs tile by tile (block-uniform trip counts around the barriers)
    const bool tiled = n_obbs > kMaxObbLds;
    if (!tiled) {
        stage(0, n_obbs);
        __syncthreads();
    }
    const long n = (long)nx * ny * nz;
    for (long base = (long)blockIdx.x * blockDim.x; base < n; base += (long)gridDim.x * blockDim.x) {
        const long i = base + threadIdx.x;
        const bool in = i < n;
        const int ix = (int)(i / ((long)ny * nz));
        const int iy = (int)((i / nz) % ny);
        const int iz = (int)(i % nz);
        const float lx = ix * prm.w - 0.5f * prm.x - gpose[0];
        const float ly = iy * prm.w - 0.5f * prm.y - gpose[1];
        const float lz = iz * prm.w - 0.5f * prm.z - gpose[2];
        // grid-local -> world with the conjugate of the stored rotation
        const float qw = gpose[3], qx = -gpose[4], qy = -gpose[5], qz = -gpose[6];
        const float tx = 2.0f * (qy * lz - qz * ly), ty = 2.0f * (qz * lx - qx * lz), tz = 2.0f * (qx * ly - qy * lx);
        const float px = lx + qw * tx + (qy * tz - qz * ty);
        const float py = ly + qw * ty + (qz * tx - qx * tz);
        const float pz = lz + qw * tz + (qx * ty - qy * tx);
        float best = FLT_MAX;
        for (int t0 = 0; t0 < n_obbs; t0 += kMaxObbLds) {
            const int t1 = min(n_obbs, t0 + kMaxObbLds);
            if (tiled) {
                __syncthreads();
                stage(t0, t1);
                __syncthreads();
            }
            for (int o = 0; o < t1 - t0; ++o) {
                float bx, by, bz;
                to_box(s_obb[4 * o], s_obb[4 * o + 1], s_obb[4 * o + 2], px, py, pz, bx, by, bz);
                best = fminf(best, box_key(bx, by, bz, s_obb[4 * o + 3]));
            }
        }
        if (in) out[i] = -key_to_sdf(best);
    }
}

int g_cus = 0;

}  // namespace

int device_cus() {
    if (g_cus == 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        (void)hipGetDeviceProperties(&prop, dev);
        g_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    }
    return g_cus;
}

namespace {

int occupancy_grid(const void* fn, size_t dyn_smem) {
    (void)device_cus();
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, dyn_smem) != hipSuccess || per_cu <= 0) per_cu = 4;
    return per_cu * g_cus;
}

}  // namespace

size_t lds_bytes(int obb_cap, int n_dof) {
    return (size_t)obb_cap * 64 + (size_t)(n_dof + 2) * kWave * sizeof(float);
}

size_t lds_bytes_cost(int obb_cap, int n_dof, int n_joints, int G, bool grad) {
    size_t b = lds_bytes(obb_cap, n_dof);
    if (grad) b += (size_t)(kWave / G) * (rec_stride(n_joints) + dq_stride(n_dof)) * sizeof(float);
    return b;
}

int grid_collide(size_t lds) {
    static size_t last = (size_t)-1;
    static int g = 0;
    if (lds != last) { g = occupancy_grid((const void*)k_collide, lds); last = lds; }
    return g;
}

int grid_cost(int n_joints, bool grad, size_t lds) {
    (void)n_joints;
    static size_t last[2] = {(size_t)-1, (size_t)-1};
    static int g[2] = {0, 0};
    const int idx = grad ? 1 : 0;
    if (lds != last[idx]) {
        g[idx] = occupancy_grid(grad ? (const void*)k_cost<true> : (const void*)k_cost<false>, lds);
        last[idx] = lds;
    }
    return g[idx];
}

// A launch with its argument block; with t0/t1 the runtime records the two
// events at the kernel's own start and end (timing without event packets of
// their own around the launch).
static hipError_t launch_args(const void* k, int grid, size_t lds, hipStream_t s, void* arg, hipEvent_t t0,
                              hipEvent_t t1) {
    void* params[] = {arg};
    if (t0) return hipExtLaunchKernel(k, dim3(grid), dim3(64), params, lds, s, t0, t1, 0);
    return hipLaunchKernel(k, dim3(grid), dim3(64), params, lds, s);
}

hipError_t launch_collide(const CollideArgs& a, int grid, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
    CollideArgs arg = a;
    const void* k = a.R.scene_gate_fused ? (const void*)k_collide : (const void*)k_collide_gated;
    return launch_args(k, grid, lds_bytes(a.C.obb_cap, a.R.n_dof), s, &arg, t0, t1);
}

hipError_t launch_cost(const CostArgs& a, int grid, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
    if (a.C.tiled) return hipErrorInvalidValue;   // k_cost stages one OBB tile: tiled problems take k_cost_grp
    const size_t lds = lds_bytes_cost(a.C.obb_cap, a.R.n_dof, a.R.n_joints, a.group, a.grad != 0);
    CostArgs arg = a;
    const void* k = a.grad ? (const void*)k_cost<true> : (const void*)k_cost<false>;
    return launch_args(k, grid, lds, s, &arg, t0, t1);
}

size_t lds_bytes_grp(int obb_cap, int n_dof, int n_joints, int n_sph, int n_clu, bool grad, bool tiled) {
    return lds_floats_grp(obb_cap, n_dof, n_joints, n_sph, n_clu, grad, tiled) * sizeof(float);
}

// Variant table: [grad][world - 1][tiled] (ESDF-only launches stage no OBBs).
const void* grp_kernel(bool grad, int world, bool tiled) {
    static const void* const tab[2][3][2] = {
        {{(const void*)k_cost_grp<false, 1, false>, (const void*)k_cost_grp<false, 1, true>},
         {(const void*)k_cost_grp_esdf<false>, (const void*)k_cost_grp_esdf<false>},
         {(const void*)k_cost_grp<false, 3, false>, (const void*)k_cost_grp<false, 3, true>}},
        {{(const void*)k_cost_grp<true, 1, false>, (const void*)k_cost_grp<true, 1, true>},
         {(const void*)k_cost_grp_esdf<true>, (const void*)k_cost_grp_esdf<true>},
         {(const void*)k_cost_grp<true, 3, false>, (const void*)k_cost_grp<true, 3, true>}}};
    return tab[grad ? 1 : 0][world - 1][tiled ? 1 : 0];
}

int grp_world(const CostArgs& a) {
    const bool obb = a.P.use_obb != 0 && !a.clearance_only && a.S.n_obbs > 0;
    const bool esdf = (a.P.use_esdf != 0 || a.clearance_only) && a.S.n_voxel_layers > 0;
    return obb && esdf ? 3 : esdf ? 2 : 1;
}

int grid_cost_grp(bool grad, int world, bool tiled, size_t lds) {
    static std::mutex mu;
    static size_t last[2][3][2];
    static int g[2][3][2];
    static bool init = false;
    std::lock_guard<std::mutex> lk(mu);
    if (!init) {
        for (auto& x : last) for (auto& y : x) for (auto& z : y) z = (size_t)-1;
        init = true;
    }
    const int i = grad ? 1 : 0, w = world - 1, t = tiled ? 1 : 0;
    if (lds != last[i][w][t]) {
        g[i][w][t] = occupancy_grid(grp_kernel(grad, world, tiled), lds);
        last[i][w][t] = lds;
    }
    return g[i][w][t];
}

hipError_t launch_cost_grp(const CostArgs& a, int grid, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
    const size_t lds = lds_bytes_grp(a.C.obb_cap, a.R.n_dof, a.R.n_joints, a.R.n_sph, a.R.n_clu, a.grad != 0,
                                     a.C.tiled != 0);
    const int world = grp_world(a);
    const bool tiled = a.C.tiled != 0 && world != 2;
    grid = std::max(1, std::min(grid, grid_cost_grp(a.grad != 0, world, tiled, lds)));   // persistent: occupancy x CUs
    CostArgs arg = a;
    return launch_args(grp_kernel(a.grad != 0, world, tiled), grid, lds, s, &arg, t0, t1);
}

hipError_t launch_self(const SelfArgs& a, hipStream_t s) {
    // X.n_states bounds the state count (a device-side count may be smaller,
    // then a bounded grid strides over it)
    long blocks = a.C.tab ? (long)a.C.n_chunks : (a.X.n_states + kSelfStates - 1) / kSelfStates;
    if (a.X.n_dyn && blocks > 16384) blocks = 16384;
    if (blocks <= 0) return hipSuccess;
    const bool gate = a.mode == 0 && !a.R.self_gate_implied;   // approximate pair positions
    const size_t smem = sizeof(float4) * (size_t)kSelfStates * (a.R.n_sph + a.R.n_clu) +
                        sizeof(float) * (size_t)kSelfStates * (a.R.n_joints * 12 + 4) +
                        sizeof(float) * (size_t)kMaxDof * kSelfStates + sizeof(int) * (size_t)kSelfStates * a.R.n_joints +
                        sizeof(int) * 2 * kWave + (gate ? sizeof(float4) * (size_t)kSelfStates * a.R.n_apx : 0);
    hipLaunchKernelGGL(k_self, dim3((unsigned)blocks), dim3(64), smem, s, a);
    return hipGetLastError();
}

hipError_t launch_fk_out(const DRobot& R, const DStates& X, float* frames, float* spheres, hipStream_t s) {
    const long blocks = (X.n_states + kWave - 1) / kWave;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fk_out, dim3((unsigned)blocks), dim3(64), 0, s, R, X, frames, spheres);
    return hipGetLastError();
}

hipError_t launch_chomp_nsub(const ChompArgs& a, hipStream_t s) {
    if (a.X.n_traj <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_chomp_nsub, dim3(a.X.n_traj), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_chomp_scan, dim3(1), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_chomp_book(const ChompArgs& a, int initial, hipStream_t s) {
    if (a.X.n_traj <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_chomp_book, dim3(a.X.n_traj), dim3(64), 0, s, a, initial);
    return hipGetLastError();
}

hipError_t launch_chomp_update(const ChompArgs& a, int max_t, hipStream_t s) {
    if (a.X.n_traj <= 0) return hipSuccess;
    const size_t smem = sizeof(float) * (3 * (size_t)max_t + 12);
    hipLaunchKernelGGL(k_chomp_update, dim3(a.R.n_dof, a.X.n_traj), dim3(256), smem, s, a);
    return hipGetLastError();
}

hipError_t launch_stomp_reuse(const StompArgs& a, int, hipStream_t s) {
    if (a.B <= 0 || a.n_old <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stomp_select, dim3(a.B), dim3(64), sizeof(float) * (size_t)a.n_roll, s, a);
    hipLaunchKernelGGL(k_stomp_gather, dim3(a.n_old, a.B), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_stomp_sample(const StompArgs& a, int max_t, hipStream_t s) {
    if (a.B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stomp_seed, dim3(a.B), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_stomp_sample, dim3(a.R.n_dof, a.cur.nf, a.B), dim3(64), sizeof(float) * (size_t)max_t, s, a);
    return hipGetLastError();
}

size_t stomp_update_lds(int n_roll, int max_t) { return sizeof(float) * ((size_t)n_roll + 1) * (size_t)max_t; }

hipError_t launch_stomp_update(const StompArgs& a, int max_t, hipStream_t s) {
    if (a.B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stomp_update, dim3(a.R.n_dof, a.B), dim3(256), stomp_update_lds(a.n_roll, max_t), s, a);
    return hipGetLastError();
}

hipError_t launch_bfs_count(const float* nodes, int dim, const int* src, const int* dst, const int* done, int n,
                            float seg_len, int sparse, int* k0, int* cnt, int* off, long long* total, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bfs_count, dim3((n + 255) / 256), dim3(256), 0, s, nodes, dim, src, dst, done, n, seg_len,
                       sparse, k0, cnt);
    hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, s, cnt, n, off, total);
    return hipGetLastError();
}

hipError_t launch_unit_hint(const int* off, int n, int* hint, int n_hint, hipStream_t s) {
    if (n <= 0 || n_hint <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unit_hint, dim3((n_hint + 255) / 256), dim3(256), 0, s, off, n, hint, n_hint);
    return hipGetLastError();
}

hipError_t launch_bfs_finish(const int* first_hit, const int* cnt, int n, uint8_t* cc, int* new_checks, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bfs_finish, dim3((n + 255) / 256), dim3(256), 0, s, first_hit, cnt, n, cc, new_checks);
    return hipGetLastError();
}

hipError_t launch_sample_clearance(const uint8_t* cc, const float* clr, int n, float weight, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sample_clearance, dim3((n + 255) / 256), dim3(256), 0, s, cc, clr, n, weight, out);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const float* src, int dim, const int* idx, int base, int n, float* out, hipStream_t s) {
    const long total = (long)n * dim;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dim, idx, base, n, out);
    return hipGetLastError();
}

hipError_t launch_scatter(const float* v, const int* idx, int base, int n, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3((n + 255) / 256), dim3(256), 0, s, v, idx, base, n, out);
    return hipGetLastError();
}

hipError_t launch_add_u32(unsigned* p, unsigned v, hipStream_t s) {
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, int v, long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_segment_sum(const float* x, const int* off, int n_seg, float, float* out, hipStream_t s) {
    if (n_seg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_segment_sum, dim3(n_seg), dim3(64), 0, s, x, off, n_seg, out);
    return hipGetLastError();
}

hipError_t launch_control_cost(const float* theta, const int* traj_off, int n_traj, int D, float dt, float w, float* out,
                               int accumulate, hipStream_t s, int t_max, const float* state) {
    if (n_traj <= 0) return hipSuccess;
    // LDS staging when the longest trajectory's padded block fits 32 KiB
    const size_t lds = t_max > 0 ? (size_t)D * (t_max + 8) * sizeof(float) : 0;
    const int stage = lds > 0 && lds <= 32768 ? 1 : 0;
    hipLaunchKernelGGL(k_control, dim3(n_traj), dim3(64), stage ? lds : 0, s, theta, traj_off, n_traj, D, dt, w, state,
                       out, accumulate, stage);
    return hipGetLastError();
}

hipError_t launch_esdf_build(const float* obb_dims, const float* obb_pose, int n_obbs, const float* voxel_params,
                             const float* voxel_pose, float* out, hipStream_t s) {
    // voxel_params / voxel_pose are host arrays here
    const float inv = 1.0f / voxel_params[3];
    auto rf = [](float x) {
        const float nearest = roundf(x);
        if (x - nearest >= 1e-4f) return (int)floorf(x);
        return (int)nearest;
    };
    const int nx = rf(voxel_params[0] * inv) + 1, ny = rf(voxel_params[1] * inv) + 1, nz = rf(voxel_params[2] * inv) + 1;
    const long n = (long)nx * ny * nz;
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_esdf_build, dim3((unsigned)blocks), dim3(256), 0, s, obb_dims, obb_pose, n_obbs,
                       make_float4(voxel_params[0], voxel_params[1], voxel_params[2], voxel_params[3]),
                       make_float4(voxel_pose[0], voxel_pose[1], voxel_pose[2], voxel_pose[3]),
                       make_float4(voxel_pose[4], voxel_pose[5], voxel_pose[6], voxel_pose[7]), nx, ny, nz, out);
    return hipGetLastError();
}

}  // namespace ppl
