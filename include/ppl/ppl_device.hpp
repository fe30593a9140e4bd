// ppl_device.hpp -- device-inline substrate API for planner kernels (HIP, gfx950, wave64).
//
// The reference compiles its substrate into every planner translation unit as
// __device__ functions: fk_runtime, two_phase_cc, scene_min_clearance_runtime,
// scene_sum_{obb,esdf}_hinge_runtime and self_min_hinge_runtime
// (planning/runtime_kinematics.cuh:16-88, 294-329; collision/two_phase_cc.cuh:57-138;
// collision/scene_collision.cuh:339-495). This header gives planner kernels
// those names with the same results, over the robot tables of a ppl_robot
// (ppl_robot_get_device_view) and a ppl_scene_view.
//
// Calling convention. The reference dedicates 4 threads of a 32-wide warp to
// one configuration (tid % 4) and synchronises the block. Here a configuration
// belongs to an aligned group of G lanes of one wave64 (G a power of two,
// 1..64; `tid` is the thread index, the group lane is tid % G). Every lane of
// a group must make the same call; results come back to all G lanes, and
// groups of one wave run independently (no __syncthreads, no block scratch).
// sphere_pos / approx_sphere_pos are this configuration's [S][3] / [A][3]
// arrays by original sphere index, in LDS or global memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ppl/ppl.h"

namespace ppl {
namespace device {

// ---------------------------------------------------------------------------
// Group collectives (G aligned lanes of a wave64)
// ---------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ bool group_any(bool p, int tid) {
    static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "G must be a power of two <= 64");
    const unsigned long long m = __ballot(p);
    if (G == 64) return m != 0ull;
    const int base = (tid & 63) & ~(G - 1);
    return ((m >> base) & ((1ull << G) - 1ull)) != 0ull;
}

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

template <int G>
__device__ __forceinline__ float group_min(float v) {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

template <int G>
__device__ __forceinline__ unsigned group_or(unsigned v) {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) v |= (unsigned)__shfl_xor((int)v, o);
    return v;
}

// Lanes of one wave see each other's preceding LDS / global writes.
__device__ __forceinline__ void group_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Geometry primitives (3x4 row-major affine maps, implicit last row 0 0 0 1)
// ---------------------------------------------------------------------------
enum { kFixedJoint = -1, kXPrismJoint = 0, kZPrismJoint = 2, kXRotJoint = 3, kYRotJoint = 4, kZRotJoint = 5 };

__device__ __forceinline__ void aff_identity(float* T) {
#pragma unroll
    for (int i = 0; i < 12; ++i) T[i] = 0.0f;
    T[0] = T[5] = T[10] = 1.0f;
}

// S = FT * M(q): rotation about the joint's local axis or translation along it
// (planning/utils.cuh:560-702 semantics); fixed joints move nothing.
__device__ __forceinline__ void joint_step(const float* __restrict__ ft, int type, float q, float* S) {
#pragma unroll
    for (int i = 0; i < 12; ++i) S[i] = ft[i];
    if (type >= kXRotJoint) {
        float s, c;
        sincosf(q, &s, &c);
        const int a = (type == kXRotJoint) ? 1 : (type == kYRotJoint ? 2 : 0);
        const int b = (type == kXRotJoint) ? 2 : (type == kYRotJoint ? 0 : 1);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float fa = ft[r * 4 + a], fb = ft[r * 4 + b];
            S[r * 4 + a] = fmaf(fa, c, fb * s);
            S[r * 4 + b] = fmaf(fb, c, -fa * s);
        }
    } else if (type >= kXPrismJoint) {
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

// World point -> frame of an inverse pose [tx, ty, tz, qw, qx, qy, qz, pad]:
// v + w t + u x t with t = 2 u x v, then + translation
// (collision/scene_collision.cuh:24-51).
__device__ __forceinline__ void inverse_pose_point(const float* pose, float x, float y, float z, float& lx, float& ly,
                                                   float& lz) {
    const float w = pose[3], ux = pose[4], uy = pose[5], uz = pose[6];
    const float tx = 2.0f * (uy * z - uz * y), ty = 2.0f * (uz * x - ux * z), tz = 2.0f * (ux * y - uy * x);
    lx = x + w * tx + (uy * tz - uz * ty) + pose[0];
    ly = y + w * ty + (uz * tx - ux * tz) + pose[1];
    lz = z + w * tz + (ux * ty - uy * tx) + pose[2];
}

// Sphere vs one OBB, dims = full extents: collision iff |max(|l| - h, 0)|^2 < r^2
// (collision/scene_collision.cuh:59-78).
__device__ __forceinline__ bool sphere_obb_collide(const float* dims, const float* pose, float x, float y, float z,
                                                   float r) {
    float lx, ly, lz;
    inverse_pose_point(pose, x, y, z, lx, ly, lz);
    const float cx = fmaxf(fabsf(lx) - 0.5f * dims[0], 0.0f);
    const float cy = fmaxf(fabsf(ly) - 0.5f * dims[1], 0.0f);
    const float cz = fmaxf(fabsf(lz) - 0.5f * dims[2], 0.0f);
    return fmaf(cz, cz, fmaf(cy, cy, cx * cx)) < r * r;
}

// Box signed distance (Quilez form) minus r; positive = free
// (collision/scene_collision.cuh:88-112).
__device__ __forceinline__ float sphere_obb_clearance(const float* dims, const float* pose, float x, float y, float z,
                                                      float r) {
    float lx, ly, lz;
    inverse_pose_point(pose, x, y, z, lx, ly, lz);
    const float qx = fabsf(lx) - 0.5f * dims[0], qy = fabsf(ly) - 0.5f * dims[1], qz = fabsf(lz) - 0.5f * dims[2];
    const float ox = fmaxf(qx, 0.0f), oy = fmaxf(qy, 0.0f), oz = fmaxf(qz, 0.0f);
    const float outside = sqrtf(fmaf(oz, oz, fmaf(oy, oy, ox * ox)));
    const float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
    return outside + inside - r;
}

__device__ __forceinline__ int robust_floor(float x) {
    const float nearest = roundf(x);
    if (x - nearest >= 1e-4f) return (int)floorf(x);
    return (int)nearest;
}

// Nearest-voxel ESDF clearance -dist - r over enabled layers, 1000 when no
// layer applies (collision/scene_collision.cuh:186-331: truncation index,
// boundary offset 2, sentinel <= -999, positive inside). sphere < 0 skips the ACM.
__device__ __forceinline__ float sphere_esdf_clearance(const ppl_scene_view& S, float x, float y, float z, float r,
                                                       int sphere) {
    float best = 1000.0f;
    for (int layer = 0; layer < S.n_voxel_layers; ++layer) {
        if (S.voxel_enable[layer] == 0) continue;
        if (sphere >= 0 && S.sphere_acm_mask && S.sphere_acm_mask[sphere * S.n_voxel_layers + layer] == 0) continue;
        const float* prm = S.voxel_params + 4 * layer;
        const float vs = prm[3];
        if (vs <= 0.0f) continue;
        const float inv = 1.0f / vs;
        const int nx = robust_floor(prm[0] * inv) + 1, ny = robust_floor(prm[1] * inv) + 1,
                  nz = robust_floor(prm[2] * inv) + 1;
        float lx, ly, lz;
        inverse_pose_point(S.voxel_pose + 8 * layer, x, y, z, lx, ly, lz);
        const int ix = (int)((lx + prm[0] * 0.5f) * inv), iy = (int)((ly + prm[1] * 0.5f) * inv),
                  iz = (int)((lz + prm[2] * 0.5f) * inv);
        if (ix >= nx - 2 || iy >= ny - 2 || iz >= nz - 2 || ix <= 2 || iy <= 2 || iz <= 2) continue;
        const long lin = (long)ix * ny * nz + (long)iy * nz + iz;
        if (lin >= S.max_voxels_per_layer) continue;
        const float d = S.voxel_features[(size_t)layer * S.max_voxels_per_layer + lin];
        if (d <= -999.0f) continue;
        best = fminf(best, -d - r);
    }
    return best;
}

// Any enabled OBB or ESDF layer in collision with the sphere; sphere < 0 =
// approximate sphere (no ACM, collision/scene_collision.cuh:574-582).
__device__ __forceinline__ bool sphere_scene_collide(const ppl_scene_view& S, float x, float y, float z, float r,
                                                     int sphere) {
    for (int o = 0; o < S.n_obbs; ++o) {
        if (S.obb_enable[o] == 0) continue;
        if (sphere >= 0 && S.sphere_obb_acm_mask && S.sphere_obb_acm_mask[(size_t)sphere * S.n_obbs + o] == 0) continue;
        if (sphere_obb_collide(S.obb_dims + 4 * o, S.obb_pose + 8 * o, x, y, z, r)) return true;
    }
    return S.n_voxel_layers > 0 && sphere_esdf_clearance(S, x, y, z, r, sphere) < 0.0f;
}

// ---------------------------------------------------------------------------
// Forward kinematics (planning/runtime_kinematics.cuh:16-88, 176-243)
// ---------------------------------------------------------------------------
// Walks the DFS-ordered tree with the live transforms in registers (every lane
// of the group walks it; lanes split the spheres) and writes world sphere
// centres to pos[3 * original index]. approx selects the approximate spheres.
// frames (nullable, [n_joints][12]) receives every joint's world frame, rows
// 0..2 of the 4x4, for spheres_from_frames.
template <int G>
__device__ void fk_spheres(const ppl_robot_device_view& R, const float* q, float* pos, int tid, bool approx,
                           float* frames = nullptr) {
    const int lane = tid & (G - 1);
    float slot[2][12];
    aff_identity(slot[0]);
    aff_identity(slot[1]);
    // the approximate spheres ride their own chain when the model has one
    const bool own = approx && R.aft != nullptr;
    const float* ftab = own ? R.aft : R.ft;
    for (int jj = 0; jj < R.n_joints; ++jj) {
        const int dof = R.jmeta[4 * jj + 1], sl = R.jmeta[4 * jj + 2], ps = R.jmeta[4 * jj + 3];
        const int type = own ? R.atype[jj] : R.jmeta[4 * jj];
        float T[12];
        if (jj == 0) {
            aff_identity(T);
        } else {
            float S[12];
            joint_step(ftab + 12 * jj, type, dof >= 0 ? q[dof] : 0.0f, S);
            aff_compose(ps == 0 ? slot[0] : slot[1], S, T);
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            if (sl == 0) slot[0][k] = T[k];
            else slot[1][k] = T[k];
        }
        if (frames && lane == 0) {
#pragma unroll
            for (int k = 0; k < 12; ++k) frames[12 * jj + k] = T[k];
        }
        const int* off = approx ? R.apx_off : R.sph_off;
        const float* tab = approx ? R.apx : R.sph;
        const int* id = approx ? R.apx_id : R.sph_id;
        for (int s = off[jj] + lane; s < off[jj + 1]; s += G) {
            float x, y, z;
            aff_point(T, tab[4 * s], tab[4 * s + 1], tab[4 * s + 2], x, y, z);
            const int o = id[s];
            pos[3 * o] = x; pos[3 * o + 1] = y; pos[3 * o + 2] = z;
        }
    }
    group_sync();
}

// World sphere centres from stored joint frames (fk_spheres' `frames`): no
// tree walk. Lanes take spheres in turn across all joints.
template <int G>
__device__ void spheres_from_frames(const ppl_robot_device_view& R, const float* frames, float* pos, int tid,
                                    bool approx) {
    const int lane = tid & (G - 1);
    const int n = approx ? R.n_approx_spheres : R.n_spheres;
    const int* off = approx ? R.apx_off : R.sph_off;
    const float* tab = approx ? R.apx : R.sph;
    const int* id = approx ? R.apx_id : R.sph_id;
    int jj = 0;
    for (int s = lane; s < n; s += G) {
        while (off[jj + 1] <= s) ++jj;
        float x, y, z;
        aff_point(frames + 12 * jj, tab[4 * s], tab[4 * s + 1], tab[4 * s + 2], x, y, z);
        const int o = id[s];
        pos[3 * o] = x; pos[3 * o + 1] = y; pos[3 * o + 2] = z;
    }
    group_sync();
}

template <int G>
__device__ void fk_runtime(const ppl_robot_device_view& R, const float* q, float* sphere_pos, int tid) {
    fk_spheres<G>(R, q, sphere_pos, tid, false);
}

template <int G>
__device__ void fk_approx_runtime(const ppl_robot_device_view& R, const float* q, float* approx_sphere_pos, int tid) {
    fk_spheres<G>(R, q, approx_sphere_pos, tid, true);
}

// ---------------------------------------------------------------------------
// Reductions over computed sphere positions
// ---------------------------------------------------------------------------
// min over spheres of the ESDF clearance with radius + margin
// (scene_min_clearance_runtime, collision/scene_collision.cuh:339-384: ESDF only).
template <int G>
__device__ float scene_min_clearance_runtime(const ppl_robot_device_view& R, const float* sphere_pos,
                                             const ppl_scene_view& S, int tid, float margin) {
    float m = 1000.0f;
    for (int s = tid & (G - 1); s < R.n_spheres; s += G)
        m = fminf(m, sphere_esdf_clearance(S, sphere_pos[3 * s], sphere_pos[3 * s + 1], sphere_pos[3 * s + 2],
                                           R.sph_orig[4 * s + 3] + margin, s));
    return group_min<G>(m);
}

// sum over spheres of relu(m - min over enabled OBBs of the clearance), raw
// radius (scene_sum_obb_hinge_runtime, collision/scene_collision.cuh:449-495).
template <int G>
__device__ float scene_sum_obb_hinge_runtime(const ppl_robot_device_view& R, const float* sphere_pos,
                                             const ppl_scene_view& S, int tid, float m) {
    float acc = 0.0f;
    for (int s = tid & (G - 1); s < R.n_spheres; s += G) {
        float best = 1000.0f;
        for (int o = 0; o < S.n_obbs; ++o) {
            if (S.obb_enable[o] == 0) continue;
            if (S.sphere_obb_acm_mask && S.sphere_obb_acm_mask[(size_t)s * S.n_obbs + o] == 0) continue;
            best = fminf(best, sphere_obb_clearance(S.obb_dims + 4 * o, S.obb_pose + 8 * o, sphere_pos[3 * s],
                                                    sphere_pos[3 * s + 1], sphere_pos[3 * s + 2], R.sph_orig[4 * s + 3]));
        }
        acc += fmaxf(m - best, 0.0f);
    }
    return group_sum<G>(acc);
}

// sum over spheres of relu(m - ESDF clearance), raw radius
// (scene_sum_esdf_hinge_runtime, collision/scene_collision.cuh:394-438).
template <int G>
__device__ float scene_sum_esdf_hinge_runtime(const ppl_robot_device_view& R, const float* sphere_pos,
                                              const ppl_scene_view& S, int tid, float m) {
    float acc = 0.0f;
    for (int s = tid & (G - 1); s < R.n_spheres; s += G)
        acc += fmaxf(m - sphere_esdf_clearance(S, sphere_pos[3 * s], sphere_pos[3 * s + 1], sphere_pos[3 * s + 2],
                                               R.sph_orig[4 * s + 3], s),
                     0.0f);
    return group_sum<G>(acc);
}

// relu(m_self - min over self pairs of |pi - pj| - ri - rj)
// (self_min_hinge_runtime, planning/runtime_kinematics.cuh:294-329).
template <int G>
__device__ float self_min_hinge_runtime(const ppl_robot_device_view& R, const float* sphere_pos, int tid,
                                        float self_margin) {
    float m = 1000.0f;
    for (int p = tid & (G - 1); p < R.n_pairs; p += G) {
        const int i = R.pairs[2 * p], j = R.pairs[2 * p + 1];
        const float dx = sphere_pos[3 * i] - sphere_pos[3 * j], dy = sphere_pos[3 * i + 1] - sphere_pos[3 * j + 1],
                    dz = sphere_pos[3 * i + 2] - sphere_pos[3 * j + 2];
        m = fminf(m, sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx))) - (R.sph_orig[4 * i + 3] + R.sph_orig[4 * j + 3]));
    }
    return fmaxf(0.0f, self_margin - group_min<G>(m));
}

// Any self pair of the (approximate) model with |pi - pj|^2 < (ri + rj)^2
// (planning/utils.cuh:239-267, planning/runtime_kinematics.cuh:248-283, 382-416).
template <int G>
__device__ bool self_collide(const ppl_robot_device_view& R, const float* pos, int tid, bool approx) {
    const int n = approx ? R.n_approx_pairs : R.n_pairs;
    const int* pairs = approx ? R.apx_pairs : R.pairs;
    const float* sph = approx ? R.apx_orig : R.sph_orig;
    bool hit = false;
    for (int p = tid & (G - 1); p < n && !hit; p += G) {
        const int i = pairs[2 * p], j = pairs[2 * p + 1];
        const float dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1], dz = pos[3 * i + 2] - pos[3 * j + 2];
        const float rs = sph[4 * i + 3] + sph[4 * j + 3];
        hit = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) - rs * rs < 0.0f;
    }
    return group_any<G>(hit, tid);
}

// ---------------------------------------------------------------------------
// Halton sampling (planning/pRRTC.cu:955-995)
// ---------------------------------------------------------------------------
// One dimension of the incremental Halton generator: base b, numerator n and
// denominator d as exact small integers in FP32 (start n = 0, d = 1). The k-th
// call returns the radical inverse of k in base b, n / d, and advances (n, d).
// Exact while d < 2^24.
__device__ __forceinline__ float halton_step(float b, float& n, float& d) {
    const float x = d - n;
    if (x == 1.0f) {
        d = floorf(d * b);
        n = 1.0f;
    } else {
        float y = floorf(d / b);
        while (x <= y) y = floorf(y / b);
        n = floorf((b + 1.0f) * y) - x;
    }
    return n / d;
}

// ---------------------------------------------------------------------------
// Two-phase any-collision (collision/two_phase_cc.cuh:57-138)
// ---------------------------------------------------------------------------
struct TwoPhaseResult {
    bool collision;
    bool scene_hit;      // the collision came from the scene check (not self)
    bool did_full_fk;    // sphere_pos holds this configuration's exact spheres
};

// Approximate FK and scene pass flag joints; the exact spheres of flagged
// joints are then checked against the scene; with check_self, an approximate
// self hit triggers the exact self check. The verdict is this configuration's
// (the reference's block-wide "any slot" is group_any / __ballot over it).
template <int G>
__device__ TwoPhaseResult two_phase_cc(const ppl_robot_device_view& R, const ppl_scene_view& S, const float* q,
                                       float* sphere_pos, float* approx_sphere_pos, int tid, float collision_margin,
                                       bool check_self) {
    const int lane = tid & (G - 1);
    TwoPhaseResult out{false, false, false};
    fk_approx_runtime<G>(R, q, approx_sphere_pos, tid);
    unsigned flags = 0u;   // bit j: joint at DFS position j flagged by the approximate pass
    for (int jj = 0; jj < R.n_joints; ++jj)
        for (int a = R.apx_off[jj] + lane; a < R.apx_off[jj + 1]; a += G) {
            const int o = R.apx_id[a];
            if (sphere_scene_collide(S, approx_sphere_pos[3 * o], approx_sphere_pos[3 * o + 1],
                                     approx_sphere_pos[3 * o + 2], R.apx[4 * a + 3] + collision_margin, -1))
                flags |= 1u << jj;
        }
    flags = group_or<G>(flags);
    if (flags) {
        fk_runtime<G>(R, q, sphere_pos, tid);
        out.did_full_fk = true;
        bool hit = false;
        for (int jj = 0; jj < R.n_joints && !hit; ++jj) {
            if (!(flags & (1u << jj))) continue;
            for (int s = R.sph_off[jj] + lane; s < R.sph_off[jj + 1] && !hit; s += G) {
                const int o = R.sph_id[s];
                hit = sphere_scene_collide(S, sphere_pos[3 * o], sphere_pos[3 * o + 1], sphere_pos[3 * o + 2],
                                           R.sph[4 * s + 3] + collision_margin, o);
            }
        }
        out.collision = out.scene_hit = group_any<G>(hit, tid);
    }
    if (check_self && !out.collision && self_collide<G>(R, approx_sphere_pos, tid, true)) {
        if (!out.did_full_fk) {
            fk_runtime<G>(R, q, sphere_pos, tid);
            out.did_full_fk = true;
        }
        out.collision = self_collide<G>(R, sphere_pos, tid, false);
    }
    return out;
}

// The same verdict with one tree walk: T ([n_joints][12] per configuration,
// the reference's FK scratch `T`) keeps the joint frames of the approximate
// pass, and the exact spheres are placed from them. Lanes take the
// approximate spheres, then the exact spheres of flagged joints, in turn
// across all joints. A model whose approximate spheres have a chain of their
// own (R.aft) takes the two-walk form above.
template <int G>
__device__ TwoPhaseResult two_phase_cc(const ppl_robot_device_view& R, const ppl_scene_view& S, const float* q,
                                       float* sphere_pos, float* approx_sphere_pos, float* T, int tid,
                                       float collision_margin, bool check_self) {
    // an approximate chain of its own: the frames of one walk do not serve both sphere sets
    if (R.aft) return two_phase_cc<G>(R, S, q, sphere_pos, approx_sphere_pos, tid, collision_margin, check_self);
    const int lane = tid & (G - 1);
    TwoPhaseResult out{false, false, false};
    fk_spheres<G>(R, q, approx_sphere_pos, tid, true, T);
    unsigned flags = 0u;
    int jj = 0;
    for (int a = lane; a < R.n_approx_spheres; a += G) {
        while (R.apx_off[jj + 1] <= a) ++jj;
        const int o = R.apx_id[a];
        if (sphere_scene_collide(S, approx_sphere_pos[3 * o], approx_sphere_pos[3 * o + 1], approx_sphere_pos[3 * o + 2],
                                 R.apx[4 * a + 3] + collision_margin, -1))
            flags |= 1u << jj;
    }
    flags = group_or<G>(flags);
    if (flags) {
        spheres_from_frames<G>(R, T, sphere_pos, tid, false);
        out.did_full_fk = true;
        bool hit = false;
        int js = 0;
        for (int s = lane; s < R.n_spheres && !hit; s += G) {
            while (R.sph_off[js + 1] <= s) ++js;
            if (!(flags & (1u << js))) continue;
            const int o = R.sph_id[s];
            hit = sphere_scene_collide(S, sphere_pos[3 * o], sphere_pos[3 * o + 1], sphere_pos[3 * o + 2],
                                       R.sph[4 * s + 3] + collision_margin, o);
        }
        out.collision = out.scene_hit = group_any<G>(hit, tid);
    }
    if (check_self && !out.collision && self_collide<G>(R, approx_sphere_pos, tid, true)) {
        if (!out.did_full_fk) {
            spheres_from_frames<G>(R, T, sphere_pos, tid, false);
            out.did_full_fk = true;
        }
        out.collision = self_collide<G>(R, sphere_pos, tid, false);
    }
    return out;
}

}  // namespace device
}  // namespace ppl
