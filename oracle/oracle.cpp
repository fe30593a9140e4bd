// oracle.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// CPU restatement of PrismPlan's collision-kinematics substrate. Each function
// cites the reference lines (paths relative to code/gpu_planning/cpp/src/) whose
// behaviour it restates. Templated on the arithmetic type: double is the
// parity oracle, float is the timed CPU path.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kFixed = -1, kXPrism = 0, kZPrism = 2, kXRot = 3, kYRot = 4, kZRot = 5;
constexpr int kMaxJoints = 64;
constexpr float kUnobserved = -1000.0f;   // scene_collision.cuh:186

template <typename T> struct V3 { T x, y, z; };

// ---------------------------------------------------------------------------
// Frames: 3x4 row-major affine, last row implicitly [0 0 0 1].
// ---------------------------------------------------------------------------
template <typename T>
inline void frame_identity(T* f) {
    for (int i = 0; i < 12; ++i) f[i] = 0;
    f[0] = f[5] = f[10] = 1;
}

// C = A * B for 3x4 affines.
template <typename T>
inline void frame_compose(const T* A, const T* B, T* C) {
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 4; ++c) {
            T acc = A[r * 4 + 0] * B[0 * 4 + c] + A[r * 4 + 1] * B[1 * 4 + c] + A[r * 4 + 2] * B[2 * 4 + c];
            if (c == 3) acc += A[r * 4 + 3];
            C[r * 4 + c] = acc;
        }
    }
}

// Joint step S = FT * M(q), utils.cuh:560-702 (xrot_fn, yrot_fn, zrot_fn,
// prism_fn). A FIXED (or unknown) non-root joint moves nothing: S = FT. The
// reference would index ft[-1] / q[-1] there (runtime_kinematics.cuh:50-52);
// no shipped robot has such a joint.
template <typename T>
inline void joint_step(const float* ft, int type, T q, T* S) {
    T F[12];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) F[r * 4 + c] = (T)ft[r * 4 + c];
    for (int i = 0; i < 12; ++i) S[i] = F[i];
    if (type >= kXRot && type <= kZRot) {
        const T c = std::cos(q), s = std::sin(q);
        int a, b;  // the two FT columns mixed by the rotation
        if (type == kXRot) { a = 1; b = 2; }
        else if (type == kYRot) { a = 2; b = 0; }   // Ry: col0 = c*f0 - s*f2, col2 = s*f0 + c*f2
        else { a = 0; b = 1; }
        for (int r = 0; r < 3; ++r) {
            const T fa = F[r * 4 + a], fb = F[r * 4 + b];
            S[r * 4 + a] = fa * c + fb * s;
            S[r * 4 + b] = -fa * s + fb * c;
        }
    } else if (type >= kXPrism && type <= kZPrism) {
        for (int r = 0; r < 3; ++r) S[r * 4 + 3] = F[r * 4 + type] * q + F[r * 4 + 3];
    }
}

// Tree walk in DFS order, runtime_kinematics.cuh:42-72: the root frame is the
// identity (its fixed transform is never applied), every other joint composes
// onto its parent's frame.
template <typename T>
void fk_frames(const orc_robot& R, const float* q, T* frames, bool approx_chain) {
    // A model without approximate transforms gets the full chain, types
    // included (python/bindings.cpp:157-164 copies both).
    const bool own = approx_chain && R.approx_fixed_transforms;
    const float* fts = own ? R.approx_fixed_transforms : R.fixed_transforms;
    const int* types = (own && R.approx_joint_types) ? R.approx_joint_types : R.joint_types;
    for (int jj = 0; jj < R.n_joints; ++jj) {
        const int i = R.dfs_order[jj];
        T* out = frames + 12 * i;
        if (jj == 0) { frame_identity(out); continue; }
        const int dof = R.joint_id_to_dof[i];
        const T qi = dof >= 0 ? (T)q[dof] : (T)0;
        T S[12];
        joint_step<T>(fts + 16 * i, types[i], qi, S);
        frame_compose(frames + 12 * R.joint_parents[i], S, out);
    }
}

template <typename T>
inline V3<T> frame_apply(const T* F, T x, T y, T z) {
    return {F[0] * x + F[1] * y + F[2] * z + F[3],
            F[4] * x + F[5] * y + F[6] * z + F[7],
            F[8] * x + F[9] * y + F[10] * z + F[11]};
}

template <typename T>
void sphere_positions(const orc_robot& R, const T* frames, bool approx, V3<T>* out) {
    const int n = approx ? R.n_approx : R.n_spheres;
    const float* sp = approx ? R.approx_spheres : R.spheres;
    const int* s2j = approx ? R.approx_sphere_to_joint : R.sphere_to_joint;
    for (int s = 0; s < n; ++s)
        out[s] = frame_apply(frames + 12 * s2j[s], (T)sp[4 * s], (T)sp[4 * s + 1], (T)sp[4 * s + 2]);
}

// ---------------------------------------------------------------------------
// World -> OBB-local via the stored inverse pose, scene_collision.cuh:24-51.
// ---------------------------------------------------------------------------
template <typename T>
inline V3<T> quat_rotate(T qw, T qx, T qy, T qz, T sx, T sy, T sz) {
    const T tx = 2 * (qy * sz - qz * sy);
    const T ty = 2 * (qz * sx - qx * sz);
    const T tz = 2 * (qx * sy - qy * sx);
    return {sx + qw * tx + (qy * tz - qz * ty),
            sy + qw * ty + (qz * tx - qx * tz),
            sz + qw * tz + (qx * ty - qy * tx)};
}

template <typename T>
inline V3<T> to_local(const float* pose, T x, T y, T z) {
    V3<T> r = quat_rotate<T>(pose[3], pose[4], pose[5], pose[6], x, y, z);
    return {r.x + (T)pose[0], r.y + (T)pose[1], r.z + (T)pose[2]};
}

// Inverse of to_local's rotation, applied to a local-frame vector.
template <typename T>
inline V3<T> local_vec_to_world(const float* pose, T x, T y, T z) {
    return quat_rotate<T>(pose[3], -(T)pose[4], -(T)pose[5], -(T)pose[6], x, y, z);
}

// scene_collision.cuh:59-78 -- strict '<' against r^2.
template <typename T>
inline bool obb_collide(const float* dims, const float* pose, T x, T y, T z, T r) {
    V3<T> l = to_local<T>(pose, x, y, z);
    const T cx = std::max((T)0, std::fabs(l.x) - (T)dims[0] * (T)0.5);
    const T cy = std::max((T)0, std::fabs(l.y) - (T)dims[1] * (T)0.5);
    const T cz = std::max((T)0, std::fabs(l.z) - (T)dims[2] * (T)0.5);
    return cx * cx + cy * cy + cz * cz < r * r;
}

// scene_collision.cuh:88-112 -- box signed distance minus radius.
template <typename T>
inline T obb_clearance(const float* dims, const float* pose, T x, T y, T z, T r) {
    V3<T> l = to_local<T>(pose, x, y, z);
    const T qx = std::fabs(l.x) - (T)dims[0] * (T)0.5;
    const T qy = std::fabs(l.y) - (T)dims[1] * (T)0.5;
    const T qz = std::fabs(l.z) - (T)dims[2] * (T)0.5;
    const T ox = std::max(qx, (T)0), oy = std::max(qy, (T)0), oz = std::max(qz, (T)0);
    const T outside = std::sqrt(ox * ox + oy * oy + oz * oz);
    const T inside = std::min(std::max(qx, std::max(qy, qz)), (T)0);
    return (outside + inside) - r;
}

// d clearance / d world position for the box SDF; flags kinks within tol.
template <typename T>
inline V3<T> obb_clearance_grad(const float* dims, const float* pose, T x, T y, T z, T tol, bool* kink) {
    V3<T> l = to_local<T>(pose, x, y, z);
    const T L[3] = {l.x, l.y, l.z};
    T q[3], sg[3];
    for (int k = 0; k < 3; ++k) {
        q[k] = std::fabs(L[k]) - (T)dims[k] * (T)0.5;
        sg[k] = L[k] >= 0 ? (T)1 : (T)-1;
    }
    T g[3] = {0, 0, 0};
    const T qmax = std::max(q[0], std::max(q[1], q[2]));
    if (qmax > 0) {
        T o[3], n2 = 0;
        for (int k = 0; k < 3; ++k) { o[k] = std::max(q[k], (T)0); n2 += o[k] * o[k]; }
        const T n = std::sqrt(n2);
        for (int k = 0; k < 3; ++k) g[k] = o[k] / n * sg[k];
        for (int k = 0; k < 3; ++k) {
            if (std::fabs(q[k]) < tol) *kink = true;                       // face/edge boundary
            if (q[k] > -tol && std::fabs(L[k]) < tol) *kink = true;        // sign flip
        }
    } else {
        int km = 0;
        for (int k = 1; k < 3; ++k) if (q[k] > q[km]) km = k;
        g[km] = sg[km];
        for (int k = 0; k < 3; ++k)
            if (k != km && q[km] - q[k] < tol) *kink = true;               // inside Voronoi tie
        if (qmax > -tol || std::fabs(L[km]) < tol) *kink = true;
    }
    return local_vec_to_world<T>(pose, g[0], g[1], g[2]);
}

// scene_collision.cuh:192-200
inline int robust_floor(float x, float threshold = 1e-4f) {
    const float nearest = std::round(x);
    const float diff = x - nearest;
    if (diff >= threshold) return (int)std::floor(x);
    return (int)nearest;
}

// One ESDF layer lookup. Returns false when the layer is skipped for this
// point (disabled, ACM-masked, out of the guarded grid, unobserved). On
// success sets the signed distance (positive inside) and, for trilinear
// mode, its gradient in world coordinates. `amb` is set when the lookup sits
// within a hair of a voxel or validity boundary (decision ambiguous).
// `vamb` (nullable) is set when the looked-up VALUE is discontinuous within
// that hair: a nearest lookup next to a voxel face whose neighbour holds
// another value or is skipped, or either mode at the guarded grid boundary.
// *dalt then receives the largest distance the other side can give (-inf
// when that side is skipped, +inf when it is not known).
template <typename T>
bool esdf_layer(const orc_scene& S, int layer, T x, T y, T z, int sphere_idx, T* dist, V3<T>* grad,
                bool* amb, T amb_r = 0, bool* vamb = nullptr, T* dalt = nullptr) {
    if (S.voxel_enable[layer] == 0) return false;
    if (sphere_idx >= 0 && S.sphere_acm && S.sphere_acm[sphere_idx * S.n_voxel_layers + layer] == 0)
        return false;
    const float* prm = S.voxel_params + 4 * layer;
    const float* pose = S.voxel_pose + 8 * layer;
    const float vs = prm[3];
    if (vs <= 0.0f) return false;
    const float inv_vs = 1.0f / vs;
    const int nx = robust_floor(prm[0] * inv_vs) + 1;
    const int ny = robust_floor(prm[1] * inv_vs) + 1;
    const int nz = robust_floor(prm[2] * inv_vs) + 1;
    V3<T> l = to_local<T>(pose, x, y, z);
    const T u[3] = {(l.x + (T)prm[0] * (T)0.5) * (T)inv_vs, (l.y + (T)prm[1] * (T)0.5) * (T)inv_vs,
                    (l.z + (T)prm[2] * (T)0.5) * (T)inv_vs};
    const int n[3] = {nx, ny, nz};
    constexpr int kOff = 2;   // scene_collision.cuh:245
    const float* feat = S.voxel_features + (size_t)layer * S.max_voxels_per_layer;
    if (S.esdf_interp == 0) {
        // Nearest lookup by C truncation, scene_collision.cuh:238-261.
        auto lookup = [&](const int* idx, float* d) {
            for (int k = 0; k < 3; ++k)
                if (idx[k] >= n[k] - kOff || idx[k] <= kOff) return false;
            const long lin = (long)idx[0] * ny * nz + (long)idx[1] * nz + idx[2];
            if (lin >= S.max_voxels_per_layer) return false;
            *d = feat[lin];
            return *d > kUnobserved + 1.0f;
        };
        int idx[3];
        for (int k = 0; k < 3; ++k) idx[k] = (int)u[k];
        float d0 = 0.0f;
        const bool ok0 = lookup(idx, &d0);
        // A lookup within a hair of a voxel face is ambiguous when the
        // neighbouring voxel would flip the collision decision.
        for (int k = 0; k < 3; ++k) {
            const T fr = u[k] - std::floor(u[k]);
            if (!(fr < (T)1e-4 || fr > (T)(1 - 1e-4))) continue;
            int alt[3] = {idx[0], idx[1], idx[2]};
            alt[k] += (fr < (T)0.5) ? -1 : 1;
            float d1 = 0.0f;
            const bool ok1 = lookup(alt, &d1);
            if ((ok0 && d0 > -amb_r) != (ok1 && d1 > -amb_r)) *amb = true;
            if (vamb && (ok0 != ok1 || (ok0 && d0 != d1))) {
                *vamb = true;
                *dalt = std::max(*dalt, ok1 ? (T)d1 : -std::numeric_limits<T>::infinity());
            }
        }
        if (!ok0) return false;
        *dist = (T)d0;
        if (grad) *grad = {0, 0, 0};
        return true;
    }
    // Trilinear over the node grid (node i at i*vs - extent/2), same guard on
    // the lower corner as the nearest lookup.
    int i0[3];
    T f[3];
    for (int k = 0; k < 3; ++k) {
        const T fl = std::floor(u[k]);
        i0[k] = (int)fl;
        f[k] = u[k] - fl;
        if ((i0[k] == kOff + 1 || i0[k] == kOff || i0[k] == n[k] - kOff - 1 || i0[k] == n[k] - kOff) &&
            (f[k] < (T)1e-4 || f[k] > (T)(1 - 1e-4))) {
            *amb = true;   // the layer switches between valid and skipped here
            if (vamb) {
                *vamb = true;
                *dalt = std::numeric_limits<T>::infinity();
            }
        }
    }
    for (int k = 0; k < 3; ++k)
        if (i0[k] >= n[k] - kOff || i0[k] <= kOff) return false;
    // The interpolant is continuous across a cell face but its gradient is
    // not: the normal component is the secant slope of the cell on either
    // side. A gradient taken within a hair of a face is ambiguous.
    if (grad)
        for (int k = 0; k < 3; ++k)
            if (f[k] < (T)1e-4 || f[k] > (T)(1 - 1e-4)) *amb = true;
    T c[8];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e) {
                const long lin = (long)(i0[0] + a) * ny * nz + (long)(i0[1] + b) * nz + (i0[2] + e);
                if (lin >= S.max_voxels_per_layer) return false;
                const float v = feat[lin];
                if (v <= kUnobserved + 1.0f) return false;
                c[a * 4 + b * 2 + e] = (T)v;
            }
    const T fx = f[0], fy = f[1], fz = f[2];
    const T c00 = c[0] + (c[4] - c[0]) * fx, c01 = c[1] + (c[5] - c[1]) * fx;
    const T c10 = c[2] + (c[6] - c[2]) * fx, c11 = c[3] + (c[7] - c[3]) * fx;
    const T c0 = c00 + (c10 - c00) * fy, c1 = c01 + (c11 - c01) * fy;
    *dist = c0 + (c1 - c0) * fz;
    if (grad) {
        // d/du of the trilinear interpolant, then d/dl = d/du * inv_vs.
        const T gx = ((c[4] - c[0]) * (1 - fy) * (1 - fz) + (c[5] - c[1]) * (1 - fy) * fz +
                      (c[6] - c[2]) * fy * (1 - fz) + (c[7] - c[3]) * fy * fz);
        const T gy = (c10 - c00) * (1 - fz) + (c11 - c01) * fz;
        const T gz = c1 - c0;
        const T s = (T)inv_vs;
        *grad = local_vec_to_world<T>(pose, gx * s, gy * s, gz * s);
    }
    return true;
}

// scene_collision.cuh:202-266 / 276-331 over every layer. Returns the minimum
// clearance (-dist - r) and its layer, 1000 when no layer applies.
// With `calt` (and `vamb`): *calt = the lowest clearance the other side of
// a value jump can give (1000 when it gives none), for deciding whether the
// jump can change a minimum or a hinge.
template <typename T>
T esdf_clearance(const orc_scene& S, T x, T y, T z, T r, int sphere_idx, V3<T>* grad, bool* amb,
                 bool* vamb = nullptr, T* calt = nullptr) {
    T best = (T)1000;
    if (calt) *calt = (T)1000;
    for (int layer = 0; layer < S.n_voxel_layers; ++layer) {
        T d; V3<T> g;
        bool va = false;
        T da = -std::numeric_limits<T>::infinity();
        const bool ok = esdf_layer<T>(S, layer, x, y, z, sphere_idx, &d, grad ? &g : nullptr, amb, r,
                                      vamb ? &va : nullptr, &da);
        if (va) {
            *vamb = true;
            if (calt) *calt = std::min(*calt, std::isinf(da) && da > 0 ? -std::numeric_limits<T>::infinity()
                                                                       : (std::isinf(da) ? (T)1000 : -da - r));
        }
        if (!ok) continue;
        const T c = -d - r;
        if (c < best) { best = c; if (grad) *grad = {-g.x, -g.y, -g.z}; }
    }
    return best;
}

template <typename T>
bool esdf_collide(const orc_scene& S, T x, T y, T z, T r, int sphere_idx) {
    bool amb = false;
    for (int layer = 0; layer < S.n_voxel_layers; ++layer) {
        T d;
        if (!esdf_layer<T>(S, layer, x, y, z, sphere_idx, &d, nullptr, &amb)) continue;
        if (d > -r) return true;
    }
    return false;
}

inline bool obb_allowed(const orc_scene& S, int o, int sphere_idx, const uint8_t* acm) {
    if (S.obb_enable[o] == 0) return false;
    if (sphere_idx >= 0 && acm != nullptr && acm[sphere_idx * S.n_obbs + o] == 0) return false;
    return true;
}

// sphere_scene_in_collision, scene_collision.cuh:501-510 (OBBs, then ESDF).
template <typename T>
bool sphere_scene_hit(const orc_scene& S, V3<T> p, T r, int sphere_idx, const uint8_t* acm) {
    for (int o = 0; o < S.n_obbs; ++o) {
        if (!obb_allowed(S, o, sphere_idx, acm)) continue;
        if (obb_collide<T>(S.obb_dims + 4 * o, S.obb_pose + 8 * o, p.x, p.y, p.z, r)) return true;
    }
    return esdf_collide<T>(S, p.x, p.y, p.z, r, sphere_idx);
}

// sphere_sphere_self_collision, utils.cuh:239-267 (strict '<').
template <typename T>
inline bool self_pair_hit(V3<T> a, T ra, V3<T> b, T rb) {
    const T dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const T rs = ra + rb;
    return (dx * dx + dy * dy + dz * dz) - rs * rs < 0;
}

template <typename T>
bool self_any(const int* ranges, int n_ranges, const V3<T>* P, const float* sp) {
    for (int e = 0; e < n_ranges; ++e) {
        const int i = ranges[3 * e], a = ranges[3 * e + 1], b = ranges[3 * e + 2];
        for (int j = a; j <= b; ++j)
            if (self_pair_hit<T>(P[i], (T)sp[4 * i + 3], P[j], (T)sp[4 * j + 3])) return true;
    }
    return false;
}

// Minimum self clearance over the full ranges (runtime_kinematics.cuh:294-329);
// 1000 when there are no pairs. Reports the argmin pair and a tie flag.
template <typename T>
T self_min_clearance(const orc_robot& R, const V3<T>* P, int* bi, int* bj, bool* tie, T tol) {
    T best = (T)1000, second = (T)1000;
    int ai = -1, aj = -1;
    for (int e = 0; e < R.n_self_ranges; ++e) {
        const int i = R.self_ranges[3 * e], a = R.self_ranges[3 * e + 1], b = R.self_ranges[3 * e + 2];
        for (int j = a; j <= b; ++j) {
            const T dx = P[i].x - P[j].x, dy = P[i].y - P[j].y, dz = P[i].z - P[j].z;
            const T c = std::sqrt(dx * dx + dy * dy + dz * dz) - ((T)R.spheres[4 * i + 3] + (T)R.spheres[4 * j + 3]);
            if (c < best) { second = best; best = c; ai = i; aj = j; }
            else if (c < second) second = c;
        }
    }
    if (bi) *bi = ai;
    if (bj) *bj = aj;
    if (tie) *tie = (second - best) < tol;
    return best;
}

// Per-configuration scratch.
template <typename T>
struct Work {
    std::vector<T> frames, aframes;
    std::vector<V3<T>> P, A;
    Work(const orc_robot& R) : frames(12 * R.n_joints), aframes(12 * R.n_joints), P(R.n_spheres), A(R.n_approx) {}
};

// two_phase_cc (two_phase_cc.cuh:57-138) with the pRRTC check-kernel and edge
// callers' conventions: approx FK -> approx scene check flags joints (ACMs are
// not applied because the approx pass passes sphere_idx = -1,
// scene_collision.cuh:598-610 via :127) -> exact FK and exact check of the
// spheres on flagged joints -> optional approx and exact self checks (no
// collision margin on self checks). mode 1 = exhaustive exact check instead.
template <typename T>
bool check_one(const orc_robot& R, const orc_scene& S, const float* q, T margin, bool check_self, int mode,
               Work<T>& w, T* clr) {
    fk_frames<T>(R, q, w.frames.data(), false);
    sphere_positions<T>(R, w.frames.data(), false, w.P.data());
    bool hit = false;
    if (mode == 1) {
        for (int s = 0; s < R.n_spheres && !hit; ++s)
            hit = sphere_scene_hit<T>(S, w.P[s], (T)R.spheres[4 * s + 3] + margin, s, S.sphere_obb_acm);
        if (!hit && check_self) hit = self_any<T>(R.self_ranges, R.n_self_ranges, w.P.data(), R.spheres);
    } else {
        fk_frames<T>(R, q, w.aframes.data(), true);
        sphere_positions<T>(R, w.aframes.data(), true, w.A.data());
        int flags[kMaxJoints] = {0};
        bool any = false;
        for (int a = 0; a < R.n_approx; ++a) {
            if (sphere_scene_hit<T>(S, w.A[a], (T)R.approx_spheres[4 * a + 3] + margin, -1, S.sphere_obb_acm_approx)) {
                flags[R.approx_sphere_to_joint[a]]++;
                any = true;
            }
        }
        if (any) {
            for (int s = 0; s < R.n_spheres && !hit; ++s) {
                if (flags[R.sphere_to_joint[s]] <= 0) continue;
                hit = sphere_scene_hit<T>(S, w.P[s], (T)R.spheres[4 * s + 3] + margin, s, S.sphere_obb_acm);
            }
        }
        if (check_self && !hit) {
            if (self_any<T>(R.approx_self_ranges, R.n_approx_self_ranges, w.A.data(), R.approx_spheres))
                hit = self_any<T>(R.self_ranges, R.n_self_ranges, w.P.data(), R.spheres);
        }
    }
    if (clr) {
        T best = (T)1000;
        bool amb = false;
        for (int s = 0; s < R.n_spheres; ++s) {
            const T r = (T)R.spheres[4 * s + 3] + margin;
            for (int o = 0; o < S.n_obbs; ++o) {
                if (!obb_allowed(S, o, s, S.sphere_obb_acm)) continue;
                best = std::min(best, obb_clearance<T>(S.obb_dims + 4 * o, S.obb_pose + 8 * o, w.P[s].x, w.P[s].y, w.P[s].z, r));
            }
            if (S.n_voxel_layers > 0)
                best = std::min(best, esdf_clearance<T>(S, w.P[s].x, w.P[s].y, w.P[s].z, r, s, nullptr, &amb));
        }
        if (check_self && R.n_self_ranges > 0)
            best = std::min(best, self_min_clearance<T>(R, w.P.data(), nullptr, nullptr, nullptr, (T)0));
        *clr = amb ? (T)0 : best;
    }
    return hit;
}

// Summed-hinge state cost, chomp.cu:151-188 / stomp.cu:206-246 with
// scene_sum_obb_hinge_runtime (scene_collision.cuh:449-495),
// scene_sum_esdf_hinge_runtime (:394-438) and self_min_hinge_runtime
// (runtime_kinematics.cuh:294-329). Raw radii, no collision margin. When
// `grad` is given, also the analytic gradient through the argmin pairs.
template <typename T>
T hinge_one(const orc_robot& R, const orc_scene& S, const float* q, const orc_cost_params& p, Work<T>& w,
            T* grad, bool* kink, T tol, bool* vamb = nullptr, int* nterm = nullptr) {
    int terms = 0;
    fk_frames<T>(R, q, w.frames.data(), false);
    sphere_positions<T>(R, w.frames.data(), false, w.P.data());
    const T m = (T)p.world_margin;
    T world = 0;
    std::vector<V3<T>> g;
    if (grad) g.assign(R.n_spheres, V3<T>{0, 0, 0});
    for (int s = 0; s < R.n_spheres; ++s) {
        const T r = (T)R.spheres[4 * s + 3];
        const V3<T> P = w.P[s];
        if (p.use_obb) {
            T best = (T)1000, second = (T)1000;
            int arg = -1;
            for (int o = 0; o < S.n_obbs; ++o) {
                if (!obb_allowed(S, o, s, S.sphere_obb_acm)) continue;
                const T c = obb_clearance<T>(S.obb_dims + 4 * o, S.obb_pose + 8 * o, P.x, P.y, P.z, r);
                if (c < best) { second = best; best = c; arg = o; }
                else if (c < second) second = c;
            }
            const T h = m - best;
            if (h > 0) {
                world += h;
                ++terms;
                if (grad && arg >= 0) {
                    bool kk = false;
                    V3<T> gc = obb_clearance_grad<T>(S.obb_dims + 4 * arg, S.obb_pose + 8 * arg, P.x, P.y, P.z, tol, &kk);
                    g[s].x -= gc.x; g[s].y -= gc.y; g[s].z -= gc.z;
                    if (kk || second - best < tol) *kink = true;
                }
            }
            if (kink && std::fabs(h) < tol) *kink = true;
        }
        if (p.use_esdf && S.n_voxel_layers > 0) {
            V3<T> gc{0, 0, 0};
            bool amb = false;
            bool va = false;
            T ca = (T)1000;
            const T c = esdf_clearance<T>(S, P.x, P.y, P.z, r, s, grad ? &gc : nullptr, &amb, &va, &ca);
            const T h = m - c;
            if (vamb && va && (c < m || ca < m)) *vamb = true;   // either side inside the band
            if (h > 0) {
                world += h;
                ++terms;
                if (grad) { g[s].x -= gc.x; g[s].y -= gc.y; g[s].z -= gc.z; }
            }
            if (kink && (amb || std::fabs(h) < tol)) *kink = true;
        }
    }
    T self = 0;
    if (p.use_self && R.n_self_ranges > 0) {
        int bi, bj;
        bool tie = false;
        const T c = self_min_clearance<T>(R, w.P.data(), &bi, &bj, &tie, tol);
        const T h = (T)p.self_margin - c;
        if (h > 0) {
            self = h;
            ++terms;
            if (grad) {
                V3<T> d{w.P[bi].x - w.P[bj].x, w.P[bi].y - w.P[bj].y, w.P[bi].z - w.P[bj].z};
                const T n = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
                g[bi].x -= d.x / n; g[bi].y -= d.y / n; g[bi].z -= d.z / n;
                g[bj].x += d.x / n; g[bj].y += d.y / n; g[bj].z += d.z / n;
                if (tie) *kink = true;
            }
        }
        if (kink && std::fabs(h) < tol) *kink = true;
    }
    if (nterm) *nterm = terms;
    const T wgt = (T)p.weight;
    if (grad) {
        // dp_s/dq for every ancestor joint of sphere s's link: revolute
        // z x (p - o), prismatic z, with z the joint axis column of the
        // joint's world frame and o its origin.
        for (int d = 0; d < R.n_dof; ++d) grad[d] = 0;
        for (int s = 0; s < R.n_spheres; ++s) {
            if (g[s].x == 0 && g[s].y == 0 && g[s].z == 0) continue;
            int j = R.sphere_to_joint[s];
            int guard = 0;
            while (guard++ < kMaxJoints) {
                const int dof = R.joint_id_to_dof[j];
                const int type = R.joint_types[j];
                if (dof >= 0 && j != R.dfs_order[0]) {
                    const T* F = w.frames.data() + 12 * j;
                    const int ax = (type >= kXRot) ? type - kXRot : type;
                    if (type >= kXRot && type <= kZRot) {
                        const V3<T> z{F[ax], F[4 + ax], F[8 + ax]};
                        const V3<T> r{w.P[s].x - F[3], w.P[s].y - F[7], w.P[s].z - F[11]};
                        const V3<T> v{z.y * r.z - z.z * r.y, z.z * r.x - z.x * r.z, z.x * r.y - z.y * r.x};
                        grad[dof] += g[s].x * v.x + g[s].y * v.y + g[s].z * v.z;
                    } else if (type >= kXPrism && type <= kZPrism) {
                        grad[dof] += g[s].x * F[ax] + g[s].y * F[4 + ax] + g[s].z * F[8 + ax];
                    }
                }
                if (j == R.joint_parents[j]) break;
                j = R.joint_parents[j];
            }
        }
        for (int d = 0; d < R.n_dof; ++d) grad[d] *= wgt;
    }
    return wgt * (world + self);
}

// Clearance used by scene_min_clearance_runtime (scene_collision.cuh:339-384):
// ESDF only, radius + margin. include_obb adds the OBB clearance.
template <typename T>
T min_clearance_one(const orc_robot& R, const orc_scene& S, const float* q, T margin, bool include_obb, Work<T>& w,
                    bool* vamb = nullptr, T* lo = nullptr) {
    fk_frames<T>(R, q, w.frames.data(), false);
    sphere_positions<T>(R, w.frames.data(), false, w.P.data());
    T best = (T)1000;
    bool amb = false;
    // spheres whose lookup sits at a value jump: (own clearance, other side's lowest)
    std::vector<std::pair<T, T>> jumps;
    for (int s = 0; s < R.n_spheres; ++s) {
        const T r = (T)R.spheres[4 * s + 3] + margin;
        bool va = false;
        T ca = (T)1000;
        const T c = esdf_clearance<T>(S, w.P[s].x, w.P[s].y, w.P[s].z, r, s, nullptr, &amb, vamb ? &va : nullptr, &ca);
        if (va) jumps.push_back({c, ca});
        best = std::min(best, c);
        if (include_obb) {
            for (int o = 0; o < S.n_obbs; ++o) {
                if (!obb_allowed(S, o, s, S.sphere_obb_acm)) continue;
                best = std::min(best, obb_clearance<T>(S.obb_dims + 4 * o, S.obb_pose + 8 * o, w.P[s].x, w.P[s].y, w.P[s].z, r));
            }
        }
    }
    // the jump matters when that sphere holds the minimum or its other side undercuts it
    if (lo) *lo = best;
    if (vamb)
        for (const auto& j : jumps)
            if (j.first <= best || j.second < best) {
                *vamb = true;
                if (lo) *lo = std::min(*lo, j.second);
            }
    return best;
}

inline void set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// Interpolated FP32 state, q = fma(t, to - from, from) (edge_eval.cuh:119-121).
inline void lerp_state(const float* from, const float* to, int dim, float t, float* q) {
    for (int i = 0; i < dim; ++i) q[i] = std::fma(t, to[i] - from[i], from[i]);
}

// Control cost of one [D][T] block as theta^T R theta per dim.
double control_one(const float* th, int D, int T, const double* Rm) {
    double acc = 0;
    for (int d = 0; d < D; ++d) {
        const float* x = th + (size_t)d * T;
        for (int i = 0; i < T; ++i) {
            double ri = 0;
            for (int j = 0; j < T; ++j) ri += Rm[(size_t)i * T + j] * (double)x[j];
            acc += (double)x[i] * ri;
        }
    }
    return acc;
}

// 5-point acceleration metric, chomp_precompute.cpp:20-38: A on a grid padded
// by 6 on both sides, R = dt * A^T A, central T x T block.
void control_R(int T, double dt, double* R) {
    const int pad = 6, Np = T + 2 * pad;
    const double coeff[7] = {0.0, -1.0 / 12, 16.0 / 12, -30.0 / 12, 16.0 / 12, -1.0 / 12, 0.0};
    const double inv_dt2 = 1.0 / (dt * dt);
    std::vector<double> A((size_t)Np * Np, 0.0);
    for (int i = 0; i < Np; ++i)
        for (int j = -3; j <= 3; ++j) {
            const int idx = i + j;
            if (idx < 0 || idx >= Np) continue;
            A[(size_t)i * Np + idx] = inv_dt2 * coeff[j + 3];
        }
    for (int a = 0; a < T; ++a)
        for (int b = 0; b < T; ++b) {
            double s = 0;
            for (int k = 0; k < Np; ++k) s += A[(size_t)k * Np + a + pad] * A[(size_t)k * Np + b + pad];
            R[(size_t)a * T + b] = dt * s;
        }
}

template <typename T>
void check_collision_t(const orc_robot* r, const orc_scene* s, const float* q, int n, float margin, int check_self,
                       int mode, uint8_t* out, T* clr, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel
    {
        Work<T> w(*r);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            T c;
            out[i] = check_one<T>(*r, *s, q + (size_t)i * r->n_dof, (T)margin, check_self != 0, mode, w,
                                  clr ? &c : nullptr) ? 1 : 0;
            if (clr) clr[i] = c;
        }
    }
}

template <typename T>
void hinge_cost_t(const orc_robot* r, const orc_scene* s, const float* q, int n, const orc_cost_params* p, T* cost,
                  int nthreads) {
    set_threads(nthreads);
#pragma omp parallel
    {
        Work<T> w(*r);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i)
            cost[i] = hinge_one<T>(*r, *s, q + (size_t)i * r->n_dof, *p, w, nullptr, nullptr, (T)0);
    }
}

template <typename T>
void chomp_fd_t(const orc_robot* r, const orc_scene* s, const float* theta, const int* off, int B,
                const orc_cost_params* p, float eps, T* grad, int nthreads) {
    set_threads(nthreads);
    const int D = r->n_dof;
    orc_cost_params pu = *p;
    pu.weight = 1.0f;
#pragma omp parallel
    {
        Work<T> w(*r);
        float q[kMaxJoints];
#pragma omp for schedule(dynamic, 4)
        for (int b = 0; b < B; ++b) {
            const int T0 = off[b], Tn = off[b + 1] - off[b];
            const float* th = theta + (size_t)T0 * D;
            T* g = grad + (size_t)T0 * D;
            for (int t = 0; t < Tn; ++t) {
                for (int d = 0; d < D; ++d) {
                    if (t == 0 || t == Tn - 1) { g[(size_t)d * Tn + t] = 0; continue; }
                    for (int k = 0; k < D; ++k) q[k] = th[(size_t)k * Tn + t];
                    // chomp.cu:108-133: q[d] += eps, then q[d] -= 2 eps, in FP32.
                    q[d] += eps;
                    const T cp = hinge_one<T>(*r, *s, q, pu, w, nullptr, nullptr, (T)0);
                    q[d] -= 2.0f * eps;
                    const T cm = hinge_one<T>(*r, *s, q, pu, w, nullptr, nullptr, (T)0);
                    g[(size_t)d * Tn + t] = (T)p->weight * (cp - cm) / ((T)2 * (T)eps);
                }
            }
        }
    }
}

template <typename T>
void edge_batch_t(const orc_robot* r, const orc_scene* s, const float* from, const float* to, const int* n_states,
                  int E, int skip_first, float margin, int check_self, int* first_hit, int nthreads) {
    set_threads(nthreads);
    const int D = r->n_dof;
#pragma omp parallel
    {
        Work<T> w(*r);
        float q[kMaxJoints];
#pragma omp for schedule(dynamic, 16)
        for (int e = 0; e < E; ++e) {
            const int n = n_states[e];
            int fh = -1;
            for (int k = skip_first ? 1 : 0; k < n; ++k) {
                const float t = n > 1 ? (float)k / (float)(n - 1) : 0.0f;
                lerp_state(from + (size_t)e * D, to + (size_t)e * D, D, t, q);
                if (check_one<T>(*r, *s, q, (T)margin, check_self != 0, 0, w, nullptr)) { fh = k; break; }
            }
            first_hit[e] = fh;
        }
    }
}

orc_scene sub_scene(const orc_scene& all, int o0, int o1) {
    orc_scene s = all;
    s.n_obbs = o1 - o0;
    s.obb_dims = all.obb_dims + 4 * (size_t)o0;
    s.obb_pose = all.obb_pose + 8 * (size_t)o0;
    s.obb_enable = all.obb_enable + o0;
    s.sphere_obb_acm = nullptr;
    s.sphere_obb_acm_approx = nullptr;
    return s;
}

template <typename T>
void stomp_costs_t(const orc_robot* r, const orc_scene* all, const int* obb_off, const float* rollouts,
                   const int* T_p, const int* roff, int P, int R, const orc_cost_params* p, double w_u, double dt,
                   T* state, T* total, int nthreads, int* nterm = nullptr) {
    set_threads(nthreads);
    const int D = r->n_dof;
    // Waypoint-level state costs.
    struct Item { int prob, roll, t; };
    std::vector<Item> items;
    for (int pi = 0; pi < P; ++pi)
        for (int ri = 0; ri < R; ++ri)
            for (int t = 0; t < T_p[pi]; ++t) items.push_back({pi, ri, t});
#pragma omp parallel
    {
        Work<T> w(*r);
        float q[kMaxJoints];
#pragma omp for schedule(static)
        for (long it = 0; it < (long)items.size(); ++it) {
            const Item& I = items[it];
            const int Tn = T_p[I.prob];
            const orc_scene sc = sub_scene(*all, obb_off[I.prob], obb_off[I.prob + 1]);
            const float* th = rollouts + ((size_t)roff[I.prob] + (size_t)I.roll * Tn) * D;
            for (int d = 0; d < D; ++d) q[d] = th[(size_t)d * Tn + I.t];
            const size_t at = (size_t)roff[I.prob] + (size_t)I.roll * Tn + I.t;
            state[at] = hinge_one<T>(*r, sc, q, *p, w, nullptr, nullptr, (T)0, nullptr, nterm ? nterm + at : nullptr);
        }
    }
    for (int pi = 0; pi < P; ++pi) {
        const int Tn = T_p[pi];
        std::vector<double> Rm((size_t)Tn * Tn);
        control_R(Tn, dt, Rm.data());
        for (int ri = 0; ri < R; ++ri) {
            const size_t base = (size_t)roff[pi] + (size_t)ri * Tn;
            double acc = 0;
            for (int t = 0; t < Tn; ++t) acc += (double)state[base + t];
            acc += w_u * control_one(rollouts + base * D, D, Tn, Rm.data());
            total[(size_t)pi * R + ri] = (T)acc;
        }
    }
}

}  // namespace

extern "C" {

void orc_fk_spheres_f64(const orc_robot* r, const float* q, int n, double* out) {
    Work<double> w(*r);
    for (int i = 0; i < n; ++i) {
        fk_frames<double>(*r, q + (size_t)i * r->n_dof, w.frames.data(), false);
        sphere_positions<double>(*r, w.frames.data(), false, w.P.data());
        for (int s = 0; s < r->n_spheres; ++s) {
            double* o = out + ((size_t)i * r->n_spheres + s) * 3;
            o[0] = w.P[s].x; o[1] = w.P[s].y; o[2] = w.P[s].z;
        }
    }
}

void orc_fk_spheres_f32(const orc_robot* r, const float* q, int n, float* out) {
    Work<float> w(*r);
    for (int i = 0; i < n; ++i) {
        fk_frames<float>(*r, q + (size_t)i * r->n_dof, w.frames.data(), false);
        sphere_positions<float>(*r, w.frames.data(), false, w.P.data());
        for (int s = 0; s < r->n_spheres; ++s) {
            float* o = out + ((size_t)i * r->n_spheres + s) * 3;
            o[0] = w.P[s].x; o[1] = w.P[s].y; o[2] = w.P[s].z;
        }
    }
}

void orc_fk_frames_f64(const orc_robot* r, const float* q, int n, double* out) {
    for (int i = 0; i < n; ++i) fk_frames<double>(*r, q + (size_t)i * r->n_dof, out + (size_t)i * 12 * r->n_joints, false);
}

int orc_sphere_obb_collide_f64(const float* dims, const float* pose, double x, double y, double z, double r) {
    return obb_collide<double>(dims, pose, x, y, z, r) ? 1 : 0;
}

double orc_sphere_obb_clearance_f64(const float* dims, const float* pose, double x, double y, double z, double r) {
    return obb_clearance<double>(dims, pose, x, y, z, r);
}

double orc_sphere_esdf_clearance_f64(const orc_scene* s, double x, double y, double z, double r, int sphere_idx) {
    bool amb = false;
    return esdf_clearance<double>(*s, x, y, z, r, sphere_idx, nullptr, &amb);
}

double orc_sphere_esdf_grad_f64(const orc_scene* s, double x, double y, double z, double r, int sphere_idx,
                                double* grad3, int* amb) {
    bool a = false;
    V3<double> g{0, 0, 0};
    const double c = esdf_clearance<double>(*s, x, y, z, r, sphere_idx, &g, &a);
    grad3[0] = g.x;
    grad3[1] = g.y;
    grad3[2] = g.z;
    *amb = a ? 1 : 0;
    return c;
}

void orc_check_collision_f64(const orc_robot* r, const orc_scene* s, const float* q, int n, float margin,
                             int check_self, int mode, uint8_t* out, double* clr, int nthreads) {
    check_collision_t<double>(r, s, q, n, margin, check_self, mode, out, clr, nthreads);
}

void orc_check_collision_f32(const orc_robot* r, const orc_scene* s, const float* q, int n, float margin,
                             int check_self, int mode, uint8_t* out, float* clr, int nthreads) {
    check_collision_t<float>(r, s, q, n, margin, check_self, mode, out, clr, nthreads);
}

void orc_min_clearance_f64(const orc_robot* r, const orc_scene* s, const float* q, int n, float margin,
                           int include_obb, double* out, int nthreads) {
    orc_min_clearance_amb_f64(r, s, q, n, margin, include_obb, out, nullptr, nullptr, nthreads);
}

void orc_min_clearance_amb_f64(const orc_robot* r, const orc_scene* s, const float* q, int n, float margin,
                               int include_obb, double* out, uint8_t* vamb, double* lo, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel
    {
        Work<double> w(*r);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            bool va = false;
            double l = 0.0;
            out[i] = min_clearance_one<double>(*r, *s, q + (size_t)i * r->n_dof, margin, include_obb != 0, w, &va, &l);
            if (vamb) vamb[i] = va ? 1 : 0;
            if (lo) lo[i] = l;
        }
    }
}

void orc_hinge_cost_amb_f64(const orc_robot* r, const orc_scene* s, const float* q, int n, const orc_cost_params* p,
                            double* cost, uint8_t* vamb, int* nterm, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel
    {
        Work<double> w(*r);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            bool va = false;
            cost[i] = hinge_one<double>(*r, *s, q + (size_t)i * r->n_dof, *p, w, nullptr, nullptr, 0.0, &va,
                                        nterm ? nterm + i : nullptr);
            if (vamb) vamb[i] = va ? 1 : 0;
        }
    }
}

void orc_hinge_cost_f64(const orc_robot* r, const orc_scene* s, const float* q, int n, const orc_cost_params* p,
                        double* cost, int nthreads) {
    hinge_cost_t<double>(r, s, q, n, p, cost, nthreads);
}

void orc_hinge_cost_f32(const orc_robot* r, const orc_scene* s, const float* q, int n, const orc_cost_params* p,
                        float* cost, int nthreads) {
    hinge_cost_t<float>(r, s, q, n, p, cost, nthreads);
}

void orc_hinge_grad_f64(const orc_robot* r, const orc_scene* s, const float* q, int n, const orc_cost_params* p,
                        double* cost, double* grad, int* kink, double tol, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel
    {
        Work<double> w(*r);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            bool k = false;
            cost[i] = hinge_one<double>(*r, *s, q + (size_t)i * r->n_dof, *p, w, grad + (size_t)i * r->n_dof, &k, tol);
            if (kink) kink[i] = k ? 1 : 0;
        }
    }
}

void orc_hinge_grad_f32(const orc_robot* r, const orc_scene* s, const float* q, int n, const orc_cost_params* p,
                        float* cost, float* grad, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel
    {
        Work<float> w(*r);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            bool k = false;
            cost[i] = hinge_one<float>(*r, *s, q + (size_t)i * r->n_dof, *p, w, grad + (size_t)i * r->n_dof, &k, 0.0f);
        }
    }
}

void orc_traj_cost_grad_f64(const orc_robot* r, const orc_scene* s, const float* theta, const int* off, int B,
                            const orc_cost_params* p, double* total, double* grad, int* kink, double tol,
                            int nthreads) {
    set_threads(nthreads);
    const int D = r->n_dof;
#pragma omp parallel
    {
        Work<double> w(*r);
        float q[kMaxJoints];
        double g[kMaxJoints];
#pragma omp for schedule(dynamic, 4)
        for (int b = 0; b < B; ++b) {
            const int T0 = off[b], Tn = off[b + 1] - off[b];
            const float* th = theta + (size_t)T0 * D;
            double acc = 0;
            for (int t = 0; t < Tn; ++t) {
                for (int d = 0; d < D; ++d) q[d] = th[(size_t)d * Tn + t];
                bool k = false;
                acc += hinge_one<double>(*r, *s, q, *p, w, grad ? g : nullptr, &k, tol);
                const bool end = (t == 0 || t == Tn - 1);
                if (grad)
                    for (int d = 0; d < D; ++d) grad[(size_t)T0 * D + (size_t)d * Tn + t] = end ? 0.0 : g[d];
                if (kink) kink[T0 + t] = (k && !end) ? 1 : 0;
            }
            total[b] = acc;
        }
    }
}

void orc_chomp_fd_grad_f64(const orc_robot* r, const orc_scene* s, const float* theta, const int* off, int B,
                           const orc_cost_params* p, float eps, double* grad, int nthreads) {
    chomp_fd_t<double>(r, s, theta, off, B, p, eps, grad, nthreads);
}

void orc_chomp_fd_grad_f32(const orc_robot* r, const orc_scene* s, const float* theta, const int* off, int B,
                           const orc_cost_params* p, float eps, float* grad, int nthreads) {
    chomp_fd_t<float>(r, s, theta, off, B, p, eps, grad, nthreads);
}

int orc_edge_granularity(const float* from, const float* to, int dim, int granularity, float resolution,
                         float* len_out) {
    float len2 = 0.0f;
    for (int i = 0; i < dim; ++i) {
        const float d = to[i] - from[i];
        len2 = std::fma(d, d, len2);
    }
    const float len = std::sqrt(len2);
    if (len_out) *len_out = len;
    if (!(resolution > 0.0f)) return granularity;   // edge_eval.cuh:54-60
    int g = (int)std::ceil(len / resolution);
    if (g < 2) g = 2;
    if (g > 1024) g = 1024;
    return g;
}

void orc_edge_eval_f64(const orc_robot* r, const orc_scene* s, const float* from, const float* to, int E,
                       const orc_edge_params* p, uint8_t* valid, float* cost, double* clearance, int nthreads) {
    set_threads(nthreads);
    const int D = r->n_dof;
#pragma omp parallel
    {
        Work<double> w(*r);
        float q[kMaxJoints];
#pragma omp for schedule(dynamic, 8)
        for (int e = 0; e < E; ++e) {
            const float* a = from + (size_t)e * D;
            const float* b = to + (size_t)e * D;
            float len;
            const int G = orc_edge_granularity(a, b, D, p->granularity, p->resolution, &len);
            bool bad = false;
            double clr = (double)FLT_MAX;
            for (int j = 1; j <= G; ++j) {
                const float t = (float)j / (float)G;
                lerp_state(a, b, D, t, q);
                if (!bad && check_one<double>(*r, *s, q, (double)p->collision_margin, p->check_self != 0, 0, w, nullptr))
                    bad = true;
                if (p->compute_clearance)
                    clr = std::min(clr, min_clearance_one<double>(*r, *s, q, (double)p->collision_margin, false, w));
                if (bad && !p->compute_clearance) break;
            }
            valid[e] = bad ? 0 : 1;
            cost[e] = bad ? FLT_MAX : len;
            if (clearance) clearance[e] = p->compute_clearance ? clr : (double)FLT_MAX;
        }
    }
}

void orc_edge_batch_f64(const orc_robot* r, const orc_scene* s, const float* from, const float* to,
                        const int* n_states, int E, int skip_first, float margin, int check_self, int* first_hit,
                        int nthreads) {
    edge_batch_t<double>(r, s, from, to, n_states, E, skip_first, margin, check_self, first_hit, nthreads);
}

void orc_edge_batch_f32(const orc_robot* r, const orc_scene* s, const float* from, const float* to,
                        const int* n_states, int E, int skip_first, float margin, int check_self, int* first_hit,
                        int nthreads) {
    edge_batch_t<float>(r, s, from, to, n_states, E, skip_first, margin, check_self, first_hit, nthreads);
}

void orc_edge_states_f32(const float* from, const float* to, const int* n_states, int E, int dim, int skip_first,
                         float* out) {
    size_t o = 0;
    for (int e = 0; e < E; ++e) {
        const int n = n_states[e];
        for (int k = skip_first ? 1 : 0; k < n; ++k) {
            const float t = n > 1 ? (float)k / (float)(n - 1) : 0.0f;
            lerp_state(from + (size_t)e * dim, to + (size_t)e * dim, dim, t, out + o * dim);
            ++o;
        }
    }
}

void orc_control_R_f64(int T, double dt, double* R) { control_R(T, dt, R); }

void orc_control_cost_f64(const float* theta, const int* off, int B, int D, double dt, double w, double* out) {
    for (int b = 0; b < B; ++b) {
        const int Tn = off[b + 1] - off[b];
        std::vector<double> Rm((size_t)Tn * Tn);
        control_R(Tn, dt, Rm.data());
        out[b] = w * control_one(theta + (size_t)off[b] * D, D, Tn, Rm.data());
    }
}

void orc_stomp_costs_f64(const orc_robot* r, const orc_scene* all, const int* obb_off, const float* rollouts,
                         const int* T_p, const int* roff, int P, int R, const orc_cost_params* p, double w_u,
                         double dt, double* state, double* total, int nthreads, int* nterm) {
    stomp_costs_t<double>(r, all, obb_off, rollouts, T_p, roff, P, R, p, w_u, dt, state, total, nthreads, nterm);
}

void orc_stomp_costs_f32(const orc_robot* r, const orc_scene* all, const int* obb_off, const float* rollouts,
                         const int* T_p, const int* roff, int P, int R, const orc_cost_params* p, double w_u,
                         double dt, float* state, float* total, int nthreads) {
    stomp_costs_t<float>(r, all, obb_off, rollouts, T_p, roff, P, R, p, w_u, dt, state, total, nthreads);
}

void orc_gated_flops_f64(const orc_robot* r, const orc_scene* s, const float* q, int n, float margin, double* flops,
                         int nthreads) {
    set_threads(nthreads);
    const orc_robot& R = *r;
    const orc_scene& S = *s;
#pragma omp parallel
    {
        Work<double> w(R);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            fk_frames<double>(R, q + (size_t)i * R.n_dof, w.frames.data(), false);
            sphere_positions<double>(R, w.frames.data(), false, w.P.data());
            sphere_positions<double>(R, w.frames.data(), true, w.A.data());
            double f = 0;
            bool hit = false;
            std::vector<char> seen(R.n_spheres, 0);
            for (int k = 0; k < R.n_joints && !hit; ++k) {
                const int j = R.dfs_order[k];
                const int t = R.joint_types[j];
                if (k > 0) f += (t >= kXRot) ? 83 : (t >= 0 ? 69 : 0);
                int a = -1;
                for (int x = 0; x < R.n_approx; ++x)
                    if (R.approx_sphere_to_joint[x] == j) { a = x; break; }
                bool has = false;
                for (int x = 0; x < R.n_spheres; ++x) has |= R.sphere_to_joint[x] == j;
                if (a < 0 || !has) continue;
                f += 18;
                const double ar = (double)R.approx_spheres[4 * a + 3] + margin;
                for (int o = 0; o < S.n_obbs && !hit; ++o) {
                    if (S.obb_enable[o] == 0) continue;
                    f += 30;
                    if (!obb_collide<double>(S.obb_dims + 4 * o, S.obb_pose + 8 * o, w.A[a].x, w.A[a].y, w.A[a].z, ar))
                        continue;
                    for (int x = 0; x < R.n_spheres; ++x) {
                        if (R.sphere_to_joint[x] != j) continue;
                        if (!seen[x]) { seen[x] = 1; f += 18; }
                        f += 30;
                        if (obb_collide<double>(S.obb_dims + 4 * o, S.obb_pose + 8 * o, w.P[x].x, w.P[x].y, w.P[x].z,
                                                (double)R.spheres[4 * x + 3] + margin))
                            hit = true;
                    }
                }
            }
            flops[i] = f;
        }
    }
}

void orc_esdf_from_obbs_f64(const float* obb_dims, const float* obb_pose, int n_obbs, const float* voxel_params,
                            const float* voxel_pose, double* out, int nthreads) {
    set_threads(nthreads);
    const float vs = voxel_params[3];
    const float inv_vs = 1.0f / vs;
    const int nx = robust_floor(voxel_params[0] * inv_vs) + 1;
    const int ny = robust_floor(voxel_params[1] * inv_vs) + 1;
    const int nz = robust_floor(voxel_params[2] * inv_vs) + 1;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                // Node in grid-local coordinates, then back to world with the
                // inverse of the stored (world -> grid) pose.
                const double lx = ix * (double)vs - 0.5 * voxel_params[0];
                const double ly = iy * (double)vs - 0.5 * voxel_params[1];
                const double lz = iz * (double)vs - 0.5 * voxel_params[2];
                V3<double> p = local_vec_to_world<double>(voxel_pose, lx - voxel_pose[0], ly - voxel_pose[1], lz - voxel_pose[2]);
                double best = 1e30;
                for (int o = 0; o < n_obbs; ++o)
                    best = std::min(best, obb_clearance<double>(obb_dims + 4 * o, obb_pose + 8 * o, p.x, p.y, p.z, 0.0));
                out[(size_t)ix * ny * nz + (size_t)iy * nz + iz] = -best;
            }
}

}  // extern "C"
