"""Pins the oracle's kinematics and distance primitives with closed forms,
an independent matrix formulation and library routines (CPU)."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import pyoracle as PO
from prismplan_amd.planning_scene import HostScene, inverse_pose, obb_arrays
from prismplan_amd.robot_model_loader import ROBOTS, robot_arrays

X_PRISM, Y_PRISM, Z_PRISM, X_ROT, Y_ROT, Z_ROT = range(6)


def joint_motion(jtype, q, f32_trig=False):
    """Textbook homogeneous joint motion (the reference's own test oracle,
    tests/test_tree_fk_vs_cpu.py:25-36, restated). f32_trig: cos and sin
    rounded to FP32, as any FP32 evaluation must hold them."""
    T = np.eye(4)
    c, s = np.cos(q), np.sin(q)
    if f32_trig:
        c, s = float(np.float32(c)), float(np.float32(s))
    if jtype == X_ROT:
        T[1:3, 1:3] = [[c, -s], [s, c]]
    elif jtype == Y_ROT:
        T[0, 0], T[0, 2], T[2, 0], T[2, 2] = c, s, -s, c
    elif jtype == Z_ROT:
        T[0:2, 0:2] = [[c, -s], [s, c]]
    elif jtype in (X_PRISM, Y_PRISM, Z_PRISM):
        T[jtype, 3] = q
    return T


def matrix_fk(r, q, f32_trig=False):
    ft = r.fixed_transforms.astype(np.float64).reshape(r.n_joints, 4, 4)
    frames = [None] * r.n_joints
    for k, i in enumerate(r.dfs_order):
        if k == 0:
            frames[i] = np.eye(4)
            continue
        dof = r.joint_id_to_dof[i]
        frames[i] = frames[r.joint_parents[i]] @ ft[i] @ joint_motion(r.joint_types[i], float(q[dof]) if dof >= 0 else 0.0,
                                                                      f32_trig)
    P = np.empty((r.n_spheres, 3))
    for s in range(r.n_spheres):
        P[s] = (frames[r.sphere_to_joint[s]] @ np.append(r.spheres[s, :3].astype(np.float64), 1.0))[:3]
    return P, frames


def rand_q(r, rng, n):
    lo, hi = r.joint_lower.astype(np.float64), r.joint_upper.astype(np.float64)
    return (lo + rng.random((n, r.n_dof)) * (hi - lo)).astype(np.float32)


@pytest.mark.parametrize("name", ROBOTS)
def test_fk_matches_matrix_formulation(name):
    r = robot_arrays(name)
    o = PO.Oracle(r)
    q = rand_q(r, np.random.default_rng(5), 20)
    P = o.fk_spheres(q)
    for i in range(q.shape[0]):
        ref, _ = matrix_fk(r, q[i])
        np.testing.assert_allclose(P[i], ref, atol=1e-12, rtol=0)


@pytest.mark.parametrize("name", ROBOTS)
def test_fp32_position_floor(name):
    """The FP32 error model behind the GPU parity bars (tests/parity_common.py,
    TERM_FLOOR = 0.05 m, i.e. 5e-7 m at 1e-5 relative). Both FP32 figures
    are pinned on 1,000 random configurations:
      * even with exact arithmetic after FP32-rounded sin and cos, and only
        the final sphere position rounded to FP32, spheres land more than
        5e-8 m from FP64 somewhere: an FP32 evaluation cannot be held much
        tighter than a few 1e-7 m;
      * the FP32 restatement (ordinary FP32 arithmetic) stays within
        5e-7 m = 1e-5 x TERM_FLOOR: FP32 arithmetic meets the bar."""
    r = robot_arrays(name)
    o = PO.Oracle(r)
    q = rand_q(r, np.random.default_rng(77), 1000)
    p64 = o.fk_spheres(q)
    p32 = o.fk_spheres(q, precision="f32").astype(np.float64)
    ideal = np.stack([matrix_fk(r, q[i], f32_trig=True)[0] for i in range(len(q))]).astype(np.float32)
    e_ideal = np.linalg.norm(ideal.astype(np.float64) - p64, axis=2).max()
    e32 = np.linalg.norm(p32 - p64, axis=2).max()
    assert 5e-8 < e_ideal < 5e-7, e_ideal
    assert e32 < 5e-7, e32


def test_panda_zero_pose_closed_form():
    r = robot_arrays("panda")
    o = PO.Oracle(r)
    F = o.fk_frames(np.zeros(7))[0]
    # link origins of the zero pose: z = 0.333 (joint 1), shoulder stack 0.333+0.316, elbow offset 0.0825 ...
    np.testing.assert_allclose(F[1][:, 3], [0, 0, 0.333], atol=1e-12)
    np.testing.assert_allclose(F[3][:, 3], [0, 0, 0.649], atol=1e-12)
    np.testing.assert_allclose(F[4][:, 3], [0.0825, 0, 0.649], atol=1e-12)
    np.testing.assert_allclose(F[5][:, 3], [0.0, 0, 1.033], atol=1e-12)
    np.testing.assert_allclose(F[7][:, 3], [0.088, 0, 1.033], atol=1e-12)


def test_base_rotation_rotates_descendants():
    r = robot_arrays("panda")
    o = PO.Oracle(r)
    rng = np.random.default_rng(1)
    q = rand_q(r, rng, 1)[0]
    dq = np.float32(0.3)
    q2 = q.copy()
    q2[0] += dq
    P1, P2 = o.fk_spheres(q)[0], o.fk_spheres(q2)[0]
    d = float(q2[0]) - float(q[0])
    Rz = Rotation.from_euler("z", d).as_matrix()
    moving = r.sphere_to_joint >= 1
    np.testing.assert_allclose(P2[moving], P1[moving] @ Rz.T, atol=1e-12)
    np.testing.assert_allclose(P2[~moving], P1[~moving], atol=0)


def test_fetch_prismatic_torso_translates():
    r = robot_arrays("fetch")
    assert r.joint_types[1] == Z_PRISM
    o = PO.Oracle(r)
    q = rand_q(r, np.random.default_rng(2), 1)[0]
    q[0] = np.float32(0.1)
    q2 = q.copy()
    q2[0] = np.float32(0.3)
    P1, P2 = o.fk_spheres(q)[0], o.fk_spheres(q2)[0]
    moving = r.sphere_to_joint >= 1
    delta = float(q2[0]) - float(q[0])
    np.testing.assert_allclose(P2[moving] - P1[moving], np.tile([0, 0, delta], (moving.sum(), 1)), atol=1e-12)


def test_baxter_arms_are_independent():
    r = robot_arrays("baxter")
    o = PO.Oracle(r)
    rng = np.random.default_rng(3)
    q = rand_q(r, rng, 1)[0]
    q2 = q.copy()
    q2[:7] = rand_q(r, rng, 1)[0][:7]
    P1, P2 = o.fk_spheres(q)[0], o.fk_spheres(q2)[0]
    right = r.sphere_to_joint >= 8
    left = (r.sphere_to_joint >= 1) & (r.sphere_to_joint <= 7)
    np.testing.assert_array_equal(P1[right], P2[right])
    assert np.abs(P1[left] - P2[left]).max() > 1e-3


# ---------------------------------------------------------------------------
# sphere vs OBB
# ---------------------------------------------------------------------------
def box(dims, pos=(0, 0, 0), quat=(1, 0, 0, 0)):
    d, p, _ = obb_arrays(np.asarray(dims)[None], np.asarray(pos)[None], np.asarray(quat)[None])
    return d[0], p[0]


@pytest.mark.parametrize("p,expect", [
    ((0.5, 0.0, 0.0), 0.5 - 0.1),                                   # face
    ((0.0, -0.7, 0.0), 0.7 - 0.2),
    ((0.4, 0.5, 0.0), np.hypot(0.4 - 0.1, 0.5 - 0.2)),              # edge
    ((0.4, 0.5, 0.6), np.sqrt(0.3 ** 2 + 0.3 ** 2 + 0.3 ** 2)),     # corner
    ((0.05, 0.0, 0.0), -min(0.1 - 0.05, 0.2, 0.3)),                 # inside, nearest face x
    ((0.0, 0.15, 0.01), -min(0.1, 0.2 - 0.15, 0.3 - 0.01)),
])
def test_axis_aligned_box_sdf(p, expect):
    d, pose = box((0.2, 0.4, 0.6))
    r = 0.03
    # dims are stored as float32 (0.2 -> 0.2000000030): 1e-7 m covers that cast
    assert PO.sphere_obb_clearance(d, pose, p, r) == pytest.approx(expect - r, abs=1e-7)


def test_rotated_box_against_scipy_rotation():
    rng = np.random.default_rng(7)
    for _ in range(200):
        dims = rng.uniform(0.05, 0.6, 3)
        pos = rng.uniform(-1, 1, 3)
        rot = Rotation.random(random_state=rng)
        x, y, z, w = rot.as_quat()
        d, pose = box(dims, pos, (w, x, y, z))
        p = rng.uniform(-1.5, 1.5, 3)
        r = float(rng.uniform(0.0, 0.1))
        # textbook: local = R^T (p - pos), using the float32-stored pose values' rotation
        qs = pose[3:7].astype(np.float64)
        R_store = Rotation.from_quat([qs[1], qs[2], qs[3], qs[0]]).as_matrix()
        local = R_store @ p + pose[:3].astype(np.float64)
        q = np.abs(local) - d[:3].astype(np.float64) / 2
        sdf = np.linalg.norm(np.maximum(q, 0)) + min(q.max(), 0.0)
        assert PO.sphere_obb_clearance(d, pose, p, r) == pytest.approx(sdf - r, abs=1e-6)
        # and the world->local map really is the inverse of the box pose
        np.testing.assert_allclose(rot.apply(local) + pos, p, atol=1e-6)


def test_quaternion_sign_invariance_and_predicate():
    rng = np.random.default_rng(8)
    for _ in range(300):
        dims = rng.uniform(0.05, 0.6, 3)
        quat = rng.standard_normal(4)
        quat /= np.linalg.norm(quat)
        d, pose = box(dims, rng.uniform(-0.5, 0.5, 3), quat)
        pose_neg = pose.copy()
        pose_neg[3:7] *= -1
        p = rng.uniform(-1, 1, 3)
        r = float(rng.uniform(0, 0.2))
        c1 = PO.sphere_obb_clearance(d, pose, p, r)
        c2 = PO.sphere_obb_clearance(d, pose_neg, p, r)
        assert c1 == pytest.approx(c2, abs=1e-12)
        hit = PO.sphere_obb_collide(d, pose, p, r)
        if abs(c1) > 1e-9:
            assert hit == (c1 < 0)


def test_brute_force_distance():
    """Sampled surface distance agrees with the analytic SDF outside the box."""
    rng = np.random.default_rng(9)
    dims = np.array([0.3, 0.2, 0.5])
    d, pose = box(dims, (0.1, -0.2, 0.3), (np.cos(0.4), 0, np.sin(0.4), 0))
    # dense surface sample of the box in world coordinates
    rot = Rotation.from_quat([0, np.sin(0.4), 0, np.cos(0.4)])
    g = np.linspace(-0.5, 0.5, 41)
    pts = []
    for ax in range(3):
        for sgn in (-0.5, 0.5):
            u, v = np.meshgrid(g, g)
            f = np.zeros((u.size, 3))
            f[:, ax] = sgn
            f[:, (ax + 1) % 3] = u.ravel()
            f[:, (ax + 2) % 3] = v.ravel()
            pts.append(f * dims)
    surf = rot.apply(np.concatenate(pts)) + np.array([0.1, -0.2, 0.3])
    for _ in range(50):
        p = rng.uniform(-1, 1, 3)
        c = PO.sphere_obb_clearance(d, pose, p, 0.0)
        if c <= 0.05:
            continue
        bf = np.min(np.linalg.norm(surf - p, axis=1))
        assert bf >= c - 1e-6   # pose stored in float32
        assert bf <= c + 0.02


def test_disabled_obb_equals_removal_and_order_irrelevant():
    from prismplan_amd import workloads as W
    w = W.any_collision(n_configs=256, n_obbs=12, seed=11)[0]
    o = PO.Oracle(w.robot, w.scene)
    hit_all, clr_all = o.check_collision(w.configs)
    perm = np.random.default_rng(0).permutation(w.scene.n_obbs)
    s2 = HostScene(obb_dims=w.scene.obb_dims[perm], obb_pose=w.scene.obb_pose[perm], obb_enable=w.scene.obb_enable[perm])
    o.set_scene(s2)
    hit_p, clr_p = o.check_collision(w.configs)
    np.testing.assert_array_equal(hit_all, hit_p)
    np.testing.assert_allclose(clr_all, clr_p, atol=0)
    en = w.scene.obb_enable.copy()
    en[:5] = 0
    o.set_scene(HostScene(obb_dims=w.scene.obb_dims, obb_pose=w.scene.obb_pose, obb_enable=en))
    hit_d, _ = o.check_collision(w.configs)
    o.set_scene(HostScene(obb_dims=w.scene.obb_dims[5:], obb_pose=w.scene.obb_pose[5:], obb_enable=en[5:]))
    hit_r, _ = o.check_collision(w.configs)
    np.testing.assert_array_equal(hit_d, hit_r)
