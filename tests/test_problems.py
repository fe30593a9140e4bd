"""Benchmark problem normalisation (prismplan_amd/problems.py) against the
mathematics and the reference's bundled MotionBenchMaker sample, mirroring
the reference's tests/test_benchmark.py cases."""
import json
from pathlib import Path

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import pyoracle as PO
from prismplan_amd import problems as P
from prismplan_amd.planning_scene import HostScene
from prismplan_amd.robot_model_loader import robot_arrays

GOLDEN = Path(__file__).resolve().parent / "golden"
SAMPLE = GOLDEN / "sample_mbm_problems.json"


def quat_R(q):
    return Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()


def obb_contains(obb, pts, tol=1e-9):
    R = quat_R(obb["quaternion"])
    local = (np.asarray(pts) - np.asarray(obb["position"])) @ R
    return np.all(np.abs(local) <= 0.5 * np.asarray(obb["dims"]) + tol)


def test_euler_identity_and_convention():
    np.testing.assert_allclose(P.euler_xyz_to_quat_wxyz([0, 0, 0]), [1, 0, 0, 0])
    rng = np.random.default_rng(0)
    for e in rng.uniform(-np.pi, np.pi, size=(50, 3)):
        q = P.euler_xyz_to_quat_wxyz(e)
        assert abs(np.linalg.norm(q) - 1.0) < 1e-12
        a, b, c = e
        Rx = Rotation.from_rotvec([a, 0, 0]).as_matrix()
        Ry = Rotation.from_rotvec([0, b, 0]).as_matrix()
        Rz = Rotation.from_rotvec([0, 0, c]).as_matrix()
        np.testing.assert_allclose(quat_R(q), Rz @ Ry @ Rx, atol=1e-12)
        np.testing.assert_allclose(quat_R(q), Rotation.from_euler("xyz", e).as_matrix(), atol=1e-12)


def test_box_is_exact_and_sphere_cylinder_enclose():
    q = P.euler_xyz_to_quat_wxyz([0.3, -0.2, 0.9])
    box = P.obstacle_to_obb(P.BenchObstacle("box", "b", np.array([0.1, 0.2, 0.3]), q, np.array([0.2, 0.3, 0.4])))
    np.testing.assert_allclose(box["dims"], [0.4, 0.6, 0.8])
    np.testing.assert_allclose(box["quaternion"], q)
    rng = np.random.default_rng(1)
    r = 0.17
    sph = P.obstacle_to_obb(P.BenchObstacle("sphere", "s", np.array([0.5, -0.1, 0.7]), radius=r))
    np.testing.assert_allclose(sph["dims"], [2 * r] * 3)
    d = rng.normal(size=(500, 3))
    pts = np.array(sph["position"]) + r * d / np.linalg.norm(d, axis=1, keepdims=True)
    assert obb_contains(sph, pts)
    # the cube is tight: its face centres touch the sphere
    assert not obb_contains(sph, np.array(sph["position"]) + [[r * 1.001, 0, 0]])
    qc = P.euler_xyz_to_quat_wxyz([0.4, 0.2, -0.6])
    L = 0.5
    cyl = P.obstacle_to_obb(P.BenchObstacle("cylinder", "c", np.array([0.2, 0.3, 0.1]), qc, radius=r, length=L))
    np.testing.assert_allclose(cyl["dims"], [2 * r, 2 * r, L])
    ang = rng.uniform(0, 2 * np.pi, 500)
    h = rng.uniform(-L / 2, L / 2, 500)
    local = np.stack([r * np.cos(ang), r * np.sin(ang), h], axis=1)
    pts = local @ quat_R(qc).T + np.array([0.2, 0.3, 0.1])
    assert obb_contains(cyl, pts)
    with pytest.raises(ValueError):
        P.obstacle_to_obb(P.BenchObstacle("cone"))


def test_names_are_made_unique():
    obs = [P.BenchObstacle("box", "a"), P.BenchObstacle("sphere", "a"), P.BenchObstacle("sphere", ""),
           P.BenchObstacle("box", "a")]
    assert [o["name"] for o in P.obstacles_to_obbs(obs)] == ["a_0", "a_1", "sphere_0", "a_2"]


def test_sample_loader():
    probs = P.load_vamp_problems(str(SAMPLE))
    assert len(probs) == 3 and {p.scene for p in probs} == {"box", "cage"}
    for p in probs:
        assert p.start.shape == (7,) and p.goal.shape == (7,) and p.source == "vamp"
    cage = [p for p in probs if p.scene == "cage"][0]
    assert sorted(o.kind for o in cage.obstacles) == ["box", "cylinder", "sphere"]
    one = P.load_vamp_problems(str(SAMPLE), scenes=["box"], max_per_scene=1)
    assert len(one) == 1 and one[0].scene == "box"
    raw = json.loads(SAMPLE.read_text())
    np.testing.assert_allclose(one[0].start, np.float32(raw["problems"]["box"][0]["start"]))


def test_loader_refuses_pickles_and_missing(tmp_path):
    pk = tmp_path / "x.pkl"
    pk.write_bytes(b"not read")
    with pytest.raises(ValueError):
        P.load_vamp_problems(str(pk))
    with pytest.raises(FileNotFoundError):
        P.load_vamp_problems(str(tmp_path / "missing.json"))


def test_sample_worlds_keep_start_and_goal_free():
    """Every valid sample problem's start and goal stay free in the
    conservative box world (the FP64 oracle, environment check)."""
    r = robot_arrays("panda")
    for p in P.load_vamp_problems(str(SAMPLE)):
        obbs = P.obstacles_to_obbs(p.obstacles)
        scene = HostScene.from_boxes([o["dims"] for o in obbs], [o["position"] for o in obbs],
                                     [o["quaternion"] for o in obbs])
        hit, clr = PO.Oracle(r, scene).check_collision(np.stack([p.start, p.goal]))
        assert not hit.any(), (p.scene, p.index, clr)
