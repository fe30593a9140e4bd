"""The reference planner plugins' call sequence, end to end on the GPU
(needs an MI355X).

The sequence is the one the reference's plugin base and its pRRTC, CHOMP and
STOMP plugins run (_substrate_planner_base.py:43, 200-219, 257-292;
prrtc_plugin.py:31-91; chomp_plugin.py:31-85; stomp_plugin.py:31-87),
driven here through the top-level ``prrtc`` name with the package's
``pythonpath`` directory on sys.path:

  robot = build_robot_model(asset)                   # a built RobotModelBuilder
  bufs = prrtc.<X>Buffers.create(...)                # the plugins' keyword arguments
  with scene._voxel_plan_lock:
      builder = scene.build_scene_collision_data()
      flags = prrtc.check_collision_scene(configs=, scene=, robot=, collision_margin=0.0)
      settings = prrtc.<X>Settings(); setattr(settings, key, value) per config key
      result = prrtc.<solver>(start, goal(s), builder, settings, robot, bufs=bufs)

The world and the query are the bundled MotionBenchMaker sample problems
(tests/golden/sample_mbm_problems.json). Every solved path is checked by the
FP64 oracle: it starts at the start and ends at the goal, and every state of
the path densified to 0.01 rad steps is collision-free with the self check,
except states within 1e-5 m of contact (counted).
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import pyoracle as PO
from parity_common import banner
from prismplan_amd import problems as PR
from prismplan_amd.planning_scene import PlanningScene
from prismplan_amd.robot_model_loader import build_robot_model, robot_arrays

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
PROBLEMS = ROOT / "tests" / "golden" / "sample_mbm_problems.json"
BAND = 1e-5

if str(ROOT / "prismplan_amd" / "pythonpath") not in sys.path:
    sys.path.insert(0, str(ROOT / "prismplan_amd" / "pythonpath"))

# flat plugin configs as the plugin base's YAML flattening yields them
CHOMP_CFG = {"chomp_num_timesteps": 51, "chomp_num_iterations": 200, "chomp_time_limit_ms": 2000.0,
             "collision_margin": 0.0}
STOMP_CFG = {"stomp_num_batch": 1, "stomp_num_timesteps": 51, "stomp_num_iterations": 200,
             "stomp_num_iterations_after_valid": 3, "stomp_time_limit_ms": 2000.0, "stomp_num_rollouts_new": 30,
             "stomp_num_rollouts_old": 20, "collision_margin": 0.0}
PRRTC_CFG = {"num_new_configs": 600, "granularity": 16, "range": 0.5, "max_samples": 100000,
             "time_limit_ms": 2000.0, "collision_margin": 0.0}


def problem_world(k=0):
    p = PR.load_vamp_problems(str(PROBLEMS))[k]
    scene = PlanningScene.from_robot_asset("panda")
    scene.update_world(PR.obstacles_to_obbs(p.obstacles))
    return p, scene


def densify(path, step=0.01):
    out = [path[0]]
    for a, b in zip(path[:-1], path[1:]):
        n = max(1, int(np.ceil(np.max(np.abs(b - a)) / step)))
        for k in range(1, n + 1):
            out.append(a + (b - a) * (k / n))
    return np.asarray(out, np.float32)


def check_path(path, p, scene, what):
    path = np.asarray(path, np.float32)
    assert path.ndim == 2 and path.shape[1] == 7 and path.shape[0] >= 2
    np.testing.assert_allclose(path[0], p.start, atol=1e-5)
    np.testing.assert_allclose(path[-1], p.goal, atol=1e-5)
    o = PO.Oracle(robot_arrays("panda"), scene.host_scene(), nthreads=8)
    states = densify(path)
    hit, clr = o.check_collision(states, check_self=True)
    amb = np.abs(clr) < BAND
    banner(f"{what}: {path.shape[0]} waypoints, {len(states)} densified states, colliding outside the band "
           f"{int((hit & ~amb).sum())}, in band {int(amb.sum())}")
    assert not (hit & ~amb).any()


def settings_from(cls, cfg, prefix):
    s = cls()
    for key, val in cfg.items():
        attr = key[len(prefix):] if prefix and key.startswith(prefix) else key
        if hasattr(s, attr):
            setattr(s, attr, val)
    return s


def test_plugin_sequence_chomp_stomp_prrtc():
    import prrtc   # the top-level name the plugins import

    robot = build_robot_model(None)
    assert robot.n_dof == 7 and robot.n_joints == 8
    p, scene = problem_world()
    banner(f"plugin sequence on '{p.scene}' problem {p.index}: {scene.n_obbs or len(scene.current_obstacles)} boxes")
    chomp_bufs = prrtc.CHOMPBuffers.create(n_dof=robot.n_dof, num_timesteps=CHOMP_CFG["chomp_num_timesteps"])
    stomp_bufs = prrtc.STOMPBuffers.create(n_dof=robot.n_dof, num_timesteps=STOMP_CFG["stomp_num_timesteps"],
                                           num_rollouts_new=STOMP_CFG["stomp_num_rollouts_new"],
                                           num_rollouts_old=STOMP_CFG["stomp_num_rollouts_old"],
                                           num_batch=STOMP_CFG["stomp_num_batch"])
    rrtc_bufs = prrtc.SolverBuffers.create(max_samples=PRRTC_CFG["max_samples"], n_dof=robot.n_dof,
                                           num_new_configs=PRRTC_CFG["num_new_configs"],
                                           granularity=PRRTC_CFG["granularity"], n_joints=robot.n_joints,
                                           enable_mesh=False)
    with scene._voxel_plan_lock:
        builder = scene.build_scene_collision_data()
        assert builder is not None
        flags = prrtc.check_collision_scene(configs=np.vstack([p.start, p.goal]).astype(np.float32), scene=builder,
                                            robot=robot, collision_margin=0.0)
        assert flags.shape == (2,) and not flags.any(), "the sample problems' starts and goals are free"

        res = prrtc.solve_scene(p.start, p.goal.reshape(1, -1), builder, settings_from(prrtc.Settings, PRRTC_CFG, ""),
                                robot, bufs=rrtc_bufs)
        assert res.solved
        check_path(res.path, p, scene, f"pRRTC (iters {res.iters}, {res.wall_time_ms:.2f} ms)")

        res = prrtc.chomp_solve_scene(p.start, p.goal, builder, settings_from(prrtc.CHOMPSettings, CHOMP_CFG, "chomp_"),
                                      robot, bufs=chomp_bufs)
        banner(f"CHOMP: solved {res.solved}, iterations {res.iterations_run}, state cost {res.final_state_cost:.4g}, "
               f"total {res.final_total_cost:.4g}, {res.wall_time_ms:.2f} ms")
        assert res.iterations_run >= 1 and np.isfinite(res.final_total_cost)
        if res.solved:
            check_path(res.path, p, scene, "CHOMP")

        res = prrtc.stomp_solve_scene(p.start, p.goal, builder, settings_from(prrtc.STOMPSettings, STOMP_CFG, "stomp_"),
                                      robot, bufs=stomp_bufs)
        banner(f"STOMP: solved {res.solved}, iterations {res.iterations_run}, state cost {res.final_state_cost:.4g}, "
               f"total {res.final_total_cost:.4g}, {res.wall_time_ms:.2f} ms")
        assert res.iterations_run >= 1 and np.isfinite(res.final_total_cost)
        if res.solved:
            check_path(res.path, p, scene, "STOMP")
    for b in (chomp_bufs, stomp_bufs, rrtc_bufs):
        b.destroy()
