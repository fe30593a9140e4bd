"""The names the reference's planner plugins reach for exist here (CPU: no
device call is made).

The plugin base and its six plugins (python/gpu_planning/plugins/) do
``import prrtc`` with the compiled module's directory on PYTHONPATH, build
the robot with ``robot_model_loader.build_robot_model`` and hold
``PlanningScene._voxel_plan_lock`` while a scene builder is alive
(_substrate_planner_base.py:43, 208-219, 257-292); each plugin copies its
YAML section's keys onto the settings object (chomp_plugin.py:40-50,
stomp_plugin.py:41-52, prrtc_plugin.py:42-53)."""
import subprocess
import sys
import threading
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PYPATH = ROOT / "prismplan_amd" / "pythonpath"

# keys of the reference's config/planners/{chomp,stomp,prrtc}.yaml sections
CHOMP_KEYS = ("num_timesteps", "num_iterations", "num_iterations_after_valid", "time_limit_ms", "step_size",
              "smoothness_weight", "collision_cost_weight", "fd_epsilon", "world_collision_margin",
              "self_collision_margin", "delta_t", "convergence_eps", "collision_free_tol", "collision_check_step",
              "collision_margin")
STOMP_KEYS = ("num_batch", "num_timesteps", "num_iterations", "num_iterations_after_valid", "time_limit_ms",
              "num_rollouts_new", "num_rollouts_old", "delta_t", "control_cost_weight", "collision_cost_weight",
              "exponentiated_cost_sensitivity", "world_collision_margin", "self_collision_margin",
              "convergence_eps", "collision_free_tol", "collision_check_step", "noise_scale", "collision_margin")
PRRTC_KEYS = ("max_samples", "max_iters", "time_limit_ms", "num_new_configs", "granularity", "range", "balance",
              "tree_ratio", "dynamic_domain", "dd_alpha", "dd_radius", "dd_min_radius", "shortcut_path",
              "valid_segment_length", "collision_margin", "shortcut_collision_margin")


def test_top_level_prrtc_import_from_the_pythonpath_directory():
    """A fresh interpreter with only that directory on PYTHONPATH (run from
    elsewhere) imports ``prrtc`` and finds the plugins' names."""
    code = ("import prrtc; names = ['RobotModelBuilder', 'SceneCollisionDataBuilder', 'check_collision_scene', "
            "'Settings', 'SolverBuffers', 'solve_scene', 'CHOMPSettings', 'CHOMPBuffers', 'chomp_solve_scene', "
            "'STOMPSettings', 'STOMPBuffers', 'stomp_solve_scene']; "
            "missing = [n for n in names if not hasattr(prrtc, n)]; assert not missing, missing; print('ok')")
    env = {"PYTHONPATH": str(PYPATH), "PATH": "/usr/bin:/bin", "HOME": "/tmp"}
    out = subprocess.run([sys.executable, "-c", code], cwd="/tmp", env=env, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.strip().endswith("ok")


def test_build_robot_model_and_scene_lock_names():
    from prismplan_amd import robot_model_loader as L
    from prismplan_amd.planning_scene import PlanningScene
    assert callable(L.build_robot_model)
    sc = PlanningScene.from_robot_asset("panda")
    lk = sc._voxel_plan_lock
    assert isinstance(lk, type(threading.RLock()))
    with lk:
        with lk:   # re-entrant, as the base takes it around calls that take it again
            sc.add_cuboid("box", [0.1, 0.1, 0.1], [0.5, 0.0, 0.3])
    assert sc.current_obstacles[0]["name"] == "box"


@pytest.mark.parametrize("cls,keys", [("CHOMPSettings", CHOMP_KEYS), ("STOMPSettings", STOMP_KEYS),
                                      ("Settings", PRRTC_KEYS)])
def test_settings_carry_every_key_the_plugins_set(cls, keys):
    from prismplan_amd import prrtc
    s = getattr(prrtc, cls)()
    missing = [k for k in keys if not hasattr(s, k)]
    assert not missing, f"{cls} lacks {missing}"
    for k in keys:   # writable with the YAML value's type
        v = getattr(s, k)
        setattr(s, k, v)
