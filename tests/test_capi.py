"""C ABI surface without a GPU: the library loads, exports every entry point
include/ppl/ppl.h declares, and rejects bad inputs before any HIP call."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from prismplan_amd import _ppl
from prismplan_amd.robot_model_loader import robot_arrays

HEADER = Path(__file__).resolve().parents[1] / "include" / "ppl" / "ppl.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ppl_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 25
    L = _ppl.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    bound = {n for n, _, _ in _ppl.SIGNATURES}
    assert set(declared()) == bound


def test_version_and_error_string():
    assert b"gfx950" in _ppl.lib().ppl_version()
    assert isinstance(_ppl.lib().ppl_last_error(), bytes)


def _desc(r, **over):
    keep = []
    f32 = lambda a: keep.append(np.ascontiguousarray(a, np.float32)) or keep[-1].ctypes.data_as(C.POINTER(C.c_float))
    i32 = lambda a: keep.append(np.ascontiguousarray(a, np.int32)) or keep[-1].ctypes.data_as(C.POINTER(C.c_int))
    d = _ppl.RobotDesc()
    d.n_dof, d.n_joints = r.n_dof, r.n_joints
    ft = over.get("ft", r.fixed_transforms)
    d.fixed_transforms = f32(ft)
    d.joint_types = i32(over.get("types", r.joint_types))
    d.n_spheres, d.spheres, d.sphere_to_joint = r.n_spheres, f32(r.spheres), i32(r.sphere_to_joint)
    d.n_self_cc_ranges, d.self_cc_ranges = len(r.self_ranges), i32(r.self_ranges)
    d.n_approx_spheres, d.approx_spheres = r.n_approx, f32(r.approx_spheres)
    d.approx_sphere_to_joint = i32(r.approx_sphere_to_joint)
    d.n_approx_self_cc_ranges, d.approx_self_cc_ranges = len(r.approx_self_ranges), i32(r.approx_self_ranges)
    d.joint_lower, d.joint_upper = f32(r.joint_lower), f32(r.joint_upper)
    d.joint_parents = i32(over.get("parents", r.joint_parents))
    d.dfs_order = i32(over.get("dfs", r.dfs_order))
    d.joint_id_to_dof = i32(r.joint_id_to_dof)
    d.t_memory_idx = i32(over.get("tmem", r.t_memory_idx))
    if "aft" in over:
        d.approx_fixed_transforms = f32(over["aft"])
    if "atypes" in over:
        d.approx_joint_types = i32(over["atypes"])
    if "n_dof" in over:
        d.n_dof = over["n_dof"]
    return d, keep


@pytest.mark.parametrize("over,status", [
    ({"n_dof": 0}, _ppl.PPL_ERR_INVALID),
    ({"n_dof": 99}, _ppl.PPL_ERR_INVALID),
    ({"dfs": [1, 0, 2, 3, 4, 5, 6, 7]}, _ppl.PPL_ERR_INVALID),           # root not first
    ({"dfs": [0, 0, 2, 3, 4, 5, 6, 7]}, _ppl.PPL_ERR_INVALID),           # not a permutation
    ({"parents": [0, 0, 1, 2, 3, 4, 7, 6]}, _ppl.PPL_ERR_INVALID),       # child before parent
    ({"types": [3, 5, 5, 9, 5, 5, 5, 5]}, _ppl.PPL_ERR_INVALID),         # unknown joint type
    ({"tmem": [0, 1, 2, 3, 1, 1, 1, 1]}, _ppl.PPL_ERR_UNSUPPORTED),      # more FK slots than supported
    ({"aft": np.where(np.arange(128) == 28, 0.5, robot_arrays("panda").fixed_transforms)},
     _ppl.PPL_ERR_INVALID),                                             # approximate chain not affine
    ({"aft": robot_arrays("panda").fixed_transforms, "atypes": [3, 5, 5, 5, 9, 5, 5, 5]},
     _ppl.PPL_ERR_INVALID),                                             # unknown approximate joint type
])
def test_robot_descriptor_validation(over, status):
    r = robot_arrays("panda")
    d, keep = _desc(r, **over)
    h = C.c_void_p()
    st = _ppl.lib().ppl_robot_create(C.byref(d), C.byref(h))
    assert st == status
    assert h.value is None
    assert len(_ppl.lib().ppl_last_error()) > 0


def test_non_affine_fixed_transform_rejected():
    r = robot_arrays("panda")
    ft = r.fixed_transforms.copy()
    ft[16 + 12] = 0.5   # joint 1 bottom row
    d, keep = _desc(r, ft=ft)
    h = C.c_void_p()
    assert _ppl.lib().ppl_robot_create(C.byref(d), C.byref(h)) == _ppl.PPL_ERR_INVALID
    assert b"last row" in _ppl.lib().ppl_last_error()


def test_null_arguments_rejected():
    L = _ppl.lib()
    assert L.ppl_robot_create(None, None) == _ppl.PPL_ERR_INVALID
    assert L.ppl_check_collision_scene(None, None, None, 0, C.c_float(0), None) == _ppl.PPL_ERR_INVALID
    assert L.ppl_min_clearance_scene(None, None, None, 0, C.c_float(0), None) == _ppl.PPL_ERR_INVALID
    assert L.ppl_hinge_cost_scene(None, None, None, 0, None, None) == _ppl.PPL_ERR_INVALID
    assert L.ppl_plan_states(None) == -1
    assert L.ppl_robot_destroy(None) == _ppl.PPL_OK
    assert L.ppl_plan_destroy(None) == _ppl.PPL_OK
    assert L.ppl_graph_destroy(None) == _ppl.PPL_OK


def test_prrtc_builder_fields():
    from prismplan_amd import prrtc
    b = prrtc.RobotModelBuilder()
    for f in ("n_dof", "n_joints", "fixed_transforms", "joint_types", "spheres_flat", "sphere_to_joint",
              "self_cc_ranges_flat", "approx_spheres_flat", "approx_sphere_to_joint", "approx_self_cc_ranges_flat",
              "approx_fixed_transforms", "approx_joint_types", "joint_lower", "joint_upper", "joint_parents",
              "joint_id_to_dof", "t_memory_idx", "dfs_order"):
        assert hasattr(b, f)
    s = prrtc.SceneCollisionDataBuilder()
    for f in ("obb_dims_ptr", "obb_pose_ptr", "obb_enable_ptr", "n_obbs", "voxel_params_ptr", "voxel_pose_ptr",
              "voxel_enable_ptr", "voxel_features_ptr", "n_voxel_layers", "max_voxels_per_layer",
              "sphere_acm_mask_ptr", "sphere_obb_acm_mask_ptr", "sphere_obb_acm_mask_approx_ptr"):
        assert hasattr(s, f)


def test_prrtc_chomp_surface_and_defaults():
    """CHOMPSettings carries the reference's field names and defaults
    (planning/CHOMP_settings.hh:21-57, pRRTC_settings.hh:6,25)."""
    from prismplan_amd import prrtc
    st = prrtc.CHOMPSettings()
    want = dict(time_limit_ms=4000.0, collision_margin=0.0, num_timesteps=51, num_iterations=100,
                num_iterations_after_valid=5, step_size=0.3, smoothness_weight=1.0, collision_cost_weight=20.0,
                fd_epsilon=1e-3, world_collision_margin=0.02, self_collision_margin=0.01, delta_t=0.1,
                convergence_eps=1e-4, collision_free_tol=1e-6, collision_check_step=0.05)
    for k, v in want.items():
        assert abs(getattr(st, k) - v) <= 1e-7 * max(1.0, abs(v)), k
    for f in ("solved", "path", "iterations_run", "final_state_cost", "final_control_cost", "final_total_cost",
              "wall_time_ms"):
        assert hasattr(prrtc.CHOMPResult, f)
    for f in ("owns_memory", "n_dof", "num_timesteps", "create", "destroy"):
        assert hasattr(prrtc.CHOMPBuffers, f)
    assert callable(prrtc.chomp_solve_scene)


def test_prrtc_planner_surface_and_defaults():
    """pRRTC's Settings carry the reference's names and defaults
    (planning/pRRTC_settings.hh:3-32, python/bindings.cpp:919-938)."""
    from prismplan_amd import prrtc
    st = prrtc.Settings()
    want = dict(max_samples=1000000, max_iters=1000000, time_limit_ms=4000.0, gpu_clock_rate_khz=0,
                num_new_configs=600, granularity=16, range=0.5, balance=1, tree_ratio=0.5, dynamic_domain=True,
                dd_alpha=1e-4, dd_radius=4.0, dd_min_radius=1.0, enable_mesh_collision=False, shortcut_path=False,
                valid_segment_length=0.025, collision_margin=0.0, shortcut_collision_margin=5e-4)
    for k, v in want.items():
        assert abs(getattr(st, k) - v) <= 1e-7 * max(1.0, abs(v)), k
    for f in ("solved", "path", "iters", "cost", "wall_time_ms"):
        assert hasattr(prrtc.PlannerResult, f)
    for f in ("owns_memory", "max_samples", "n_dof", "create", "destroy"):
        assert hasattr(prrtc.SolverBuffers, f)
    with pytest.raises(RuntimeError):
        prrtc.SolverBuffers.create(max_samples=100, n_dof=7, num_new_configs=8, granularity=16, n_joints=9,
                                   enable_mesh=True)
    with pytest.raises(RuntimeError):
        prrtc.SolverBuffers.create(max_samples=100, n_dof=0, num_new_configs=8, granularity=16, n_joints=9,
                                   enable_mesh=False)
    assert callable(prrtc.solve_scene)


def test_chomp_solve_rejects_bad_arguments():
    L = _ppl.lib()
    out = C.c_void_p()
    assert L.ppl_chomp_buffers_create(0, 51, C.byref(out)) == _ppl.PPL_ERR_INVALID
    assert L.ppl_chomp_buffers_create(7, 1, C.byref(out)) == _ppl.PPL_ERR_INVALID
    assert L.ppl_chomp_buffers_destroy(None) == _ppl.PPL_OK
    res = _ppl.ChompResult()
    assert L.ppl_chomp_solve(None, None, None, None, None, 51, 10, C.c_float(0), None, None, C.byref(res),
                             None) == _ppl.PPL_ERR_INVALID


def test_edge_granularities_follow_the_rule_at_the_extremes():
    """The host-side granularity rule of the edge evaluator (no device work)
    against the oracle's restatement, including lengths whose ratio to the
    resolution overflows an int and a NaN endpoint (both clamp to 1024)."""
    import sys
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
    import pyoracle as PO
    from prismplan_amd import substrate as S
    rng = np.random.default_rng(3)
    a = rng.uniform(-2, 2, (200, 7)).astype(np.float32)
    b = (a + rng.normal(0, 0.3, (200, 7))).astype(np.float32)
    b[0] = a[0]                                          # zero length
    G, ln = S.edge_granularities(a, b, 32, 0.025)
    for k in range(200):
        g_ref, l_ref = PO.edge_granularity(a[k], b[k], 32, 0.025)
        assert G[k] == g_ref and ln[k] == pytest.approx(l_ref, rel=1e-6)
    assert G[0] == 2
    G, _ = S.edge_granularities(a[:3], b[:3], 32, 1e-30)   # len / res far above 2^31
    assert list(G) == [2, 1024, 1024]                      # the zero-length edge stays at 2
    nan = b[:2].copy()
    nan[1, 0] = np.nan
    G, _ = S.edge_granularities(a[:2], nan, 32, 0.025)
    assert G[1] == 1024
    G, _ = S.edge_granularities(a[:2], b[:2], 17, 0.0)     # fixed granularity
    assert list(G) == [17, 17]
