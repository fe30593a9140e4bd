"""ctypes binding of libppl.so (include/ppl/ppl.h).

torch is imported first: its bundled HIP runtime carries the same SONAME as
the one libppl.so links, so the process keeps a single runtime and device
pointers from torch tensors are valid in the library. Loading fails loudly
when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch  # noqa: F401  (must precede the library load, see above)

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libppl.so"

PPL_OK, PPL_ERR_INVALID, PPL_ERR_HIP, PPL_ERR_ALLOC, PPL_ERR_UNSUPPORTED, PPL_ERR_CAPACITY = range(6)
PPL_COST_GROUPED = -1   # ppl_plan_set_cost_group: the grouped cost kernel
PLAN_CONFIGS, PLAN_EDGES, PLAN_TRAJ = 0, 1, 2

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p


class PplError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where} failed with status {status}: {msg}")
        self.status = status


class RobotDesc(C.Structure):
    _fields_ = [("n_dof", C.c_int), ("n_joints", C.c_int),
                ("fixed_transforms", _f32p), ("joint_types", _i32p),
                ("n_spheres", C.c_int), ("spheres", _f32p), ("sphere_to_joint", _i32p),
                ("n_self_cc_ranges", C.c_int), ("self_cc_ranges", _i32p),
                ("n_approx_spheres", C.c_int), ("approx_spheres", _f32p), ("approx_sphere_to_joint", _i32p),
                ("n_approx_self_cc_ranges", C.c_int), ("approx_self_cc_ranges", _i32p),
                ("approx_fixed_transforms", _f32p), ("approx_joint_types", _i32p),
                ("joint_lower", _f32p), ("joint_upper", _f32p),
                ("joint_parents", _i32p), ("joint_id_to_dof", _i32p), ("t_memory_idx", _i32p),
                ("dfs_order", _i32p)]


class RobotInfo(C.Structure):
    _fields_ = [("n_dof", C.c_int), ("n_joints", C.c_int), ("n_spheres", C.c_int), ("n_approx_spheres", C.c_int),
                ("n_self_pairs", C.c_int), ("n_t_slots", C.c_int)]


class RobotDeviceView(C.Structure):
    _fields_ = [("n_dof", C.c_int), ("n_joints", C.c_int), ("n_spheres", C.c_int), ("n_approx_spheres", C.c_int),
                ("n_pairs", C.c_int), ("n_approx_pairs", C.c_int),
                ("ft", _vp), ("jmeta", _vp), ("sph_off", _vp), ("sph", _vp), ("sph_id", _vp), ("apx_off", _vp),
                ("apx", _vp), ("apx_id", _vp), ("sph_orig", _vp), ("apx_orig", _vp), ("pairs", _vp),
                ("apx_pairs", _vp), ("lim", _vp), ("aft", _vp), ("atype", _vp)]


class SceneView(C.Structure):
    _fields_ = [("obb_dims", _vp), ("obb_pose", _vp), ("obb_enable", _vp), ("n_obbs", C.c_int),
                ("voxel_params", _vp), ("voxel_pose", _vp), ("voxel_enable", _vp), ("voxel_features", _vp),
                ("n_voxel_layers", C.c_int), ("max_voxels_per_layer", C.c_int),
                ("sphere_acm_mask", _vp), ("sphere_obb_acm_mask", _vp), ("sphere_obb_acm_mask_approx", _vp),
                ("esdf_interp", C.c_int)]


class EdgeParams(C.Structure):
    _fields_ = [("granularity", C.c_int), ("collision_margin", C.c_float), ("check_self", C.c_int),
                ("compute_clearance", C.c_int), ("enable_mesh", C.c_int), ("resolution", C.c_float)]


class CostParams(C.Structure):
    _fields_ = [("world_margin", C.c_float), ("self_margin", C.c_float), ("weight", C.c_float),
                ("use_obb", C.c_int), ("use_esdf", C.c_int), ("use_self", C.c_int)]


class PlanDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("n_units", C.c_int), ("unit_states", _i32p), ("skip_first", C.c_int),
                ("n_problems", C.c_int), ("problem_units", _i32p), ("obb_offsets", _i32p),
                ("waypoint_major", C.c_int)]


class ChompSettings(C.Structure):
    _fields_ = [("step_size", C.c_float), ("smoothness_weight", C.c_float), ("collision_cost_weight", C.c_float),
                ("world_margin", C.c_float), ("self_margin", C.c_float), ("dt", C.c_float),
                ("collision_check_step", C.c_float), ("collision_margin", C.c_float), ("convergence_eps", C.c_float),
                ("iterations_after_valid", C.c_int), ("use_self", C.c_int), ("fd_gradient", C.c_int),
                ("fd_eps", C.c_float)]


class StompSettings(C.Structure):
    _fields_ = [("num_rollouts_new", C.c_int), ("num_rollouts_old", C.c_int), ("control_cost_weight", C.c_float),
                ("collision_cost_weight", C.c_float), ("exponentiated_cost_sensitivity", C.c_float),
                ("world_margin", C.c_float), ("self_margin", C.c_float), ("dt", C.c_float), ("noise_scale", C.c_float),
                ("collision_check_step", C.c_float), ("collision_margin", C.c_float), ("convergence_eps", C.c_float),
                ("iterations_after_valid", C.c_int), ("use_self", C.c_int), ("seed", C.c_ulonglong)]


class ChompResult(C.Structure):
    _fields_ = [("solved", C.c_int), ("iterations_run", C.c_int), ("final_state_cost", C.c_float),
                ("final_control_cost", C.c_float), ("final_total_cost", C.c_float), ("wall_time_ms", C.c_double)]


CHOMP_BEST_COST, CHOMP_PREV_COST, CHOMP_SOLVED, CHOMP_STREAK, CHOMP_POLISH_LEFT, CHOMP_DONE, CHOMP_ITERS, \
    CHOMP_LAST_COST = range(8)


class RrtcSettings(C.Structure):
    _fields_ = [("max_samples", C.c_int), ("max_iters", C.c_int), ("time_limit_ms", C.c_float),
                ("num_new_configs", C.c_int), ("granularity", C.c_int), ("range", C.c_float), ("balance", C.c_int),
                ("tree_ratio", C.c_float), ("dynamic_domain", C.c_int), ("dd_alpha", C.c_float),
                ("dd_radius", C.c_float), ("dd_min_radius", C.c_float), ("shortcut_path", C.c_int),
                ("valid_segment_length", C.c_float), ("collision_margin", C.c_float),
                ("shortcut_collision_margin", C.c_float), ("seed", C.c_ulonglong)]


class RrtcResult(C.Structure):
    _fields_ = [("solved", C.c_int), ("iters", C.c_int), ("start_tree_size", C.c_int), ("goal_tree_size", C.c_int),
                ("path_length", C.c_int), ("n_waypoints", C.c_int), ("reached_goal", C.c_int), ("cost", C.c_float),
                ("wall_ms", C.c_double), ("kernel_ms", C.c_double)]


class EdgeBatchParams(C.Structure):
    _fields_ = [("collision_margin", C.c_float), ("check_self", C.c_int), ("skip_resolved", C.c_int)]


# (name, restype, argtypes) of every exported entry point of ppl.h
SIGNATURES = [
    ("ppl_last_error", C.c_char_p, []),
    ("ppl_version", C.c_char_p, []),
    ("ppl_robot_create", C.c_int, [C.POINTER(RobotDesc), C.POINTER(_vp)]),
    ("ppl_robot_destroy", C.c_int, [_vp]),
    ("ppl_robot_get_info", C.c_int, [_vp, C.POINTER(RobotInfo)]),
    ("ppl_robot_update_joint_limits", C.c_int, [_vp, _f32p, _f32p]),
    ("ppl_robot_get_device_view", C.c_int, [_vp, C.POINTER(RobotDeviceView)]),
    ("ppl_check_collision_scene", C.c_int, [_vp, C.POINTER(SceneView), _f32p, C.c_int, C.c_float, _u8p]),
    ("ppl_edge_eval_create", C.c_int, [_vp, C.POINTER(SceneView), C.c_int, C.c_int, C.POINTER(_vp)]),
    ("ppl_edge_eval_destroy", C.c_int, [_vp]),
    ("ppl_edge_eval_evaluate", C.c_int, [_vp, _f32p, _f32p, C.c_int, C.POINTER(EdgeParams), _u8p, _f32p, _f32p]),
    ("ppl_edge_granularities", C.c_int, [_f32p, _f32p, C.c_int, C.c_int, C.POINTER(EdgeParams), _i32p, _f32p]),
    ("ppl_chomp_state_cost", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.c_int, C.POINTER(CostParams), _vp, _vp]),
    ("ppl_chomp_fd_grad", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.c_int, C.POINTER(CostParams), C.c_float, _vp,
                                    _vp]),
    ("ppl_plan_create", C.c_int, [_vp, C.POINTER(PlanDesc), C.POINTER(_vp)]),
    ("ppl_plan_destroy", C.c_int, [_vp]),
    ("ppl_plan_states", C.c_long, [_vp]),
    ("ppl_fk_batched", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("ppl_check_collision_batched", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.c_float, C.c_int, _vp, _vp]),
    ("ppl_edge_check_batched", C.c_int, [_vp, C.POINTER(SceneView), _vp, _vp, C.POINTER(EdgeBatchParams), _vp, _vp,
                                         _vp]),
    ("ppl_min_clearance_batched", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.c_float, _vp, _vp]),
    ("ppl_hinge_cost_batched", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.POINTER(CostParams), _vp, _vp, _vp, _vp]),
    ("ppl_plan_autotune_cost", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.POINTER(CostParams), C.c_int, C.c_int,
                                         _i32p, _f32p, _vp]),
    ("ppl_plan_set_cost_group", C.c_int, [_vp, C.c_int, C.c_int]),
    ("ppl_tune_cache_file", C.c_int, [C.c_char_p]),
    ("ppl_plan_autotune_collide", C.c_int, [_vp, C.POINTER(SceneView), _vp, _vp, C.c_int, C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.POINTER(C.c_float), _vp]),
    ("ppl_plan_collide_shape", C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("ppl_plan_set_collide_shape", C.c_int, [_vp, C.c_int, C.c_int]),
    ("ppl_plan_time_kernels", C.c_int, [_vp, C.c_int]),
    ("ppl_plan_kernel_ms", C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    ("ppl_plan_set_optimizer_graphs", C.c_int, [_vp, C.c_int]),
    ("ppl_chomp_fd_grad_batched", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.POINTER(CostParams), C.c_float, _vp,
                                            _vp]),
    ("ppl_stomp_rollout_costs", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.POINTER(CostParams), C.c_float, C.c_float,
                                          _vp, _vp, _vp]),
    ("ppl_chomp_optimize", C.c_int, [_vp, C.POINTER(SceneView), C.POINTER(ChompSettings), _vp, _vp, _vp, C.c_int,
                                     C.c_int, _vp]),
    ("ppl_chomp_buffers_create", C.c_int, [C.c_int, C.c_int, C.POINTER(_vp)]),
    ("ppl_chomp_buffers_destroy", C.c_int, [_vp]),
    ("ppl_chomp_solve", C.c_int, [_vp, _vp, C.POINTER(SceneView), _f32p, _f32p, C.c_int, C.c_int, C.c_float,
                                  C.POINTER(ChompSettings), _f32p, C.POINTER(ChompResult), _vp]),
    ("ppl_stomp_optimize", C.c_int, [_vp, C.POINTER(SceneView), C.POINTER(StompSettings), _vp, _vp, _vp, C.c_int,
                                     C.c_int, _vp]),
    ("ppl_stomp_buffers_create", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    ("ppl_stomp_buffers_destroy", C.c_int, [_vp]),
    ("ppl_stomp_solve", C.c_int, [_vp, _vp, C.POINTER(SceneView), _f32p, _f32p, C.c_int, C.c_int, C.c_float,
                                  C.POINTER(StompSettings), _f32p, C.POINTER(ChompResult), _vp]),
    ("ppl_min_clearance_scene", C.c_int, [_vp, C.POINTER(SceneView), _f32p, C.c_int, C.c_float, _f32p]),
    ("ppl_hinge_cost_scene", C.c_int, [_vp, C.POINTER(SceneView), _f32p, C.c_int, C.POINTER(CostParams), _f32p]),
    ("ppl_mitstar_batch_cc", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.c_int, C.c_float, _vp, _vp, C.c_float,
                                       _vp]),
    ("ppl_mitstar_batch_cc_static", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.c_int, C.c_float, _vp, _vp, C.c_float,
                                       _vp]),
    ("ppl_mitstar_node_clearance", C.c_int, [_vp, C.POINTER(SceneView), _vp, _vp, C.c_int, C.c_int, C.c_float, _vp,
                                             _vp]),
    ("ppl_mitstar_edge_cc", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.c_int, _vp, _vp, _vp, C.c_int, C.c_float,
                                      C.c_int, C.c_float, _vp, _vp, _vp]),
    ("ppl_rrtc_default_settings", None, [C.POINTER(RrtcSettings)]),
    ("ppl_rrtc_halton_bases", C.c_int, [C.c_ulonglong, C.c_int, C.c_int, _f32p]),
    ("ppl_rrtc_buffers_create", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    ("ppl_rrtc_buffers_destroy", C.c_int, [_vp]),
    ("ppl_rrtc_solve", C.c_int, [_vp, C.POINTER(SceneView), _f32p, _f32p, C.c_int, C.POINTER(RrtcSettings), _vp,
                                 _f32p, C.c_int, C.POINTER(RrtcResult), _vp]),
    ("ppl_shortcut_path", C.c_int, [_vp, C.POINTER(SceneView), _f32p, C.c_int, C.c_float, C.c_float, _f32p, _i32p,
                                    _vp]),
    ("ppl_esdf_from_obbs", C.c_int, [C.POINTER(SceneView), _f32p, _f32p, _vp, _vp]),
    ("ppl_graph_capture_stomp", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.POINTER(CostParams), C.c_float,
                                          C.c_float, _vp, _vp, _vp, C.POINTER(_vp)]),
    ("ppl_graph_capture_hinge", C.c_int, [_vp, C.POINTER(SceneView), _vp, C.POINTER(CostParams), _vp, _vp, _vp, _vp,
                                          C.POINTER(_vp)]),
    ("ppl_graph_launch", C.c_int, [_vp, _vp]),
    ("ppl_graph_destroy", C.c_int, [_vp]),
    ("ppl_algorithmic_flops", C.c_double, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
]
WORK_ANY_COLLISION, WORK_HINGE_STATE, WORK_CHOMP_ADJOINT, WORK_CHOMP_FD, WORK_STOMP_STATE, WORK_ESDF_WAYPOINT = range(6)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libppl.so not built at {LIB_PATH}; run `python -m prismplan_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, where: str) -> None:
    if status != PPL_OK:
        msg = lib().ppl_last_error().decode(errors="replace")
        raise PplError(status, where, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
