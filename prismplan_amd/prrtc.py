"""The reference's Python boundary names, from the compiled pybind11 module.

`prismplan_amd/_lib/_prrtc*.so` (prismplan_amd/csrc/prrtc_module.cpp) mirrors
the builder and check surface of the reference's pybind module `prrtc`
(cpp/src/python/bindings.cpp:118-222, 225-233, 363-437, 493-530, 764-902,
910-1037, 1163-1176, 1191-1282): RobotModelBuilder with the same fields,
SceneCollisionDataBuilder with int64 device-pointer fields taken from torch
tensors, check_collision_scene(configs, scene, robot, collision_margin) ->
bool[N], pRRTC's solve_scene(start, goals, scene, settings, robot, bufs=None)
-> PlannerResult with Settings and SolverBuffers, and chomp_solve_scene /
stomp_solve_scene(start, goal, scene, settings, robot, bufs=None) ->
CHOMPResult / STOMPResult with their settings and buffers.
Errors surface as RuntimeError. Importing fails if the module is not built.
"""
from __future__ import annotations

import importlib.util
import sysconfig

import torch  # noqa: F401  (HIP runtime before libppl.so, see _ppl)

from . import _ppl

_ppl.lib()
_PATH = _ppl.LIB_PATH.parent / ("_prrtc" + sysconfig.get_config_var("EXT_SUFFIX"))
if not _PATH.exists():
    raise ImportError(f"{_PATH} not built; run `python -m prismplan_amd.build`")
_spec = importlib.util.spec_from_file_location("prismplan_amd._prrtc", _PATH)
_m = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_m)

RobotModelBuilder = _m.RobotModelBuilder
SceneCollisionDataBuilder = _m.SceneCollisionDataBuilder
check_collision_scene = _m.check_collision_scene
CHOMPSettings = _m.CHOMPSettings
CHOMPResult = _m.CHOMPResult
CHOMPBuffers = _m.CHOMPBuffers
chomp_solve_scene = _m.chomp_solve_scene
STOMPSettings = _m.STOMPSettings
STOMPResult = _m.STOMPResult
STOMPBuffers = _m.STOMPBuffers
stomp_solve_scene = _m.stomp_solve_scene
Settings = _m.Settings
PlannerResult = _m.PlannerResult
SolverBuffers = _m.SolverBuffers
solve_scene = _m.solve_scene

__all__ = ["RobotModelBuilder", "SceneCollisionDataBuilder", "check_collision_scene", "CHOMPSettings", "CHOMPResult",
           "CHOMPBuffers", "chomp_solve_scene", "STOMPSettings", "STOMPResult", "STOMPBuffers", "stomp_solve_scene",
           "Settings", "PlannerResult", "SolverBuffers", "solve_scene", "scene_builder"]


def scene_builder(ds) -> SceneCollisionDataBuilder:
    """SceneCollisionDataBuilder over a substrate DeviceScene's tensors; the
    builder holds a reference to `ds` so the pointers stay valid."""
    v = ds.view()
    b = SceneCollisionDataBuilder()
    b.obb_dims_ptr = v.obb_dims or 0
    b.obb_pose_ptr = v.obb_pose or 0
    b.obb_enable_ptr = v.obb_enable or 0
    b.n_obbs = v.n_obbs
    b.voxel_params_ptr = v.voxel_params or 0
    b.voxel_pose_ptr = v.voxel_pose or 0
    b.voxel_enable_ptr = v.voxel_enable or 0
    b.voxel_features_ptr = v.voxel_features or 0
    b.n_voxel_layers = v.n_voxel_layers
    b.max_voxels_per_layer = v.max_voxels_per_layer
    b.sphere_acm_mask_ptr = v.sphere_acm_mask or 0
    b.sphere_obb_acm_mask_ptr = v.sphere_obb_acm_mask or 0
    b.sphere_obb_acm_mask_approx_ptr = v.sphere_obb_acm_mask_approx or 0
    b.esdf_interp = v.esdf_interp
    b.owner = ds
    return b
