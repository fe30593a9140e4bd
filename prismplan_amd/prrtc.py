"""Reference-compatible names of the substrate's Python boundary.

Mirrors the builder and check entry points of the reference's pybind module
`prrtc` (cpp/src/python/bindings.cpp:118-222, 363-402, 493-530, 960-1037):
RobotModelBuilder with the same fields, SceneCollisionDataBuilder with
int64 device-pointer fields taken from torch tensors, and
check_collision_scene(configs, scene, robot, collision_margin) -> bool[N].
Errors surface as RuntimeError, as the pybind module's std::runtime_error did.
"""
from __future__ import annotations

import ctypes as C
from typing import List

import numpy as np

from . import _ppl
from .substrate import DeviceScene, Robot, check_collision_scene as _check


class RobotModelBuilder:
    _LIST_FIELDS = ("fixed_transforms", "joint_types", "spheres_flat", "sphere_to_joint", "self_cc_ranges_flat",
                    "approx_spheres_flat", "approx_sphere_to_joint", "approx_self_cc_ranges_flat",
                    "approx_fixed_transforms", "approx_joint_types", "joint_lower", "joint_upper",
                    "joint_parents", "joint_id_to_dof", "t_memory_idx", "dfs_order")

    def __init__(self):
        self.n_dof = 0
        self.n_joints = 0
        for f in self._LIST_FIELDS:
            setattr(self, f, [])
        self._robot = None

    def _asset(self) -> dict:
        J = int(self.n_joints)
        if self.approx_fixed_transforms and list(self.approx_fixed_transforms) != list(self.fixed_transforms):
            raise RuntimeError("approx_fixed_transforms differing from fixed_transforms are not supported")
        return {
            "name": "robot", "n_dof": int(self.n_dof), "n_joints": J,
            "joint_types": list(self.joint_types), "fixed_transforms": list(self.fixed_transforms),
            "spheres_flat": list(self.spheres_flat), "sphere_to_joint": list(self.sphere_to_joint),
            "self_cc_ranges_flat": list(self.self_cc_ranges_flat),
            "joint_lower": list(self.joint_lower), "joint_upper": list(self.joint_upper),
            "joint_parents": list(self.joint_parents) or [max(i - 1, 0) for i in range(J)],
            "joint_id_to_dof": list(self.joint_id_to_dof) or [i - 1 for i in range(J)],
            "t_memory_idx": list(self.t_memory_idx) or [0] + [1] * (J - 1),
            "dfs_order": list(self.dfs_order) or list(range(J)),
            "approx_spheres_flat": list(self.approx_spheres_flat),
            "approx_sphere_to_joint": list(self.approx_sphere_to_joint),
            "approx_self_cc_ranges_flat": list(self.approx_self_cc_ranges_flat),
        }

    def build(self) -> None:
        """Materialise the device model now (idempotent)."""
        if self._robot is None:
            try:
                self._robot = Robot(self._asset())
            except _ppl.PplError as e:
                raise RuntimeError(str(e)) from e

    def build_on_device(self) -> Robot:
        self.build()
        return self._robot

    def update_joint_limits(self) -> None:
        if self._robot is None:
            self.build()
            return
        if len(self.joint_lower) != self.n_dof or len(self.joint_upper) != self.n_dof:
            raise RuntimeError("joint_lower/joint_upper size != n_dof")
        lo = np.asarray(self.joint_lower, np.float32)
        hi = np.asarray(self.joint_upper, np.float32)
        _ppl.call("ppl_robot_update_joint_limits", self._robot.handle, lo.ctypes.data_as(C.POINTER(C.c_float)),
                  hi.ctypes.data_as(C.POINTER(C.c_float)))
        self._robot.arrays.joint_lower[:] = lo
        self._robot.arrays.joint_upper[:] = hi

    def destroy(self) -> None:
        if self._robot is not None:
            self._robot.close()
            self._robot = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class SceneCollisionDataBuilder:
    """int64 device pointers of the scene tensors (non-owning)."""

    def __init__(self):
        self.obb_dims_ptr = 0
        self.obb_pose_ptr = 0
        self.obb_enable_ptr = 0
        self.n_obbs = 0
        self.voxel_params_ptr = 0
        self.voxel_pose_ptr = 0
        self.voxel_enable_ptr = 0
        self.voxel_features_ptr = 0
        self.n_voxel_layers = 0
        self.max_voxels_per_layer = 0
        self.sphere_acm_mask_ptr = 0
        self.sphere_obb_acm_mask_ptr = 0
        self.sphere_obb_acm_mask_approx_ptr = 0
        self.esdf_interp = 0
        self._owner = None

    @staticmethod
    def from_device_scene(ds: DeviceScene) -> "SceneCollisionDataBuilder":
        b = SceneCollisionDataBuilder()
        v = ds.view()
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
        b._owner = ds
        return b

    def view(self) -> _ppl.SceneView:
        v = _ppl.SceneView()
        v.obb_dims, v.obb_pose, v.obb_enable = self.obb_dims_ptr or None, self.obb_pose_ptr or None, self.obb_enable_ptr or None
        v.n_obbs = int(self.n_obbs)
        v.voxel_params, v.voxel_pose = self.voxel_params_ptr or None, self.voxel_pose_ptr or None
        v.voxel_enable, v.voxel_features = self.voxel_enable_ptr or None, self.voxel_features_ptr or None
        v.n_voxel_layers = int(self.n_voxel_layers)
        v.max_voxels_per_layer = int(self.max_voxels_per_layer)
        v.sphere_acm_mask = self.sphere_acm_mask_ptr or None
        v.sphere_obb_acm_mask = self.sphere_obb_acm_mask_ptr or None
        v.sphere_obb_acm_mask_approx = self.sphere_obb_acm_mask_approx_ptr or None
        v.esdf_interp = int(self.esdf_interp)
        return v


class _ViewScene:
    def __init__(self, b: SceneCollisionDataBuilder):
        self._b = b

    def view(self):
        return self._b.view()


def check_collision_scene(configs, scene: SceneCollisionDataBuilder, robot: RobotModelBuilder,
                          collision_margin: float = 0.0) -> np.ndarray:
    """Batch environment any-collision with the two-phase gate; configs [N, n_dof]."""
    rob = robot.build_on_device()
    q = np.asarray(configs, dtype=np.float32)
    if q.ndim != 2 or q.shape[1] != rob.n_dof:
        raise RuntimeError("configs columns mismatch with robot n_dof")
    try:
        return _check(rob, _ViewScene(scene), q, collision_margin)
    except _ppl.PplError as e:
        raise RuntimeError(str(e)) from e
