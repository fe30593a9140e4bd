"""Obstacle world in the kernel layout.

Tensor conventions follow the reference's SceneCollisionData
(cpp/src/collision/scene_collision_data.hh:10-47) and its producer
PlanningScene._rebuild_tensors (python/gpu_planning/substrate/
planning_scene.py:189-217):

* ``obb_dims[i]  = [dx, dy, dz, 0]`` full extents;
* ``obb_pose[i]  = [tx, ty, tz, qw, qx, qy, qz, 0]`` the INVERSE pose
  (world -> box): ``p_local = R(q) p_world + t``;
* ``obb_enable[i]`` in {0, 1};
* voxel layers: ``params[l] = [dx, dy, dz, voxel_size]``, ``pose[l]`` inverse
  pose of the grid, ``features[l, ix*ny*nz + iy*nz + iz]`` signed distance,
  positive inside, <= -999 unobserved.
"""
from __future__ import annotations

import threading
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np


def quat_rotate(q_wxyz: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Rotate v by the unit quaternion (w, x, y, z): v + w t + u x t, t = 2 u x v."""
    w = q_wxyz[..., :1]
    u = q_wxyz[..., 1:4]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def inverse_pose(position: np.ndarray, quat_wxyz: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """(t, q_stored) of the world->local transform of a box at `position` with
    orientation `quat_wxyz` (local->world): q_stored = conj(Q), t = -R(q_stored) P
    (planning_scene.py:52-63)."""
    q = np.asarray(quat_wxyz, dtype=np.float64)
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    q_conj = q * np.array([1.0, -1.0, -1.0, -1.0])
    t = -quat_rotate(q_conj, np.asarray(position, dtype=np.float64))
    return t, q_conj


def obb_arrays(dims: np.ndarray, positions: np.ndarray, quats_wxyz: np.ndarray):
    """Full extents [O,3], world positions [O,3], orientations [O,4] ->
    (dims[O,4] f32, pose[O,8] f32, enable[O] u8) in the kernel layout."""
    dims = np.asarray(dims, dtype=np.float64).reshape(-1, 3)
    positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    quats = np.asarray(quats_wxyz, dtype=np.float64).reshape(-1, 4)
    n = dims.shape[0]
    d = np.zeros((n, 4), np.float32)
    p = np.zeros((n, 8), np.float32)
    d[:, :3] = dims
    if n:
        t, qs = inverse_pose(positions, quats)
        p[:, 0:3] = t
        p[:, 3:7] = qs
    return d, p, np.ones((n,), np.uint8)


@dataclass
class HostScene:
    """Host (numpy) copy of a scene in the kernel layout."""
    obb_dims: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    obb_pose: np.ndarray = field(default_factory=lambda: np.zeros((0, 8), np.float32))
    obb_enable: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint8))
    voxel_params: Optional[np.ndarray] = None     # [L,4] f32
    voxel_pose: Optional[np.ndarray] = None       # [L,8] f32
    voxel_enable: Optional[np.ndarray] = None     # [L] u8
    voxel_features: Optional[np.ndarray] = None   # [L, max_voxels] f32
    sphere_obb_acm: Optional[np.ndarray] = None   # [S, O] u8
    sphere_obb_acm_approx: Optional[np.ndarray] = None
    sphere_acm: Optional[np.ndarray] = None       # [S, L] u8
    esdf_interp: int = 0                          # 0 nearest voxel, 1 trilinear

    @property
    def n_obbs(self) -> int:
        return int(self.obb_dims.shape[0])

    @property
    def n_voxel_layers(self) -> int:
        return 0 if self.voxel_params is None else int(self.voxel_params.shape[0])

    @property
    def max_voxels_per_layer(self) -> int:
        return 0 if self.voxel_features is None else int(self.voxel_features.shape[1])

    @staticmethod
    def from_boxes(dims, positions, quats) -> "HostScene":
        d, p, e = obb_arrays(dims, positions, quats)
        return HostScene(obb_dims=d, obb_pose=p, obb_enable=e)


def robust_floor(x: float, threshold: float = 1e-4) -> int:
    """Grid-size rounding of scene_collision.cuh:192-200, in float32."""
    x = np.float32(x)
    nearest = np.float32(np.round(x))
    if np.float32(x - nearest) >= np.float32(threshold):
        return int(np.floor(x))
    return int(nearest)


def grid_dims(params: Sequence[float]) -> Tuple[int, int, int]:
    inv = np.float32(1.0) / np.float32(params[3])
    return tuple(robust_floor(np.float32(params[k]) * inv) + 1 for k in range(3))


class PlanningScene:
    """OBB world that hands out device tensors in the kernel layout.

    Same obstacle API as the reference's PlanningScene
    (planning_scene.py:131-184); tensors are built lazily on the scene's device
    and retained on the object for as long as a kernel may use them.
    """

    def __init__(self, num_joints: int, robot_id: str = "robot", device: str = "cuda"):
        self._num_joints = int(num_joints)
        self._robot_id = str(robot_id)
        self._device = device
        self._cuboids: Dict[str, Tuple[np.ndarray, np.ndarray, np.ndarray]] = {}
        # the reference's name (planning_scene.py:79): its planner base holds it
        # while a SceneCollisionDataBuilder's device pointers are in use
        self._voxel_plan_lock = threading.RLock()
        self._dirty = True
        self._tensors = None
        self._n_obbs = 0

    @classmethod
    def from_robot_asset(cls, asset_path=None, device: str = "cuda") -> "PlanningScene":
        from .robot_model_loader import load_robot_asset
        a = load_robot_asset(asset_path)
        return cls(num_joints=a["n_dof"], robot_id=a["name"], device=device)

    @property
    def num_joints(self) -> int:
        return self._num_joints

    @property
    def robot_id(self) -> str:
        return self._robot_id

    @property
    def n_obbs(self) -> int:
        return self._n_obbs

    @property
    def lock(self):
        return self._voxel_plan_lock

    @property
    def current_obstacles(self) -> List[Dict[str, Any]]:
        return [{"name": n, "type": "cuboid", "dims": d.tolist(), "position": p.tolist(),
                 "quaternion": q.tolist()} for n, (d, p, q) in self._cuboids.items()]

    def add_cuboid(self, name: str, dims: Sequence[float], position: Sequence[float],
                   quaternion=(1.0, 0.0, 0.0, 0.0)) -> None:
        q = np.asarray(quaternion, np.float64).reshape(4)
        nrm = np.linalg.norm(q)
        if nrm > 0:
            q = q / nrm
        with self._voxel_plan_lock:
            self._cuboids[name] = (np.asarray(dims, np.float64).reshape(3),
                                   np.asarray(position, np.float64).reshape(3), q)
            self._dirty = True

    def remove_cuboid(self, name: str) -> bool:
        with self._voxel_plan_lock:
            existed = self._cuboids.pop(name, None) is not None
            self._dirty |= existed
            return existed

    def clear(self) -> None:
        with self._voxel_plan_lock:
            if self._cuboids:
                self._cuboids.clear()
                self._dirty = True

    def update_world(self, obstacles: List[Dict[str, Any]]) -> None:
        with self._voxel_plan_lock:
            self._cuboids.clear()
            for i, obs in enumerate(obstacles):
                if "pose" in obs:
                    pos, quat = obs["pose"][0:3], obs["pose"][3:7]
                else:
                    pos, quat = obs["position"], obs.get("quaternion", (1.0, 0.0, 0.0, 0.0))
                self.add_cuboid(obs.get("name", f"obs_{i}"), obs["dims"], pos, quat)
            self._dirty = True

    def host_scene(self) -> HostScene:
        with self._voxel_plan_lock:
            if not self._cuboids:
                return HostScene()
            d = np.stack([v[0] for v in self._cuboids.values()])
            p = np.stack([v[1] for v in self._cuboids.values()])
            q = np.stack([v[2] for v in self._cuboids.values()])
            return HostScene.from_boxes(d, p, q)

    def device_scene(self):
        """DeviceScene (torch tensors on the scene device), rebuilt when dirty."""
        from .substrate import DeviceScene
        with self._voxel_plan_lock:
            if self._dirty or self._tensors is None:
                hs = self.host_scene()
                self._tensors = DeviceScene.from_host(hs, device=self._device)
                self._n_obbs = hs.n_obbs
                self._dirty = False
            return self._tensors

    def build_scene_collision_data(self):
        """A SceneCollisionDataBuilder over this scene's device tensors, or None
        when the world is empty (planning_scene.py:219-242)."""
        from .prrtc import scene_builder
        ds = self.device_scene()
        if self._n_obbs == 0:
            return None
        return scene_builder(ds)
