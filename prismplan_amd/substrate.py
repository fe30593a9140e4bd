"""Torch-facing API of the collision-kinematics substrate.

Thin wrappers over the C ABI (include/ppl/ppl.h): the robot model, a device
scene, the plain per-request forms (host arrays, synchronous), the batched
device-resident forms over a plan, and `HingeCostFn`, a torch.autograd.Function
whose backward is the fused adjoint kernel. Device tensors are passed by
pointer on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _ppl
from ._ppl import call, lib
from .planning_scene import HostScene, grid_dims
from .robot_model_loader import RobotArrays, load_robot_asset

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)


def _np(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def _p(a: Optional[np.ndarray], t):
    return None if a is None else a.ctypes.data_as(t)


def _dp(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream=None):
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


class Robot:
    """Runtime robot model on the device (RobotModelDevice equivalent)."""

    def __init__(self, asset="panda"):
        if isinstance(asset, RobotArrays):
            arr = asset
        elif isinstance(asset, dict):
            arr = RobotArrays(asset)
        else:
            arr = RobotArrays(load_robot_asset(asset))
        self.arrays = arr
        keep = [arr.fixed_transforms, arr.joint_types, arr.spheres, arr.sphere_to_joint, arr.self_ranges,
                arr.approx_spheres, arr.approx_sphere_to_joint, arr.approx_self_ranges, arr.joint_lower,
                arr.joint_upper, arr.joint_parents, arr.joint_id_to_dof, arr.t_memory_idx, arr.dfs_order]
        self._keep = keep
        d = _ppl.RobotDesc()
        d.n_dof, d.n_joints = arr.n_dof, arr.n_joints
        d.fixed_transforms = _p(arr.fixed_transforms, _f32p)
        d.joint_types = _p(arr.joint_types, _i32p)
        d.n_spheres = arr.n_spheres
        d.spheres = _p(arr.spheres, _f32p)
        d.sphere_to_joint = _p(arr.sphere_to_joint, _i32p)
        d.n_self_cc_ranges = int(arr.self_ranges.shape[0])
        d.self_cc_ranges = _p(arr.self_ranges, _i32p)
        d.n_approx_spheres = arr.n_approx
        d.approx_spheres = _p(arr.approx_spheres, _f32p)
        d.approx_sphere_to_joint = _p(arr.approx_sphere_to_joint, _i32p)
        d.n_approx_self_cc_ranges = int(arr.approx_self_ranges.shape[0])
        d.approx_self_cc_ranges = _p(arr.approx_self_ranges, _i32p)
        d.joint_lower = _p(arr.joint_lower, _f32p)
        d.joint_upper = _p(arr.joint_upper, _f32p)
        d.joint_parents = _p(arr.joint_parents, _i32p)
        d.joint_id_to_dof = _p(arr.joint_id_to_dof, _i32p)
        d.t_memory_idx = _p(arr.t_memory_idx, _i32p)
        d.dfs_order = _p(arr.dfs_order, _i32p)
        if arr.approx_fixed_transforms is not None:
            keep.append(arr.approx_fixed_transforms)
            d.approx_fixed_transforms = _p(arr.approx_fixed_transforms, _f32p)
        if arr.approx_joint_types is not None:
            keep.append(arr.approx_joint_types)
            d.approx_joint_types = _p(arr.approx_joint_types, _i32p)
        h = C.c_void_p()
        call("ppl_robot_create", C.byref(d), C.byref(h))
        self.handle = h

    @property
    def n_dof(self) -> int:
        return self.arrays.n_dof

    def info(self) -> dict:
        i = _ppl.RobotInfo()
        call("ppl_robot_get_info", self.handle, C.byref(i))
        return {f: getattr(i, f) for f, _ in _ppl.RobotInfo._fields_}

    def flops(self, kind: int, n_obbs=0, n_waypoints=0, use_self=False) -> float:
        """Algorithmic flops of one work unit (include/ppl/roofline.h)."""
        v = lib().ppl_algorithmic_flops(self.handle, int(kind), int(n_obbs), int(n_waypoints), int(bool(use_self)))
        if v < 0:
            raise ValueError(f"no flop count for work kind {kind}")
        return float(v)

    def close(self):
        if getattr(self, "handle", None):
            lib().ppl_robot_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceScene:
    """Scene tensors on the device in the SceneCollisionData layout."""

    def __init__(self, obb_dims, obb_pose, obb_enable, voxel_params=None, voxel_pose=None, voxel_enable=None,
                 voxel_features=None, sphere_acm=None, sphere_obb_acm=None, sphere_obb_acm_approx=None,
                 esdf_interp: int = 0):
        self.obb_dims, self.obb_pose, self.obb_enable = obb_dims, obb_pose, obb_enable
        self.voxel_params, self.voxel_pose, self.voxel_enable = voxel_params, voxel_pose, voxel_enable
        self.voxel_features = voxel_features
        self.sphere_acm, self.sphere_obb_acm, self.sphere_obb_acm_approx = sphere_acm, sphere_obb_acm, sphere_obb_acm_approx
        self.esdf_interp = int(esdf_interp)

    @staticmethod
    def from_host(hs: HostScene, device="cuda") -> "DeviceScene":
        t = lambda a, dt: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dt)
        feats = hs.voxel_features
        return DeviceScene(t(hs.obb_dims, torch.float32), t(hs.obb_pose, torch.float32),
                           t(hs.obb_enable, torch.uint8), t(hs.voxel_params, torch.float32),
                           t(hs.voxel_pose, torch.float32), t(hs.voxel_enable, torch.uint8),
                           t(feats, torch.float32), t(hs.sphere_acm, torch.uint8), t(hs.sphere_obb_acm, torch.uint8),
                           t(hs.sphere_obb_acm_approx, torch.uint8), hs.esdf_interp)

    @property
    def n_obbs(self) -> int:
        return 0 if self.obb_dims is None else int(self.obb_dims.shape[0])

    def view(self) -> _ppl.SceneView:
        v = _ppl.SceneView()
        v.obb_dims, v.obb_pose, v.obb_enable = _dp(self.obb_dims), _dp(self.obb_pose), _dp(self.obb_enable)
        v.n_obbs = self.n_obbs
        v.voxel_params, v.voxel_pose = _dp(self.voxel_params), _dp(self.voxel_pose)
        v.voxel_enable, v.voxel_features = _dp(self.voxel_enable), _dp(self.voxel_features)
        v.n_voxel_layers = 0 if self.voxel_params is None else int(self.voxel_params.shape[0])
        v.max_voxels_per_layer = 0 if self.voxel_features is None else int(self.voxel_features.reshape(
            v.n_voxel_layers, -1).shape[1])
        v.sphere_acm_mask = _dp(self.sphere_acm)
        v.sphere_obb_acm_mask = _dp(self.sphere_obb_acm)
        v.sphere_obb_acm_mask_approx = _dp(self.sphere_obb_acm_approx)
        v.esdf_interp = self.esdf_interp
        return v


def cost_params(world_margin=0.02, self_margin=0.01, weight=1.0, use_obb=True, use_esdf=True,
                use_self=True) -> _ppl.CostParams:
    return _ppl.CostParams(world_margin, self_margin, weight, int(use_obb), int(use_esdf), int(use_self))


# ---------------------------------------------------------------------------
# Plain per-request forms
# ---------------------------------------------------------------------------
def check_collision_scene(robot: Robot, scene: DeviceScene, configs, collision_margin: float = 0.0) -> np.ndarray:
    """pRRTC::check_collision_scene: host [N, D] float32 -> host bool [N]."""
    q = _np(configs, np.float32).reshape(-1, robot.n_dof)
    out = np.zeros(q.shape[0], np.uint8)
    v = scene.view()
    call("ppl_check_collision_scene", robot.handle, C.byref(v), _p(q, _f32p), q.shape[0],
         C.c_float(collision_margin), _p(out, _u8p))
    return out.astype(bool)


def edge_granularities(frm, to, granularity=32, resolution=0.0):
    """States per edge (j = 1..G) and FP32 lengths under the reference rule."""
    a = _np(frm, np.float32)
    b = _np(to, np.float32)
    n, dim = a.shape
    G = np.zeros(n, np.int32)
    ln = np.zeros(n, np.float32)
    p = _ppl.EdgeParams(int(granularity), 0.0, 0, 0, 0, float(resolution))
    call("ppl_edge_granularities", _p(a, _f32p), _p(b, _f32p), n, dim, C.byref(p), _p(G, _i32p), _p(ln, _f32p))
    return G, ln


class EdgeEvaluator:
    """EdgeEvalGpu: persistent buffers for up to max_edges edges per call."""

    def __init__(self, robot: Robot, scene: DeviceScene, max_edges: int):
        self.robot, self.scene = robot, scene
        self._view = scene.view()
        h = C.c_void_p()
        call("ppl_edge_eval_create", robot.handle, C.byref(self._view), robot.n_dof, int(max_edges), C.byref(h))
        self.handle = h

    def evaluate(self, frm, to, granularity=32, collision_margin=0.0, check_self=False, compute_clearance=False,
                 resolution=0.0) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        a = _np(frm, np.float32).reshape(-1, self.robot.n_dof)
        b = _np(to, np.float32).reshape(-1, self.robot.n_dof)
        n = a.shape[0]
        valid = np.zeros(n, np.uint8)
        cost = np.zeros(n, np.float32)
        clr = np.zeros(n, np.float32)
        p = _ppl.EdgeParams(int(granularity), float(collision_margin), int(check_self), int(compute_clearance), 0,
                            float(resolution))
        call("ppl_edge_eval_evaluate", self.handle, _p(a, _f32p), _p(b, _f32p), n, C.byref(p), _p(valid, _u8p),
             _p(cost, _f32p), _p(clr, _f32p))
        return valid.astype(bool), cost, clr

    def close(self):
        if getattr(self, "handle", None):
            lib().ppl_edge_eval_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Plans and batched forms
# ---------------------------------------------------------------------------
class Plan:
    """A flattened state space (configs, edge states or waypoints) split into
    64-state chunks, optionally with one OBB range per problem."""

    def __init__(self, robot: Robot, kind: int, n_units: int, unit_states=None, skip_first=False,
                 problem_units=None, obb_offsets=None, waypoint_major=False):
        self.robot = robot
        self.kind = kind
        self.n_units = int(n_units)
        us = None if unit_states is None else _np(unit_states, np.int32)
        pu = None if problem_units is None else _np(problem_units, np.int32)
        oo = None if obb_offsets is None else _np(obb_offsets, np.int32)
        n_prob = 1 if pu is None else int(pu.shape[0] - 1)
        d = _ppl.PlanDesc(int(kind), self.n_units, _p(us, _i32p), int(bool(skip_first)), n_prob, _p(pu, _i32p),
                          _p(oo, _i32p), int(bool(waypoint_major)))
        h = C.c_void_p()
        call("ppl_plan_create", robot.handle, C.byref(d), C.byref(h))
        self.handle = h
        self.unit_states = us
        self.n_states = int(lib().ppl_plan_states(h))
        if kind == _ppl.PLAN_TRAJ:
            self.offsets = np.concatenate([[0], np.cumsum(us)]).astype(np.int32)

    @staticmethod
    def configs(robot, n, problem_units=None, obb_offsets=None) -> "Plan":
        return Plan(robot, _ppl.PLAN_CONFIGS, n, None, False, problem_units, obb_offsets)

    @staticmethod
    def edges(robot, n_states, skip_first=False, problem_units=None, obb_offsets=None) -> "Plan":
        ns = _np(n_states, np.int32)
        return Plan(robot, _ppl.PLAN_EDGES, ns.shape[0], ns, skip_first, problem_units, obb_offsets)

    @staticmethod
    def trajectories(robot, lengths, problem_units=None, obb_offsets=None, waypoint_major=False) -> "Plan":
        """waypoint_major: cut the work waypoint-major inside each problem
        (rollout sets of equal length, e.g. STOMP), so a wave's states are
        near-identical configurations. Results are the same either way."""
        ls = _np(lengths, np.int32)
        return Plan(robot, _ppl.PLAN_TRAJ, ls.shape[0], ls, False, problem_units, obb_offsets, waypoint_major)

    def autotune_cost(self, scene: "DeviceScene", x: torch.Tensor, params: _ppl.CostParams, grad=False, reps=5,
                      stream=None):
        """Time the cost kernel at every lane-group size and in the grouped
        layout on these inputs and keep the fastest for this plan; returns
        (lanes per state or _ppl.PPL_COST_GROUPED, ms per call)."""
        _dev(x)
        g = C.c_int(0)
        ms = C.c_float(0.0)
        v = scene.view()
        call("ppl_plan_autotune_cost", self.handle, C.byref(v), _dp(x), C.byref(params), int(bool(grad)), int(reps),
             C.byref(g), C.byref(ms), _stream(stream))
        return g.value, ms.value

    def set_optimizer_graphs(self, enable: bool):
        """Run optimiser iterations as captured hipGraphs (default) or eagerly."""
        call("ppl_plan_set_optimizer_graphs", self.handle, int(bool(enable)))

    def set_cost_group(self, lanes_per_state: int, grad=False):
        """Pin the cost kernel's lanes per state (a power of two <= 16),
        _ppl.PPL_COST_GROUPED for the grouped kernel, or 0 for the built-in
        choice (grouped for the adjoint)."""
        call("ppl_plan_set_cost_group", self.handle, int(bool(grad)), int(lanes_per_state))

    def autotune_collide(self, scene: "DeviceScene", x: torch.Tensor, x2: Optional[torch.Tensor] = None, reps=5,
                         stream=None):
        """Time the any-collision kernel's launch shapes (OBB staging tile x
        grid form) on these inputs (configurations; or edge starts x and ends
        x2), keep the fastest for this plan and record it in the tuning cache;
        returns (tile, grid form, ms per call)."""
        _dev(x)
        if x2 is not None:
            _dev(x2)
        t, g, ms = C.c_int(0), C.c_int(0), C.c_float(0.0)
        v = scene.view()
        call("ppl_plan_autotune_collide", self.handle, C.byref(v), _dp(x), _dp(x2), int(reps), C.byref(t), C.byref(g),
             C.byref(ms), _stream(stream))
        return t.value, g.value, ms.value

    def collide_shape(self, max_obbs: int):
        """(tile, grid form, source) of the any-collision launches for problems
        of up to max_obbs OBBs; source 0 built-in, 1 tuning cache, 2 this plan."""
        t, g, src = C.c_int(0), C.c_int(0), C.c_int(0)
        call("ppl_plan_collide_shape", self.handle, int(max_obbs), C.byref(t), C.byref(g), C.byref(src))
        return t.value, g.value, src.value

    def set_collide_shape(self, tile: int, grid: int):
        call("ppl_plan_set_collide_shape", self.handle, int(tile), int(grid))

    def time_kernels(self, enable: bool = True):
        """Record a HIP event pair around the main kernel of each later batched
        check or cost evaluation on this plan (read with kernel_ms)."""
        call("ppl_plan_time_kernels", self.handle, int(bool(enable)))

    def kernel_ms(self):
        """(summed ms, launches) of the kernels timed since the last read;
        waits for them and clears the record."""
        ms, n = C.c_double(0.0), C.c_int(0)
        call("ppl_plan_kernel_ms", self.handle, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def close(self):
        if getattr(self, "handle", None):
            lib().ppl_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_cache_file(path: Optional[str]):
    """Where the autotuners' choices are kept (None: in memory only)."""
    call("ppl_tune_cache_file", None if path is None else str(path).encode())


def _dev(t: torch.Tensor, dtype=torch.float32) -> torch.Tensor:
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"expected a contiguous {dtype} tensor")
    return t


def check_collision_batched(plan: Plan, scene: DeviceScene, configs: torch.Tensor, collision_margin=0.0,
                            check_self=False, out: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    _dev(configs)
    if out is None:
        out = torch.empty(plan.n_states, dtype=torch.uint8, device=configs.device)
    v = scene.view()
    call("ppl_check_collision_batched", plan.handle, C.byref(v), _dp(configs), C.c_float(collision_margin),
         int(check_self), _dp(out), _stream(stream))
    return out


def edge_check_batched(plan: Plan, scene: DeviceScene, frm: torch.Tensor, to: torch.Tensor, collision_margin=0.0,
                       check_self=False, skip_resolved=True, state_hits=False, first_hit=None, stream=None):
    _dev(frm)
    _dev(to)
    if first_hit is None:
        first_hit = torch.empty(plan.n_units, dtype=torch.int32, device=frm.device)
    hits = torch.empty(plan.n_states, dtype=torch.uint8, device=frm.device) if state_hits else None
    p = _ppl.EdgeBatchParams(float(collision_margin), int(check_self), int(skip_resolved))
    v = scene.view()
    call("ppl_edge_check_batched", plan.handle, C.byref(v), _dp(frm), _dp(to), C.byref(p), _dp(first_hit),
         _dp(hits), _stream(stream))
    return (first_hit, hits) if state_hits else first_hit


def chomp_state_cost(robot: Robot, scene: DeviceScene, theta: torch.Tensor, T: int, params: _ppl.CostParams,
                     stream=None) -> torch.Tensor:
    """Plain per-trajectory state cost (chomp_state_cost_kernel semantics):
    theta [D*T] dim-major -> [D, T], the per-waypoint cost broadcast over D."""
    _dev(theta)
    out = torch.empty((robot.n_dof, T), dtype=torch.float32, device=theta.device)
    v = scene.view()
    call("ppl_chomp_state_cost", robot.handle, C.byref(v), _dp(theta), int(T), C.byref(params), _dp(out),
         _stream(stream))
    return out


def chomp_fd_grad_plain(robot: Robot, scene: DeviceScene, theta: torch.Tensor, T: int, params: _ppl.CostParams,
                        eps=1e-3, stream=None) -> torch.Tensor:
    """Plain per-trajectory central-FD obstacle gradient (chomp_obs_grad_kernel
    semantics): theta [D*T] dim-major -> [D, T], zero at both ends."""
    _dev(theta)
    out = torch.empty((robot.n_dof, T), dtype=torch.float32, device=theta.device)
    v = scene.view()
    call("ppl_chomp_fd_grad", robot.handle, C.byref(v), _dp(theta), int(T), C.byref(params), C.c_float(eps), _dp(out),
         _stream(stream))
    return out


def chomp_settings(step_size=0.3, smoothness_weight=1.0, collision_cost_weight=20.0, world_margin=0.02,
                   self_margin=0.01, dt=0.1, collision_check_step=0.05, collision_margin=0.0, convergence_eps=1e-4,
                   iterations_after_valid=5, use_self=True, fd_gradient=False, fd_eps=1e-3) -> _ppl.ChompSettings:
    """CHOMP optimiser settings; defaults are the reference's CHOMP_settings."""
    return _ppl.ChompSettings(step_size, smoothness_weight, collision_cost_weight, world_margin, self_margin, dt,
                              collision_check_step, collision_margin, convergence_eps, int(iterations_after_valid),
                              int(bool(use_self)), int(bool(fd_gradient)), fd_eps)


class ChompOptimizer:
    """Batched CHOMP over the trajectories of a plan, all on the device.

    theta is optimised in place; best holds each trajectory's best valid
    snapshot and record the per-trajectory state (see _ppl.CHOMP_*)."""

    def __init__(self, plan: Plan, scene: DeviceScene, theta: torch.Tensor, settings: _ppl.ChompSettings):
        _dev(theta)
        self.plan, self.scene, self.theta, self.settings = plan, scene, theta, settings
        self.best = theta.clone()
        self.record = torch.zeros((plan.n_units, 8), dtype=torch.float32, device=theta.device)
        self._started = False

    def step(self, iterations: int, stream=None):
        v = self.scene.view()
        call("ppl_chomp_optimize", self.plan.handle, C.byref(v), C.byref(self.settings), _dp(self.theta),
             _dp(self.best), _dp(self.record), int(iterations), int(not self._started), _stream(stream))
        self._started = True
        return self

    def solved(self) -> torch.Tensor:
        return self.record[:, _ppl.CHOMP_SOLVED] != 0

    def done(self) -> torch.Tensor:
        return self.record[:, _ppl.CHOMP_DONE] != 0


def stomp_settings(num_rollouts_new=30, num_rollouts_old=20, control_cost_weight=0.1, collision_cost_weight=20.0,
                   exponentiated_cost_sensitivity=2.0, world_margin=0.02, self_margin=0.01, dt=0.1, noise_scale=1.0,
                   collision_check_step=0.05, collision_margin=0.0, convergence_eps=1e-4, iterations_after_valid=3,
                   use_self=True, seed=0x5BE5) -> _ppl.StompSettings:
    """STOMP optimiser settings; defaults are the reference's STOMP_settings."""
    return _ppl.StompSettings(int(num_rollouts_new), int(num_rollouts_old), control_cost_weight, collision_cost_weight,
                              exponentiated_cost_sensitivity, world_margin, self_margin, dt, noise_scale,
                              collision_check_step, collision_margin, convergence_eps, int(iterations_after_valid),
                              int(bool(use_self)), int(seed))


class StompOptimizer(ChompOptimizer):
    """Batched STOMP over the trajectories of a plan, all on the device; the
    same theta / best / record contract as ChompOptimizer."""

    def step(self, iterations: int, stream=None):
        v = self.scene.view()
        call("ppl_stomp_optimize", self.plan.handle, C.byref(v), C.byref(self.settings), _dp(self.theta),
             _dp(self.best), _dp(self.record), int(iterations), int(not self._started), _stream(stream))
        self._started = True
        return self


def _solve(fn: str, robot: Robot, scene: DeviceScene, start, goal, num_timesteps: int, num_iterations: int, settings,
           time_limit_ms: float = 0.0):
    D = robot.n_dof
    a = np.ascontiguousarray(start, np.float32)
    b = np.ascontiguousarray(goal, np.float32)
    if a.shape != (D,) or b.shape != (D,):
        raise ValueError(f"start and goal must have shape ({D},)")
    path = np.empty((num_timesteps, D), np.float32)
    res = _ppl.ChompResult()
    v = scene.view()
    call(fn, robot.handle, None, C.byref(v), a.ctypes.data_as(_ppl._f32p), b.ctypes.data_as(_ppl._f32p),
         int(num_timesteps), int(num_iterations), C.c_float(time_limit_ms), C.byref(settings),
         path.ctypes.data_as(_ppl._f32p), C.byref(res), None)
    return {"solved": bool(res.solved), "path": path, "iterations_run": res.iterations_run,
            "final_state_cost": res.final_state_cost, "final_control_cost": res.final_control_cost,
            "final_total_cost": res.final_total_cost, "wall_time_ms": res.wall_time_ms}


def chomp_solve(robot: Robot, scene: DeviceScene, start, goal, num_timesteps=51, num_iterations=100, settings=None,
                time_limit_ms=0.0):
    """Single-goal CHOMP (ppl_chomp_solve): path [T, D] from start to goal."""
    return _solve("ppl_chomp_solve", robot, scene, start, goal, num_timesteps, num_iterations,
                  settings if settings is not None else chomp_settings(), time_limit_ms)


def stomp_solve(robot: Robot, scene: DeviceScene, start, goal, num_timesteps=51, num_iterations=20, settings=None,
                time_limit_ms=0.0):
    """Single-goal STOMP (ppl_stomp_solve): path [T, D] from start to goal."""
    return _solve("ppl_stomp_solve", robot, scene, start, goal, num_timesteps, num_iterations,
                  settings if settings is not None else stomp_settings(), time_limit_ms)


def rrtc_settings(**kw) -> _ppl.RrtcSettings:
    """pRRTC settings: the reference defaults (pRRTC_settings.hh), overridden by keyword."""
    s = _ppl.RrtcSettings()
    _ppl.lib().ppl_rrtc_default_settings(C.byref(s))
    for k, v in kw.items():
        if not hasattr(s, k):
            raise AttributeError(f"unknown pRRTC setting {k!r}")
        setattr(s, k, v)
    return s


def rrtc_halton_bases(seed: int, n_workgroups: int, n_dof: int) -> np.ndarray:
    """Halton bases [n_workgroups, n_dof] each pRRTC workgroup samples with."""
    out = np.empty((n_workgroups, n_dof), np.float32)
    call("ppl_rrtc_halton_bases", C.c_ulonglong(seed), n_workgroups, n_dof, _p(out, _f32p))
    return out


class RrtcBuffers:
    """Reusable pRRTC trees (ppl_rrtc_buffers): SolverBuffers' role."""

    def __init__(self, n_dof: int, max_samples=1000000, num_new_configs=600, granularity=16):
        self.handle = C.c_void_p()
        call("ppl_rrtc_buffers_create", int(n_dof), int(max_samples), int(num_new_configs), int(granularity),
             C.byref(self.handle))
        self.n_dof, self.max_samples = int(n_dof), int(max_samples)

    def close(self):
        if self.handle:
            _ppl.lib().ppl_rrtc_buffers_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rrtc_solve(robot: Robot, scene: DeviceScene, start, goals, settings: Optional[_ppl.RrtcSettings] = None,
               bufs: Optional[RrtcBuffers] = None, max_waypoints=8194) -> dict:
    """pRRTC (ppl_rrtc_solve): path [N, D] start -> goal when solved."""
    D = robot.n_dof
    a = np.ascontiguousarray(start, np.float32).reshape(-1)
    g = np.ascontiguousarray(goals, np.float32).reshape(-1, D)
    if a.shape != (D,):
        raise ValueError(f"start must have shape ({D},)")
    st = settings if settings is not None else rrtc_settings()
    path = np.empty((max_waypoints, D), np.float32)
    res = _ppl.RrtcResult()
    v = scene.view()
    call("ppl_rrtc_solve", robot.handle, C.byref(v), _p(a, _f32p), _p(g, _f32p), g.shape[0], C.byref(st),
         bufs.handle if bufs is not None else None, _p(path, _f32p), max_waypoints, C.byref(res), None)
    return {"solved": bool(res.solved), "path": path[:res.n_waypoints].copy() if res.solved else path[:0].copy(),
            "iters": res.iters, "cost": res.cost, "start_tree_size": res.start_tree_size,
            "goal_tree_size": res.goal_tree_size, "path_length": res.path_length, "reached_goal": res.reached_goal,
            "wall_time_ms": res.wall_ms, "kernel_ms": res.kernel_ms}


def shortcut_path(robot: Robot, scene: DeviceScene, path, valid_segment_length=0.025, collision_margin=0.0):
    """Greedy shortcut of a host path [N, D] (ppl_shortcut_path)."""
    q = _np(path, np.float32).reshape(-1, robot.n_dof)
    out = np.empty_like(q)
    n = C.c_int(0)
    v = scene.view()
    call("ppl_shortcut_path", robot.handle, C.byref(v), _p(q, _f32p), q.shape[0], C.c_float(valid_segment_length),
         C.c_float(collision_margin), _p(out, _f32p), C.byref(n), None)
    return out[:n.value].copy()


def min_clearance_scene(robot: Robot, scene: DeviceScene, configs, collision_margin=0.0) -> np.ndarray:
    """Plain form: min ESDF clearance (+ margin) of host configurations."""
    q = _np(configs, np.float32).reshape(-1, robot.n_dof)
    out = np.empty(q.shape[0], np.float32)
    v = scene.view()
    call("ppl_min_clearance_scene", robot.handle, C.byref(v), _p(q, _f32p), q.shape[0],
         C.c_float(collision_margin), _p(out, _f32p))
    return out


def hinge_cost_scene(robot: Robot, scene: DeviceScene, configs, params: _ppl.CostParams) -> np.ndarray:
    """Plain form: weighted summed-hinge state cost of host configurations."""
    q = _np(configs, np.float32).reshape(-1, robot.n_dof)
    out = np.empty(q.shape[0], np.float32)
    v = scene.view()
    call("ppl_hinge_cost_scene", robot.handle, C.byref(v), _p(q, _f32p), q.shape[0], C.byref(params),
         _p(out, _f32p))
    return out


def mitstar_batch_cc(robot: Robot, scene: DeviceScene, samples: torch.Tensor, collision_margin=0.0,
                     clearance: Optional[torch.Tensor] = None, clearance_weight=0.0, stream=None) -> torch.Tensor:
    """MIT*'s batch_cc: uint8 verdicts (1 = collision, self check included);
    `clearance` (float32 [n], optional) receives free samples' clearances."""
    _dev(samples)
    n = samples.shape[0]
    cc = torch.empty(n, dtype=torch.uint8, device=samples.device)
    v = scene.view()
    call("ppl_mitstar_batch_cc", robot.handle, C.byref(v), _dp(samples), int(n), C.c_float(collision_margin), _dp(cc),
         _dp(clearance), C.c_float(clearance_weight), _stream(stream))
    return cc


def mitstar_node_clearance(robot: Robot, scene: DeviceScene, nodes: torch.Tensor, clearance: torch.Tensor,
                           node_indices: Optional[torch.Tensor] = None, node_start=0, n_nodes=None,
                           collision_margin=0.0, stream=None) -> torch.Tensor:
    """MIT*'s compute_node_clearance: clearance[node] for the listed nodes."""
    _dev(nodes)
    n = int(node_indices.shape[0]) if node_indices is not None else int(n_nodes)
    v = scene.view()
    call("ppl_mitstar_node_clearance", robot.handle, C.byref(v), _dp(nodes), _dp(node_indices), int(node_start), n,
         C.c_float(collision_margin), _dp(clearance), _stream(stream))
    return clearance


def mitstar_edge_cc(robot: Robot, scene: DeviceScene, nodes: torch.Tensor, src: torch.Tensor, dst: torch.Tensor,
                    valid_segment_length: float, num_sparse_checks=-1, checks_done: Optional[torch.Tensor] = None,
                    collision_margin=0.0, stream=None):
    """MIT*'s sparse_cc / full_cc (num_sparse_checks < 0): per-edge uint8
    verdicts and the midpoints verified free."""
    _dev(nodes)
    E = int(src.shape[0])
    cc = torch.empty(E, dtype=torch.uint8, device=nodes.device)
    new = torch.empty(E, dtype=torch.int32, device=nodes.device)
    v = scene.view()
    call("ppl_mitstar_edge_cc", robot.handle, C.byref(v), _dp(nodes), int(nodes.shape[1]), _dp(src), _dp(dst),
         _dp(checks_done), E, C.c_float(valid_segment_length), int(num_sparse_checks), C.c_float(collision_margin),
         _dp(cc), _dp(new), _stream(stream))
    return cc, new


def fk_batched(plan: Plan, configs: torch.Tensor, frames=True, spheres=True, stream=None):
    """Joint frames [n, J, 4, 4] (row-major view of the column-major output)
    and sphere centres + radii [n, S, 4] of every configuration."""
    _dev(configs)
    info = plan.robot.info()
    f = torch.empty((plan.n_states, info["n_joints"], 16), dtype=torch.float32, device=configs.device) if frames else None
    s = torch.empty((plan.n_states, info["n_spheres"], 4), dtype=torch.float32, device=configs.device) if spheres else None
    call("ppl_fk_batched", plan.handle, _dp(configs), _dp(f), _dp(s), _stream(stream))
    if f is not None:
        f = f.view(plan.n_states, info["n_joints"], 4, 4).transpose(-1, -2)
    return f, s


def min_clearance_batched(plan: Plan, scene: DeviceScene, configs: torch.Tensor, collision_margin=0.0,
                          stream=None) -> torch.Tensor:
    _dev(configs)
    out = torch.empty(plan.n_states, dtype=torch.float32, device=configs.device)
    v = scene.view()
    call("ppl_min_clearance_batched", plan.handle, C.byref(v), _dp(configs), C.c_float(collision_margin), _dp(out),
         _stream(stream))
    return out


def hinge_cost(plan: Plan, scene: DeviceScene, x: torch.Tensor, params: _ppl.CostParams, grad=False, stream=None):
    """Per-state weighted summed-hinge cost, per-trajectory totals (trajectory
    plans) and, with grad=True, d cost / d x from the fused adjoint."""
    _dev(x)
    state = torch.empty(plan.n_states, dtype=torch.float32, device=x.device)
    total = torch.empty(plan.n_units, dtype=torch.float32, device=x.device) if plan.kind == _ppl.PLAN_TRAJ else None
    g = torch.empty_like(x) if grad else None
    v = scene.view()
    call("ppl_hinge_cost_batched", plan.handle, C.byref(v), _dp(x), C.byref(params), _dp(state), _dp(total), _dp(g),
         _stream(stream))
    return state, total, g


def chomp_fd_grad(plan: Plan, scene: DeviceScene, theta: torch.Tensor, params: _ppl.CostParams, eps=1e-3,
                  stream=None) -> torch.Tensor:
    _dev(theta)
    g = torch.empty_like(theta)
    v = scene.view()
    call("ppl_chomp_fd_grad_batched", plan.handle, C.byref(v), _dp(theta), C.byref(params), C.c_float(eps), _dp(g),
         _stream(stream))
    return g


def stomp_rollout_costs(plan: Plan, scene: DeviceScene, rollouts: torch.Tensor, params: _ppl.CostParams,
                        control_weight=0.1, dt=0.1, stream=None):
    _dev(rollouts)
    state = torch.empty(plan.n_states, dtype=torch.float32, device=rollouts.device)
    total = torch.empty(plan.n_units, dtype=torch.float32, device=rollouts.device)
    v = scene.view()
    call("ppl_stomp_rollout_costs", plan.handle, C.byref(v), _dp(rollouts), C.byref(params),
         C.c_float(control_weight), C.c_float(dt), _dp(state), _dp(total), _stream(stream))
    return state, total


def esdf_from_obbs(obbs: DeviceScene, voxel_params, voxel_pose, stream=None) -> torch.Tensor:
    prm = _np(voxel_params, np.float32).reshape(4)
    pose = _np(voxel_pose, np.float32).reshape(8)
    nx, ny, nz = grid_dims(prm)
    out = torch.empty(nx * ny * nz, dtype=torch.float32, device=obbs.obb_dims.device)
    v = obbs.view()
    call("ppl_esdf_from_obbs", C.byref(v), _p(prm, _f32p), _p(pose, _f32p), _dp(out), _stream(stream))
    return out


class Graph:
    """A captured hipGraph of one plan's cost evaluation."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream=None):
        call("ppl_graph_launch", self.handle, _stream(stream))

    def close(self):
        if getattr(self, "handle", None):
            lib().ppl_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def capture_stomp(plan: Plan, scene: DeviceScene, rollouts: torch.Tensor, params, control_weight, dt, state, total,
                  stream=None) -> Graph:
    v = scene.view()
    h = C.c_void_p()
    call("ppl_graph_capture_stomp", plan.handle, C.byref(v), _dp(rollouts), C.byref(params), C.c_float(control_weight),
         C.c_float(dt), _dp(state), _dp(total), _stream(stream), C.byref(h))
    g = Graph(h)
    g._keep = (v, rollouts, state, total, scene)
    return g


def capture_hinge(plan: Plan, scene: DeviceScene, x: torch.Tensor, params, state, total, grad, stream=None) -> Graph:
    v = scene.view()
    h = C.c_void_p()
    call("ppl_graph_capture_hinge", plan.handle, C.byref(v), _dp(x), C.byref(params), _dp(state), _dp(total), _dp(grad),
         _stream(stream), C.byref(h))
    g = Graph(h)
    g._keep = (v, x, state, total, grad, scene)
    return g


class HingeCostFn(torch.autograd.Function):
    """Per-trajectory summed-hinge cost with the fused adjoint as backward.

    forward(theta, plan, scene, params) -> totals [B]; theta is the flat
    ragged trajectory tensor of a trajectory plan. The gradient is computed
    in the forward launch and scaled by grad_output per trajectory."""

    @staticmethod
    def forward(ctx, theta, plan, scene, params):
        _, total, g = hinge_cost(plan, scene, theta.detach().contiguous(), params, grad=True)
        ctx.save_for_backward(g)
        ctx.plan = plan
        return total

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        plan = ctx.plan
        D = plan.robot.n_dof
        lens = torch.as_tensor(plan.unit_states, device=g.device, dtype=torch.int64)
        per_wp = torch.repeat_interleave(grad_out, lens * D)
        return g * per_wp, None, None, None
