"""Seeded synthetic workloads for the substrate benchmarks and parity tests.

Every array is generated in FP64 with numpy's PCG64 and cast to FP32; the FP32
values are what both the GPU path and the FP64 oracle consume. One seed per
workload (0..4). `digest()` gives a SHA-256 over a workload's arrays so runs
can prove they used identical inputs.

Workloads (names describe the benchmark, sizes are the defaults):
  any_collision      Panda, 4,096 configs uniform in limits, 20 random OBBs
  edges              Panda, 64 OBBs, 16,384 straight edges, 8..256 states each
  chomp              Panda, 40 OBBs, 256 trajectories x 64 waypoints
  stomp              Fetch, 128 problems (32..128 waypoints, 30..80 OBBs) x 32 rollouts
  esdf               Baxter, 128^3 ESDF at 1 cm from 60 OBBs, 1,024 trajectories of 32..128
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .planning_scene import HostScene, grid_dims, inverse_pose
from .robot_model_loader import RobotArrays, robot_arrays

PANDA_BOX_LO = (0.2, -0.6, 0.0)
PANDA_BOX_HI = (0.9, 0.6, 1.0)


def random_unit_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform rotations by Shoemake's subgroup algorithm, (w, x, y, z)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    x = a * np.sin(2 * np.pi * u2)
    y = a * np.cos(2 * np.pi * u2)
    z = b * np.sin(2 * np.pi * u3)
    w = b * np.cos(2 * np.pi * u3)
    return np.stack([w, x, y, z], axis=1)


def random_obbs(rng: np.random.Generator, n: int, lo=PANDA_BOX_LO, hi=PANDA_BOX_HI,
                half_lo: float = 0.02, half_hi: float = 0.30) -> HostScene:
    centers = rng.uniform(lo, hi, size=(n, 3))
    half = rng.uniform(half_lo, half_hi, size=(n, 3))
    quats = random_unit_quaternions(rng, n)
    return HostScene.from_boxes(2.0 * half, centers, quats)


def uniform_configs(rng: np.random.Generator, robot: RobotArrays, n: int) -> np.ndarray:
    lo = robot.joint_lower.astype(np.float64)
    hi = robot.joint_upper.astype(np.float64)
    return (lo + rng.random((n, robot.n_dof)) * (hi - lo)).astype(np.float32)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        if a is None:
            continue
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


@dataclass
class AnyCollisionWorkload:
    robot: RobotArrays
    scene: HostScene
    configs: np.ndarray          # [N, D] f32

    def sha256(self) -> str:
        return digest(self.configs, self.scene.obb_dims, self.scene.obb_pose)


def any_collision(n_configs: int = 4096, n_obbs: int = 20, seed: int = 0, robot: str = "panda",
                  n_requests: int = 1) -> List[AnyCollisionWorkload]:
    """`n_requests` independent requests, each with its own configs and OBBs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rob = robot_arrays(robot)
    out = []
    for _ in range(n_requests):
        scene = random_obbs(rng, n_obbs)
        out.append(AnyCollisionWorkload(rob, scene, uniform_configs(rng, rob, n_configs)))
    return out


@dataclass
class EdgeWorkload:
    robot: RobotArrays
    scene: HostScene
    frm: np.ndarray              # [E, D] f32
    to: np.ndarray               # [E, D] f32
    n_states: np.ndarray         # [E] i32, states k = 0..n-1 at t = k/(n-1)

    @property
    def offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.n_states)]).astype(np.int32)

    @property
    def total_states(self) -> int:
        return int(np.sum(self.n_states))

    def sha256(self) -> str:
        return digest(self.frm, self.to, self.n_states, self.scene.obb_dims, self.scene.obb_pose)


def edges(n_edges: int = 16384, n_obbs: int = 64, seed: int = 1, max_step: float = 0.02,
          min_states: int = 8, max_states: int = 256, robot: str = "panda") -> EdgeWorkload:
    """Straight edges with log-uniform state counts. q0 ~ U(limits), direction
    uniform on the sphere scaled so the largest joint step is max_step*(n-1),
    endpoint clamped to limits, then n recomputed from the FP32 endpoints as
    ceil(max|dq| / max_step) + 1 so every step stays within max_step."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rob = robot_arrays(robot)
    scene = random_obbs(rng, n_obbs)
    lo = rob.joint_lower.astype(np.float64)
    hi = rob.joint_upper.astype(np.float64)
    q0 = lo + rng.random((n_edges, rob.n_dof)) * (hi - lo)
    n = np.round(np.exp(rng.uniform(np.log(min_states), np.log(max_states), n_edges))).astype(np.int64)
    d = rng.standard_normal((n_edges, rob.n_dof))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= (max_step * (n - 1) / np.max(np.abs(d), axis=1))[:, None]
    q1 = np.clip(q0 + d, lo, hi)
    f0, f1 = q0.astype(np.float32), q1.astype(np.float32)
    span = np.max(np.abs(f1.astype(np.float64) - f0.astype(np.float64)), axis=1)
    ns = np.maximum(np.ceil(span / max_step - 1e-9).astype(np.int64) + 1, 2)
    return EdgeWorkload(rob, scene, f0, f1, ns.astype(np.int32))


@dataclass
class TrajWorkload:
    robot: RobotArrays
    scene: HostScene
    theta: np.ndarray            # flat ragged [B] x [D][T_b] f32
    offsets: np.ndarray          # [B+1] i32 waypoint offsets

    @property
    def n_traj(self) -> int:
        return int(self.offsets.shape[0] - 1)

    @property
    def n_waypoints(self) -> int:
        return int(self.offsets[-1])

    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets)

    def sha256(self) -> str:
        return digest(self.theta, self.offsets, self.scene.obb_dims, self.scene.obb_pose,
                      self.scene.voxel_features)


def _perturbed_lines(rng, rob: RobotArrays, lengths: np.ndarray, amp: float) -> np.ndarray:
    lo = rob.joint_lower.astype(np.float64)
    hi = rob.joint_upper.astype(np.float64)
    D = rob.n_dof
    blocks = []
    for Tn in lengths:
        start = lo + rng.random(D) * (hi - lo)
        goal = lo + rng.random(D) * (hi - lo)
        k = rng.integers(1, 4, size=D)
        t = np.arange(Tn, dtype=np.float64)
        s = t / (Tn - 1)
        th = start[:, None] + s[None, :] * (goal - start)[:, None] \
            + amp * np.sin(np.pi * k[:, None] * t[None, :] / (Tn - 1))
        th = np.clip(th, lo[:, None], hi[:, None])
        blocks.append(th.astype(np.float32).reshape(-1))
    return np.concatenate(blocks)


def chomp(n_traj: int = 256, n_wp: int = 64, n_obbs: int = 40, seed: int = 2, amp: float = 0.3,
          robot: str = "panda", lengths=None) -> TrajWorkload:
    """Straight start->goal lines plus 0.3 rad sinusoids that vanish at both
    ends; `lengths` (optional) gives each trajectory its own waypoint count."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rob = robot_arrays(robot)
    scene = random_obbs(rng, n_obbs)
    lengths = np.full(n_traj, n_wp, np.int64) if lengths is None else np.asarray(lengths, np.int64)
    theta = _perturbed_lines(rng, rob, lengths, amp)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    return TrajWorkload(rob, scene, theta, off)


def smoothness_cholesky(T: int, dt: float = 0.1) -> np.ndarray:
    """Lower Cholesky factor of R^-1 (R the 5-point acceleration metric),
    scaled so the largest diagonal entry of L L^T is 1."""
    R = control_R_closed_form(T, dt)
    Rinv = np.linalg.inv(R + 1e-8 * np.eye(T))
    Rinv = 0.5 * (Rinv + Rinv.T)
    Rinv /= np.max(np.diag(Rinv))
    return np.linalg.cholesky(Rinv)


def control_R_closed_form(T: int, dt: float) -> np.ndarray:
    """R = dt * A^T A restricted to the central block, A the 5-point
    acceleration stencil on a grid padded by 6 on each side."""
    pad = 6
    Np = T + 2 * pad
    c = np.array([-1.0, 16.0, -30.0, 16.0, -1.0]) / 12.0 / (dt * dt)
    A = np.zeros((Np, Np))
    for i in range(Np):
        for j, v in zip(range(i - 2, i + 3), c):
            if 0 <= j < Np:
                A[i, j] = v
    return dt * (A.T @ A)[pad:pad + T, pad:pad + T]


@dataclass
class StompWorkload:
    robot: RobotArrays
    scene: HostScene              # all problems' OBBs concatenated
    obb_offsets: np.ndarray       # [P+1] i32
    rollouts: np.ndarray          # flat: problem p at roll_offsets[p], [R][D][T_p]
    lengths: np.ndarray           # [P] i32
    roll_offsets: np.ndarray      # [P+1] i32 waypoint offsets (R*T_p per problem)
    n_rollouts: int

    @property
    def n_problems(self) -> int:
        return int(self.lengths.shape[0])

    @property
    def n_states(self) -> int:
        return int(self.roll_offsets[-1])

    def sha256(self) -> str:
        return digest(self.rollouts, self.lengths, self.obb_offsets, self.scene.obb_dims, self.scene.obb_pose)


FETCH_BOX_LO = (0.1, -0.8, 0.0)
FETCH_BOX_HI = (1.1, 0.8, 1.6)


def stomp(n_problems: int = 128, n_rollouts: int = 32, t_lo: int = 32, t_hi: int = 128, o_lo: int = 30,
          o_hi: int = 80, sigma: float = 0.05, seed: int = 3, robot: str = "fetch") -> StompWorkload:
    """Rollouts = clamp(mean + sigma * L z) with L the normalised Cholesky
    factor of R^-1 (zero noise at both ends), mean a straight start->goal line."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rob = robot_arrays(robot)
    lo = rob.joint_lower.astype(np.float64)
    hi = rob.joint_upper.astype(np.float64)
    D = rob.n_dof
    lengths = rng.integers(t_lo, t_hi + 1, size=n_problems)
    n_obb = rng.integers(o_lo, o_hi + 1, size=n_problems)
    scenes, rolls = [], []
    chol = {}
    for p in range(n_problems):
        scenes.append(random_obbs(rng, int(n_obb[p]), lo=FETCH_BOX_LO, hi=FETCH_BOX_HI))
        Tn = int(lengths[p])
        if Tn not in chol:
            chol[Tn] = smoothness_cholesky(Tn)
        L = chol[Tn]
        start = lo + rng.random(D) * (hi - lo)
        goal = lo + rng.random(D) * (hi - lo)
        s = np.arange(Tn) / (Tn - 1)
        mean = start[:, None] + s[None, :] * (goal - start)[:, None]
        z = rng.standard_normal((n_rollouts, D, Tn))
        noise = sigma * np.einsum("tk,rdk->rdt", L, z)
        noise[:, :, 0] = 0.0
        noise[:, :, -1] = 0.0
        r = np.clip(mean[None] + noise, lo[None, :, None], hi[None, :, None])
        rolls.append(r.astype(np.float32).reshape(-1))
    scene = HostScene(obb_dims=np.concatenate([s.obb_dims for s in scenes]),
                      obb_pose=np.concatenate([s.obb_pose for s in scenes]),
                      obb_enable=np.concatenate([s.obb_enable for s in scenes]))
    obb_off = np.concatenate([[0], np.cumsum(n_obb)]).astype(np.int32)
    roff = np.concatenate([[0], np.cumsum(lengths * n_rollouts)]).astype(np.int32)
    return StompWorkload(rob, scene, obb_off, np.concatenate(rolls), lengths.astype(np.int32), roff, n_rollouts)


BAXTER_BOX_LO = (0.25, -0.55, -0.25)
BAXTER_BOX_HI = (1.25, 0.55, 0.75)
BAXTER_GRID_CENTER = (0.75, 0.0, 0.25)


def esdf_grid(center=BAXTER_GRID_CENTER, n: int = 128, voxel: float = 0.01):
    """Voxel params/pose of an axis-aligned n^3 node grid of spacing `voxel`
    centred at `center` (extent (n-1)*voxel)."""
    ext = (n - 1) * voxel
    params = np.array([[ext, ext, ext, voxel]], np.float32)
    t, q = inverse_pose(np.asarray(center, np.float64), np.array([1.0, 0.0, 0.0, 0.0]))
    pose = np.zeros((1, 8), np.float32)
    pose[0, :3] = t
    pose[0, 3:7] = q
    assert grid_dims(params[0]) == (n, n, n)
    return params, pose


def esdf(n_traj: int = 1024, t_lo: int = 32, t_hi: int = 128, n_obbs: int = 60, grid_n: int = 128,
         voxel: float = 0.01, seed: int = 4, amp: float = 0.3, robot: str = "baxter",
         features: Optional[np.ndarray] = None) -> TrajWorkload:
    """Trajectories plus a 1-layer trilinear ESDF of the given OBBs. The ESDF
    values come from `features` when given (e.g. computed on the GPU), else
    they are left None for the caller to fill."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rob = robot_arrays(robot)
    obbs = random_obbs(rng, n_obbs, lo=BAXTER_BOX_LO, hi=BAXTER_BOX_HI, half_lo=0.02, half_hi=0.15)
    params, pose = esdf_grid(n=grid_n, voxel=voxel)
    lengths = rng.integers(t_lo, t_hi + 1, size=n_traj)
    theta = _perturbed_lines(rng, rob, lengths, amp)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    scene = HostScene(voxel_params=params, voxel_pose=pose, voxel_enable=np.ones(1, np.uint8),
                      voxel_features=features, esdf_interp=1)
    scene.source_obbs = obbs  # the OBBs the ESDF is built from
    return TrajWorkload(rob, scene, theta, off)
