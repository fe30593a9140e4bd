"""Benchmark problems in normalised form and their conservative OBB worlds.

The reference's benchmark layer (python/gpu_planning/benchmark/problem.py:
78-177, loaders.py:34-118) reads MotionBenchMaker problems whose obstacles
are boxes, spheres and cylinders, while the collision substrate takes
oriented boxes only. Every primitive therefore becomes one box that contains
it: a box stays itself, a sphere becomes the axis-aligned cube of side 2r,
and a cylinder (axis = local +z) becomes the box [2r, 2r, length] in its own
frame. A containing box can only report a collision early, never miss one.

Names, fields and the box rules follow the reference so its scripts run
unchanged. Orientation triples are extrinsic XYZ Euler angles (R = Rz Ry Rx,
SciPy's "xyz"). Problem files are read as JSON; pickled files are refused,
because unpickling runs code from the file.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np


def euler_xyz_to_quat_wxyz(euler: Sequence[float]) -> np.ndarray:
    """Unit quaternion (w, x, y, z) of R = Rz(c) Ry(b) Rx(a) for (a, b, c):
    the half-angle product qz * qy * qx written out, with w >= 0 for small
    angles (problem.py:33-59 convention)."""
    a, b, c = (0.5 * float(v) for v in np.asarray(euler, dtype=np.float64).reshape(3))
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    q = np.array([
        ca * cb * cc + sa * sb * sc,
        sa * cb * cc - ca * sb * sc,
        ca * sb * cc + sa * cb * sc,
        ca * cb * sc - sa * sb * cc,
    ])
    n = float(np.linalg.norm(q))
    return q / n if n > 0.0 else np.array([1.0, 0.0, 0.0, 0.0])


@dataclass
class BenchObstacle:
    """A box ("box": position, quat_wxyz, half_extents), sphere ("sphere":
    position, radius) or cylinder ("cylinder": position, quat_wxyz, radius,
    length along local +z) -- problem.py:78-98."""

    kind: str
    name: str = ""
    position: np.ndarray = field(default_factory=lambda: np.zeros(3))
    quat_wxyz: np.ndarray = field(default_factory=lambda: np.array([1.0, 0.0, 0.0, 0.0]))
    half_extents: np.ndarray = field(default_factory=lambda: np.zeros(3))
    radius: float = 0.0
    length: float = 0.0


@dataclass
class BenchProblem:
    """A joint-space query (start, goal in radians) and its obstacles
    (problem.py:101-117)."""

    scene: str
    index: int
    start: np.ndarray
    goal: np.ndarray
    obstacles: List[BenchObstacle] = field(default_factory=list)
    source: str = ""
    meta: Dict[str, object] = field(default_factory=dict)


def obstacle_to_obb(obs: BenchObstacle) -> Dict[str, object]:
    """One containing box as PlanningScene.update_world takes it: name, full
    extents `dims`, `position`, `quaternion` (w, x, y, z) -- problem.py:123-159."""
    pos = np.asarray(obs.position, dtype=np.float64).reshape(3)
    if obs.kind == "box":
        dims = 2.0 * np.asarray(obs.half_extents, dtype=np.float64).reshape(3)
        quat = np.asarray(obs.quat_wxyz, dtype=np.float64).reshape(4)
    elif obs.kind == "sphere":
        d = 2.0 * float(obs.radius)
        dims, quat = np.array([d, d, d]), np.array([1.0, 0.0, 0.0, 0.0])
    elif obs.kind == "cylinder":
        d = 2.0 * float(obs.radius)
        dims = np.array([d, d, float(obs.length)])
        quat = np.asarray(obs.quat_wxyz, dtype=np.float64).reshape(4)
    else:
        raise ValueError(f"unknown obstacle kind: {obs.kind!r}")
    return {"name": obs.name or obs.kind, "dims": dims.tolist(), "position": pos.tolist(),
            "quaternion": quat.tolist()}


def obstacles_to_obbs(obstacles: Sequence[BenchObstacle]) -> List[Dict[str, object]]:
    """Boxes for a list of primitives. update_world keys obstacles by name, so
    each name gets a running suffix: the first `n` becomes `n_0`, the next
    `n_1`; an empty name becomes `obs_<position in the list>`
    (problem.py:162-177)."""
    out: List[Dict[str, object]] = []
    count: Dict[str, int] = {}
    for i, obs in enumerate(obstacles):
        d = obstacle_to_obb(obs)
        base = str(d["name"])
        k = count.get(base, -1) + 1
        count[base] = k
        d["name"] = f"{base}_{k}" if (base or k) else f"obs_{i}"
        out.append(d)
    return out


def _vamp_obstacles(prob: Dict[str, object]) -> List[BenchObstacle]:
    """Obstacles of one VAMP problem dict (loaders.py:47-72)."""
    out: List[BenchObstacle] = []
    v3 = lambda x: np.asarray(x, dtype=np.float64).reshape(3)
    for i, b in enumerate(prob.get("box") or []):
        out.append(BenchObstacle("box", b.get("name", f"box_{i}"), v3(b["position"]),
                                 euler_xyz_to_quat_wxyz(b["orientation_euler_xyz"]), v3(b["half_extents"])))
    for i, s in enumerate(prob.get("sphere") or []):
        out.append(BenchObstacle("sphere", s.get("name", f"sphere_{i}"), v3(s["position"]), radius=float(s["radius"])))
    for i, c in enumerate(prob.get("cylinder") or []):
        out.append(BenchObstacle("cylinder", c.get("name", f"cylinder_{i}"), v3(c["position"]),
                                 euler_xyz_to_quat_wxyz(c["orientation_euler_xyz"]), radius=float(c["radius"]),
                                 length=float(c["length"])))
    return out


def load_vamp_problems(path: str, scenes: Optional[Sequence[str]] = None,
                       max_per_scene: Optional[int] = None) -> List[BenchProblem]:
    """VAMP MotionBenchMaker problems from a JSON file (loaders.py:75-118):
    raw["problems"][scene] (or the top level) maps a scene to problem dicts
    with valid, start, goals and box / sphere / cylinder lists. Invalid
    problems and problems without a joint goal are skipped; the first goal is
    used; `scenes` filters and `max_per_scene` caps per scene."""
    p = Path(path)
    if not p.is_file():
        raise FileNotFoundError(f"benchmark file not found: {path}")
    if p.suffix != ".json":
        raise ValueError(f"{path}: only JSON problem files are read (pickled files would run code when loaded)")
    raw = json.loads(p.read_text())
    by_scene = raw.get("problems", raw)
    if not isinstance(by_scene, dict):
        raise ValueError("unexpected VAMP file layout: expected a dict of scene -> problems")
    wanted = set(scenes) if scenes else None
    out: List[BenchProblem] = []
    for scene, probs in by_scene.items():
        if wanted is not None and scene not in wanted:
            continue
        kept = 0
        for idx, prob in enumerate(probs or []):
            goals = prob.get("goals") or []
            if not prob.get("valid", True) or not goals:
                continue
            out.append(BenchProblem(scene=scene, index=idx,
                                    start=np.asarray(prob["start"], dtype=np.float32).reshape(-1),
                                    goal=np.asarray(goals[0], dtype=np.float32).reshape(-1),
                                    obstacles=_vamp_obstacles(prob), source="vamp",
                                    meta={"problem_type": prob.get("problem", scene), "n_goals": len(goals)}))
            kept += 1
            if max_per_scene is not None and kept >= max_per_scene:
                break
    return out
