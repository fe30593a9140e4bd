"""Robot asset loader: YAML -> flat tables in the runtime robot-model layout.

Mirrors the reference loader's contract (python/gpu_planning/substrate/
robot_model_loader.py:118-190): the YAML already stores row-major fixed
transforms per joint, spheres with their joint index and inclusive-end
self-collision ranges, so loading is flattening plus deriving the approximate
gate model (one bounding sphere per joint that has spheres, and approximate
self-collision ranges that cover every joint pair the full ranges test).

The assets under ``assets/robots`` are the reference's Apache-2.0 robot data
files (see the NOTICE there), unchanged.
"""
from __future__ import annotations

from collections import defaultdict
from pathlib import Path
from typing import Dict, List, Sequence, Tuple, Union

import numpy as np
import yaml

ASSET_DIR = Path(__file__).resolve().parent / "assets" / "robots"
ROBOTS = ("panda", "fetch", "baxter")


def asset_path(name_or_path: Union[str, Path, None]) -> Path:
    """Resolve a robot name ('panda', 'fetch', 'baxter') or a YAML path."""
    if name_or_path is None:
        return ASSET_DIR / "panda.yaml"
    p = Path(name_or_path)
    if p.suffix in (".yaml", ".yml"):
        return p
    return ASSET_DIR / f"{name_or_path}.yaml"


def build_approx_spheres(spheres_flat: Sequence[float], sphere_to_joint: Sequence[int],
                         n_joints: int) -> Tuple[List[float], List[int]]:
    """One bounding sphere per joint: centroid of the joint's sphere centres,
    radius = max over those spheres of (distance to centroid + radius).

    Same derivation as robot_model_loader.py:32-64, in FP64 with the same
    summation order; spheres with radius <= 0 are ignored.
    """
    per_joint: Dict[int, List[Tuple[float, float, float, float]]] = defaultdict(list)
    for si, j in enumerate(sphere_to_joint):
        x, y, z, r = (float(v) for v in spheres_flat[4 * si:4 * si + 4])
        if r <= 0.0:
            continue
        per_joint[int(j)].append((x, y, z, r))
    flat: List[float] = []
    joints: List[int] = []
    for j in range(n_joints):
        group = per_joint.get(j)
        if not group:
            continue
        n = len(group)
        cx = sum(s[0] for s in group) / n
        cy = sum(s[1] for s in group) / n
        cz = sum(s[2] for s in group) / n
        rad = 0.0
        for x, y, z, r in group:
            rad = max(rad, ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) ** 0.5 + r)
        flat.extend([cx, cy, cz, rad])
        joints.append(j)
    return flat, joints


def build_approx_self_cc_ranges(full_ranges: Sequence[int], sphere_to_joint: Sequence[int],
                                approx_sphere_to_joint: Sequence[int]) -> List[int]:
    """Approximate self-collision ranges over approx-sphere indices covering
    every (joint, joint) pair present in the full ranges, as
    [i, start, end] triples with inclusive ends and partners > i, grouped
    into maximal consecutive runs (robot_model_loader.py:67-115).
    """
    if len(full_ranges) == 0:
        return []
    joint_to_approx = {int(j): a for a, j in enumerate(approx_sphere_to_joint)}
    partners: Dict[int, set] = defaultdict(set)
    for e in range(len(full_ranges) // 3):
        si, start, end = (int(v) for v in full_ranges[3 * e:3 * e + 3])
        ji = int(sphere_to_joint[si])
        for sj in range(start, end + 1):
            jj = int(sphere_to_joint[sj])
            if ji == jj:
                continue
            ai, aj = joint_to_approx.get(ji), joint_to_approx.get(jj)
            if ai is None or aj is None:
                continue
            partners[ai].add(aj)
            partners[aj].add(ai)
    out: List[int] = []
    for ai in sorted(partners):
        ps = sorted(p for p in partners[ai] if p > ai)
        if not ps:
            continue
        lo = hi = ps[0]
        for p in ps[1:]:
            if p == hi + 1:
                hi = p
            else:
                out.extend([ai, lo, hi])
                lo = hi = p
        out.extend([ai, lo, hi])
    return out


def load_robot_asset(path=None) -> dict:
    """Parse a robot YAML (name or path) into flat host arrays plus the derived
    approximate gate model. Field names follow the reference builder
    (bindings.cpp:118-160)."""
    path = asset_path(path)
    with open(path, "r") as f:
        doc = yaml.safe_load(f)
    n_dof = int(doc["n_dof"])
    n_joints = int(doc["n_joints"])
    fixed = [float(v) for row in doc["fixed_transforms"] for v in row]
    if len(fixed) != 16 * n_joints:
        raise ValueError(f"{path}: fixed_transforms has {len(fixed)} values, expected {16 * n_joints}")
    spheres = [float(v) for s in doc["spheres"] for v in s]
    s2j = [int(v) for v in doc["sphere_to_joint"]]
    if len(spheres) != 4 * len(s2j):
        raise ValueError(f"{path}: spheres/sphere_to_joint size mismatch")
    ranges = [int(v) for c in doc["self_cc_ranges"] for v in c]
    parents = [int(v) for v in doc.get("joint_parents", [max(i - 1, 0) for i in range(n_joints)])]
    id2dof = [int(v) for v in doc.get("joint_id_to_dof", [i - 1 for i in range(n_joints)])]
    tmem = [int(v) for v in doc.get("t_memory_idx", [0] + [1] * (n_joints - 1))]
    dfs = [int(v) for v in doc.get("dfs_order", list(range(n_joints)))]
    for name, arr in (("joint_parents", parents), ("joint_id_to_dof", id2dof),
                      ("t_memory_idx", tmem), ("dfs_order", dfs)):
        if len(arr) != n_joints:
            raise ValueError(f"{path}: {name} has {len(arr)} entries, expected {n_joints}")
    approx, approx_s2j = build_approx_spheres(spheres, s2j, n_joints)
    approx_ranges = build_approx_self_cc_ranges(ranges, s2j, approx_s2j)
    return {
        "name": doc.get("name", path.stem),
        "n_dof": n_dof,
        "n_joints": n_joints,
        "joint_types": [int(v) for v in doc["joint_types"]],
        "fixed_transforms": fixed,
        "spheres_flat": spheres,
        "sphere_to_joint": s2j,
        "self_cc_ranges_flat": ranges,
        "joint_lower": [float(v) for v in doc["joint_lower"]],
        "joint_upper": [float(v) for v in doc["joint_upper"]],
        "joint_parents": parents,
        "joint_id_to_dof": id2dof,
        "t_memory_idx": tmem,
        "dfs_order": dfs,
        "n_t_slots": max(tmem) + 1,
        "approx_spheres_flat": approx,
        "approx_sphere_to_joint": approx_s2j,
        "approx_self_cc_ranges_flat": approx_ranges,
    }


class RobotArrays:
    """Typed numpy views of an asset (float32 / int32, C-contiguous)."""

    def __init__(self, asset: dict):
        f32 = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.float64).astype(np.float32))
        i32 = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.int32))
        self.name = asset["name"]
        self.n_dof = int(asset["n_dof"])
        self.n_joints = int(asset["n_joints"])
        self.fixed_transforms = f32(asset["fixed_transforms"])
        self.joint_types = i32(asset["joint_types"])
        self.spheres = f32(asset["spheres_flat"]).reshape(-1, 4)
        self.sphere_to_joint = i32(asset["sphere_to_joint"])
        self.self_ranges = i32(asset["self_cc_ranges_flat"]).reshape(-1, 3)
        self.approx_spheres = f32(asset["approx_spheres_flat"]).reshape(-1, 4)
        self.approx_sphere_to_joint = i32(asset["approx_sphere_to_joint"])
        self.approx_self_ranges = i32(asset["approx_self_cc_ranges_flat"]).reshape(-1, 3)
        self.joint_lower = f32(asset["joint_lower"])
        self.joint_upper = f32(asset["joint_upper"])
        self.joint_parents = i32(asset["joint_parents"])
        self.joint_id_to_dof = i32(asset["joint_id_to_dof"])
        self.t_memory_idx = i32(asset["t_memory_idx"])
        self.dfs_order = i32(asset["dfs_order"])
        # the approximate spheres' own chain (RobotModel approx_fixed_transforms /
        # approx_joint_types); None: they ride the full chain
        aft, aty = asset.get("approx_fixed_transforms"), asset.get("approx_joint_types")
        self.approx_fixed_transforms = f32(aft) if aft is not None and len(aft) else None
        self.approx_joint_types = i32(aty) if aty is not None and len(aty) else None

    @property
    def n_spheres(self) -> int:
        return int(self.spheres.shape[0])

    @property
    def n_approx(self) -> int:
        return int(self.approx_spheres.shape[0])

    @property
    def n_self_pairs(self) -> int:
        r = self.self_ranges
        return int(np.sum(r[:, 2] - r[:, 1] + 1)) if len(r) else 0


def robot_arrays(name_or_path=None) -> RobotArrays:
    return RobotArrays(load_robot_asset(name_or_path))


def build_robot_model(path=None):
    """The asset as a built ``prrtc.RobotModelBuilder``, the object the solve
    entry points take (robot_model_loader.py:193-243): every builder field is
    filled from load_robot_asset, the device tables are created at once (so
    an allocation or table error surfaces here, not at the first solve), and
    the builder is returned. `path` is a robot name or YAML path (default:
    the bundled Panda)."""
    from . import prrtc

    asset = load_robot_asset(path)
    b = prrtc.RobotModelBuilder()
    for name in ("n_dof", "n_joints", "fixed_transforms", "joint_types", "spheres_flat", "sphere_to_joint",
                 "self_cc_ranges_flat", "approx_spheres_flat", "approx_sphere_to_joint",
                 "approx_self_cc_ranges_flat", "joint_lower", "joint_upper", "joint_parents", "joint_id_to_dof",
                 "t_memory_idx", "dfs_order"):
        setattr(b, name, asset[name])
    b.build()
    return b
