"""Robot assets and the derived approximate gate model (CPU)."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from prismplan_amd.robot_model_loader import ASSET_DIR, ROBOTS, load_robot_asset, robot_arrays

GOLDEN = json.loads((Path(__file__).parent / "golden" / "loader_golden.json").read_text())

# Sizes counted from the YAML (spheres, approx spheres, self pairs).
SIZES = {"panda": (7, 8, 59, 8, 663), "fetch": (8, 9, 111, 9, 2585), "baxter": (14, 15, 75, 15, 1845)}


@pytest.mark.parametrize("name", ROBOTS)
def test_asset_is_the_reference_data(name):
    data = (ASSET_DIR / f"{name}.yaml").read_bytes()
    assert hashlib.sha256(data).hexdigest() == GOLDEN[name]["yaml_sha256"]


@pytest.mark.parametrize("name", ROBOTS)
def test_sizes(name):
    r = robot_arrays(name)
    assert (r.n_dof, r.n_joints, r.n_spheres, r.n_approx, r.n_self_pairs) == SIZES[name]


@pytest.mark.parametrize("name", ROBOTS)
def test_approx_model_matches_reference_loader(name):
    a = load_robot_asset(name)
    g = GOLDEN[name]
    assert a["approx_sphere_to_joint"] == g["approx_sphere_to_joint"]
    assert a["approx_self_cc_ranges_flat"] == g["approx_self_cc_ranges_flat"]
    # bit-exact FP64 derivation
    assert a["approx_spheres_flat"] == g["approx_spheres_flat"]
    assert a["n_t_slots"] == g["n_t_slots"]


@pytest.mark.parametrize("name", ROBOTS)
def test_approx_spheres_enclose_joint_spheres(name):
    r = robot_arrays(name)
    for a, j in enumerate(r.approx_sphere_to_joint):
        c, R = r.approx_spheres[a, :3].astype(np.float64), float(r.approx_spheres[a, 3])
        for s in np.nonzero(r.sphere_to_joint == j)[0]:
            d = np.linalg.norm(r.spheres[s, :3] - c) + r.spheres[s, 3]
            assert d <= R + 1e-6


@pytest.mark.parametrize("name", ROBOTS)
def test_approx_ranges_cover_joint_pairs(name):
    r = robot_arrays(name)
    j2a = {int(j): a for a, j in enumerate(r.approx_sphere_to_joint)}
    covered = set()
    for i, s, e in r.approx_self_ranges:
        for j in range(s, e + 1):
            covered.add((min(i, j), max(i, j)))
    for i, s, e in r.self_ranges:
        for j in range(s, e + 1):
            ji, jj = int(r.sphere_to_joint[i]), int(r.sphere_to_joint[j])
            if ji != jj:
                ai, aj = j2a[ji], j2a[jj]
                assert (min(ai, aj), max(ai, aj)) in covered


def test_tree_topology_of_baxter():
    r = robot_arrays("baxter")
    # the right arm hangs off the root, not off the left arm's last joint
    assert int(r.joint_parents[8]) == 0
    assert list(r.dfs_order) == list(range(15))
