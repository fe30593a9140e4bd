"""MIT*'s substrate calls (ppl_mitstar_*) against the FP64 oracle.

batch_cc and compute_node_clearance follow planning/MITStar.cu:212-329;
sparse_cc / full_cc's collision part follows MITStar.cu:832-1375: breadth-
first midpoints t = (2 (k + 1 - 2^L) + 1) / 2^(L + 1), q = q_src (1 - t) +
q_dst t, checked with the two-phase test and the self check. Verdicts must
match exactly except for states within 1e-5 m of contact (counted);
clearances agree to 1e-5 relative (denominator max(|ref|, 0.05 m), the FP32
error model of parity_common) at every result the oracle does not flag as
value-ambiguous (a nearest-voxel lookup on a face; counted).
"""
import numpy as np
import pytest
import torch

import pyoracle as PO
from prismplan_amd import substrate as S
from prismplan_amd import workloads as W
from prismplan_amd.planning_scene import HostScene
from parity_common import assert_values, clearance_den

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
BAND = 1e-5


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def world(seed=40):
    rng = np.random.default_rng(seed)
    obbs = W.random_obbs(rng, 12, lo=(0.2, -0.4, 0.1), hi=(0.7, 0.4, 0.8), half_hi=0.12)
    params, pose = W.esdf_grid(center=(0.45, 0.0, 0.45), n=48, voxel=0.02)
    vals = PO.esdf_from_obbs(obbs.obb_dims, obbs.obb_pose, params, pose).astype(np.float32)
    extra = W.random_obbs(rng, 6, lo=(0.3, -0.8, 0.5), hi=(0.9, 0.8, 1.2), half_hi=0.1)
    return HostScene(obb_dims=extra.obb_dims, obb_pose=extra.obb_pose, obb_enable=extra.obb_enable,
                     voxel_params=params, voxel_pose=pose, voxel_enable=np.ones(1, np.uint8),
                     voxel_features=vals.reshape(1, -1), esdf_interp=0)


def bfs_t(k):
    kk = np.asarray(k) + 1
    lvl = np.floor(np.log2(kk)).astype(np.int64)
    return ((2 * (kk - (1 << lvl)) + 1) / (1 << (lvl + 1)).astype(np.float64)).astype(np.float32)


def test_bfs_order_closed_form():
    np.testing.assert_array_equal(bfs_t(np.arange(7)), np.float32([1 / 2, 1 / 4, 3 / 4, 1 / 8, 3 / 8, 5 / 8, 7 / 8]))


def test_batch_cc_and_node_clearance():
    panda = S.Robot("panda")
    r = panda.arrays
    scene = world()
    ds = S.DeviceScene.from_host(scene)
    q = W.uniform_configs(np.random.default_rng(41), r, 3000)
    o = PO.Oracle(r, scene, nthreads=8)
    ref, clr = o.check_collision(q, margin=0.01, check_self=True)
    mref, vamb = o.min_clearance(q, margin=0.01, with_amb=True)
    x = torch.from_numpy(q).to(DEV)
    out = torch.full((len(q),), -7.0, device=DEV)
    cc = S.mitstar_batch_cc(panda, ds, x, 0.01, out, clearance_weight=1.0).cpu().numpy()
    out = out.cpu().numpy()
    amb = np.abs(clr) < BAND
    bad = (cc.astype(bool) != ref) & ~amb
    banner(f"batch_cc: {int(ref.sum())}/{len(q)} colliding, {int(amb.sum())} ambiguous, {int(bad.sum())} mismatches")
    assert not bad.any()
    assert 0.2 < ref.mean() < 0.8
    assert np.all(out[cc == 1] == -7.0)                 # colliding samples keep their entry
    free = cc == 0
    assert_values(out[free], mref[free], vamb[free], clearance_den(mref[free]),
                  "  free-sample clearance")
    out2 = torch.full((len(q),), -7.0, device=DEV)
    cc2 = S.mitstar_batch_cc(panda, ds, x, 0.01, out2, clearance_weight=0.0).cpu().numpy()
    np.testing.assert_array_equal(cc2, cc)
    out2 = out2.cpu().numpy()
    assert np.all(out2[cc2 == 0] == 1000.0) and np.all(out2[cc2 == 1] == -7.0)
    # node clearance: gathered by index, scattered back; others untouched
    idx = np.random.default_rng(42).permutation(len(q))[:700].astype(np.int32)
    dst = torch.full((len(q),), -3.0, device=DEV)
    S.mitstar_node_clearance(panda, ds, x, dst, node_indices=torch.from_numpy(idx).to(DEV), collision_margin=0.01)
    dst = dst.cpu().numpy()
    rest = np.ones(len(q), bool)
    rest[idx] = False
    assert np.all(dst[rest] == -3.0)
    assert_values(dst[idx], mref[idx], vamb[idx], clearance_den(mref[idx]), "  gathered node clearance")
    seq = torch.full((len(q),), -3.0, device=DEV)
    S.mitstar_node_clearance(panda, ds, x, seq, node_start=100, n_nodes=50, collision_margin=0.01)
    seq = seq.cpu().numpy()
    assert np.all(seq[:100] == -3.0) and np.all(seq[150:] == -3.0)
    assert_values(seq[100:150], mref[100:150], vamb[100:150], clearance_den(mref[100:150]),
                  "  node clearance range")


@pytest.mark.parametrize("sparse", [3, -1])
def test_edge_cc_breadth_first_midpoints(sparse):
    panda = S.Robot("panda")
    r = panda.arrays
    scene = world(43)
    ds = S.DeviceScene.from_host(scene)
    rng = np.random.default_rng(44 + sparse)
    nodes = W.uniform_configs(rng, r, 400)
    # near neighbours: short edges, like a planner's graph
    nodes[200:] = (nodes[:200] + rng.normal(scale=0.15, size=(200, 7))).astype(np.float32)
    nodes = np.clip(nodes, r.joint_lower, r.joint_upper).astype(np.float32)
    E = 1500
    src = rng.integers(0, 200, E).astype(np.int32)
    dst = np.where(rng.random(E) < 0.7, src + 200, rng.integers(0, 400, E)).astype(np.int32)
    dst[:5] = src[:5]                                     # zero-length edges
    seg = 0.05
    # FP32 length as the kernels form it: a running fma over the dofs
    diff = (nodes[src] - nodes[dst]).astype(np.float32)
    d2 = np.zeros(E, np.float32)
    for d in range(7):
        d2 = (diff[:, d].astype(np.float64) ** 2 + d2.astype(np.float64)).astype(np.float32)
    length = np.sqrt(d2)
    total = np.maximum(1, np.ceil(length / np.float32(seg))).astype(np.int64)
    target = np.minimum(sparse, total - 1) if sparse >= 0 else total - 1
    done = np.minimum(rng.integers(0, 3, E), np.maximum(target, 0)).astype(np.int32)
    cnt = np.maximum(0, target - done)
    # oracle states, interpolated as the kernels write it (FP32, no contraction)
    e_of = np.repeat(np.arange(E), cnt)
    k = np.concatenate([np.arange(d, d + c) for d, c in zip(done, cnt)]).astype(np.int64)
    t = bfs_t(k)[:, None]
    a, b = nodes[src[e_of]], nodes[dst[e_of]]
    states = ((a * (np.float32(1) - t)).astype(np.float32) + (b * t).astype(np.float32)).astype(np.float32)
    hit, clr = PO.Oracle(r, scene, nthreads=8).check_collision(states, margin=0.005, check_self=True)
    ref = np.zeros(E, bool)
    np.logical_or.at(ref, e_of, hit)
    amb = np.zeros(E, bool)
    np.logical_or.at(amb, e_of, np.abs(clr) < BAND)
    cc, new = S.mitstar_edge_cc(panda, ds, torch.from_numpy(nodes).to(DEV), torch.from_numpy(src).to(DEV),
                                torch.from_numpy(dst).to(DEV), seg, num_sparse_checks=sparse,
                                checks_done=torch.from_numpy(done).to(DEV), collision_margin=0.005)
    cc, new = cc.cpu().numpy().astype(bool), new.cpu().numpy()
    bad = (cc != ref) & ~amb
    banner(f"edge_cc sparse={sparse}: {len(states)} midpoints, {int(ref.sum())}/{E} colliding edges, "
           f"{int(amb.sum())} ambiguous, {int(bad.sum())} mismatches")
    assert not bad.any()
    ok = ~amb
    np.testing.assert_array_equal(new[ok], np.where(ref, 0, cnt)[ok])
    assert np.all(cc[:5] == False) and np.all(new[:5] == 0)
