"""Scenes with more OBBs than one LDS staging holds (256 per problem), against
the FP64 oracle (needs an MI355X).

The reference loops over any number of OBBs (collision/scene_collision.cuh:
118-167). Here a problem's OBBs pass through LDS in tiles of 256: the
any-collision kernel carries each state's verdict and each joint's
approximate-pass flag across tiles, and the grouped cost kernel folds each
sphere's argmin box across tiles (ties keep the lowest OBB index, as one
pass would). Every entry point that stages OBBs is covered: plain and batched
any-collision (with self), edge first hits, min clearance, summed hinge with
its adjoint and ACM masks, the FD gradient, per-problem OBB sets mixing tiled
and untiled problems in one launch, and the ESDF build.

Bars as in test_gpu_parity: verdicts exact outside |clearance| < 1e-5 m
(counted), costs 1e-5 relative with parity_common's FP32 error model,
gradients 1e-4 norm-wise away from oracle-flagged kinks.
"""
import numpy as np
import pytest
import torch

import pyoracle as PO
from parity_common import assert_values, banner, clearance_den, cost_den, fd_bar, traj_rows
from prismplan_amd import substrate as S
from prismplan_amd import workloads as W
from prismplan_amd.planning_scene import HostScene

pytestmark = pytest.mark.gpu
BAND = 1e-5
DEV = "cuda"
LO, HI = (0.3, -1.2, 0.0), (1.5, 1.2, 1.5)


def many_boxes(rng, n, disabled=0.1):
    sc = W.random_obbs(rng, n, lo=LO, hi=HI, half_lo=0.01, half_hi=0.03)
    sc.obb_enable = (rng.random(n) >= disabled).astype(np.uint8)
    return sc


def assert_verdicts(gpu, ref, clr, what):
    gpu = np.asarray(gpu).astype(bool)
    amb = np.abs(clr) < BAND
    bad = (gpu != ref) & ~amb
    banner(f"{what}: n={gpu.size} colliding={ref.mean():.3f} ambiguous={int(amb.sum())} mismatches={int(bad.sum())}")
    assert not bad.any()


def grad_rel_err(g, ref, w):
    return np.abs(g - ref).max(axis=1) / np.maximum(np.abs(ref).max(axis=1), 1e-3 * w)


@pytest.fixture(scope="module")
def panda():
    return S.Robot("panda")


def test_many_obbs_verdicts_and_clearance(panda):
    rng = np.random.default_rng(90)
    scene = many_boxes(rng, 700)
    r = panda.arrays
    q = W.uniform_configs(rng, r, 3000)
    banner(f"700 boxes ({int(scene.obb_enable.sum())} enabled), 3000 Panda configurations")
    o = PO.Oracle(r, scene, nthreads=8)
    ds = S.DeviceScene.from_host(scene)
    ref, clr = o.check_collision(q)
    assert 0.05 < ref.mean() < 0.5
    assert_verdicts(S.check_collision_scene(panda, ds, q), ref, clr, "plain any-collision, 3 tiles")
    plan = S.Plan.configs(panda, len(q))
    x = torch.from_numpy(q).to(DEV)
    ref_s, clr_s = o.check_collision(q, check_self=True, margin=0.01)
    got = S.check_collision_batched(plan, ds, x, check_self=True, collision_margin=0.01).cpu().numpy()
    assert_verdicts(got, ref_s, clr_s, "batched any-collision + self, margin 0.01")
    mref, vamb = o.min_clearance(q, margin=0.01, with_amb=True)
    assert_values(S.min_clearance_batched(plan, ds, x, collision_margin=0.01).cpu().numpy(), mref, vamb,
                  clearance_den(mref), "  batched min clearance")
    assert_values(S.min_clearance_scene(panda, ds, q, collision_margin=0.01), mref, vamb, clearance_den(mref),
                  "  plain min clearance")
    # the clearance is the ESDF's (scene_min_clearance_runtime): 1000 without layers, tiled or not
    assert np.all(mref == 1000.0)
    few = S.DeviceScene.from_host(many_boxes(rng, 30))
    np.testing.assert_array_equal(S.min_clearance_batched(plan, few, x).cpu().numpy(), np.full(len(q), 1000.0, np.float32))
    np.testing.assert_array_equal(S.min_clearance_scene(panda, few, q), np.full(len(q), 1000.0, np.float32))


def test_many_obbs_edges_first_hit(panda):
    rng = np.random.default_rng(91)
    scene = many_boxes(rng, 600)
    r = panda.arrays
    n_edges = 600
    frm = W.uniform_configs(rng, r, n_edges)
    to = np.clip(frm + rng.normal(scale=0.3, size=frm.shape), r.joint_lower, r.joint_upper).astype(np.float32)
    n_states = rng.integers(8, 64, n_edges).astype(np.int32)
    banner(f"edges first hit: {n_edges} edges, {int(n_states.sum())} states, 600 boxes")
    plan = S.Plan.edges(panda, n_states)
    ds = S.DeviceScene.from_host(scene)
    fh, hits = S.edge_check_batched(plan, ds, torch.from_numpy(frm).to(DEV), torch.from_numpy(to).to(DEV),
                                    skip_resolved=False, state_hits=True)
    fh, hits = fh.cpu().numpy(), hits.cpu().numpy()
    states = PO.edge_states(frm, to, n_states)
    ref, clr = PO.Oracle(r, scene, nthreads=8).check_collision(states)
    assert_verdicts(hits, ref, clr, "per-state edge verdicts")
    off = np.concatenate([[0], np.cumsum(n_states)])
    for k in range(n_edges):
        hk = ref[off[k]:off[k + 1]]
        want = int(np.argmax(hk)) if hk.any() else -1
        if fh[k] != want:
            lo = min(x for x in (fh[k], want) if x >= 0)
            assert abs(clr[off[k] + lo]) < BAND, f"edge {k}: first hit {fh[k]} vs {want}"


@pytest.mark.parametrize("with_acm", [False, True])
def test_many_obbs_hinge_and_adjoint(panda, with_acm):
    rng = np.random.default_rng(92)
    scene = many_boxes(rng, 520)
    r = panda.arrays
    if with_acm:
        scene.sphere_obb_acm = (rng.random((r.n_spheres, scene.n_obbs)) > 0.3).astype(np.uint8)
    q = W.uniform_configs(rng, r, 2000)
    banner(f"hinge + adjoint, 520 boxes, ACM {with_acm}")
    o = PO.Oracle(r, scene, nthreads=8)
    wt = 2.0
    prm = S.cost_params(world_margin=0.05, self_margin=0.01, weight=wt)
    prm_o = PO.cost_params(world_margin=0.05, self_margin=0.01, weight=wt)
    plan = S.Plan.configs(panda, len(q))
    ds = S.DeviceScene.from_host(scene)
    st, _, g = S.hinge_cost(plan, ds, torch.from_numpy(q).to(DEV), prm, grad=True)
    c_ref, g_ref, kink = o.hinge_grad(q, prm_o)
    rel = np.abs(st.cpu().numpy() - c_ref) / cost_den(c_ref, wt, o.hinge_terms(q, prm_o))
    banner(f"  cost max rel err {rel.max():.2e} (active {np.mean(c_ref > 0):.2f})")
    assert np.mean(c_ref > 0) > 0.1 and rel.max() < 1e-5
    gerr = grad_rel_err(g.cpu().numpy(), g_ref, wt)[~kink]
    banner(f"  grad max rel err {gerr.max():.2e} (norm-wise), kink-excluded {int(kink.sum())}")
    assert gerr.max() < 1e-4
    # the cost alone (no adjoint) and the plain form take the same tiled path
    st2 = S.hinge_cost(plan, ds, torch.from_numpy(q).to(DEV), prm)[0].cpu().numpy()
    np.testing.assert_array_equal(st2, st.cpu().numpy())
    c3 = S.hinge_cost_scene(panda, ds, q, prm)
    rel3 = np.abs(c3 - c_ref) / cost_den(c_ref, wt, o.hinge_terms(q, prm_o))
    assert rel3.max() < 1e-5


def test_many_obbs_fd_gradient(panda):
    w = W.chomp(n_traj=8, n_wp=32, seed=93)
    rng = np.random.default_rng(93)
    scene = many_boxes(rng, 400)
    banner("FD gradient kernel, 8 x 32 waypoints, 400 boxes")
    plan = S.Plan.trajectories(panda, w.lengths())
    ds = S.DeviceScene.from_host(scene)
    prm = S.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    prm_o = PO.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    o = PO.Oracle(w.robot, scene, nthreads=8)
    th = torch.from_numpy(w.theta).to(DEV)
    gfd = traj_rows(S.chomp_fd_grad(plan, ds, th, prm, eps=1e-3).cpu().numpy(), w.offsets, 7)
    gfd_ref = traj_rows(o.chomp_fd_grad(w.theta, w.offsets, prm_o, eps=1e-3), w.offsets, 7)
    tol = fd_bar(o, traj_rows(w.theta, w.offsets, 7), prm_o, 1.0, 1e-3)
    diff = np.abs(gfd - gfd_ref)
    interior = np.ones(len(gfd), bool)
    interior[w.offsets[:-1]] = False
    interior[w.offsets[1:] - 1] = False
    banner(f"  max |diff| {diff.max():.2e}, max diff / bar {np.max(diff / tol):.2f}")
    assert not ((diff > tol) & interior[:, None]).any()


def test_tiled_and_untiled_problems_in_one_launch():
    """Per-problem OBB sets of 300, 20, 0 and 530 boxes in one STOMP plan:
    problems 0 and 3 pass through LDS in tiles, 1 in one staging, 2 has none."""
    fetch = S.Robot("fetch")
    wk = W.stomp(n_problems=4, n_rollouts=5, seed=94)
    rng = np.random.default_rng(94)
    counts = [300, 20, 0, 530]
    parts = [many_boxes(rng, n, disabled=0.05) for n in counts]
    scene = HostScene(obb_dims=np.concatenate([p.obb_dims for p in parts]),
                      obb_pose=np.concatenate([p.obb_pose for p in parts]),
                      obb_enable=np.concatenate([p.obb_enable for p in parts]))
    oo = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    R = wk.n_rollouts
    banner(f"STOMP costs, problems with {counts} boxes")
    plan = S.Plan.trajectories(fetch, np.repeat(wk.lengths, R), problem_units=np.arange(5) * R, obb_offsets=oo)
    ds = S.DeviceScene.from_host(scene)
    wc = 20.0
    prm = S.cost_params(world_margin=0.02, self_margin=0.01, weight=wc)
    state, total = S.stomp_rollout_costs(plan, ds, torch.from_numpy(wk.rollouts).to(DEV), prm, control_weight=0.1,
                                         dt=0.1)
    o = PO.Oracle(wk.robot, nthreads=8)
    s_ref, t_ref, nt = o.stomp_costs(scene, oo, wk.rollouts, wk.lengths, wk.roll_offsets, R,
                                     PO.cost_params(world_margin=0.02, self_margin=0.01, weight=wc), 0.1, 0.1,
                                     with_terms=True)
    srel = np.abs(state.cpu().numpy() - s_ref) / cost_den(s_ref, wc, nt)
    trel = np.abs(total.cpu().numpy() - t_ref) / np.abs(t_ref)
    banner(f"  state cost max rel err {srel.max():.2e} (active {np.mean(s_ref > 0):.2f}), totals {trel.max():.2e}")
    assert srel.max() < 1e-5 and trel.max() < 1e-5
    # any-collision over the same per-problem sets (rollout states as configurations)
    n_roll_states = len(wk.rollouts) // fetch.n_dof
    cfg = traj_rows(wk.rollouts, np.concatenate([[0], np.cumsum(np.repeat(wk.lengths, R))]), fetch.n_dof)
    assert len(cfg) == n_roll_states
    per_prob = np.repeat(wk.lengths, R).reshape(4, R).sum(axis=1)
    pu = np.concatenate([[0], np.cumsum(per_prob)])
    cplan = S.Plan.configs(fetch, len(cfg), problem_units=pu, obb_offsets=oo)
    got = S.check_collision_batched(cplan, ds, torch.from_numpy(np.ascontiguousarray(cfg)).to(DEV)).cpu().numpy()
    for p in range(4):
        sub = HostScene(obb_dims=parts[p].obb_dims, obb_pose=parts[p].obb_pose, obb_enable=parts[p].obb_enable)
        ref, clr = PO.Oracle(wk.robot, sub, nthreads=8).check_collision(cfg[pu[p]:pu[p + 1]])
        assert_verdicts(got[pu[p]:pu[p + 1]], ref, clr, f"  problem {p} ({counts[p]} boxes)")


def test_esdf_build_many_obbs():
    rng = np.random.default_rng(95)
    obbs = W.random_obbs(rng, 600, lo=(0.2, -0.4, 0.1), hi=(0.7, 0.4, 0.8), half_lo=0.01, half_hi=0.04)
    params, pose = W.esdf_grid(center=(0.45, 0.0, 0.45), n=32, voxel=0.02)
    banner("ESDF build, 32^3 nodes, 600 boxes")
    feats = S.esdf_from_obbs(S.DeviceScene.from_host(obbs), params[0], pose[0]).cpu().numpy()
    ref = PO.esdf_from_obbs(obbs.obb_dims, obbs.obb_pose, params, pose)
    err = np.abs(feats - ref).max()
    banner(f"  max |diff| {err:.2e}")
    assert err < 2e-6
