"""HIP path vs the FP64 oracle, through the C ABI (needs an MI355X).

Tolerances (stated per test):
  * Boolean verdicts match exactly except where the FP64 clearance of the
    configuration satisfies |clr| < 1e-5 m (ambiguous band, counted);
  * costs and clearances within 1e-5 relative, with the denominators of
    parity_common's FP32 error model: max(|ref|, 0.05 m) for a clearance,
    max(|ref|, w * 0.05 m * n) for a weighted sum of n non-zero hinge terms
    (n from the FP64 oracle);
  * gradients within 1e-4 relative of the FP64 analytic adjoint, norm-wise
    per waypoint: max_d |g - ref| / max(max_d |ref|, 1e-3 * w), excluding
    waypoints within 1e-5 m of a hinge kink, an argmin tie or a box-SDF kink
    (counted). Element-wise relative errors of small components carry FP32
    cancellation of the large per-sphere terms (the FP32 CPU restatement
    shows the same ~3e-4 element-wise worst case), so the vector is the
    unit of comparison.
"""
from pathlib import Path

import numpy as np
import pytest
import torch

import pyoracle as PO
from prismplan_amd import _ppl
from prismplan_amd import substrate as S
from prismplan_amd import workloads as W
from prismplan_amd.planning_scene import HostScene
from parity_common import assert_values, clearance_den, cost_den, fd_bar, traj_rows, traj_terms

pytestmark = pytest.mark.gpu
BAND = 1e-5
DEV = "cuda"


def grad_rel_err(g, ref, w=1.0):
    """Norm-wise relative error of each waypoint gradient (rows)."""
    num = np.abs(g - ref).max(axis=1)
    den = np.maximum(np.abs(ref).max(axis=1), 1e-3 * w)
    return num / den


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def assert_verdicts(gpu, ref, clr, what):
    gpu = np.asarray(gpu).astype(bool)
    ref = np.asarray(ref).astype(bool)
    amb = np.abs(clr) < BAND
    bad = (gpu != ref) & ~amb
    banner(f"{what}: n={gpu.size} collide={ref.mean():.3f} ambiguous={int(amb.sum())} "
           f"mismatch_in_band={int(((gpu != ref) & amb).sum())}")
    assert not bad.any(), f"{int(bad.sum())} verdict mismatches outside the ambiguous band"


@pytest.fixture(scope="module")
def robots():
    return {n: S.Robot(n) for n in ("panda", "fetch", "baxter")}


def test_plain_check_collision_any_collision_workload(robots):
    w = W.any_collision()[0]
    banner(f"plain check_collision_scene panda N={len(w.configs)} O={w.scene.n_obbs}")
    ds = S.DeviceScene.from_host(w.scene)
    got = S.check_collision_scene(robots["panda"], ds, w.configs)
    ref, clr = PO.Oracle(w.robot, w.scene, nthreads=8).check_collision(w.configs)
    assert_verdicts(got, ref, clr, "plain any-collision")
    # margin variant
    got = S.check_collision_scene(robots["panda"], ds, w.configs, collision_margin=0.03)
    ref, clr = PO.Oracle(w.robot, w.scene, nthreads=8).check_collision(w.configs, margin=0.03)
    assert_verdicts(got, ref, clr, "plain any-collision margin 0.03")


@pytest.mark.parametrize("name", ["fetch", "baxter"])
def test_plain_check_collision_tree_and_prismatic(robots, name):
    rng = np.random.default_rng(21)
    r = robots[name].arrays
    # boxes in front of the base and torso, so some configurations collide and some do not
    lo, hi = ((0.45, -1.0, 0.3), (1.2, 1.0, 1.4)) if name == "fetch" else ((0.45, -1.0, -0.2), (1.2, 1.0, 0.9))
    scene = W.random_obbs(rng, 25, lo=lo, hi=hi, half_hi=0.15)
    q = W.uniform_configs(rng, r, 3000)
    banner(f"plain check {name} N=3000")
    got = S.check_collision_scene(robots[name], S.DeviceScene.from_host(scene), q)
    ref, clr = PO.Oracle(r, scene, nthreads=8).check_collision(q)
    assert 0.05 < ref.mean() < 0.95
    assert_verdicts(got, ref, clr, f"plain any-collision {name}")


def test_empty_inputs(robots):
    w = W.any_collision(n_configs=64)[0]
    ds = S.DeviceScene.from_host(w.scene)
    assert S.check_collision_scene(robots["panda"], ds, np.zeros((0, 7), np.float32)).shape == (0,)
    empty = S.DeviceScene.from_host(HostScene())
    assert not S.check_collision_scene(robots["panda"], empty, w.configs).any()


def test_batched_requests_with_own_scenes(robots):
    reqs = W.any_collision(n_configs=700, n_obbs=20, seed=5, n_requests=6)
    banner("batched any-collision, 6 requests with own OBB sets")
    cfg = np.concatenate([r.configs for r in reqs])
    scene = HostScene(obb_dims=np.concatenate([r.scene.obb_dims for r in reqs]),
                      obb_pose=np.concatenate([r.scene.obb_pose for r in reqs]),
                      obb_enable=np.concatenate([r.scene.obb_enable for r in reqs]))
    pu = np.arange(7) * 700
    oo = np.arange(7) * 20
    plan = S.Plan.configs(robots["panda"], len(cfg), problem_units=pu, obb_offsets=oo)
    out = S.check_collision_batched(plan, S.DeviceScene.from_host(scene), torch.from_numpy(cfg).to(DEV))
    got = out.cpu().numpy()
    for k, rq in enumerate(reqs):
        ref, clr = PO.Oracle(rq.robot, rq.scene, nthreads=8).check_collision(rq.configs)
        assert_verdicts(got[k * 700:(k + 1) * 700], ref, clr, f"request {k}")


@pytest.mark.parametrize("name", ["panda", "fetch", "baxter"])
def test_check_self(robots, name):
    rng = np.random.default_rng(22)
    r = robots[name].arrays
    scene = W.random_obbs(rng, 10, lo=(0.4, -1.0, 0.0), hi=(1.2, 1.0, 1.4), half_hi=0.1)
    q = W.uniform_configs(rng, r, 2000)
    banner(f"batched any-collision + self {name}")
    plan = S.Plan.configs(robots[name], len(q))
    got = S.check_collision_batched(plan, S.DeviceScene.from_host(scene), torch.from_numpy(q).to(DEV),
                                    check_self=True).cpu().numpy()
    ref, clr = PO.Oracle(r, scene, nthreads=8).check_collision(q, check_self=True)
    assert_verdicts(got, ref, clr, f"any-collision+self {name}")


def esdf_scene(n=48, vs=0.02, center=(0.45, 0.0, 0.45), interp=0, with_obbs=True, seed=30, extra_half=0.3):
    rng = np.random.default_rng(seed)
    obbs = W.random_obbs(rng, 12, lo=(0.2, -0.4, 0.1), hi=(0.7, 0.4, 0.8), half_hi=0.12)
    params, pose = W.esdf_grid(center=center, n=n, voxel=vs)
    vals = PO.esdf_from_obbs(obbs.obb_dims, obbs.obb_pose, params, pose).astype(np.float32)
    extra = W.random_obbs(rng, 6, half_hi=extra_half) if with_obbs else HostScene()
    s = HostScene(obb_dims=extra.obb_dims, obb_pose=extra.obb_pose, obb_enable=extra.obb_enable,
                  voxel_params=params, voxel_pose=pose, voxel_enable=np.ones(1, np.uint8),
                  voxel_features=vals.reshape(1, -1), esdf_interp=interp)
    return s, obbs


@pytest.mark.parametrize("interp", [0, 1])
def test_esdf_any_collision_and_min_clearance(robots, interp):
    s, _ = esdf_scene(interp=interp, extra_half=0.1)   # ~25 % of the configurations collide
    r = robots["panda"].arrays
    q = W.uniform_configs(np.random.default_rng(31), r, 3000)
    banner(f"ESDF any-collision interp={interp}")
    ds = S.DeviceScene.from_host(s)
    got = S.check_collision_scene(robots["panda"], ds, q)
    o = PO.Oracle(r, s, nthreads=8)
    ref, clr = o.check_collision(q)
    assert_verdicts(got, ref, clr, f"ESDF any-collision interp={interp}")
    plan = S.Plan.configs(robots["panda"], len(q))
    mc = S.min_clearance_batched(plan, ds, torch.from_numpy(q).to(DEV), collision_margin=0.01).cpu().numpy()
    mref, vamb = o.min_clearance(q, margin=0.01, with_amb=True)
    # a lookup at a voxel face may take either voxel in FP32: the oracle flags exactly those
    assert_values(mc, mref, vamb, clearance_den(mref), f"min clearance interp={interp}")


def test_plain_clearance_and_hinge_forms(robots):
    """Host-array forms of min clearance and the summed hinge cost
    (scene_collision.cuh:339-384, 394-495): the FP64 oracle's values to the
    batched bar, and the device-resident batched forms' values."""
    s, _ = esdf_scene(interp=1, seed=33)
    ds = S.DeviceScene.from_host(s)
    for name in ("panda", "fetch"):
        r = robots[name].arrays
        q = W.uniform_configs(np.random.default_rng(34), r, 1500)
        banner(f"plain min_clearance / hinge_cost forms {name} N={len(q)}")
        o = PO.Oracle(r, s, nthreads=8)
        mc = S.min_clearance_scene(robots[name], ds, q, collision_margin=0.01)
        mref, vamb = o.min_clearance(q, margin=0.01, with_amb=True)
        assert_values(mc, mref, vamb, clearance_den(mref), f"  plain clearance {name}")
        plan = S.Plan.configs(robots[name], len(q))
        x = torch.from_numpy(q).to(DEV)
        np.testing.assert_array_equal(mc, S.min_clearance_batched(plan, ds, x, collision_margin=0.01).cpu().numpy())
        for use_self in (False, True):
            prm = S.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0, use_self=use_self)
            c = S.hinge_cost_scene(robots[name], ds, q, prm)
            c_ref, vamb, nt = o.hinge_cost_amb(q, PO.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0,
                                                                 use_self=use_self), with_terms=True)
            banner(f"  hinge use_self={use_self}: active {np.mean(c_ref > 0):.2f}")
            assert_values(c, c_ref, vamb, cost_den(c_ref, 2.0, nt), f"  plain hinge {name}")
            # the batched form may split a state's spheres over lanes: same terms, other summation order
            np.testing.assert_allclose(c, S.hinge_cost(plan, ds, x, prm)[0].cpu().numpy(), rtol=2e-6, atol=1e-7)
    assert S.min_clearance_scene(robots["panda"], ds, np.zeros((0, 7), np.float32)).shape == (0,)
    assert S.hinge_cost_scene(robots["panda"], ds, np.zeros((0, 7), np.float32), S.cost_params()).shape == (0,)


def edge_verdict_faults(o, frm, to, valid, rv, granularity, resolution, check_self, margin=0.0):
    """Edges whose verdict differs from the oracle's without an explanation:
    a mismatch is explained only when some state the edge checks (j = 1..G)
    lies within BAND of contact in the FP64 oracle."""
    faults, explained = [], 0
    for k in np.nonzero(np.asarray(valid) != np.asarray(rv))[0]:
        G, _ = PO.edge_granularity(frm[k], to[k], granularity, resolution)
        t = np.arange(1, G + 1, dtype=np.float32) / np.float32(G)
        st = (frm[k] + t[:, None] * (to[k] - frm[k])).astype(np.float32)
        _, clr = o.check_collision(st, margin=margin, check_self=check_self)
        if np.any(np.abs(clr) < BAND):
            explained += 1
        else:
            faults.append(int(k))
    return faults, explained


def test_plain_edge_evaluator(robots):
    e = W.edges(n_edges=600, n_obbs=12, seed=40)
    ds = S.DeviceScene.from_host(e.scene)
    ev = S.EdgeEvaluator(robots["panda"], ds, max_edges=1024)
    o = PO.Oracle(e.robot, e.scene, nthreads=8)
    for check_self in (False, True):
        banner(f"EdgeEvalGpu resolution=0.025 check_self={check_self}")
        valid, cost, _ = ev.evaluate(e.frm, e.to, granularity=32, resolution=0.025, check_self=check_self)
        rv, rc, _ = o.edge_eval(e.frm, e.to, granularity=32, resolution=0.025, check_self=check_self)
        faults, explained = edge_verdict_faults(o, e.frm, e.to, valid, rv, 32, 0.025, check_self)
        banner(f"  valid frac {rv.mean():.3f}, in-band mismatches {explained}, unexplained {faults}")
        assert not faults, f"edges {faults}: verdict mismatch with no state in the ambiguous band"
        ok = valid == rv
        np.testing.assert_array_equal(cost[ok], rc[ok])
    # fixed granularity (resolution 0): states j = 1..20
    valid, cost, _ = ev.evaluate(e.frm, e.to, granularity=20, resolution=0.0)
    rv, rc, _ = o.edge_eval(e.frm, e.to, granularity=20, resolution=0.0)
    faults, explained = edge_verdict_faults(o, e.frm, e.to, valid, rv, 20, 0.0, False)
    banner(f"EdgeEvalGpu granularity 20: in-band mismatches {explained}, unexplained {faults}")
    assert not faults, f"edges {faults}: verdict mismatch with no state in the ambiguous band"
    np.testing.assert_array_equal(cost[valid == rv], rc[valid == rv])


def test_edge_eval_flush_pattern(robots):
    """The MHA* / wPA*SE callers' use of EdgeEvalGpu (mhastar_search.hh:256-345,
    wpase_search.hh:192-240): one evaluator, many synchronous calls of varying
    size -- lazy flushes of the pending batch (up to batch_edges) and eager
    expansions of 2 * dim successor edges from one state -- each answered
    before the next, and an over-size batch refused without side effects."""
    e = W.edges(n_edges=400, n_obbs=16, seed=42)
    ds = S.DeviceScene.from_host(e.scene)
    ev = S.EdgeEvaluator(robots["panda"], ds, max_edges=256)
    o = PO.Oracle(e.robot, e.scene, nthreads=8)
    rng = np.random.default_rng(43)
    at, calls, bad = 0, 0, 0
    sizes = [1, 14, 2, 256, 14, 7, 64, 14, 1, 30]
    banner(f"EdgeEvalGpu flush pattern: {len(sizes)} calls of {sizes} edges, max_edges 256")
    for m in sizes:
        if m == 14:   # eager expansion: 2 * dim successors of one state at +-step along each joint
            q = e.frm[rng.integers(len(e.frm))]
            step = np.float32(0.08)
            frm = np.repeat(q[None], 14, axis=0)
            to = frm.copy()
            for d in range(7):
                to[2 * d, d] += step
                to[2 * d + 1, d] -= step
            to = np.clip(to, e.robot.joint_lower, e.robot.joint_upper).astype(np.float32)
        else:
            idx = (np.arange(m) + at) % len(e.frm)
            at += m
            frm, to = e.frm[idx], e.to[idx]
        for check_self in (False, True):
            valid, cost, _ = ev.evaluate(frm, to, granularity=32, resolution=0.025, check_self=check_self)
            rv, rc, _ = o.edge_eval(frm, to, granularity=32, resolution=0.025, check_self=check_self)
            faults, _ = edge_verdict_faults(o, frm, to, valid, rv, 32, 0.025, check_self)
            bad += len(faults)
            ok = valid == rv
            np.testing.assert_array_equal(cost[ok], rc[ok])
            calls += 1
    assert bad == 0, f"{bad} verdict mismatches outside the ambiguous band"
    with pytest.raises(_ppl.PplError) as err:
        ev.evaluate(e.frm[:257], e.to[:257], resolution=0.025)
    assert err.value.status in (_ppl.PPL_ERR_INVALID, _ppl.PPL_ERR_CAPACITY)
    # the evaluator still answers after the refused call
    valid, cost, _ = ev.evaluate(e.frm[:5], e.to[:5], granularity=32, resolution=0.025)
    rv, rc, _ = o.edge_eval(e.frm[:5], e.to[:5], granularity=32, resolution=0.025)
    faults, _ = edge_verdict_faults(o, e.frm[:5], e.to[:5], valid, rv, 32, 0.025, False)
    assert not faults, f"edges {faults} after the refused call"
    banner(f"  {calls} calls checked against the oracle")


def test_plain_edge_clearance_with_esdf(robots):
    s, _ = esdf_scene(interp=0, with_obbs=False)
    e = W.edges(n_edges=200, n_obbs=1, seed=41)
    ds = S.DeviceScene.from_host(s)
    ev = S.EdgeEvaluator(robots["panda"], ds, max_edges=256)
    valid, cost, clr = ev.evaluate(e.frm, e.to, resolution=0.05, compute_clearance=True, collision_margin=0.01)
    o = PO.Oracle(e.robot, s, nthreads=8)
    rv, rc, rclr = o.edge_eval(e.frm, e.to, resolution=0.05, compute_clearance=True, margin=0.01)
    # an edge's clearance is the min over its states j = 1..G: flagged when a state's lookup sits at a jump
    amb = np.zeros(len(e.frm), bool)
    for k in range(len(e.frm)):
        G, _ = PO.edge_granularity(e.frm[k], e.to[k], 32, 0.05)
        t = np.arange(1, G + 1, dtype=np.float32) / np.float32(G)
        st = (e.frm[k] + t[:, None] * (e.to[k] - e.frm[k])).astype(np.float32)
        c_st, a_st, lo_st = o.min_clearance(st, margin=0.01, with_low=True)
        m = c_st.min()
        assert m == pytest.approx(rclr[k], abs=1e-12)
        # a flagged state moves the edge's minimum when it holds it or its other side undercuts it
        amb[k] = bool(np.any(a_st & ((c_st <= m) | (lo_st < m))))
    assert_values(clr, rclr, amb, clearance_den(rclr), "edge clearance")


@pytest.mark.parametrize("n_obbs,seed", [(64, 1), (10, 42)])
def test_batched_edges_first_hit(robots, n_obbs, seed):
    e = W.edges(n_edges=3000, n_obbs=n_obbs, seed=seed)
    banner(f"batched edges E=3000 O={n_obbs} states={e.total_states}")
    ds = S.DeviceScene.from_host(e.scene)
    plan = S.Plan.edges(robots["panda"], e.n_states)
    frm, to = torch.from_numpy(e.frm).to(DEV), torch.from_numpy(e.to).to(DEV)
    fh_skip = S.edge_check_batched(plan, ds, frm, to, skip_resolved=True).cpu().numpy()
    fh_all, hits = S.edge_check_batched(plan, ds, frm, to, skip_resolved=False, state_hits=True)
    fh_all, hits = fh_all.cpu().numpy(), hits.cpu().numpy()
    np.testing.assert_array_equal(fh_skip, fh_all)
    o = PO.Oracle(e.robot, e.scene, nthreads=8)
    states = PO.edge_states(e.frm, e.to, e.n_states)
    ref, clr = o.check_collision(states)
    assert_verdicts(hits, ref, clr, "per-state edge verdicts")
    off = e.offsets
    for k in range(len(e.n_states)):
        hk = ref[off[k]:off[k + 1]]
        want = int(np.argmax(hk)) if hk.any() else -1
        if fh_all[k] != want:
            lo = min(x for x in (fh_all[k], want) if x >= 0)
            assert abs(clr[off[k] + lo]) < BAND, f"edge {k}: first hit {fh_all[k]} vs {want}"


def test_first_hits_across_calls_and_plan_forms(robots):
    """First hits with every state checked equal those of the skipping form
    (which stops at a known hit) over consecutive calls on one plan with
    alternating inputs and scenes, a three-problem plan with its own OBB
    ranges, skip_first, the self pass, and the gated kernel (a model whose
    approximate spheres do not enclose the exact ones)."""
    from parity_common import model_variant
    e1 = W.edges(n_edges=1500, n_obbs=40, seed=61)
    e2 = W.edges(n_edges=1500, n_obbs=12, seed=62)
    n_states = np.maximum(e1.n_states, 1)
    for name, rob in (("panda", robots["panda"]), ("panda gated", S.Robot(model_variant("panda", "shrunk")))):
        for skip_first in (False, True):
            # three problems with their own OBB ranges
            pu = np.array([0, 400, 900, len(n_states)], np.int32)
            plan = S.Plan.edges(rob, n_states + int(skip_first), skip_first=skip_first, problem_units=pu,
                                obb_offsets=np.array([0, 4, 8, 12], np.int32))
            ds1, ds2 = S.DeviceScene.from_host(e1.scene), S.DeviceScene.from_host(e2.scene)
            runs = [(ds1, e1), (ds2, e2), (ds1, e2), (ds2, e1), (ds1, e1)]
            for k, (ds, e) in enumerate(runs):
                frm, to = torch.from_numpy(e.frm).to(DEV), torch.from_numpy(e.to).to(DEV)
                for check_self in (False, True):
                    ref = S.edge_check_batched(plan, ds, frm, to, check_self=check_self, skip_resolved=True)
                    got = S.edge_check_batched(plan, ds, frm, to, check_self=check_self, skip_resolved=False)
                    torch.testing.assert_close(got, ref, rtol=0, atol=0,
                                               msg=f"{name} skip_first={skip_first} call {k} self={check_self}")
            banner(f"{name} skip_first={skip_first}: first hits equal with and without skipping over "
                   f"{len(runs) * 4} calls (hit edges {float((ref >= 0).float().mean()):.2f})")


def test_self_culling_without_enclosing_approx_spheres():
    """The self pass culls pair groups with spheres it builds around each
    joint's spheres, not with the approximate model: with approximate spheres
    shrunk so they no longer enclose, the min-hinge self cost still equals the
    oracle's minimum over every pair (the gated any-collision verdicts of such
    models are covered by test_gpu_general_models)."""
    from prismplan_amd.robot_model_loader import load_robot_asset
    from prismplan_amd.robot_model_loader import RobotArrays
    a = dict(load_robot_asset("panda"))
    ap = np.array(a["approx_spheres_flat"], np.float64).reshape(-1, 4)
    ap[:, 3] *= 0.3
    a["approx_spheres_flat"] = ap.reshape(-1).tolist()
    rob = S.Robot(a)
    r = RobotArrays(a)
    rng = np.random.default_rng(51)
    scene = W.random_obbs(rng, 0)
    q = W.uniform_configs(rng, r, 3000)
    prm = S.cost_params(world_margin=0.05, self_margin=0.05, weight=1.0)
    plan = S.Plan.configs(rob, len(q))
    ds = S.DeviceScene.from_host(scene)
    c = S.hinge_cost(plan, ds, torch.from_numpy(q).to(DEV), prm)[0].cpu().numpy()
    c_ref, _, nt = PO.Oracle(r, scene, nthreads=8).hinge_cost_amb(
        q, PO.cost_params(world_margin=0.05, self_margin=0.05, weight=1.0), with_terms=True)
    rel = np.abs(c - c_ref) / cost_den(c_ref, 1.0, nt)
    banner(f"self hinge with shrunk approx spheres: max rel err {rel.max():.2e}, active {np.mean(c_ref > 0):.2f}")
    assert np.mean(c_ref > 0) > 0.05
    assert rel.max() < 1e-5


def test_hinge_cost_configs(robots):
    rng = np.random.default_rng(50)
    for name in ("panda", "fetch", "baxter"):
        r = robots[name].arrays
        scene = W.random_obbs(rng, 30, lo=(-0.2, -0.9, 0.0), hi=(1.0, 0.9, 1.4))
        q = W.uniform_configs(rng, r, 2000)
        banner(f"hinge cost + self {name}")
        prm_o = PO.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0)
        prm = S.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0)
        plan = S.Plan.configs(robots[name], len(q))
        st, _, g = S.hinge_cost(plan, S.DeviceScene.from_host(scene), torch.from_numpy(q).to(DEV), prm, grad=True)
        o = PO.Oracle(r, scene, nthreads=8)
        c_ref, g_ref, kink = o.hinge_grad(q, prm_o)
        c = st.cpu().numpy()
        rel = np.abs(c - c_ref) / cost_den(c_ref, 2.0, o.hinge_terms(q, prm_o))
        banner(f"  cost max rel err {rel.max():.2e} (active {np.mean(c_ref > 0):.2f})")
        assert rel.max() < 1e-5
        g = g.cpu().numpy()
        gerr = grad_rel_err(g, g_ref, 2.0)[~kink]
        banner(f"  grad max rel err {gerr.max():.2e} (norm-wise), kink-excluded {int(kink.sum())}")
        assert gerr.max() < 1e-4


def test_chomp_adjoint_vs_oracle_and_fd(robots):
    w = W.chomp(n_traj=48, n_wp=64)
    banner("CHOMP adjoint: 48 trajectories x 64 waypoints")
    ds = S.DeviceScene.from_host(w.scene)
    plan = S.Plan.trajectories(robots["panda"], w.lengths())
    th = torch.from_numpy(w.theta).to(DEV)
    prm = S.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    prm_o = PO.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    st, total, g = S.hinge_cost(plan, ds, th, prm, grad=True)
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    tot_ref, g_ref, kink = o.traj_cost_grad(w.theta, w.offsets, prm_o)
    total = total.cpu().numpy()
    rel = np.abs(total - tot_ref) / cost_den(tot_ref, 1.0, traj_terms(o, w.theta, w.offsets, 7, prm_o))
    banner(f"  total max rel err {rel.max():.2e}")
    assert rel.max() < 1e-5
    g = traj_rows(g.cpu().numpy(), w.offsets, 7)
    gr = traj_rows(g_ref, w.offsets, 7)
    interior = np.ones(len(gr), bool)
    interior[w.offsets[:-1]] = False
    interior[w.offsets[1:] - 1] = False
    ok = interior & ~kink
    e = grad_rel_err(g, gr)[ok]
    banner(f"  grad max rel err {e.max():.2e} (norm-wise), kink waypoints {int(kink.sum())}")
    assert e.max() < 1e-4
    assert np.all(g[~interior] == 0)
    # the finite-difference kernel (chomp_obs_grad_kernel semantics) vs the FP64 oracle FD
    gfd = traj_rows(S.chomp_fd_grad(plan, ds, th, prm, eps=1e-3).cpu().numpy(), w.offsets, 7)
    gfd_ref = traj_rows(o.chomp_fd_grad(w.theta, w.offsets, prm_o, eps=1e-3), w.offsets, 7)
    tol = fd_bar(o, traj_rows(w.theta, w.offsets, 7), PO.cost_params(world_margin=0.05, weight=1.0, use_self=False),
                 1.0, 1e-3)
    diff = np.abs(gfd - gfd_ref)
    bad = (diff > tol) & interior[:, None]
    banner(f"  FD kernel: max |diff| {diff.max():.2e}, max diff / bar {np.max(diff / tol):.2f}, over bar "
           f"{int(bad.sum())}")
    assert not bad.any()


def test_autograd_function(robots):
    w = W.chomp(n_traj=8, n_wp=32, seed=7)
    ds = S.DeviceScene.from_host(w.scene)
    plan = S.Plan.trajectories(robots["panda"], w.lengths())
    th = torch.from_numpy(w.theta).to(DEV).requires_grad_(True)
    prm = S.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    total = S.HingeCostFn.apply(th, plan, ds, prm)
    wts = torch.arange(1, 9, dtype=torch.float32, device=DEV)
    (total * wts).sum().backward()
    _, _, g = S.hinge_cost(plan, ds, th.detach(), prm, grad=True)
    expect = g * torch.repeat_interleave(wts, torch.as_tensor(w.lengths(), device=DEV) * 7)
    torch.testing.assert_close(th.grad, expect, rtol=0, atol=0)


@pytest.mark.parametrize("wmajor", [False, True])
def test_stomp_rollout_costs_and_graph(robots, wmajor):
    wk = W.stomp(n_problems=8, n_rollouts=6, seed=3)
    banner(f"STOMP rollouts: {wk.n_problems} problems x {wk.n_rollouts} rollouts, states {wk.n_states}, "
           f"waypoint-major {wmajor}")
    R = wk.n_rollouts
    traj_lengths = np.repeat(wk.lengths, R)
    pu = np.arange(wk.n_problems + 1) * R
    plan = S.Plan.trajectories(robots["fetch"], traj_lengths, problem_units=pu, obb_offsets=wk.obb_offsets,
                               waypoint_major=wmajor)
    ds = S.DeviceScene.from_host(wk.scene)
    ro = torch.from_numpy(wk.rollouts).to(DEV)
    prm = S.cost_params(world_margin=0.02, self_margin=0.01, weight=20.0)
    state, total = S.stomp_rollout_costs(plan, ds, ro, prm, control_weight=0.1, dt=0.1)
    o = PO.Oracle(wk.robot)
    s_ref, t_ref, nt = o.stomp_costs(wk.scene, wk.obb_offsets, wk.rollouts, wk.lengths, wk.roll_offsets, R,
                                     PO.cost_params(world_margin=0.02, self_margin=0.01, weight=20.0), 0.1, 0.1,
                                     with_terms=True)
    st = state.cpu().numpy()
    rel = np.abs(st - s_ref) / cost_den(s_ref, 20.0, nt)
    banner(f"  state cost max rel err {rel.max():.2e}")
    assert rel.max() < 1e-5
    tt = total.cpu().numpy()
    trel = np.abs(tt - t_ref) / np.abs(t_ref)
    banner(f"  rollout total max rel err {trel.max():.2e}")
    assert trel.max() < 1e-5
    # graph replay == eager, several replays with eager work and syncs between
    s2 = torch.zeros_like(state)
    t2 = torch.zeros_like(total)
    graph = S.capture_stomp(plan, ds, ro, prm, 0.1, 0.1, s2, t2)
    for k in range(3):
        s2.zero_()
        t2.zero_()
        graph.launch()
        torch.cuda.synchronize()
        torch.testing.assert_close(t2, total, rtol=0, atol=0)
        torch.testing.assert_close(s2, state, rtol=0, atol=0)
        _ = S.stomp_rollout_costs(plan, ds, ro, prm, control_weight=0.1, dt=0.1)
        torch.cuda.synchronize()


def test_waypoint_major_plan_matches_unit_major(robots):
    """Waypoint-major cutting changes only which states share a wave: costs
    and fused-adjoint gradients equal the unit-major plan's."""
    wk = W.stomp(n_problems=5, n_rollouts=7, seed=11)
    R = wk.n_rollouts
    lengths = np.repeat(wk.lengths, R)
    pu = np.arange(wk.n_problems + 1) * R
    ds = S.DeviceScene.from_host(wk.scene)
    ro = torch.from_numpy(wk.rollouts).to(DEV)
    prm = S.cost_params(world_margin=0.02, self_margin=0.01, weight=20.0)
    outs = []
    for wm in (False, True):
        plan = S.Plan.trajectories(robots["fetch"], lengths, problem_units=pu, obb_offsets=wk.obb_offsets,
                                   waypoint_major=wm)
        outs.append(S.hinge_cost(plan, ds, ro, prm, grad=True))
    (s0, t0, g0), (s1, t1, g1) = outs
    torch.testing.assert_close(s1, s0, rtol=1e-6, atol=1e-7)
    torch.testing.assert_close(t1, t0, rtol=1e-6, atol=1e-6)
    torch.testing.assert_close(g1, g0, rtol=1e-5, atol=1e-6)
    banner(f"  waypoint-major == unit-major over {wk.n_states} states")
    bad = lengths.copy()
    bad[1] += 1
    with pytest.raises(_ppl.PplError) as ei:
        S.Plan.trajectories(robots["fetch"], bad, problem_units=pu, obb_offsets=wk.obb_offsets, waypoint_major=True)
    assert ei.value.status == _ppl.PPL_ERR_INVALID


def test_esdf_build_and_baxter_trilinear_adjoint(robots):
    wk = W.esdf(n_traj=24, t_lo=16, t_hi=40, grid_n=64, voxel=0.02, seed=9)
    obbs = wk.scene.source_obbs
    banner("ESDF build 64^3 + Baxter trilinear adjoint")
    params, pose = wk.scene.voxel_params, wk.scene.voxel_pose
    feats = S.esdf_from_obbs(S.DeviceScene.from_host(obbs), params[0], pose[0]).cpu().numpy()
    ref = PO.esdf_from_obbs(obbs.obb_dims, obbs.obb_pose, params, pose)
    assert np.abs(feats - ref).max() < 2e-6
    wk.scene.voxel_features = feats.reshape(1, -1)
    ds = S.DeviceScene.from_host(wk.scene)
    plan = S.Plan.trajectories(robots["baxter"], wk.lengths())
    th = torch.from_numpy(wk.theta).to(DEV)
    prm = S.cost_params(world_margin=0.05, self_margin=0.01, weight=1.0, use_obb=False, use_self=False)
    st, total, g = S.hinge_cost(plan, ds, th, prm, grad=True)
    o = PO.Oracle(wk.robot, wk.scene, nthreads=8)
    prm_o = PO.cost_params(world_margin=0.05, self_margin=0.01, weight=1.0, use_obb=False, use_self=False)
    tot_ref, g_ref, kink = o.traj_cost_grad(wk.theta, wk.offsets, prm_o)
    rel = np.abs(total.cpu().numpy() - tot_ref) / cost_den(tot_ref, 1.0, traj_terms(o, wk.theta, wk.offsets, 14, prm_o))
    banner(f"  total max rel err {rel.max():.2e} active={np.mean(tot_ref > 0):.2f}")
    assert rel.max() < 1e-5
    g = traj_rows(g.cpu().numpy(), wk.offsets, 14)
    gr = traj_rows(g_ref, wk.offsets, 14)
    interior = np.ones(len(gr), bool)
    interior[wk.offsets[:-1]] = False
    interior[wk.offsets[1:] - 1] = False
    worst = grad_rel_err(g, gr)[interior & ~kink].max()
    banner(f"  grad max rel err {worst:.2e} (norm-wise), kink waypoints {int(kink.sum())}")
    assert worst < 1e-4


def test_esdf_adjoint_at_full_size(robots):
    """The benchmark's ESDF workload at its full size (1,024 Baxter
    trajectories, 81,740 waypoints, 128^3 grid at 1 cm built on the GPU):
    totals within 1e-5 and gradients within 1e-4 of FP64, away from kinks
    (hinge boundaries and voxel faces, where the trilinear gradient jumps)."""
    wk = W.esdf()
    banner(f"ESDF adjoint, full size: {wk.n_traj} trajectories, {wk.n_waypoints} waypoints")
    params, pose = wk.scene.voxel_params, wk.scene.voxel_pose
    feats = S.esdf_from_obbs(S.DeviceScene.from_host(wk.scene.source_obbs), params[0], pose[0])
    wk.scene.voxel_features = feats.cpu().numpy().reshape(1, -1)
    plan = S.Plan.trajectories(robots["baxter"], wk.lengths())
    th = torch.from_numpy(wk.theta).to(DEV)
    prm = S.cost_params(world_margin=0.02, weight=1.0, use_obb=False, use_self=False)
    _, total, g = S.hinge_cost(plan, S.DeviceScene.from_host(wk.scene), th, prm, grad=True)
    o = PO.Oracle(wk.robot, wk.scene, nthreads=16)
    prm_o = PO.cost_params(world_margin=0.02, weight=1.0, use_obb=False, use_self=False)
    tot_ref, g_ref, kink = o.traj_cost_grad(wk.theta, wk.offsets, prm_o)
    rel = np.abs(total.cpu().numpy() - tot_ref) / cost_den(tot_ref, 1.0, traj_terms(o, wk.theta, wk.offsets, 14, prm_o))
    g = traj_rows(g.cpu().numpy(), wk.offsets, 14)
    gr = traj_rows(g_ref, wk.offsets, 14)
    interior = np.ones(len(gr), bool)
    interior[wk.offsets[:-1]] = False
    interior[wk.offsets[1:] - 1] = False
    worst = grad_rel_err(g, gr)[interior & ~kink].max()
    banner(f"  total max rel err {rel.max():.2e}, grad max rel err {worst:.2e}, kinks {int(kink.sum())}")
    assert rel.max() < 1e-5
    assert worst < 1e-4
    assert kink.mean() < 0.05


def test_prrtc_compatible_surface():
    from prismplan_amd import prrtc
    from prismplan_amd.planning_scene import PlanningScene
    from prismplan_amd.robot_model_loader import load_robot_asset
    a = load_robot_asset("panda")
    b = prrtc.RobotModelBuilder()
    b.n_dof, b.n_joints = a["n_dof"], a["n_joints"]
    for k in ("fixed_transforms", "joint_types", "sphere_to_joint", "joint_lower", "joint_upper", "joint_parents",
              "joint_id_to_dof", "t_memory_idx", "dfs_order"):
        setattr(b, k, a[k])
    b.spheres_flat = a["spheres_flat"]
    b.self_cc_ranges_flat = a["self_cc_ranges_flat"]
    b.approx_spheres_flat = a["approx_spheres_flat"]
    b.approx_sphere_to_joint = a["approx_sphere_to_joint"]
    b.approx_self_cc_ranges_flat = a["approx_self_cc_ranges_flat"]
    b.build()
    scene = PlanningScene.from_robot_asset("panda")
    scene.add_cuboid("table", dims=(1.2, 1.2, 0.05), position=(0.5, 0.0, 0.0))
    ready = np.array([[0.0, -0.785, 0.0, -2.356, 0.0, 1.571, 0.785]], np.float32)
    assert prrtc.check_collision_scene(ready, scene.build_scene_collision_data(), b)[0]
    scene.add_cuboid("table", dims=(1.2, 1.2, 0.05), position=(0.5, 0.0, -0.1))
    assert not prrtc.check_collision_scene(ready, scene.build_scene_collision_data(), b)[0]
    with pytest.raises(RuntimeError):
        prrtc.check_collision_scene(np.zeros((2, 6), np.float32), scene.build_scene_collision_data(), b)


DEVTEST = Path(__file__).resolve().parent / "devapi" / "_build" / "libppl_devtest.so"


def devapi_scene(name):
    """An ESDF + OBB world in each robot's reach where its configurations
    collide some of the time (Panda: the shared ESDF world; Fetch and Baxter:
    obstacles moved clear of their bases)."""
    if name == "panda":
        return esdf_scene(interp=0, seed=41)[0]
    center, lo, hi = {"fetch": ((0.85, 0.0, 0.8), (0.6, -0.4, 0.5), (1.1, 0.4, 1.1)),
                      "baxter": ((0.75, 0.0, 0.2), (0.5, -0.5, -0.1), (1.0, 0.5, 0.5))}[name]
    rng = np.random.default_rng(41)
    obbs = W.random_obbs(rng, 12, lo=lo, hi=hi, half_hi=0.12)
    params, pose = W.esdf_grid(center=center, n=48, voxel=0.02)
    vals = PO.esdf_from_obbs(obbs.obb_dims, obbs.obb_pose, params, pose).astype(np.float32)
    extra = W.random_obbs(rng, 6, lo=lo, hi=hi, half_hi=0.1)
    return HostScene(obb_dims=extra.obb_dims, obb_pose=extra.obb_pose, obb_enable=extra.obb_enable,
                     voxel_params=params, voxel_pose=pose, voxel_enable=np.ones(1, np.uint8),
                     voxel_features=vals.reshape(1, -1), esdf_interp=0)


@pytest.mark.parametrize("name,G,frames", [("panda", 4, 0), ("fetch", 16, 0), ("baxter", 1, 0), ("panda", 64, 0),
                                            ("panda", 16, 1), ("fetch", 4, 1), ("baxter", 32, 1)])
def test_device_inline_api(robots, name, G, frames):
    """A planner-style kernel built only on include/ppl/ppl_device.hpp:
    two_phase_cc (+ self), fk_runtime, scene_min_clearance_runtime,
    scene_sum_{obb,esdf}_hinge_runtime and self_min_hinge_runtime, one G-lane
    group per configuration, against the FP64 oracle."""
    import ctypes as C
    assert DEVTEST.exists(), f"{DEVTEST} not built (tests/devapi/build_devtest.py)"
    lib = C.CDLL(str(DEVTEST))
    rob = robots[name]
    r = rob.arrays
    scene = devapi_scene(name)
    rng = np.random.default_rng(43)
    q = W.uniform_configs(rng, r, 1500)
    n = len(q)
    banner(f"device API {name} G={G} n={n} frames-overload={frames}")
    ds = S.DeviceScene.from_host(scene)
    view = ds.view()
    dv = _ppl.RobotDeviceView()
    _ppl.call("ppl_robot_get_device_view", rob.handle, C.byref(dv))
    per = 64 // G
    slots = (n + per - 1) // per * per
    dq = torch.from_numpy(q).to(DEV)
    hit = torch.zeros(n, dtype=torch.uint8, device=DEV)
    outs = [torch.zeros(n, dtype=torch.float32, device=DEV) for _ in range(4)]
    scratch = torch.zeros(slots * (3 * (r.n_spheres + r.n_approx) + 12 * r.n_joints), dtype=torch.float32, device=DEV)
    vp = C.c_void_p
    lib.ppl_devtest_run.restype = C.c_int
    lib.ppl_devtest_run.argtypes = [C.POINTER(_ppl.RobotDeviceView), C.POINTER(_ppl.SceneView), vp, C.c_int, C.c_int,
                                    C.c_float, C.c_int, C.c_float, C.c_float, vp, vp, vp, vp, vp, vp, vp]
    st = lib.ppl_devtest_run(C.byref(dv), C.byref(view), dq.data_ptr(), n, G, 0.0, 1 | (2 * frames), 0.03, 0.01,
                             hit.data_ptr(),
                             *[o.data_ptr() for o in outs], scratch.data_ptr(), None)
    assert st == 0
    torch.cuda.synchronize()
    mc, ho, he, sh = [o.cpu().numpy().astype(np.float64) for o in outs]
    o = PO.Oracle(r, scene, nthreads=8)
    ref, clr = o.check_collision(q, check_self=True)
    assert 0.05 < ref.mean() < 0.95, "the world should make some configurations collide and leave others free"
    assert_verdicts(hit.cpu().numpy(), ref, clr, f"device two_phase_cc {name}")
    ref_mc, vamb = o.min_clearance(q, with_amb=True)
    assert_values(mc, ref_mc, vamb, clearance_den(ref_mc), f"  device min clearance {name}")
    # cost tolerance 1e-5 relative, denominator cost_den (the FP32 error model of parity_common)
    for what, prm, got in (("obb hinge", PO.cost_params(world_margin=0.03, use_obb=True, use_esdf=False,
                                                        use_self=False), ho),
                           ("esdf hinge", PO.cost_params(world_margin=0.03, use_obb=False, use_esdf=True,
                                                         use_self=False), he),
                           ("self hinge", PO.cost_params(self_margin=0.01, use_obb=False, use_esdf=False,
                                                         use_self=True), sh)):
        exp, hamb, nt = o.hinge_cost_amb(q, prm, with_terms=True)
        den = cost_den(exp, 1.0, nt)
        err32 = np.abs(o.hinge_cost(q, prm, precision="f32") - exp) / den
        banner(f"  {what}: FP32 restatement max rel err {err32.max():.2e}, active {np.mean(exp > 0):.2f}")
        assert_values(got, exp, hamb, den, f"  device {what} {name}")


def test_allowed_collision_masks(robots):
    """Scene ACM masks (SceneCollisionData's sphere_obb_acm_mask [S x O] and
    sphere_acm_mask [S x L]; collision/scene_collision.cuh:118-167, 186-331):
    a masked (sphere, box) or (sphere, layer) pair is skipped by the exact
    checks, clearances and hinges; the approximate pass ignores the masks.
    Plain, batched, cost and device-inline paths against the FP64 oracle."""
    import ctypes as C
    rob = robots["panda"]
    r = rob.arrays
    base = devapi_scene("panda")
    rng = np.random.default_rng(77)
    scene = HostScene(obb_dims=base.obb_dims, obb_pose=base.obb_pose, obb_enable=base.obb_enable,
                      voxel_params=base.voxel_params, voxel_pose=base.voxel_pose, voxel_enable=base.voxel_enable,
                      voxel_features=base.voxel_features, esdf_interp=0,
                      sphere_obb_acm=(rng.random((r.n_spheres, base.n_obbs)) > 0.4).astype(np.uint8),
                      sphere_acm=(rng.random((r.n_spheres, base.n_voxel_layers)) > 0.4).astype(np.uint8))
    q = W.uniform_configs(np.random.default_rng(78), r, 2000)
    o = PO.Oracle(r, scene, nthreads=8)
    o_open = PO.Oracle(r, base, nthreads=8)
    ds = S.DeviceScene.from_host(scene)
    banner(f"ACM masks: {int((scene.sphere_obb_acm == 0).sum())} masked sphere-box pairs, "
           f"{int((scene.sphere_acm == 0).sum())} masked sphere-layer pairs")
    ref, clr = o.check_collision(q)
    ref_open, _ = o_open.check_collision(q)
    assert ref.mean() < ref_open.mean(), "masking pairs should free some configurations"
    assert_verdicts(S.check_collision_scene(rob, ds, q), ref, clr, "plain any-collision with ACM")
    plan = S.Plan.configs(rob, len(q))
    x = torch.from_numpy(q).to(DEV)
    ref_s, clr_s = o.check_collision(q, check_self=True)
    got = S.check_collision_batched(plan, ds, x, check_self=True).cpu().numpy()
    assert_verdicts(got, ref_s, clr_s, "batched any-collision + self with ACM")
    mc = S.min_clearance_batched(plan, ds, x, collision_margin=0.01).cpu().numpy()
    mref, vamb = o.min_clearance(q, margin=0.01, with_amb=True)
    assert_values(mc, mref, vamb, clearance_den(mref), "  min clearance with ACM")
    prm = S.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0, use_self=False)
    c = S.hinge_cost(plan, ds, x, prm)[0].cpu().numpy()
    c_ref, hamb, nt = o.hinge_cost_amb(q, PO.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0,
                                                         use_self=False), with_terms=True)
    c_open = o_open.hinge_cost(q, PO.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0, use_self=False))
    banner(f"  hinge with ACM: cost lowered on {np.mean(c_ref < c_open - 1e-6):.2f}")
    assert_values(c, c_ref, hamb, cost_den(c_ref, 2.0, nt), "  hinge with ACM")
    # the fused adjoint's gradient follows the masked argmin
    _, _, g = S.hinge_cost(plan, ds, x, prm, grad=True)
    c_g, g_ref, kink = o.hinge_grad(q, PO.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0,
                                                      use_self=False))
    gerr = grad_rel_err(g.cpu().numpy(), g_ref, 2.0)[~kink]
    banner(f"  hinge gradient with ACM: max rel err {gerr.max():.2e} (norm-wise), kink-excluded {int(kink.sum())}")
    assert gerr.max() < 1e-4
    # the device-inline API, both two_phase_cc overloads
    lib = C.CDLL(str(DEVTEST))
    dv = _ppl.RobotDeviceView()
    _ppl.call("ppl_robot_get_device_view", rob.handle, C.byref(dv))
    view = ds.view()
    n, G = len(q), 16
    slots = (n + 64 // G - 1) // (64 // G) * (64 // G)
    vp = C.c_void_p
    lib.ppl_devtest_run.restype = C.c_int
    lib.ppl_devtest_run.argtypes = [C.POINTER(_ppl.RobotDeviceView), C.POINTER(_ppl.SceneView), vp, C.c_int, C.c_int,
                                    C.c_float, C.c_int, C.c_float, C.c_float, vp, vp, vp, vp, vp, vp, vp]
    for frames in (0, 1):
        hit = torch.zeros(n, dtype=torch.uint8, device=DEV)
        outs = [torch.zeros(n, dtype=torch.float32, device=DEV) for _ in range(4)]
        scratch = torch.zeros(slots * (3 * (r.n_spheres + r.n_approx) + 12 * r.n_joints), dtype=torch.float32,
                              device=DEV)
        assert lib.ppl_devtest_run(C.byref(dv), C.byref(view), x.data_ptr(), n, G, 0.0, 1 | (2 * frames), 0.05,
                                   0.01, hit.data_ptr(), *[t.data_ptr() for t in outs], scratch.data_ptr(), None) == 0
        torch.cuda.synchronize()
        assert_verdicts(hit.cpu().numpy(), ref_s, clr_s, f"device two_phase_cc with ACM (frames overload {frames})")
        for what, got, prm in (("obb", outs[1], PO.cost_params(world_margin=0.05, use_obb=True, use_esdf=False,
                                                                use_self=False)),
                               ("esdf", outs[2], PO.cost_params(world_margin=0.05, use_obb=False, use_esdf=True,
                                                                use_self=False))):
            exp, hamb, nt = o.hinge_cost_amb(q, prm, with_terms=True)
            den = cost_den(exp, 1.0, nt)
            err32 = np.abs(o.hinge_cost(q, prm, precision="f32") - exp) / den
            banner(f"  device {what} hinge with ACM: FP32 restatement max rel err {err32.max():.2e}")
            assert_values(got.cpu().numpy(), exp, hamb, den, f"  device {what} hinge with ACM")


def varied_scene(interp, layer1_enabled, seed=90):
    """Disabled boxes, two ESDF layers (the second on a grid rotated 30 deg
    about z, coarser, possibly disabled), a slab of unobserved voxels (-1000)
    in the first, layers of different sizes in one feature table."""
    from prismplan_amd.planning_scene import inverse_pose
    rng = np.random.default_rng(seed)
    boxes = W.random_obbs(rng, 12, lo=(0.3, -0.6, 0.0), hi=(0.9, 0.6, 1.0), half_hi=0.1)
    enable = (rng.random(12) > 0.3).astype(np.uint8)
    src0 = W.random_obbs(rng, 8, lo=(0.3, -0.4, 0.2), hi=(0.8, 0.4, 0.8), half_hi=0.1)
    src1 = W.random_obbs(rng, 6, lo=(0.2, -0.3, 0.1), hi=(0.7, 0.5, 0.7), half_hi=0.08)
    p0, pose0 = W.esdf_grid(center=(0.5, 0.0, 0.5), n=48, voxel=0.02)
    n1, v1 = 40, 0.025
    ext1 = (n1 - 1) * v1
    p1 = np.array([[ext1, ext1, ext1, v1]], np.float32)
    ang = np.deg2rad(30.0)
    t1, q1 = inverse_pose(np.array([0.45, 0.1, 0.4]), np.array([np.cos(ang / 2), 0.0, 0.0, np.sin(ang / 2)]))
    pose1 = np.zeros((1, 8), np.float32)
    pose1[0, :3], pose1[0, 3:7] = t1, q1
    f0 = PO.esdf_from_obbs(src0.obb_dims, src0.obb_pose, p0, pose0).astype(np.float32)
    f0.reshape(48, 48, 48)[20:28] = -1000.0          # unobserved slab
    f1 = PO.esdf_from_obbs(src1.obb_dims, src1.obb_pose, p1, pose1).astype(np.float32)
    feats = np.full((2, f0.size), -1000.0, np.float32)
    feats[0] = f0
    feats[1, :f1.size] = f1
    return HostScene(obb_dims=boxes.obb_dims, obb_pose=boxes.obb_pose, obb_enable=enable,
                     voxel_params=np.concatenate([p0, p1]), voxel_pose=np.concatenate([pose0, pose1]),
                     voxel_enable=np.array([1, layer1_enabled], np.uint8), voxel_features=feats, esdf_interp=interp)


@pytest.mark.parametrize("interp,layer1", [(0, 1), (1, 1), (0, 0)])
def test_scene_variants(robots, interp, layer1):
    """Disabled boxes, two ESDF layers (one rotated, one disabled or not),
    unobserved voxels: plain, batched, clearance and cost paths against the
    FP64 oracle (scene_collision.cuh:118-331)."""
    rob = robots["panda"]
    r = rob.arrays
    scene = varied_scene(interp, layer1)
    q = W.uniform_configs(np.random.default_rng(91), r, 2000)
    o = PO.Oracle(r, scene, nthreads=8)
    ds = S.DeviceScene.from_host(scene)
    banner(f"scene variants interp={interp} layer1={layer1}: {int(scene.obb_enable.sum())}/12 boxes enabled")
    ref, clr = o.check_collision(q)
    assert 0.05 < ref.mean() < 0.95
    assert_verdicts(S.check_collision_scene(rob, ds, q), ref, clr, "plain any-collision, varied scene")
    plan = S.Plan.configs(rob, len(q))
    x = torch.from_numpy(q).to(DEV)
    ref_s, clr_s = o.check_collision(q, check_self=True)
    assert_verdicts(S.check_collision_batched(plan, ds, x, check_self=True).cpu().numpy(), ref_s, clr_s,
                    "batched any-collision + self, varied scene")
    mc = S.min_clearance_batched(plan, ds, x, collision_margin=0.01).cpu().numpy()
    mref, vamb = o.min_clearance(q, margin=0.01, with_amb=True)
    assert_values(mc, mref, vamb, clearance_den(mref), "  min clearance, varied scene")
    prm = S.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0, use_self=True)
    c = S.hinge_cost(plan, ds, x, prm)[0].cpu().numpy()
    c_ref, hamb, nt = o.hinge_cost_amb(q, PO.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0,
                                                         use_self=True), with_terms=True)
    banner(f"  hinge: active {np.mean(c_ref > 0):.2f}")
    assert_values(c, c_ref, hamb, cost_den(c_ref, 2.0, nt), "  hinge, varied scene")


def test_algorithmic_flop_counts(robots):
    """ppl_algorithmic_flops (include/ppl/roofline.h) reproduces the hand
    counts of the benchmark workloads (FMA = 2, sqrt/sin/cos = 1, abs = 0)."""
    P, F, B = robots["panda"], robots["fetch"], robots["baxter"]
    assert P.flops(_ppl.WORK_ANY_COLLISION, n_obbs=20) == 7 * 83 + 59 * 18 + 59 * 20 * 30 == 37043
    assert P.flops(_ppl.WORK_ANY_COLLISION, n_obbs=64) == 114923
    fwd = 581 + 1062 + 59 * 40 * 36 + 59 * 3
    assert P.flops(_ppl.WORK_CHOMP_ADJOINT, n_obbs=40, n_waypoints=64) == 64 * fwd + 62 * (59 * 35 + 7 * 25 + 1643)
    assert P.flops(_ppl.WORK_CHOMP_FD, n_obbs=40, n_waypoints=64) == (2 * 7 * 62 + 64) * fwd
    assert F.flops(_ppl.WORK_STOMP_STATE, n_obbs=55, use_self=True) == 251199
    assert B.flops(_ppl.WORK_ESDF_WAYPOINT) == 8137 + 5487
    with pytest.raises(ValueError):
        P.flops(99)


def test_autotuned_tile_keeps_results(robots):
    """The cost-kernel autotuner picks a lane-group size that fits or the
    grouped layout, and the plan's results under it equal the default
    layout's (summation order aside)."""
    w = W.chomp(n_traj=24, n_wp=40, n_obbs=30, seed=5)
    ds = S.DeviceScene.from_host(w.scene)
    plan = S.Plan.trajectories(robots["panda"], w.lengths())
    th = torch.from_numpy(w.theta).to(DEV)
    prm = S.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    s0, t0, g0 = S.hinge_cost(plan, ds, th, prm, grad=True)
    G, ms = plan.autotune_cost(ds, th, prm, grad=True, reps=3)
    banner(f"autotuned lanes per state {G} ({ms * 1e3:.1f} us per call)")
    assert G in (1, 2, 4, 8, 16, _ppl.PPL_COST_GROUPED) and ms > 0
    s1, t1, g1 = S.hinge_cost(plan, ds, th, prm, grad=True)
    torch.testing.assert_close(s1, s0, rtol=1e-5, atol=1e-6)
    torch.testing.assert_close(t1, t0, rtol=1e-5, atol=1e-5)
    torch.testing.assert_close(g1, g0, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("name", ["panda", "fetch", "baxter"])
def test_fk_frames_and_spheres(robots, name):
    """ppl_fk_batched vs the FP64 oracle: joint frames (column-major per joint,
    original joint order) and sphere centres, within 2e-6 m / 2e-6."""
    rob = robots[name]
    r = rob.arrays
    q = W.uniform_configs(np.random.default_rng(51), r, 700)
    plan = S.Plan.configs(rob, len(q))
    f, s = S.fk_batched(plan, torch.from_numpy(q).to(DEV))
    o = PO.Oracle(r)
    ref_f = o.fk_frames(q)
    ref_s = o.fk_spheres(q)
    got_f = f.cpu().numpy().astype(np.float64)
    got_s = s.cpu().numpy().astype(np.float64)
    err_f = np.abs(got_f[:, :, :3, :] - ref_f).max()
    err_s = np.abs(got_s[..., :3] - ref_s).max()
    banner(f"FK {name}: frame max err {err_f:.2e}, sphere max err {err_s:.2e}")
    assert err_f < 2e-6 and err_s < 2e-6
    np.testing.assert_array_equal(got_f[:, :, 3, :], np.broadcast_to([0, 0, 0, 1], got_f[:, :, 3, :].shape))
    np.testing.assert_array_equal(got_s[..., 3], np.broadcast_to(r.spheres.reshape(-1, 4)[:, 3], got_s[..., 3].shape))


def test_plain_chomp_forms(robots):
    """Per-trajectory forms (chomp_state_cost_kernel / chomp_obs_grad_kernel
    semantics) against the FP64 oracle and the batched forms."""
    w = W.chomp(n_traj=3, n_wp=48, n_obbs=30, seed=17)
    panda = robots["panda"]
    ds = S.DeviceScene.from_host(w.scene)
    prm = S.cost_params(world_margin=0.05, weight=2.0, use_self=True)
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    D, T = panda.n_dof, 48
    plan = S.Plan.trajectories(panda, w.lengths())
    th_all = torch.from_numpy(w.theta).to(DEV)
    fd_b = S.chomp_fd_grad(plan, ds, th_all, prm, eps=1e-3).cpu().numpy()
    for b in range(w.n_traj):
        th = th_all[b * D * T:(b + 1) * D * T]
        sc = S.chomp_state_cost(panda, ds, th, T, prm).cpu().numpy()
        q = th.cpu().numpy().reshape(D, T).T
        prm_o = PO.cost_params(world_margin=0.05, self_margin=0.01, weight=2.0)
        ref, _, nt = o.hinge_cost_amb(q, prm_o, with_terms=True)
        rel = np.abs(sc - ref[None, :]) / cost_den(ref, 2.0, nt)[None, :]
        assert rel.max() < 1e-5
        fd = S.chomp_fd_grad_plain(panda, ds, th, T, prm, eps=1e-3).cpu().numpy()
        np.testing.assert_array_equal(fd.reshape(-1), fd_b[b * D * T:(b + 1) * D * T])
        assert np.all(fd[:, 0] == 0) and np.all(fd[:, -1] == 0)
    banner(f"plain CHOMP forms: {w.n_traj} trajectories, state cost vs oracle and FD == batched FD")
