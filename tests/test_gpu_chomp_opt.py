"""Batched CHOMP optimiser (ppl_chomp_optimize) against the FP64 oracle.

The reference loop (planning/chomp.cu:358-610) is restated here from oracle
pieces: trajectory cost + control cost, densified two-phase validity with the
self check, the covariant step with A = R + 1e-8 I (R the 5-point metric,
inverted in FP64 by numpy), and the best / streak / polish bookkeeping.
Tolerances: costs 1e-5 relative (denominator max(|ref|, weight * margin)),
validity exact except trajectories with a state within 1e-5 m of contact,
the covariant step 1e-5 relative to the largest step of the trajectory.
"""
import numpy as np
import pytest
import torch

import pyoracle as PO
from prismplan_amd import _ppl
from prismplan_amd import substrate as S
from prismplan_amd import workloads as W
from opt_common import densified, prrtc_panda

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
BAND = 1e-5


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def setup(n_traj, n_wp, seed, robot="panda", lengths=None, **kw):
    w = W.chomp(n_traj=n_traj, n_wp=n_wp, n_obbs=30, seed=seed, robot=robot, lengths=lengths)
    panda = S.Robot(robot)
    plan = S.Plan.trajectories(panda, w.lengths())
    ds = S.DeviceScene.from_host(w.scene)
    th = torch.from_numpy(w.theta.copy()).to(DEV)
    st = S.chomp_settings(**kw)
    return w, panda, plan, ds, th, st


def oracle_eval(o, w, theta_flat, st):
    D = w.robot.n_dof
    off = np.concatenate([[0], np.cumsum(w.lengths())]).astype(np.int32)
    prm = PO.cost_params(world_margin=st.world_margin, self_margin=st.self_margin, weight=st.collision_cost_weight,
                         use_self=bool(st.use_self))
    total, grad, kink = o.traj_cost_grad(theta_flat, off, prm)
    total = total + PO.control_cost(theta_flat, off, D, st.dt, st.smoothness_weight)
    valid, amb = [], []
    for b in range(len(off) - 1):
        T = off[b + 1] - off[b]
        th_td = theta_flat[off[b] * D: off[b + 1] * D].reshape(D, T).T
        states, _ = densified(th_td, st.collision_check_step)
        hit, clr = o.check_collision(states, check_self=True)
        valid.append(not hit.any())
        amb.append(bool((np.abs(clr) < BAND).any()))
    return total, grad, np.array(valid), np.array(amb)


def test_initial_evaluation_matches_oracle():
    w, panda, plan, ds, th, st = setup(8, 24, 21, world_margin=0.05)
    opt = S.ChompOptimizer(plan, ds, th, st).step(0)
    rec = opt.record.cpu().numpy()
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    total, _, valid, amb = oracle_eval(o, w, w.theta, st)
    rel = np.abs(rec[:, _ppl.CHOMP_LAST_COST] - total) / np.maximum(np.abs(total), 20.0 * 0.05)
    banner(f"CHOMP initial cost max rel err {rel.max():.2e}; valid {valid.sum()}/{len(valid)}, ambiguous {amb.sum()}")
    assert rel.max() < 1e-5
    solved = rec[:, _ppl.CHOMP_SOLVED] != 0
    assert np.array_equal(solved[~amb], valid[~amb])
    best_cost = rec[:, _ppl.CHOMP_BEST_COST]
    assert np.all(np.isinf(best_cost[~solved]))
    np.testing.assert_allclose(best_cost[solved], rec[solved, _ppl.CHOMP_LAST_COST])
    assert np.all(rec[:, _ppl.CHOMP_POLISH_LEFT] == st.iterations_after_valid)
    np.testing.assert_array_equal(opt.best.cpu().numpy(), w.theta)


def test_covariant_step_matches_fp64():
    w, panda, plan, ds, th, st = setup(6, 40, 22, world_margin=0.05, step_size=0.05)
    D = w.robot.n_dof
    prm = S.cost_params(world_margin=st.world_margin, self_margin=st.self_margin, weight=st.collision_cost_weight,
                        use_self=True)
    plan2 = S.Plan.trajectories(panda, w.lengths())
    _, _, g = S.hinge_cost(plan2, ds, th, prm, grad=True)
    g = g.cpu().numpy().astype(np.float64)
    opt = S.ChompOptimizer(plan, ds, th, st).step(1)
    got = opt.theta.cpu().numpy().astype(np.float64)
    r = w.robot
    off = np.concatenate([[0], np.cumsum(w.lengths())])
    worst = 0.0
    for b in range(w.n_traj):
        T = off[b + 1] - off[b]
        x0 = w.theta[off[b] * D: off[b + 1] * D].reshape(D, T).astype(np.float64)
        gb = g[off[b] * D: off[b + 1] * D].reshape(D, T)
        Ai = np.linalg.inv(PO.control_R(T, st.dt) + 1e-8 * np.eye(T))
        R = PO.control_R(T, st.dt)
        step = (gb + st.smoothness_weight * x0 @ R.T) @ Ai.T   # row d: A^-1 (g_d + lambda R theta_d)
        x1 = x0 - st.step_size * step
        x1[:, 0], x1[:, -1] = x0[:, 0], x0[:, -1]
        x1 = np.clip(x1, r.joint_lower[:, None], r.joint_upper[:, None])
        gb1 = got[off[b] * D: off[b + 1] * D].reshape(D, T)
        scale = max(np.abs(st.step_size * step).max(), 1e-3)
        worst = max(worst, np.abs(gb1 - x1).max() / scale)
    banner(f"covariant step max err {worst:.2e} (relative to the largest step)")
    assert worst < 1e-5


@pytest.mark.parametrize("fd", [False, True])
def test_optimize_loop_bookkeeping(fd):
    w, panda, plan, ds, th, st = setup(12, 32, 23, world_margin=0.05, fd_gradient=fd)
    opt = S.ChompOptimizer(plan, ds, th, st)
    opt.step(0)
    c0 = opt.record[:, _ppl.CHOMP_LAST_COST].clone()
    opt.step(15)
    rec = opt.record.cpu().numpy()
    banner(f"CHOMP {'FD' if fd else 'adjoint'}: solved {int(rec[:, _ppl.CHOMP_SOLVED].sum())}/{w.n_traj}, "
           f"done {int(rec[:, _ppl.CHOMP_DONE].sum())}, cost {c0.mean().item():.3f} -> "
           f"{rec[:, _ppl.CHOMP_LAST_COST].mean():.3f}")
    assert np.all(np.isfinite(rec[:, _ppl.CHOMP_LAST_COST]))
    assert np.all(rec[:, _ppl.CHOMP_ITERS] <= 15)
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    best = opt.best.cpu().numpy()
    total, _, valid, amb = oracle_eval(o, w, best, st)
    solved = rec[:, _ppl.CHOMP_SOLVED] != 0
    # every reported solution is collision-free and its recorded cost is its cost
    assert np.all(valid[solved & ~amb])
    rel = np.abs(rec[solved, _ppl.CHOMP_BEST_COST] - total[solved]) / np.maximum(np.abs(total[solved]), 1.0)
    assert rel.size == 0 or rel.max() < 1e-5
    # endpoints never move; every waypoint stays within the joint limits
    D = w.robot.n_dof
    off = np.concatenate([[0], np.cumsum(w.lengths())])
    cur = opt.theta.cpu().numpy()
    for b in range(w.n_traj):
        T = off[b + 1] - off[b]
        a0 = w.theta[off[b] * D: off[b + 1] * D].reshape(D, T)
        a1 = cur[off[b] * D: off[b + 1] * D].reshape(D, T)
        np.testing.assert_array_equal(a1[:, [0, -1]], a0[:, [0, -1]])
        assert np.all(a1 >= w.robot.joint_lower[:, None]) and np.all(a1 <= w.robot.joint_upper[:, None])


def test_chomp_solve_scene():
    """chomp_solve_scene (planning/chomp.cu:358-610 through the pybind surface):
    the same trajectory as the batched optimiser run from the same straight
    line, endpoints pinned, and reported costs / validity equal to the
    oracle's for the returned path."""
    from prismplan_amd import prrtc
    w, panda, plan, ds, th, _ = setup(1, 32, 24)
    D, T = w.robot.n_dof, 32
    x0 = w.theta[:D * T].reshape(D, T)
    start, goal = x0[:, 0].copy(), x0[:, -1].copy()
    robot = prrtc_panda()
    scene = prrtc.scene_builder(ds)
    ps = prrtc.CHOMPSettings()
    ps.num_timesteps, ps.num_iterations, ps.time_limit_ms = T, 0, 0.0
    ps.world_collision_margin = 0.05
    line = prrtc.chomp_solve_scene(start, goal, scene, ps, robot)
    assert line.iterations_run == 0 and line.path.shape == (T, D)
    ps.num_iterations = 40
    res = prrtc.chomp_solve_scene(start, goal, scene, ps, robot)
    banner(f"chomp_solve_scene: solved {res.solved}, iterations {res.iterations_run}, cost "
           f"{res.final_state_cost:.4f} + {res.final_control_cost:.4f} = {res.final_total_cost:.4f}, "
           f"{res.wall_time_ms:.2f} ms")
    np.testing.assert_array_equal(res.path[0], start)
    np.testing.assert_array_equal(res.path[-1], goal)
    assert 0 < res.iterations_run <= 40
    # the batched optimiser from the same straight line
    st = S.chomp_settings(world_margin=0.05)
    theta = torch.from_numpy(np.ascontiguousarray(line.path.T).reshape(-1)).to(DEV)
    opt = S.ChompOptimizer(S.Plan.trajectories(panda, [T]), ds, theta, st).step(0).step(40)
    rec = opt.record.cpu().numpy()[0]
    assert bool(rec[_ppl.CHOMP_SOLVED]) == res.solved
    assert int(rec[_ppl.CHOMP_ITERS]) == res.iterations_run
    want = (opt.best if res.solved else opt.theta).cpu().numpy().reshape(D, T).T
    np.testing.assert_allclose(res.path[1:-1], want[1:-1], rtol=0, atol=1e-6)
    # reported costs are those of the returned path
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    flat = np.ascontiguousarray(res.path.T).reshape(-1).astype(np.float32)
    off = np.array([0, T], np.int32)
    prm = PO.cost_params(world_margin=0.05, self_margin=st.self_margin, weight=st.collision_cost_weight,
                         use_self=True)
    state, _, _ = o.traj_cost_grad(flat, off, prm)
    ctrl = PO.control_cost(flat, off, D, st.dt, st.smoothness_weight)
    for got, ref in ((res.final_state_cost, state[0]), (res.final_control_cost, ctrl[0]),
                     (res.final_total_cost, state[0] + ctrl[0])):
        assert abs(got - ref) <= 1e-5 * max(abs(ref), st.collision_cost_weight * 0.05)
    if res.solved:
        states, _ = densified(res.path, st.collision_check_step)
        hit, clr = o.check_collision(states, check_self=True)
        assert not hit.any() or (np.abs(clr) < BAND).any()
    # reused buffers give the same solve; mismatched buffers are an error
    bufs = prrtc.CHOMPBuffers.create(D, T)
    assert bufs.owns_memory and bufs.n_dof == D and bufs.num_timesteps == T
    for _ in range(2):
        again = prrtc.chomp_solve_scene(start, goal, scene, ps, robot, bufs)
        np.testing.assert_array_equal(again.path, res.path)
        assert again.iterations_run == res.iterations_run and again.final_total_cost == res.final_total_cost
    bufs.destroy()
    assert not bufs.owns_memory
    small = prrtc.CHOMPBuffers.create(D, T // 2)
    with pytest.raises(RuntimeError):
        prrtc.chomp_solve_scene(start, goal, scene, ps, robot, small)
    small.destroy()
    # a time limit is checked before every iteration
    ps.time_limit_ms = 1e-3
    assert prrtc.chomp_solve_scene(start, goal, scene, ps, robot).iterations_run <= 1
    with pytest.raises(RuntimeError):
        prrtc.chomp_solve_scene(start[:3], goal, scene, ps, robot)


@pytest.mark.parametrize("robot", ["fetch", "baxter"])
def test_covariant_step_other_robots_and_ragged(robot):
    """The covariant step on Fetch and Baxter, ragged lengths including a
    two-waypoint trajectory (no interior: unchanged)."""
    lengths = [24, 2, 40, 17]
    w, rob, plan, ds, th, st = setup(len(lengths), 0, 36, robot=robot, lengths=lengths, world_margin=0.05,
                                     step_size=0.05)
    D = w.robot.n_dof
    prm = S.cost_params(world_margin=st.world_margin, self_margin=st.self_margin, weight=st.collision_cost_weight,
                        use_self=True)
    _, _, g = S.hinge_cost(S.Plan.trajectories(rob, w.lengths()), ds, th, prm, grad=True)
    g = g.cpu().numpy().astype(np.float64)
    opt = S.ChompOptimizer(plan, ds, th, st).step(1)
    got = opt.theta.cpu().numpy().astype(np.float64)
    r = w.robot
    off = np.concatenate([[0], np.cumsum(lengths)])
    worst = 0.0
    for b, T in enumerate(lengths):
        x0 = w.theta[off[b] * D: off[b + 1] * D].reshape(D, T).astype(np.float64)
        x1g = got[off[b] * D: off[b + 1] * D].reshape(D, T)
        if T == 2:
            np.testing.assert_array_equal(x1g, x0)
            continue
        gb = g[off[b] * D: off[b + 1] * D].reshape(D, T)
        R = PO.control_R(T, st.dt)
        Ai = np.linalg.inv(R + 1e-8 * np.eye(T))
        step = (gb + st.smoothness_weight * x0 @ R.T) @ Ai.T
        x1 = x0 - st.step_size * step
        x1[:, 0], x1[:, -1] = x0[:, 0], x0[:, -1]
        x1 = np.clip(x1, r.joint_lower[:, None], r.joint_upper[:, None])
        scale = max(np.abs(st.step_size * step).max(), 1e-3)
        worst = max(worst, np.abs(x1g - x1).max() / scale)
    banner(f"covariant step {robot}: max err {worst:.2e} (relative to the largest step)")
    assert worst < 1e-5
