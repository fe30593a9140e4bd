"""Batched STOMP optimiser (ppl_stomp_optimize) against the FP64 STOMP oracle
(oracle/stomp_oracle.py, planning/stomp.cu:565-930 semantics with the same
Philox sampler).

Tolerances: the optimum after each iteration within 1e-5 rad of the oracle's
(FP32 sampling, costs and control stencil against FP64; measured 4e-7 over
four iterations, while a wrong sign, index or term moves it by ~1e-2); recorded costs 1e-5 relative (denominator
max(|ref|, 1)); validity exact except states within 1e-5 m of contact.
"""
import numpy as np
import pytest
import torch

import pyoracle as PO
import stomp_oracle as SO
from opt_common import densified, prrtc_panda
from prismplan_amd import _ppl
from prismplan_amd import substrate as S
from prismplan_amd import workloads as W

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
BAND = 1e-5
STEP_TOL = 1e-5


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def setup(lengths, seed, robot="panda", **kw):
    w = W.chomp(n_traj=len(lengths), n_obbs=30, seed=seed, lengths=lengths, robot=robot)
    panda = S.Robot(robot)
    plan = S.Plan.trajectories(panda, w.lengths())
    ds = S.DeviceScene.from_host(w.scene)
    th = torch.from_numpy(w.theta.copy()).to(DEV)
    st = S.stomp_settings(**kw)
    return w, panda, plan, ds, th, st


def oracle_fns(o, st):
    prm = PO.cost_params(world_margin=st.world_margin, self_margin=st.self_margin, weight=st.collision_cost_weight,
                         use_self=bool(st.use_self))

    def cost(x):
        return o.hinge_cost(np.ascontiguousarray(x, np.float32), prm)

    def check(th_td):
        states, _ = densified(th_td, st.collision_check_step)
        hit, clr = o.check_collision(states, check_self=True)
        return (not hit.any()), bool((np.abs(clr) < BAND).any())

    return cost, check


def blocks(flat, off, D):
    return [flat[off[b] * D: off[b + 1] * D].reshape(D, off[b + 1] - off[b]) for b in range(len(off) - 1)]


def test_stomp_iterations_match_oracle():
    lengths = [24, 17, 24, 31]
    w, panda, plan, ds, th, st = setup(lengths, 31, num_rollouts_new=10, num_rollouts_old=6, world_margin=0.05,
                                       iterations_after_valid=50)
    D = w.robot.n_dof
    off = np.concatenate([[0], np.cumsum(lengths)])
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    cost, check = oracle_fns(o, st)
    orc = [SO.StompOracle(x.astype(np.float64), w.robot.joint_lower, w.robot.joint_upper, cost,
                          lambda t: check(t)[0], Rn=10, Ro=6, dt=st.dt, w_u=st.control_cost_weight,
                          h=st.exponentiated_cost_sensitivity, seed=st.seed, b=b, polish=50)
           for b, x in enumerate(blocks(w.theta, off, D))]
    opt = S.StompOptimizer(plan, ds, th, st).step(0)
    for it in range(4):
        opt.step(1)
        for x in orc:
            x.iterate()
        got = blocks(opt.theta.cpu().numpy().astype(np.float64), off, D)
        errs = [np.abs(g - x.th).max() for g, x in zip(got, orc)]
        moved = [np.abs(x.th - x0.astype(np.float64)).max() for x, x0 in zip(orc, blocks(w.theta, off, D))]
        banner(f"STOMP iteration {it + 1}: max |theta - oracle| {max(errs):.2e} (optimum moved up to "
               f"{max(moved):.3f} rad)")
        assert max(errs) < STEP_TOL
    rec = opt.record.cpu().numpy()
    for b, x in enumerate(orc):
        assert int(rec[b, _ppl.CHOMP_ITERS]) == x.iters
        rel = abs(rec[b, _ppl.CHOMP_LAST_COST] - x.total) / max(abs(x.total), 1.0)
        assert rel < 1e-4, (b, rec[b, _ppl.CHOMP_LAST_COST], x.total)


def test_stomp_loop_bookkeeping():
    lengths = [32] * 6 + [20, 40, 28, 32, 25, 36]
    w, panda, plan, ds, th, st = setup(lengths, 32, world_margin=0.05)
    D = w.robot.n_dof
    off = np.concatenate([[0], np.cumsum(lengths)])
    opt = S.StompOptimizer(plan, ds, th, st)
    opt.step(0)
    c0 = opt.record[:, _ppl.CHOMP_LAST_COST].clone()
    opt.step(10)
    rec = opt.record.cpu().numpy()
    banner(f"STOMP: solved {int(rec[:, _ppl.CHOMP_SOLVED].sum())}/{len(lengths)}, done "
           f"{int(rec[:, _ppl.CHOMP_DONE].sum())}, cost {c0.mean().item():.3f} -> "
           f"{rec[:, _ppl.CHOMP_LAST_COST].mean():.3f}")
    assert np.all(np.isfinite(rec[:, _ppl.CHOMP_LAST_COST]))
    assert np.all(rec[:, _ppl.CHOMP_ITERS] <= 10)
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    cost, check = oracle_fns(o, st)
    best = blocks(opt.best.cpu().numpy(), off, D)
    cur = blocks(opt.theta.cpu().numpy(), off, D)
    R_ = {T: SO.control_R(T, st.dt) for T in set(lengths)}
    for b, T in enumerate(lengths):
        # endpoints never move
        a0 = w.theta[off[b] * D: off[b + 1] * D].reshape(D, T)
        np.testing.assert_array_equal(cur[b][:, [0, -1]], a0[:, [0, -1]])
        if rec[b, _ppl.CHOMP_SOLVED] == 0:
            continue
        # a reported solution is collision-free and its recorded cost is its cost
        x = best[b].astype(np.float64)
        ok, amb = check(x.T)
        assert ok or amb
        total = cost(x.T).sum() + st.control_cost_weight * np.sum(x * (x @ R_[T].T))
        rel = abs(rec[b, _ppl.CHOMP_BEST_COST] - total) / max(abs(total), 1.0)
        assert rel < 1e-5, (b, rec[b, _ppl.CHOMP_BEST_COST], total)


def test_stomp_solve_scene():
    """stomp_solve_scene (planning/stomp.cu:565-930 through the pybind
    surface): the batched optimiser's run from the same straight line, the
    reported costs those of the returned path, reusable buffers."""
    from prismplan_amd import prrtc
    w, panda, plan, ds, th, _ = setup([32], 33)
    D, T = w.robot.n_dof, 32
    x0 = w.theta.reshape(D, T)
    start, goal = x0[:, 0].copy(), x0[:, -1].copy()
    robot = prrtc_panda()
    scene = prrtc.scene_builder(ds)
    ps = prrtc.STOMPSettings()
    ps.num_timesteps, ps.num_iterations, ps.time_limit_ms = T, 0, 0.0
    ps.world_collision_margin = 0.05
    line = prrtc.stomp_solve_scene(start, goal, scene, ps, robot)
    assert line.iterations_run == 0 and line.path.shape == (T, D)
    ps.num_iterations = 12
    res = prrtc.stomp_solve_scene(start, goal, scene, ps, robot)
    banner(f"stomp_solve_scene: solved {res.solved}, iterations {res.iterations_run}, cost "
           f"{res.final_state_cost:.4f} + {res.final_control_cost:.4f} = {res.final_total_cost:.4f}, "
           f"{res.wall_time_ms:.2f} ms")
    np.testing.assert_array_equal(res.path[0], start)
    np.testing.assert_array_equal(res.path[-1], goal)
    assert 0 < res.iterations_run <= 12
    st = S.stomp_settings(world_margin=0.05)
    theta = torch.from_numpy(np.ascontiguousarray(line.path.T).reshape(-1)).to(DEV)
    opt = S.StompOptimizer(S.Plan.trajectories(panda, [T]), ds, theta, st).step(0).step(12)
    rec = opt.record.cpu().numpy()[0]
    assert bool(rec[_ppl.CHOMP_SOLVED]) == res.solved
    assert int(rec[_ppl.CHOMP_ITERS]) == res.iterations_run
    want = (opt.best if res.solved else opt.theta).cpu().numpy().reshape(D, T).T
    np.testing.assert_allclose(res.path[1:-1], want[1:-1], rtol=0, atol=1e-6)
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    cost, check = oracle_fns(o, st)
    x = res.path.astype(np.float64)
    state = cost(res.path).sum()
    ctrl = st.control_cost_weight * np.sum(x.T * (x.T @ SO.control_R(T, st.dt).T))
    for got, ref in ((res.final_state_cost, state), (res.final_control_cost, ctrl),
                     (res.final_total_cost, state + ctrl)):
        assert abs(got - ref) <= 1e-5 * max(abs(ref), 1.0), (got, ref)
    if res.solved:
        ok, amb = check(res.path)
        assert ok or amb
    bufs = prrtc.STOMPBuffers.create(D, T, ps.num_rollouts_new, ps.num_rollouts_old)
    assert bufs.owns_memory and bufs.num_rollouts_all == ps.num_rollouts_new + ps.num_rollouts_old + 1
    for _ in range(2):
        again = prrtc.stomp_solve_scene(start, goal, scene, ps, robot, bufs)
        np.testing.assert_array_equal(again.path, res.path)
        assert again.iterations_run == res.iterations_run and again.final_total_cost == res.final_total_cost
    bufs.destroy()
    assert not bufs.owns_memory
    ps.time_limit_ms = 1e-3
    assert prrtc.stomp_solve_scene(start, goal, scene, ps, robot).iterations_run <= 1
    ps.num_batch = 2
    with pytest.raises(RuntimeError):
        prrtc.stomp_solve_scene(start, goal, scene, ps, robot)


@pytest.mark.parametrize("robot,n_old", [("fetch", 6), ("baxter", 0)])
def test_stomp_other_robots_and_edge_cases(robot, n_old):
    """Fetch (prismatic torso, tree) and Baxter (two arms, 14 dof); no kept
    rollouts; a two-waypoint trajectory (no interior: it never moves)."""
    lengths = [20, 2, 33]
    w, rob, plan, ds, th, st = setup(lengths, 35, robot=robot, num_rollouts_new=9, num_rollouts_old=n_old,
                                     world_margin=0.05, iterations_after_valid=50)
    D = w.robot.n_dof
    off = np.concatenate([[0], np.cumsum(lengths)])
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    cost, check = oracle_fns(o, st)
    orc = [SO.StompOracle(x.astype(np.float64), w.robot.joint_lower, w.robot.joint_upper, cost,
                          lambda t: check(t)[0], Rn=9, Ro=n_old, dt=st.dt, w_u=st.control_cost_weight,
                          h=st.exponentiated_cost_sensitivity, seed=st.seed, b=b, polish=50)
           for b, x in enumerate(blocks(w.theta, off, D))]
    opt = S.StompOptimizer(plan, ds, th, st).step(0)
    for it in range(3):
        opt.step(1)
        for x in orc:
            x.iterate()
        got = blocks(opt.theta.cpu().numpy().astype(np.float64), off, D)
        err = max(np.abs(g - x.th).max() for g, x in zip(got, orc))
        banner(f"STOMP {robot} (old {n_old}) iteration {it + 1}: max |theta - oracle| {err:.2e}")
        assert err < STEP_TOL
    np.testing.assert_array_equal(got[1], w.theta[off[1] * D: off[2] * D].reshape(D, 2))


@pytest.mark.parametrize("polish,eps", [(3, 1e-3), (50, 1e9)])
def test_stomp_solved_bookkeeping_matches_oracle(polish, eps):
    """A world with nothing to hit: trajectories are valid from the start, so
    best snapshots, the convergence streak, the polish count and the done
    flag all act -- done by polish (3 iterations) or, with polish to spare
    and a huge eps, by a streak of two flat iterations; every record field
    follows the oracle's loop (stomp.cu:838-887)."""
    lengths = [24, 30, 24, 18, 27, 24]
    w = W.chomp(n_traj=len(lengths), n_obbs=0, seed=37, lengths=lengths, amp=0.05)
    rob = S.Robot("panda")
    plan = S.Plan.trajectories(rob, w.lengths())
    ds = S.DeviceScene.from_host(w.scene)
    th = torch.from_numpy(w.theta.copy()).to(DEV)
    st = S.stomp_settings(num_rollouts_new=8, num_rollouts_old=4, iterations_after_valid=polish, convergence_eps=eps)
    D = w.robot.n_dof
    off = np.concatenate([[0], np.cumsum(lengths)])
    o = PO.Oracle(w.robot, w.scene, nthreads=8)
    cost, check = oracle_fns(o, st)
    orc = [SO.StompOracle(x.astype(np.float64), w.robot.joint_lower, w.robot.joint_upper, cost,
                          lambda t: check(t)[0], Rn=8, Ro=4, dt=st.dt, w_u=st.control_cost_weight,
                          h=st.exponentiated_cost_sensitivity, seed=st.seed, b=b, polish=polish, eps=eps)
           for b, x in enumerate(blocks(w.theta, off, D))]
    opt = S.StompOptimizer(plan, ds, th, st).step(0)
    for _ in range(8):
        opt.step(1)
        for x in orc:
            x.iterate()
    rec = opt.record.cpu().numpy()
    best = blocks(opt.best.cpu().numpy().astype(np.float64), off, D)
    n_solved = 0
    for b, x in enumerate(orc):
        solved = x.best is not None
        n_solved += solved
        assert bool(rec[b, _ppl.CHOMP_SOLVED]) == solved, b
        assert int(rec[b, _ppl.CHOMP_ITERS]) == x.iters, (b, rec[b], x.iters)
        assert bool(rec[b, _ppl.CHOMP_DONE]) == x.done, b
        if solved:
            assert abs(rec[b, _ppl.CHOMP_BEST_COST] - x.best_cost) <= 1e-4 * max(abs(x.best_cost), 1.0)
            assert np.abs(best[b] - x.best).max() < STEP_TOL
    banner(f"STOMP solved bookkeeping: {n_solved}/{len(lengths)} solved, done {int(rec[:, _ppl.CHOMP_DONE].sum())}, "
           f"iterations {rec[:, _ppl.CHOMP_ITERS].astype(int).tolist()}")
    assert n_solved >= 4 and rec[:, _ppl.CHOMP_DONE].sum() >= 4
