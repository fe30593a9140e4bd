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

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
BAND = 1e-5


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def densified(th_td: np.ndarray, step: float):
    """FP32 densification exactly as the validity pass (chomp.cu:459-475, 297-351)."""
    th = th_td.astype(np.float32)
    T, D = th.shape
    d2 = np.zeros(T - 1, np.float32)
    for d in range(D):
        diff = (th[1:, d] - th[:-1, d]).astype(np.float32)
        d2 = (d2 + (diff * diff).astype(np.float32)).astype(np.float32)
    m = np.float32(np.sqrt(d2).max()) if T > 1 else np.float32(0)
    n_sub = int(np.ceil(np.float32(m / np.float32(step)))) if (step > 0 and m > np.float32(step)) else 1
    out = []
    for seg in range(T - 1):
        a, b = th[seg], th[seg + 1]
        for sub in range(n_sub):
            frac = np.float32(np.float32(sub) / np.float32(n_sub))
            out.append((a + ((b - a).astype(np.float32) * frac).astype(np.float32)).astype(np.float32))
    out.append(th[T - 1])
    return np.stack(out), n_sub


def setup(n_traj, n_wp, seed, **kw):
    w = W.chomp(n_traj=n_traj, n_wp=n_wp, n_obbs=30, seed=seed)
    panda = S.Robot("panda")
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
