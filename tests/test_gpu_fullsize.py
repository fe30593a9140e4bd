"""Parity at the benchmark workloads' full sizes (needs an MI355X).

  * edges: Panda, 64 boxes, 16,384 edges of 8-256 states (1,002,744 states):
    every state's verdict and every edge's first hit;
  * CHOMP: Panda, 40 boxes, 256 trajectories x 64 waypoints: per-trajectory
    cost, the fused adjoint's gradient and the central-FD gradient kernel;
  * STOMP: Fetch, 128 problems x 32 rollouts of 32-128 waypoints, 30-80
    boxes per problem, self pairs on: per-waypoint state cost and
    per-rollout totals, unit-major and waypoint-major plans.

Bars: verdicts exact outside |clearance| < 1e-5 m (counted); costs 1e-5
relative with the denominator cost_den of parity_common (its FP32 error
model: 0.05 m per non-zero hinge term); gradients 1e-4 relative, norm-wise
per waypoint, away from oracle-flagged kinks (counted). The FD kernel's bar
follows from the cost bar of its two evaluations (parity_common.fd_bar).
"""
import numpy as np
import pytest
import torch

import pyoracle as PO
from parity_common import banner, cost_den, fd_bar, traj_rows
from prismplan_amd import substrate as S
from prismplan_amd import workloads as W

pytestmark = pytest.mark.gpu
BAND = 1e-5
DEV = "cuda"
THREADS = 16


@pytest.fixture(scope="module")
def robots():
    return {n: S.Robot(n) for n in ("panda", "fetch")}


def test_edges_at_full_size(robots):
    e = W.edges()
    banner(f"edges full size: E={len(e.n_states)} states={e.total_states} O={e.scene.n_obbs}")
    ds = S.DeviceScene.from_host(e.scene)
    plan = S.Plan.edges(robots["panda"], e.n_states)
    frm, to = torch.from_numpy(e.frm).to(DEV), torch.from_numpy(e.to).to(DEV)
    fh_all, hits = S.edge_check_batched(plan, ds, frm, to, skip_resolved=False, state_hits=True)
    fh_skip = S.edge_check_batched(plan, ds, frm, to, skip_resolved=True).cpu().numpy()
    fh_all, hits = fh_all.cpu().numpy(), hits.cpu().numpy().astype(bool)
    np.testing.assert_array_equal(fh_skip, fh_all)
    o = PO.Oracle(e.robot, e.scene, nthreads=THREADS)
    states = PO.edge_states(e.frm, e.to, e.n_states)
    ref, clr = o.check_collision(states)
    amb = np.abs(clr) < BAND
    bad = (hits != ref) & ~amb
    banner(f"  state verdicts: colliding {ref.mean():.3f}, ambiguous {int(amb.sum())}, "
           f"in-band mismatches {int(((hits != ref) & amb).sum())}, others {int(bad.sum())}")
    assert not bad.any()
    off = e.offsets
    edge_of = np.repeat(np.arange(len(e.n_states)), e.n_states)
    hit_idx = np.nonzero(ref)[0]
    # first colliding state of each edge: the smallest state index per edge among hits
    first_ref = np.full(len(e.n_states), np.iinfo(np.int64).max, np.int64)
    np.minimum.at(first_ref, edge_of[hit_idx], hit_idx - off[edge_of[hit_idx]])
    first_ref[first_ref == np.iinfo(np.int64).max] = -1
    diff = np.nonzero(fh_all != first_ref)[0]
    unexplained = []
    for k in diff:
        lo = min(x for x in (fh_all[k], first_ref[k]) if x >= 0)
        if not np.any(amb[off[k] + lo:off[k] + max(fh_all[k], first_ref[k], lo) + 1]):
            unexplained.append(int(k))
    banner(f"  first hits: colliding edges {np.mean(first_ref >= 0):.3f}, in-band differences "
           f"{len(diff) - len(unexplained)}, unexplained {unexplained[:10]}")
    assert not unexplained


def test_chomp_at_full_size(robots):
    w = W.chomp()
    D = w.robot.n_dof
    banner(f"CHOMP full size: {w.n_traj} trajectories x {w.n_waypoints // w.n_traj} waypoints, "
           f"{w.scene.n_obbs} boxes")
    ds = S.DeviceScene.from_host(w.scene)
    plan = S.Plan.trajectories(robots["panda"], w.lengths())
    th = torch.from_numpy(w.theta).to(DEV)
    m, wt = 0.05, 1.0
    prm = S.cost_params(world_margin=m, weight=wt, use_self=False)
    prm_o = PO.cost_params(world_margin=m, weight=wt, use_self=False)
    st, total, g = S.hinge_cost(plan, ds, th, prm, grad=True)
    o = PO.Oracle(w.robot, w.scene, nthreads=THREADS)
    tot_ref, g_ref, kink = o.traj_cost_grad(w.theta, w.offsets, prm_o)
    rows = traj_rows(w.theta, w.offsets, D)
    st_ref, _, nt = o.hinge_cost_amb(rows, prm_o, with_terms=True)
    st_err = np.abs(st.cpu().numpy() - st_ref) / cost_den(st_ref, wt, nt)
    rel = np.abs(total.cpu().numpy() - tot_ref) / cost_den(tot_ref, wt, np.add.reduceat(nt, w.offsets[:-1]))
    banner(f"  state cost max rel err {st_err.max():.2e}, trajectory total {rel.max():.2e} "
           f"(active waypoints {np.mean(st_ref > 0):.2f})")
    assert st_err.max() < 1e-5 and rel.max() < 1e-5
    gg = traj_rows(g.cpu().numpy(), w.offsets, D)
    gr = traj_rows(g_ref, w.offsets, D)
    interior = np.ones(len(gr), bool)
    interior[w.offsets[:-1]] = False
    interior[w.offsets[1:] - 1] = False
    ok = interior & ~kink
    num = np.abs(gg - gr).max(axis=1)
    e = num / np.maximum(np.abs(gr).max(axis=1), 1e-3 * wt)
    banner(f"  adjoint gradient max rel err {e[ok].max():.2e} (norm-wise), kink waypoints {int(kink.sum())} "
           f"of {len(gr)}")
    assert e[ok].max() < 1e-4
    assert np.all(gg[~interior] == 0)
    # FD kernel vs the FP64 FD, bar from the cost bar of its two evaluations
    eps = np.float32(1e-3)
    gfd = traj_rows(S.chomp_fd_grad(plan, ds, th, prm, eps=float(eps)).cpu().numpy(), w.offsets, D)
    gfd_ref = traj_rows(o.chomp_fd_grad(w.theta, w.offsets, prm_o, eps=float(eps)), w.offsets, D)
    tol = fd_bar(o, rows, PO.cost_params(world_margin=m, weight=1.0, use_self=False), wt, eps)
    diff = np.abs(gfd - gfd_ref)
    bad = (diff > tol) & interior[:, None]
    banner(f"  FD kernel: max |diff| {diff.max():.2e}, max diff / bar {np.max(diff / tol):.2f}, over bar "
           f"{int(bad.sum())}")
    assert not bad.any()


@pytest.mark.parametrize("wmajor", [False, True])
def test_stomp_at_full_size(robots, wmajor):
    wk = W.stomp()
    R = wk.n_rollouts
    banner(f"STOMP full size: {wk.n_problems} problems x {R} rollouts, {wk.n_states} rollout states, "
           f"waypoint-major {wmajor}")
    plan = S.Plan.trajectories(robots["fetch"], np.repeat(wk.lengths, R),
                               problem_units=np.arange(wk.n_problems + 1) * R, obb_offsets=wk.obb_offsets,
                               waypoint_major=wmajor)
    ds = S.DeviceScene.from_host(wk.scene)
    ro = torch.from_numpy(wk.rollouts).to(DEV)
    wc, m = 20.0, 0.02
    prm = S.cost_params(world_margin=m, self_margin=0.01, weight=wc)
    state, total = S.stomp_rollout_costs(plan, ds, ro, prm, control_weight=0.1, dt=0.1)
    o = PO.Oracle(wk.robot, nthreads=THREADS)
    s_ref, t_ref, nt = o.stomp_costs(wk.scene, wk.obb_offsets, wk.rollouts, wk.lengths, wk.roll_offsets, R,
                                     PO.cost_params(world_margin=m, self_margin=0.01, weight=wc), 0.1, 0.1,
                                     with_terms=True)
    srel = np.abs(state.cpu().numpy() - s_ref) / cost_den(s_ref, wc, nt)
    trel = np.abs(total.cpu().numpy() - t_ref) / np.abs(t_ref)
    banner(f"  state cost max rel err {srel.max():.2e} (active {np.mean(s_ref > 0):.2f}), rollout total "
           f"{trel.max():.2e}")
    assert srel.max() < 1e-5 and trel.max() < 1e-5
