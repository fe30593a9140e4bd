"""Optimiser iterations replayed as hipGraphs (ppl_plan_set_optimizer_graphs)
against the same iterations launched eagerly (needs an MI355X).

A graph replays the launches it captured, so theta, the best snapshot and
the per-trajectory records must equal the eager run's bit for bit, over
enough iterations that several replays run (CHOMP: one iteration per
replay; STOMP: two per replay, after the population reaches its steady
shape), across calls that continue a run, and after a setting changes
(which must re-capture)."""
import numpy as np
import pytest
import torch

from prismplan_amd import substrate as S
from prismplan_amd import workloads as W

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def run(kind, graphs, steps, settings_seq):
    w = W.chomp(n_traj=12, n_obbs=30, seed=71, lengths=[24, 40, 51, 33, 51, 20, 45, 51, 38, 27, 51, 30])
    robot = S.Robot("panda")
    plan = S.Plan.trajectories(robot, w.lengths())
    plan.set_optimizer_graphs(graphs)
    ds = S.DeviceScene.from_host(w.scene)
    th = torch.from_numpy(w.theta.copy()).to(DEV)
    opt = (S.ChompOptimizer if kind == "chomp" else S.StompOptimizer)(plan, ds, th, settings_seq[0])
    opt.step(0)
    for n, st in zip(steps, settings_seq):
        opt.settings = st
        opt.step(n)
    torch.cuda.synchronize()
    return opt.theta.cpu().numpy(), opt.best.cpu().numpy(), opt.record.cpu().numpy()


@pytest.mark.parametrize("kind", ["chomp", "stomp"])
def test_graph_iterations_equal_eager(kind):
    if kind == "chomp":
        seq = [S.chomp_settings(world_margin=0.05), S.chomp_settings(world_margin=0.05),
               S.chomp_settings(world_margin=0.05, step_size=0.5)]
    else:
        seq = [S.stomp_settings(world_margin=0.05, num_rollouts_new=12, num_rollouts_old=6)] * 3
    steps = [7, 6, 5]
    banner(f"{kind}: {sum(steps)} iterations in calls of {steps}, graphs vs eager")
    eager = run(kind, False, steps, seq)
    graph = run(kind, True, steps, seq)
    for name, e, g in zip(("theta", "best", "record"), eager, graph):
        assert np.array_equal(e, g, equal_nan=True), f"{kind} {name}: max |diff| {np.nanmax(np.abs(e - g))}"
    assert eager[2][:, 6].max() > 0   # iterations ran
