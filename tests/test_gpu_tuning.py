"""Launch-shape autotuners and their tuning cache (needs an MI355X).

The any-collision kernel's shape (OBB records per LDS staging, grid form)
and the cost kernel's layout are chosen by timing on the caller's inputs,
kept for the plan and recorded in a cache file keyed by the robot's tables,
the largest problem's OBB count (cost: and the state count's power of two)
and the GPU architecture, so that later plans and processes take the choice
without timing. The fixed tile these replace is the reference's BATCH_SIZE
(collision/scene_collision.cuh:9-11). Every shape must give the same
results: first hits and verdicts bit-identical, costs to FP32 summation
order.
"""
import numpy as np
import pytest
import torch

from parity_common import banner
from prismplan_amd import _ppl
from prismplan_amd import substrate as S
from prismplan_amd import workloads as W

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture
def cache_file(tmp_path):
    path = tmp_path / "tune.txt"
    S.tune_cache_file(path)
    yield path
    S.tune_cache_file("")


def entries(path):
    lines = [l.split() for l in path.read_text().splitlines() if l and not l.startswith("#")]
    return {l[0]: l[1:] for l in lines}


def test_collide_autotune_edges_and_cache(cache_file):
    panda = S.Robot("panda")
    e = W.edges(n_edges=3000, n_obbs=64, seed=1)
    ds = S.DeviceScene.from_host(e.scene)
    frm, to = torch.from_numpy(e.frm).to(DEV), torch.from_numpy(e.to).to(DEV)
    plan = S.Plan.edges(panda, e.n_states)
    assert plan.collide_shape(64) == (64, 0, 0)          # built-in: one staging, persistent grid
    fh0, hits0 = S.edge_check_batched(plan, ds, frm, to, skip_resolved=False, state_hits=True)
    tile, grid, ms = plan.autotune_collide(ds, frm, to, reps=3)
    banner(f"edges: tuned tile {tile}, grid form {grid}, {ms * 1e3:.1f} us per call")
    assert tile in (64, 32) and grid in (0, 1, 2) and ms > 0
    assert plan.collide_shape(64) == (tile, grid, 2)
    ent = entries(cache_file)
    keys = [k for k in ent if k.startswith("collide/") and k.endswith("/64/gfx950")]
    assert len(keys) == 1 and int(ent[keys[0]][0]) in (0, 32) and int(ent[keys[0]][1]) == grid
    # every shape gives the same verdicts and first hits
    for t, g in ((32, 0), (16, 1), (8, 2), (64, 1), (0, 0)):
        plan.set_collide_shape(t, g)
        fh, hits = S.edge_check_batched(plan, ds, frm, to, skip_resolved=False, state_hits=True)
        torch.testing.assert_close(fh, fh0, rtol=0, atol=0)
        torch.testing.assert_close(hits, hits0, rtol=0, atol=0)
        fs = S.edge_check_batched(plan, ds, frm, to, skip_resolved=True)
        torch.testing.assert_close(fs, fh0, rtol=0, atol=0)
    # a new plan of the same robot and OBB count takes the cached shape, also
    # after the file is read afresh (a new process); other robots do not
    plan2 = S.Plan.edges(panda, e.n_states[:500])
    assert plan2.collide_shape(64) == (tile, grid, 1)
    # a later autotuning call adopts the cached choice (and its recorded time) without timing
    t2, g2, ms2 = plan2.autotune_collide(ds, frm[:500], to[:500], reps=3)
    assert (t2, g2) == (tile, grid) and abs(ms2 - ms) < 1e-6 * max(1.0, ms)
    S.tune_cache_file(cache_file)
    plan3 = S.Plan.configs(panda, 100)
    assert plan3.collide_shape(64) == (tile, grid, 1)
    assert plan3.collide_shape(63)[2] == 0
    assert S.Plan.configs(S.Robot("fetch"), 100).collide_shape(64)[2] == 0


def test_collide_autotune_configs_many_obbs(cache_file):
    panda = S.Robot("panda")
    rng = np.random.default_rng(5)
    scene = W.random_obbs(rng, 300, lo=(0.3, -1.2, 0.0), hi=(1.5, 1.2, 1.5), half_lo=0.01, half_hi=0.03)
    q = torch.from_numpy(W.uniform_configs(rng, panda.arrays, 20000)).to(DEV)
    ds = S.DeviceScene.from_host(scene)
    plan = S.Plan.configs(panda, q.shape[0])
    ref = S.check_collision_batched(plan, ds, q)
    tile, grid, ms = plan.autotune_collide(ds, q, reps=3)
    banner(f"configs, 300 boxes: tuned tile {tile}, grid form {grid}, {ms * 1e3:.1f} us per call")
    assert tile in (256, 128, 64, 32)
    torch.testing.assert_close(S.check_collision_batched(plan, ds, q), ref, rtol=0, atol=0)
    assert any(k.endswith("/300/gfx950") for k in entries(cache_file))


def test_cost_autotune_is_cached(cache_file):
    panda = S.Robot("panda")
    w = W.chomp(n_traj=24, n_wp=40, n_obbs=30, seed=5)
    ds = S.DeviceScene.from_host(w.scene)
    th = torch.from_numpy(w.theta).to(DEV)
    prm = S.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    plan = S.Plan.trajectories(panda, w.lengths())
    G, ms = plan.autotune_cost(ds, th, prm, grad=True, reps=3)
    G2, ms2 = S.Plan.trajectories(panda, w.lengths()).autotune_cost(ds, th, prm, grad=True, reps=3)
    assert G2 == G and abs(ms2 - ms) < 1e-6   # adopted from the cache (time kept to 1 ns)
    ent = entries(cache_file)
    key = [k for k in ent if k.startswith("adjoint/")]
    assert len(key) == 1 and int(ent[key[0]][0]) == G
    # an untuned plan of the same shape uses the cached layout: same results
    # as the tuned plan, bit for bit
    plan2 = S.Plan.trajectories(panda, w.lengths())
    s1, t1, g1 = S.hinge_cost(plan, ds, th, prm, grad=True)
    s2, t2, g2 = S.hinge_cost(plan2, ds, th, prm, grad=True)
    torch.testing.assert_close(s2, s1, rtol=0, atol=0)
    torch.testing.assert_close(g2, g1, rtol=0, atol=0)


def test_kernel_timing_events():
    """ppl_plan_time_kernels: one event pair per timed launch of the main
    kernel, none when off or under graph capture, and the kernel alone takes
    less than the whole call (which also resets the first hits)."""
    panda = S.Robot("panda")
    e = W.edges(n_edges=2048, seed=3)
    ds = S.DeviceScene.from_host(e.scene)
    plan = S.Plan.edges(panda, e.n_states)
    frm, to = torch.from_numpy(e.frm).to(DEV), torch.from_numpy(e.to).to(DEV)
    fh = torch.empty(len(e.n_states), dtype=torch.int32, device=DEV)

    def call():
        S.edge_check_batched(plan, ds, frm, to, skip_resolved=False, first_hit=fh)

    call()
    torch.cuda.synchronize()
    assert plan.kernel_ms() == (0.0, 0)
    plan.time_kernels(True)
    for _ in range(5):
        call()
    ms, n = plan.kernel_ms()
    assert n == 5 and ms > 0.0
    assert plan.kernel_ms() == (0.0, 0)   # read clears
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(5):
        call()
    t1.record()
    torch.cuda.synchronize()
    ms, n = plan.kernel_ms()
    banner(f"k_collide {ms / n * 1e3:.1f} us per launch; whole call {t0.elapsed_time(t1) / 5 * 1e3:.1f} us")
    assert n == 5 and ms < t0.elapsed_time(t1)
    # cost kernel of a trajectory plan; graph capture records nothing
    w = W.chomp(n_traj=16, n_wp=32, n_obbs=20, seed=4)
    dsc = S.DeviceScene.from_host(w.scene)
    tp = S.Plan.trajectories(panda, w.lengths())
    th = torch.from_numpy(w.theta).to(DEV)
    prm = S.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    S.hinge_cost(tp, dsc, th, prm, grad=True)
    tp.time_kernels(True)
    S.hinge_cost(tp, dsc, th, prm, grad=True)
    torch.cuda.synchronize()
    state = torch.empty(tp.n_states, device=DEV)
    total = torch.empty(len(w.lengths()), device=DEV)
    grad = torch.empty_like(th)
    g = S.capture_hinge(tp, dsc, th, prm, state, total, grad)   # one eager warm-up call, then the capture
    torch.cuda.synchronize()
    assert tp.kernel_ms()[1] == 2
    g.launch()
    torch.cuda.synchronize()
    assert tp.kernel_ms() == (0.0, 0)
    tp.time_kernels(False)
    S.hinge_cost(tp, dsc, th, prm, grad=True)
    assert tp.kernel_ms() == (0.0, 0)
