#!/usr/bin/env python3
"""Substrate benchmark on one MI355X. Prints ONE JSON line on stdout.

Headline step: the batched edge check of the edges workload (Panda 7-DoF,
64 random OBBs, 16,384 straight edges of 8..256 states at <= 0.02 rad joint
step): one persistent launch that returns every edge's any-collision verdict
and first-hit index, with every state's verdict computed (no skipping past a
known hit). value = states checked per second.

Also measured (extra fields): first-hit early-exit mode, plain-vs-batched
any-collision (64 requests of 4,096 configs), plain-vs-batched edges, the
fused CHOMP adjoint vs finite-difference gradients, STOMP rollout costs
through a hipGraph, the STOMP / CHOMP optimisers, MIT*'s collision calls,
pRRTC solves on the MotionBenchMaker sample problems, and the Baxter ESDF
adjoint; the FP32 CPU restatement
timed on this host's cores; parity of the timed outputs vs the FP64 oracle.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-extras] [--legs a,b,...] [--no-cpu]

--legs picks the extra measurements (skip, any, edges_plain, edges_free, chomp,
stomp, chomp_opt, stomp_opt, mitstar, rrtc, esdf, copy, cpu_table); the headline always runs. --no-cpu drops the
CPU baseline legs (for profiler runs, whose JSON line then carries cpu_baseline null). cpu_table times the FP32
CPU restatement of the other workloads and checks their GPU outputs against FP64.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
import traceback
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

METRIC = "collision-checked configs/s + CHOMP cost+grad traj/s, batched vs plain, 1 GPU"
PEAK_FP32 = 157.3e12          # MI355X FP32 peak (packed issue)
PEAK_FP32_UNPACKED = 78.6e12  # without packed FP32 (this build disables packed FP32 ops)
# Mean gated early-exit flops per edges-workload state, as the CPU leg counts
# them on its 20,000-state sample; used when that leg is switched off.
GATED_FLOPS_EDGES = 3191.0
FLOPS_RULE = ("achieved = gated early-exit flops per state x states / kernel time; the gated count is the FP64 "
              "oracle's mean over a 20,000-state sample of the work its two-phase check does (FK, approximate "
              "spheres, exact spheres of flagged joints against every OBB; include/ppl/roofline.h item costs)")
# The headline kernel's counters (rocprofv3 --pmc, median per dispatch over
# the timed loop; profiles/edges_collide_counters.txt): VALU instructions and
# SQ_INSTS_VALU_FLOPS_FP32 per dispatch, with the kernel time of that run.
COUNTER_VALU_INSTS_EDGES = 15_633_400
COUNTER_VALU_FLOPS_FP32_EDGES = 7_971_864
COUNTER_KERNEL_US_EDGES = 40.67
CUS, SIMDS_PER_CU, CLOCK_HZ = 256, 4, 2.4e9
# HBM bytes one headline launch moves (k_collide), from rocprofv3 FETCH_SIZE and
# WRITE_SIZE scaled by a 256 MiB copy (same profile); the algorithmic bytes are
# 16,384 edges x 69 B.
TRAFFIC_BYTES_EDGES = 2_189_004
ALGORITHMIC_BYTES_EDGES = 16384 * 69


def kernel_time(plan, fn, steps, torch):
    """Mean duration (s) of the plan's main kernel over `steps` calls of fn,
    from the HIP event pairs the runtime records at each launch's start and
    end (ppl_plan_time_kernels)."""
    torch.cuda.synchronize()
    plan.kernel_ms()
    plan.time_kernels(True)
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    ms, n = plan.kernel_ms()
    plan.time_kernels(False)
    assert n == steps, f"{n} timed launches for {steps} calls"
    return ms / n * 1e-3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_loop(fn, steps, warmup, torch):
    """Wall time per step over exactly `steps` back-to-back steps, synchronized
    on both sides, and the HIP-event time per step of the same loop (one event
    pair around all of it: events recorded around every step would add their
    own GPU work to each step)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record()
    for _ in range(steps):
        fn()
    b.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    ev = a.elapsed_time(b) / steps / 1e3
    return wall, ev


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name"):
                return ln.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--legs", default="skip,any,edges_plain,edges_free,chomp,stomp,chomp_opt,stomp_opt,mitstar,rrtc,esdf,"
                                      "cpu_table")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    legs = set(x for x in args.legs.split(",") if x)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the substrate shards by "
              f"independent requests: each GPU would take whole problems, edges or trajectories).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from prismplan_amd import substrate as S
    from prismplan_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    props = torch.cuda.get_device_properties(dev)
    log(f"device {props.name} CUs={props.multi_processor_count}")

    # ------------------------------------------------------------------ headline
    panda = S.Robot("panda")
    e = W.edges()
    ds = S.DeviceScene.from_host(e.scene)
    plan = S.Plan.edges(panda, e.n_states)
    frm, to = torch.from_numpy(e.frm).to(dev), torch.from_numpy(e.to).to(dev)
    fh = torch.empty(len(e.n_states), dtype=torch.int32, device=dev)
    total_states = e.total_states
    log(f"edges: E={len(e.n_states)} states={total_states} sha256={e.sha256()[:16]}")
    # launch shape (OBB staging tile x grid form) from the autotuner, before the timed region
    c_tile, c_grid, c_ms = plan.autotune_collide(ds, frm, to, reps=5)
    log(f"collide shape: tile {c_tile}, grid form {c_grid} ({c_ms * 1e3:.1f} us per first-hit call)")

    def step():
        S.edge_check_batched(plan, ds, frm, to, skip_resolved=False, first_hit=fh)

    wall, ev_step = timed_loop(step, args.steps, args.warmup, torch)
    value = total_states / wall
    # the roofline's kernel time: k_collide alone (the step also resets the first hits)
    ev = kernel_time(plan, step, args.steps, torch)
    log(f"headline: {wall * 1e3:.3f} ms/step wall, {ev_step * 1e3:.3f} ms step event, {ev * 1e3:.4f} ms k_collide "
        f"-> {value:.3e} states/s")
    fh_host = fh.cpu().numpy()
    hit_edges = fh_host >= 0
    # the workload's 64 boxes of up to 0.6 m fill the Panda's reach: edges collide early
    workload_stats = {"colliding_edge_frac": float(hit_edges.mean()),
                      "first_hit_state_mean": float(fh_host[hit_edges].mean()) if hit_edges.any() else None}

    # ------------------------------------------------------------------ CPU baseline leg
    if args.no_cpu:
        cpu_baseline, parity, f_state = None, None, GATED_FLOPS_EDGES
    else:
        cpu_baseline, parity, f_state = cpu_baseline_leg(e, fh_host, args)
    achieved = f_state * total_states / ev
    from prismplan_amd import _ppl
    f_exh = panda.flops(_ppl.WORK_ANY_COLLISION, n_obbs=e.scene.n_obbs)
    roofline = {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32, "traffic": TRAFFIC_BYTES_EDGES,
                "traffic_note": "HBM bytes per launch (counters, copy-scaled); algorithmic bytes "
                                f"{ALGORITHMIC_BYTES_EDGES}; achieved HBM GB/s "
                                f"{TRAFFIC_BYTES_EDGES / ev / 1e9:.1f} of 8000",
                "frac_of_unpacked_peak": achieved / PEAK_FP32_UNPACKED,
                "flops_per_state": f_state, "flops_rule": FLOPS_RULE,
                "exhaustive_flops_per_state": f_exh,
                "exhaustive_frac": f_exh * total_states / ev / PEAK_FP32, "kernel_ms": ev * 1e3,
                "kernel_ms_source": "HIP events the runtime records at each k_collide launch's start and end "
                                    "(hipExtLaunchKernel, ppl_plan_time_kernels)",
                "step_event_ms": ev_step * 1e3,
                # counter-derived, from the committed profile of this kernel: lane flops if every lane of
                # each FP32 VALU instruction were active (an upper bound), and VALU issue occupancy
                # (wave64 VALU instructions x 4 cycles over every SIMD's cycles at 2.4 GHz)
                "counter_fp32_tflops_upper": 64 * COUNTER_VALU_FLOPS_FP32_EDGES / (COUNTER_KERNEL_US_EDGES * 1e-6) / 1e12,
                "counter_valu_busy": COUNTER_VALU_INSTS_EDGES * 4 / (CUS * SIMDS_PER_CU * COUNTER_KERNEL_US_EDGES * 1e-6
                                                                     * CLOCK_HZ)}

    extras = {}
    if not args.no_extras:
        for key, name, fn in (
                ("skip", "first_hit_skip", lambda: bench_skip(plan, ds, frm, to, fh, total_states, S, torch, args)),
                ("any", "any_collision", lambda: bench_any(panda, S, W, torch, dev, args)),
                ("edges_plain", "edges_plain_vs_batched", lambda: bench_edges_plain(panda, e, ds, S, torch, dev, args)),
                ("edges_free", "edges_mostly_free_world", lambda: bench_edges_free(panda, S, W, torch, dev, args)),
                ("chomp", "chomp", lambda: bench_chomp(panda, S, W, torch, dev, args)),
                ("stomp", "stomp", lambda: bench_stomp(S, W, torch, dev, args)),
                ("chomp_opt", "chomp_optimizer", lambda: bench_chomp_opt(panda, S, W, torch, dev, args)),
                ("stomp_opt", "stomp_optimizer", lambda: bench_stomp_opt(panda, S, W, torch, dev, args)),
                ("mitstar", "mitstar", lambda: bench_mitstar(panda, S, W, torch, dev, args)),
                ("rrtc", "prrtc", lambda: bench_rrtc(panda, S, W, torch, dev, args)),
                ("esdf", "esdf", lambda: bench_esdf(S, W, torch, dev, args)),
                ("copy", "copy", lambda: bench_copy(torch, dev, args)),
                ("cpu_table", "cpu_table", lambda: cpu_table_leg(S, W, torch, dev, args))):
            if key not in legs or (key == "cpu_table" and args.no_cpu):
                continue
            try:
                extras[name] = fn()
                log(f"{name}: {json.dumps(extras[name])}")
            except Exception as ex:  # keep the headline line even if an extra fails
                traceback.print_exc()
                extras[name] = {"error": repr(ex)}

    line = {
        "metric": METRIC, "value": value, "unit": "configs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": f"synthetic (seeded PCG64, sha256 {e.sha256()[:16]})",
        "config": {"workload": "edges: Panda 7-DoF, 64 random OBBs, 16384 straight edges, 8-256 states at <=0.02 rad "
                               "joint step, per-edge any-collision + first-hit, every state checked",
                   "edges": int(len(e.n_states)), "states": int(total_states), "obbs": int(e.scene.n_obbs),
                   **workload_stats},
        "roofline": roofline, "cpu_baseline": cpu_baseline, "parity": parity,
        "collide_shape": {"obb_tile": c_tile, "grid_form": c_grid, "source": "ppl_plan_autotune_collide"},
        "device": props.name, **extras,
    }
    print(json.dumps(line), flush=True)


def host_threads():
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    return max(1, min(threads, len(os.sched_getaffinity(0))))


def cpu_edges_leg(e, args, n_max=None):
    """FP32 CPU restatement of the edge-state check timed on this host (all
    cores, warmed median of 7 on a sample sized to ~cpu_seconds / 4 per run;
    one core, warmed median of 3 on a smaller sample), and the oracle's gated
    flops per state on a 20,000-state sample."""
    import numpy as np
    import pyoracle as PO
    states = PO.edge_states(e.frm, e.to, e.n_states)
    if n_max:
        states = states[:n_max]
    rng = np.random.default_rng(0)
    o = PO.Oracle(e.robot, e.scene, nthreads=host_threads())
    samp = states[rng.choice(len(states), size=min(20000, len(states)), replace=False)]
    f_state = float(np.mean(o.gated_flops(samp)))
    threads = host_threads()
    ocpu = PO.Oracle(e.robot, e.scene, nthreads=threads)
    n = 20000
    while True:
        t0 = time.perf_counter()
        ocpu.check_collision_f32_timed(states[:n])
        dt = time.perf_counter() - t0
        if dt > args.cpu_seconds / 4 or n >= len(states):
            break
        n = min(len(states), n * 4)

    def median_rate(oo, m, reps):
        oo.check_collision_f32_timed(states[:m])   # warm
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            oo.check_collision_f32_timed(states[:m])
            ts.append(time.perf_counter() - t0)
        return m / float(np.median(ts))
    cpu_value = median_rate(ocpu, n, 7)
    n1 = min(n, 20000)
    cpu_1 = median_rate(PO.Oracle(e.robot, e.scene, nthreads=1), n1, 3)
    log(f"cpu baseline: {cpu_value:.3e} states/s on {threads} threads, {cpu_1:.3e} on 1")
    cpu = {"value": cpu_value, "unit": "configs/s", "cores": threads, "kind": "port",
           "sample": f"first {n} states of the workload, FP32 two-phase check, warmed median of 7 "
                     f"(one core: first {n1} states, warmed median of 3)",
           "one_core_value": cpu_1, "cpu_model": cpu_model()}
    return f_state, cpu


def cpu_baseline_leg(e, fh_gpu, args):
    """The only place bench.py touches the oracle (with cpu_edges_leg and
    cpu_table_leg): time the FP32 CPU restatement on this host's cores, check
    the timed GPU output against the FP64 oracle, and count the gated
    algorithmic work per state."""
    import numpy as np
    import pyoracle as PO
    o = PO.Oracle(e.robot, e.scene, nthreads=host_threads())
    n_chk = 2000
    fh_ref = o.edge_batch(e.frm[:n_chk], e.to[:n_chk], e.n_states[:n_chk])
    parity = {"edges_checked": n_chk, "first_hit_mismatches": int(np.sum(fh_gpu[:n_chk] != fh_ref))}
    f_state, cpu = cpu_edges_leg(e, args)
    return cpu, parity, f_state


def cpu_table_leg(S, W, torch, dev, args):
    """The CPU-baseline leg widened to the other workloads: each one's FP32 CPU
    restatement timed on this host (all cores, then one core on a bounded
    sample), and the GPU output's worst error against the FP64 oracle with the
    count of results inside the ambiguous band (clearance within 1e-5 of zero,
    gradient kinks). GPU rates come from the other legs."""
    import numpy as np
    import pyoracle as PO
    from prismplan_amd.planning_scene import HostScene
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    threads = max(1, min(threads, len(os.sched_getaffinity(0))))

    def rate(fn, units, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return units / float(np.median(ts))

    def rows(theta, off, D):   # ragged [B] x [D][T_b] -> [waypoints, D]
        return np.concatenate([theta[D * off[b]:D * off[b + 1]].reshape(D, -1).T for b in range(len(off) - 1)])

    out = {"threads": threads}
    # any-collision: 64 requests x 4,096 configs, 20 OBBs each
    panda = S.Robot("panda")
    reqs = W.any_collision(n_requests=64)
    o, o1 = PO.Oracle(panda.arrays, reqs[0].scene, nthreads=threads), PO.Oracle(panda.arrays, reqs[0].scene, nthreads=1)
    mism = amb = 0
    for r in reqs:
        o.set_scene(r.scene)
        got = S.check_collision_scene(panda, S.DeviceScene.from_host(r.scene), r.configs)
        ref, clr = o.check_collision(r.configs)
        near = np.abs(clr) < 1e-5
        mism += int(np.sum((got != ref) & ~near))
        amb += int(near.sum())
    o.set_scene(reqs[0].scene)
    o1.set_scene(reqs[0].scene)
    q0 = np.concatenate([r.configs for r in reqs[:16]])
    out["any_collision"] = {"cpu_configs_per_s": rate(lambda: o.check_collision_f32_timed(q0), len(q0)),
                            "cpu_1core_configs_per_s": rate(lambda: o1.check_collision_f32_timed(q0[:16384]), 16384, 1),
                            "mismatches_outside_band": mism, "ambiguous": amb, "configs": 64 * 4096}
    # CHOMP: 256 trajectories x 64 waypoints, 40 OBBs, margin 0.05
    w = W.chomp()
    D, off = panda.n_dof, w.offsets
    ds = S.DeviceScene.from_host(w.scene)
    plan = S.Plan.trajectories(panda, w.lengths())
    th = torch.from_numpy(w.theta).to(dev)
    prm = S.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    prm_o = PO.cost_params(world_margin=0.05, weight=1.0, use_obb=True, use_esdf=True, use_self=False)
    _, tot, g = S.hinge_cost(plan, ds, th, prm, grad=True)
    o = PO.Oracle(panda.arrays, w.scene, nthreads=threads)
    o1 = PO.Oracle(panda.arrays, w.scene, nthreads=1)
    tot_ref, g_ref, kink = o.traj_cost_grad(w.theta, off, prm_o)
    gr, gg = rows(g_ref, off, D), rows(g.cpu().numpy(), off, D)
    den = np.maximum(np.abs(gr).max(axis=1), 1e-3)
    gerr = (np.abs(gg - gr).max(axis=1) / den)[~kink]
    fd = S.chomp_fd_grad(plan, ds, th, prm, eps=1e-3).cpu().numpy()
    fd_ref = o.chomp_fd_grad(w.theta, off, prm_o, eps=1e-3)
    fr = rows(fd_ref, off, D)
    fd_err = np.abs(rows(fd, off, D) - fr).max(axis=1) / np.maximum(np.abs(fr).max(axis=1), 1e-3)
    qw = rows(w.theta, off, D)
    B = w.n_traj
    out["chomp"] = {"cpu_adjoint_traj_per_s": rate(lambda: o.hinge_grad_f32_timed(qw, prm_o), B),
                    "cpu_1core_adjoint_traj_per_s": rate(lambda: o1.hinge_grad_f32_timed(qw[:64 * 32], prm_o), 32, 1),
                    "cpu_fd_traj_per_s": rate(lambda: o.chomp_fd_grad(w.theta, off, prm_o, 1e-3, "f32"), B, 1),
                    "cpu_1core_fd_traj_per_s": rate(lambda: o1.chomp_fd_grad(w.theta[:D * off[8]], off[:9], prm_o,
                                                                              1e-3, "f32"), 8, 1),
                    "adjoint_grad_max_rel_err": float(gerr.max()), "kinks": int(kink.sum()),
                    "total_max_rel_err": float(np.max(np.abs(tot.cpu().numpy() - tot_ref) /
                                                      np.maximum(np.abs(tot_ref), 0.05))),
                    "fd_grad_max_rel_err": float(fd_err[~kink].max())}
    # STOMP: Fetch, 128 problems x 32 rollouts, self pairs on
    wk = W.stomp()
    fetch = S.Robot("fetch")
    R = wk.n_rollouts
    plan = S.Plan.trajectories(fetch, np.repeat(wk.lengths, R), problem_units=np.arange(wk.n_problems + 1) * R,
                               obb_offsets=wk.obb_offsets, waypoint_major=True)
    ro = torch.from_numpy(wk.rollouts).to(dev)
    prm = S.cost_params(world_margin=0.02, self_margin=0.01, weight=20.0)
    prm_o = PO.cost_params(world_margin=0.02, self_margin=0.01, weight=20.0)
    st, _ = S.stomp_rollout_costs(plan, S.DeviceScene.from_host(wk.scene), ro, prm, 0.1, 0.1)
    o = PO.Oracle(wk.robot, nthreads=threads)
    o1 = PO.Oracle(wk.robot, nthreads=1)
    s_ref, _ = o.stomp_costs(wk.scene, wk.obb_offsets, wk.rollouts, wk.lengths, wk.roll_offsets, R, prm_o, 0.1, 0.1)
    rel = np.abs(st.cpu().numpy() - s_ref) / np.maximum(np.abs(s_ref), 20.0 * 0.02)
    P1 = 4
    out["stomp"] = {"cpu_rollout_states_per_s": rate(lambda: o.stomp_costs(
                        wk.scene, wk.obb_offsets, wk.rollouts, wk.lengths, wk.roll_offsets, R, prm_o, 0.1, 0.1, "f32"),
                        wk.n_states, 1),
                    "cpu_1core_rollout_states_per_s": rate(lambda: o1.stomp_costs(
                        wk.scene, wk.obb_offsets[:P1 + 1], wk.rollouts[:int(wk.roll_offsets[P1]) * fetch.n_dof],
                        wk.lengths[:P1], wk.roll_offsets[:P1 + 1], R, prm_o, 0.1, 0.1, "f32"),
                        int(wk.roll_offsets[P1]), 1),
                    "state_cost_max_rel_err": float(rel.max()), "states": int(wk.n_states)}
    # Baxter trilinear ESDF adjoint: 1,024 trajectories, 128^3 grid from 60 OBBs
    baxter = S.Robot("baxter")
    wk = W.esdf()
    params, pose = wk.scene.voxel_params, wk.scene.voxel_pose
    feats = S.esdf_from_obbs(S.DeviceScene.from_host(wk.scene.source_obbs), params[0], pose[0])
    wk.scene.voxel_features = feats.cpu().numpy().reshape(1, -1)
    plan = S.Plan.trajectories(baxter, wk.lengths())
    th = torch.from_numpy(wk.theta).to(dev)
    prm = S.cost_params(world_margin=0.02, weight=1.0, use_obb=False, use_self=False)
    prm_o = PO.cost_params(world_margin=0.02, weight=1.0, use_obb=False, use_self=False)
    _, tot, g = S.hinge_cost(plan, S.DeviceScene.from_host(wk.scene), th, prm, grad=True)
    o = PO.Oracle(wk.robot, wk.scene, nthreads=threads)
    o1 = PO.Oracle(wk.robot, wk.scene, nthreads=1)
    tot_ref, g_ref, kink = o.traj_cost_grad(wk.theta, wk.offsets, prm_o)
    Db = baxter.n_dof
    gr, gg = rows(g_ref, wk.offsets, Db), rows(g.cpu().numpy(), wk.offsets, Db)
    interior = np.ones(len(gr), bool)
    interior[wk.offsets[:-1]] = False
    interior[wk.offsets[1:] - 1] = False
    gerr = (np.abs(gg - gr).max(axis=1) / np.maximum(np.abs(gr).max(axis=1), 1e-3))[interior & ~kink]
    qw = rows(wk.theta, wk.offsets, Db)
    out["esdf"] = {"cpu_waypoints_per_s": rate(lambda: o.hinge_grad_f32_timed(qw, prm_o), len(qw)),
                   "cpu_1core_waypoints_per_s": rate(lambda: o1.hinge_grad_f32_timed(qw[:8192], prm_o), 8192, 1),
                   "grad_max_rel_err": float(gerr.max()), "kinks": int(kink.sum()),
                   "total_max_rel_err": float(np.max(np.abs(tot.cpu().numpy() - tot_ref) /
                                                     np.maximum(np.abs(tot_ref), 0.02)))}
    return out


def bench_copy(torch, dev, args):
    """A plain 256 MiB device-to-device copy: the reference kernel whose
    FETCH_SIZE / WRITE_SIZE readings scale the HBM counters."""
    n = 64 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    wall, ev = timed_loop(lambda: b.copy_(a), 10, 2, torch)
    return {"bytes_read": 4 * n, "bytes_written": 4 * n, "GB_per_s": 8 * n / ev / 1e9}


def bench_skip(plan, ds, frm, to, fh, total_states, S, torch, args):
    def step():
        S.edge_check_batched(plan, ds, frm, to, skip_resolved=True, first_hit=fh)
    wall, ev = timed_loop(step, args.steps, args.warmup, torch)
    return {"states_resolved_per_s": total_states / wall, "edges_per_s": len(plan.unit_states) / wall,
            "ms_per_step": wall * 1e3}


def bench_any(panda, S, W, torch, dev, args):
    """64 independent any-collision requests of 4,096 configs, each with its
    own 20 OBBs: 64 plain calls (host arrays, synchronous) vs one batched
    launch over device-resident configs."""
    import numpy as np
    reqs = W.any_collision(n_requests=64)
    scenes = [S.DeviceScene.from_host(r.scene) for r in reqs]
    for _ in range(2):
        for r, sc in zip(reqs, scenes):
            S.check_collision_scene(panda, sc, r.configs)
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        outs = [S.check_collision_scene(panda, sc, r.configs) for r, sc in zip(reqs, scenes)]
    plain = 64 * 4096 * reps / (time.perf_counter() - t0)
    from prismplan_amd.planning_scene import HostScene
    cfg = torch.from_numpy(np.concatenate([r.configs for r in reqs])).to(dev)
    hs = HostScene(obb_dims=np.concatenate([r.scene.obb_dims for r in reqs]),
                   obb_pose=np.concatenate([r.scene.obb_pose for r in reqs]),
                   obb_enable=np.concatenate([r.scene.obb_enable for r in reqs]))
    ds = S.DeviceScene.from_host(hs)
    plan = S.Plan.configs(panda, 64 * 4096, problem_units=np.arange(65) * 4096, obb_offsets=np.arange(65) * 20)
    out = torch.empty(64 * 4096, dtype=torch.uint8, device=dev)
    wall, ev = timed_loop(lambda: S.check_collision_batched(plan, ds, cfg, out=out), args.steps, args.warmup, torch)
    batched = 64 * 4096 / wall
    same = bool(np.array_equal(out.cpu().numpy().astype(bool), np.concatenate(outs)))
    # sweep of one request size (single scene) for the launch-bound regime
    sweep = {}
    for n in (4096, 65536, 1048576):
        w = W.any_collision(n_configs=n, seed=0)[0]
        p2 = S.Plan.configs(panda, n)
        c2 = torch.from_numpy(w.configs).to(dev)
        d2 = S.DeviceScene.from_host(w.scene)
        o2 = torch.empty(n, dtype=torch.uint8, device=dev)
        wl, evv = timed_loop(lambda: S.check_collision_batched(p2, d2, c2, out=o2), 20, 3, torch)
        sweep[str(n)] = n / wl
    return {"plain_configs_per_s": plain, "batched_configs_per_s": batched, "batched_over_plain": batched / plain,
            "batched_equals_plain": same, "batched_kernel_ms": ev * 1e3, "single_scene_sweep_configs_per_s": sweep}


def bench_edges_free(panda, S, W, torch, dev, args):
    """The headline kernel and edge layout (16,384 edges of 8-256 states at
    <= 0.02 rad, every state checked) in a 10-box world, where most states are
    free and so are checked through every joint and candidate box."""
    import numpy as np
    e = W.edges(n_obbs=10, seed=42)
    ds = S.DeviceScene.from_host(e.scene)
    plan = S.Plan.edges(panda, e.n_states)
    frm, to = torch.from_numpy(e.frm).to(dev), torch.from_numpy(e.to).to(dev)
    fh = torch.empty(len(e.n_states), dtype=torch.int32, device=dev)
    wall, ev = timed_loop(lambda: S.edge_check_batched(plan, ds, frm, to, skip_resolved=False, first_hit=fh),
                          args.steps, args.warmup, torch)
    _, hits = S.edge_check_batched(plan, ds, frm, to, skip_resolved=False, state_hits=True)
    fh_h = fh.cpu().numpy()
    out = {"states_per_s": e.total_states / wall, "kernel_ms": ev * 1e3, "states": int(e.total_states),
           "obbs": int(e.scene.n_obbs), "colliding_state_frac": float(hits.float().mean().item()),
           "colliding_edge_frac": float(np.mean(fh_h >= 0))}
    if not args.no_cpu:
        # the same roofline rule and CPU baseline as the headline, on this world
        f_state, cpu = cpu_edges_leg(e, args, n_max=200000)
        from prismplan_amd import _ppl
        f_exh = panda.flops(_ppl.WORK_ANY_COLLISION, n_obbs=e.scene.n_obbs)
        achieved = f_state * e.total_states / ev
        out["roofline"] = {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                           "frac": achieved / PEAK_FP32, "frac_of_unpacked_peak": achieved / PEAK_FP32_UNPACKED,
                           "flops_per_state": f_state, "flops_rule": FLOPS_RULE,
                           "exhaustive_flops_per_state": f_exh,
                           "exhaustive_frac": f_exh * e.total_states / ev / PEAK_FP32}
        out["cpu_baseline"] = cpu
    return out


def bench_edges_plain(panda, e, ds, S, torch, dev, args):
    """64 plain EdgeEvalGpu calls of 256 edges (states j=1..G, G from 0.02 rad
    resolution) vs the same edges in one batched launch (first-hit skipping on
    in both); also the latency of the 14-edge calls an eager MHA* / wPA*SE
    expansion makes (2 * dim successors at +-0.08 rad)."""
    import numpy as np
    ev = S.EdgeEvaluator(panda, ds, max_edges=256)
    Gs, _ = S.edge_granularities(e.frm, e.to, resolution=0.02)
    for k in range(4):
        ev.evaluate(e.frm[k * 256:(k + 1) * 256], e.to[k * 256:(k + 1) * 256], resolution=0.02)
    t0 = time.perf_counter()
    valid = []
    for k in range(64):
        v, c, _ = ev.evaluate(e.frm[k * 256:(k + 1) * 256], e.to[k * 256:(k + 1) * 256], resolution=0.02)
        valid.append(v)
    dt = time.perf_counter() - t0
    plain_states = float(np.sum(Gs)) / dt
    # batched over the same states (n = G + 1 with the start skipped)
    plan = S.Plan.edges(panda, Gs + 1, skip_first=True)
    frm, to = torch.from_numpy(e.frm).to(dev), torch.from_numpy(e.to).to(dev)
    fh = torch.empty(len(Gs), dtype=torch.int32, device=dev)
    wall, evt = timed_loop(lambda: S.edge_check_batched(plan, ds, frm, to, skip_resolved=True, first_hit=fh),
                           args.steps, args.warmup, torch)
    batched_states = float(np.sum(Gs)) / wall
    same = bool(np.array_equal(np.concatenate(valid), fh.cpu().numpy() < 0))
    # MHA* / wPA*SE eager expansion: one synchronous call per state for its 2 * dim successor edges
    q0 = e.frm[:64]
    succ_to = np.repeat(q0, 14, axis=0).reshape(64, 14, 7)
    for d in range(7):
        succ_to[:, 2 * d, d] += 0.08
        succ_to[:, 2 * d + 1, d] -= 0.08
    succ_frm = np.repeat(q0, 14, axis=0).reshape(64, 14, 7)
    for k in range(4):
        ev.evaluate(succ_frm[k], succ_to[k], resolution=0.02)
    t0 = time.perf_counter()
    for k in range(64):
        ev.evaluate(succ_frm[k], succ_to[k], resolution=0.02)
    dt14 = time.perf_counter() - t0
    return {"plain_states_per_s": plain_states, "batched_states_per_s": batched_states,
            "batched_over_plain": batched_states / plain_states, "verdicts_equal": same,
            "plain_ms_per_256_edge_call": dt / 64 * 1e3, "plain_ms_per_14_edge_call": dt14 / 64 * 1e3}


def bench_chomp(panda, S, W, torch, dev, args):
    w = W.chomp()
    ds = S.DeviceScene.from_host(w.scene)
    plan = S.Plan.trajectories(panda, w.lengths())
    th = torch.from_numpy(w.theta).to(dev)
    prm = S.cost_params(world_margin=0.05, weight=1.0, use_self=False)
    tuned = plan.autotune_cost(ds, th, prm, grad=True)
    wall_a, ev_a = timed_loop(lambda: S.hinge_cost(plan, ds, th, prm, grad=True), args.steps, args.warmup, torch)
    k_a = kernel_time(plan, lambda: S.hinge_cost(plan, ds, th, prm, grad=True), args.steps, torch)
    wall_f, ev_f = timed_loop(lambda: S.chomp_fd_grad(plan, ds, th, prm, eps=1e-3), max(5, args.steps // 5), 2, torch)
    # the reference's call pattern: one FD launch per trajectory (plain per-trajectory form)
    D, Tw = panda.n_dof, int(w.lengths()[0])
    views = [th[b * D * Tw:(b + 1) * D * Tw] for b in range(w.n_traj)]
    S.chomp_fd_grad_plain(panda, ds, views[0], Tw, prm, eps=1e-3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for v in views:
        S.chomp_fd_grad_plain(panda, ds, v, Tw, prm, eps=1e-3)
    torch.cuda.synchronize()
    wall_fp = (time.perf_counter() - t0)
    from prismplan_amd import _ppl
    B, T, O = w.n_traj, int(w.lengths()[0]), w.scene.n_obbs
    f_adj = panda.flops(_ppl.WORK_CHOMP_ADJOINT, n_obbs=O, n_waypoints=T)
    f_fd = panda.flops(_ppl.WORK_CHOMP_FD, n_obbs=O, n_waypoints=T)
    return {"adjoint_traj_per_s": B / wall_a, "fd_traj_per_s": B / wall_f, "adjoint_over_fd": wall_f / wall_a,
            "adjoint_ms": wall_a * 1e3, "fd_ms": wall_f * 1e3,
            "fd_per_trajectory_launch_traj_per_s": B / wall_fp, "adjoint_over_fd_per_trajectory": wall_fp / wall_a,
            "adjoint_flops_per_traj": f_adj, "fd_flops_per_traj": f_fd, "adjoint_lanes_per_state": tuned[0],
            "adjoint_roofline_frac": B * f_adj / ev_a / PEAK_FP32,
            "adjoint_kernel_ms": k_a * 1e3, "adjoint_kernel_roofline_frac": B * f_adj / k_a / PEAK_FP32,
            "fd_roofline_frac": B * f_fd / ev_f / PEAK_FP32}


def bench_chomp_opt(panda, S, W, torch, dev, args):
    """The batched CHOMP optimiser on the CHOMP workload (256 trajectories x 64
    waypoints, 40 OBBs): iterations of adjoint gradient, covariant step,
    cost, densified validity and bookkeeping, all on the device."""
    from prismplan_amd import _ppl
    w = W.chomp()
    ds = S.DeviceScene.from_host(w.scene)
    plan = S.Plan.trajectories(panda, w.lengths())
    st = S.chomp_settings(world_margin=0.05)
    th = torch.from_numpy(w.theta.copy()).to(dev)
    opt = S.ChompOptimizer(plan, ds, th, st).step(0)
    opt.step(4)   # captures the iteration graph
    iters = 20
    dt = timed_steps(opt, iters, torch)
    plan.set_optimizer_graphs(False)
    dt_eager = timed_steps(opt, iters, torch)
    plan.set_optimizer_graphs(True)
    rec = opt.record.cpu()
    return {"iterations": iters, "ms_per_iteration": dt / iters * 1e3, "eager_ms_per_iteration": dt_eager / iters * 1e3,
            "trajectory_iterations_per_s": w.n_traj * iters / dt,
            "solved": int(rec[:, _ppl.CHOMP_SOLVED].sum()), "trajectories": w.n_traj,
            "done": int(rec[:, _ppl.CHOMP_DONE].sum())}


def timed_steps(opt, iters, torch):
    """Wall time of one optimiser call of `iters` iterations, synchronised."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    opt.step(iters)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def bench_stomp_opt(panda, S, W, torch, dev, args):
    """The batched STOMP optimiser: 128 Panda trajectories x 51 waypoints, 40
    OBBs, 30 new + 20 kept rollouts (the reference's STOMP_settings defaults):
    sampling, rollout costs, weighting, projection, optimum cost, densified
    validity and bookkeeping, all on the device. Beside it, the single-goal
    solve (stomp_solve_scene's contract) per iteration."""
    import numpy as np
    from prismplan_amd import _ppl
    w = W.chomp(n_traj=128, n_wp=51, seed=5)
    ds = S.DeviceScene.from_host(w.scene)
    plan = S.Plan.trajectories(panda, w.lengths())
    st = S.stomp_settings(world_margin=0.05, iterations_after_valid=1000)
    th = torch.from_numpy(w.theta.copy()).to(dev)
    opt = S.StompOptimizer(plan, ds, th, st).step(0)
    opt.step(6)   # steady population; captures the two-iteration graph
    iters = 20
    dt = timed_steps(opt, iters, torch)
    plan.set_optimizer_graphs(False)
    dt_eager = timed_steps(opt, iters, torch)
    plan.set_optimizer_graphs(True)
    rec = opt.record.cpu()
    n_fresh = w.n_traj * st.num_rollouts_new * 51
    # one request at a time through the single-goal solve (one trajectory, same settings)
    x0 = w.theta[:7 * 51].reshape(7, 51)
    start, goal = np.ascontiguousarray(x0[:, 0]), np.ascontiguousarray(x0[:, -1])
    single = S.stomp_solve(panda, ds, start, goal, 51, 10, st)
    single = S.stomp_solve(panda, ds, start, goal, 51, 10, st)
    return {"iterations": iters, "ms_per_iteration": dt / iters * 1e3, "eager_ms_per_iteration": dt_eager / iters * 1e3,
            "trajectories": w.n_traj,
            "trajectory_iterations_per_s": w.n_traj * iters / dt, "fresh_rollout_states_per_s": n_fresh * iters / dt,
            "solved": int(rec[:, _ppl.CHOMP_SOLVED].sum()), "done": int(rec[:, _ppl.CHOMP_DONE].sum()),
            "single_goal_ms_per_iteration": single["wall_time_ms"] / max(single["iterations_run"], 1),
            "batched_over_single_goal": (w.n_traj * iters / dt) /
                                        (max(single["iterations_run"], 1) / (single["wall_time_ms"] * 1e-3))}


def bench_mitstar(panda, S, W, torch, dev, args):
    """MIT*'s substrate calls on a device-resident graph: 4,096 Panda nodes
    (half uniform, half within ~0.15 rad of them), 16,384 edges, a 20-OBB
    world, self checks on: batch_cc of the nodes, then full_cc (every
    breadth-first midpoint at 0.02 rad) and sparse_cc (3 midpoints) of the
    edges."""
    import numpy as np
    ds = S.DeviceScene.from_host(W.any_collision(n_configs=16)[0].scene)
    rng = np.random.default_rng(7)
    r = panda.arrays
    nodes = W.uniform_configs(rng, r, 4096)
    nodes[2048:] = nodes[:2048] + rng.normal(scale=0.15, size=(2048, r.n_dof)).astype(np.float32)
    nodes = np.clip(nodes, r.joint_lower, r.joint_upper).astype(np.float32)
    src = rng.integers(0, 2048, 16384).astype(np.int32)
    dst = np.where(rng.random(16384) < 0.8, src + 2048, rng.integers(0, 4096, 16384)).astype(np.int32)
    xn, xs, xd = (torch.from_numpy(a).to(dev) for a in (nodes, src, dst))
    clr = torch.empty(4096, device=dev)
    wall_b, _ = timed_loop(lambda: S.mitstar_batch_cc(panda, ds, xn, 0.0, clr, 1.0), args.steps, args.warmup, torch)
    cc, new = S.mitstar_edge_cc(panda, ds, xn, xs, xd, 0.02)
    mids = int(new.sum().item())
    wall_f, _ = timed_loop(lambda: S.mitstar_edge_cc(panda, ds, xn, xs, xd, 0.02), args.steps, args.warmup, torch)
    wall_s, _ = timed_loop(lambda: S.mitstar_edge_cc(panda, ds, xn, xs, xd, 0.02, num_sparse_checks=3),
                           args.steps, args.warmup, torch)
    return {"batch_cc_samples_per_s": 4096 / wall_b, "batch_cc_ms": wall_b * 1e3,
            "full_cc_edges_per_s": 16384 / wall_f, "full_cc_ms": wall_f * 1e3,
            "full_cc_free_edges": int((cc == 0).sum().item()), "full_cc_verified_midpoints": mids,
            "sparse_cc_edges_per_s": 16384 / wall_s, "sparse_cc_ms": wall_s * 1e3}


def bench_rrtc(panda, S, W, torch, dev, args):
    """pRRTC solves on the bundled MotionBenchMaker sample problems (Panda,
    box and cage scenes, primitives as containing OBBs), five seeds each,
    the reference's default settings (600 workgroups, granularity 16, range
    0.5, dynamic domain, balance 1) with reusable trees; wall and kernel time
    per solve."""
    import numpy as np
    from prismplan_amd import problems as P
    from prismplan_amd.planning_scene import HostScene
    probs = P.load_vamp_problems(str(ROOT / "tests" / "golden" / "sample_mbm_problems.json"))
    bufs = S.RrtcBuffers(7, max_samples=1000000, num_new_configs=600, granularity=16)
    walls, kernels, iters, wps, solved = [], [], [], [], 0
    for p in probs:
        obbs = P.obstacles_to_obbs(p.obstacles)
        ds = S.DeviceScene.from_host(HostScene.from_boxes([o["dims"] for o in obbs], [o["position"] for o in obbs],
                                                          [o["quaternion"] for o in obbs]))
        S.rrtc_solve(panda, ds, p.start, p.goal[None], S.rrtc_settings(seed=99), bufs=bufs)   # warm-up
        for seed in range(5):
            r = S.rrtc_solve(panda, ds, p.start, p.goal[None], S.rrtc_settings(seed=seed + 1), bufs=bufs)
            solved += int(r["solved"])
            walls.append(r["wall_time_ms"])
            kernels.append(r["kernel_ms"])
            iters.append(r["iters"])
            wps.append(len(r["path"]))
    bufs.close()
    return {"problems": len(probs), "solves": len(walls), "solved": solved,
            "wall_ms_median": float(np.median(walls)), "wall_ms_max": float(np.max(walls)),
            "kernel_ms_median": float(np.median(kernels)), "kernel_ms_mean": float(np.mean(kernels)),
            "iters_median": float(np.median(iters)), "waypoints_median": float(np.median(wps))}


def bench_stomp(S, W, torch, dev, args):
    import numpy as np
    wk = W.stomp()
    fetch = S.Robot("fetch")
    R = wk.n_rollouts
    plan = S.Plan.trajectories(fetch, np.repeat(wk.lengths, R), problem_units=np.arange(wk.n_problems + 1) * R,
                               obb_offsets=wk.obb_offsets, waypoint_major=True)
    ds = S.DeviceScene.from_host(wk.scene)
    ro = torch.from_numpy(wk.rollouts).to(dev)
    prm = S.cost_params(world_margin=0.02, self_margin=0.01, weight=20.0)
    state = torch.empty(wk.n_states, device=dev)
    total = torch.empty(wk.n_problems * R, device=dev)
    tuned = plan.autotune_cost(ds, ro, prm, grad=False)
    g = S.capture_stomp(plan, ds, ro, prm, 0.1, 0.1, state, total)
    wall, ev = timed_loop(lambda: g.launch(), args.steps, args.warmup, torch)
    wall_e, _ = timed_loop(lambda: S.stomp_rollout_costs(plan, ds, ro, prm, 0.1, 0.1), args.steps, args.warmup, torch)
    from prismplan_amd import _ppl
    # exhaustive flops per rollout state at each problem's own OBB count, + control
    n_obb = np.diff(wk.obb_offsets)
    states_p = wk.lengths * R
    flops = sum(float(sp) * (fetch.flops(_ppl.WORK_STOMP_STATE, n_obbs=int(o), use_self=True) + 18.0 * fetch.n_dof)
                for sp, o in zip(states_p, n_obb))
    return {"rollout_states_per_s": wk.n_states / wall, "rollouts_per_s": wk.n_problems * R / wall,
            "graph_ms": wall * 1e3, "eager_ms": wall_e * 1e3, "states": wk.n_states,
            "flops_per_state_mean": flops / wk.n_states, "lanes_per_state": tuned[0],
            "roofline_frac_exhaustive": flops / ev / PEAK_FP32}


def bench_esdf(S, W, torch, dev, args):
    baxter = S.Robot("baxter")
    wk = W.esdf()
    obbs = S.DeviceScene.from_host(wk.scene.source_obbs)
    params, pose = wk.scene.voxel_params, wk.scene.voxel_pose
    wall_b, _ = timed_loop(lambda: S.esdf_from_obbs(obbs, params[0], pose[0]), 10, 2, torch)
    feats = S.esdf_from_obbs(obbs, params[0], pose[0])
    ds = S.DeviceScene(None, None, None, torch.from_numpy(params).to(dev), torch.from_numpy(pose).to(dev),
                       torch.ones(1, dtype=torch.uint8, device=dev), feats, esdf_interp=1)
    plan = S.Plan.trajectories(baxter, wk.lengths())
    th = torch.from_numpy(wk.theta).to(dev)
    prm = S.cost_params(world_margin=0.02, weight=1.0, use_obb=False, use_self=False)
    tuned = plan.autotune_cost(ds, th, prm, grad=True)
    wall, ev = timed_loop(lambda: S.hinge_cost(plan, ds, th, prm, grad=True), args.steps, args.warmup, torch)
    k_a = kernel_time(plan, lambda: S.hinge_cost(plan, ds, th, prm, grad=True), args.steps, torch)
    from prismplan_amd import _ppl
    f_wp = baxter.flops(_ppl.WORK_ESDF_WAYPOINT)
    return {"waypoints_per_s": wk.n_waypoints / wall, "trajectories_per_s": wk.n_traj / wall,
            "adjoint_ms": wall * 1e3, "esdf_build_ms": wall_b * 1e3, "waypoints": wk.n_waypoints,
            "flops_per_waypoint": f_wp, "lanes_per_state": tuned[0],
            "roofline_frac": wk.n_waypoints * f_wp / ev / PEAK_FP32,
            "adjoint_kernel_ms": k_a * 1e3, "kernel_roofline_frac": wk.n_waypoints * f_wp / k_a / PEAK_FP32}


if __name__ == "__main__":
    main()
