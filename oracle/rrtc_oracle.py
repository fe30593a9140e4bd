"""Restatement of pRRTC's deterministic parts (test infrastructure only).

Follows planning/pRRTC.cu (Halton sampling :955-995, edge states :1668-1677
and :1852-1866, the tree walk :1962-1988, the goal -> start assembly
:2167-2186) and planning/shortcut.cuh:16-274 (skip-edge states and the
greedy pass). With many workgroups the planner is concurrent and not
reproducible state for state; its outputs are checked against what these
functions and the FP64 collision oracle fix: bases, samples, edge states,
path validity, costs and the shortcut's choices. With one workgroup it is
sequential, and rrtc_single_workgroup restates the whole run.

One documented difference of the product, shared here: the reference
shuffles each workgroup's Halton bases with cuRAND (pRRTC.cu:59-66, 961);
the product draws the Fisher-Yates indices from splitmix64 (Steele, Lea,
Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014),
state seed ^ (w * 0xD1B54A32D192ED03), j = next % (i + 1) for i = 15..1.
The generator is pinned by its published outputs in the tests.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

PRIMES = (3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59)
_M64 = (1 << 64) - 1


def splitmix64(state: int):
    """One step: (new state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def halton_bases(seed: int, n_workgroups: int, n_dof: int) -> np.ndarray:
    """[n_workgroups, n_dof] bases: workgroup 0 the primes in order, the
    others a Fisher-Yates shuffle of all 16 (pRRTC.cu:955-966)."""
    out = np.empty((n_workgroups, n_dof), np.float32)
    for w in range(n_workgroups):
        p = list(PRIMES)
        if w != 0:
            x = (seed ^ ((w * 0xD1B54A32D192ED03) & _M64)) & _M64
            for i in range(15, 0, -1):
                x, r = splitmix64(x)
                j = r % (i + 1)
                p[i], p[j] = p[j], p[i]
        out[w] = np.float32(p[:n_dof])
    return out


def halton_sequence(bases, n_steps: int) -> np.ndarray:
    """[n_steps, dim] samples of the incremental generator in FP32, each
    operation rounded as the device rounds it (pRRTC.cu:982-995)."""
    b = np.asarray(bases, np.float32)
    n = np.zeros_like(b)
    d = np.ones_like(b)
    out = np.empty((n_steps, b.size), np.float32)
    one = np.float32(1.0)
    for k in range(n_steps):
        for i in range(b.size):
            x = np.float32(d[i] - n[i])
            if x == one:
                d[i] = np.floor(np.float32(d[i] * b[i]))
                n[i] = one
            else:
                y = np.floor(np.float32(d[i] / b[i]))
                while x <= y:
                    y = np.floor(np.float32(y / b[i]))
                n[i] = np.float32(np.floor(np.float32((b[i] + one) * y)) - x)
            out[k, i] = np.float32(n[i] / d[i])
    return out


def radical_inverse(k: int, base: int) -> Fraction:
    """The van der Corput value of k in `base`: digits of k mirrored about
    the radix point (the closed form the Halton generator must reproduce)."""
    r, f = Fraction(0), Fraction(1, base)
    while k > 0:
        k, digit = divmod(k, base)
        r += digit * f
        f /= base
    return r


def scale_to_limits(h, lower, upper) -> np.ndarray:
    """q = fma(h, upper - lower, lower) in FP32 (runtime_kinematics.cuh:520-527)."""
    h = np.asarray(h, np.float32)
    lo = np.asarray(lower, np.float32)
    rng = (np.asarray(upper, np.float32) - lo).astype(np.float32)
    return (h.astype(np.float64) * rng.astype(np.float64) + lo.astype(np.float64)).astype(np.float32)


def segment_states(a, b, granularity: int) -> np.ndarray:
    """States a + k (b - a) / g, k = 0..g, of one path segment in FP64: the
    points pRRTC's edge checks cover whichever end the edge grew from."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    k = np.arange(granularity + 1, dtype=np.float64)[:, None] / granularity
    return (a[None, :] + k * (b - a)[None, :]).astype(np.float32)


def segment_interior(a, b, valid_segment_length: float) -> np.ndarray:
    """Interior states of one skip edge a -> b (shortcut.cuh:67-92): total =
    max(1, ceil(|b - a| / L)) from an FP32 sequential sum, t = s / total for
    s = 1..total-1, q = a (1 - t) + b t with the two products and the sum
    rounded separately."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    L = np.float32(valid_segment_length)
    s = np.float32(0.0)
    for d in range(a.size):
        df = np.float32(a[d] - b[d])
        s = np.float32(s + np.float32(df * df))
    total = max(1, int(np.ceil(np.float32(np.sqrt(s) / L))))
    t = (np.arange(1, total, dtype=np.float32) / np.float32(total)).astype(np.float32)[:, None]
    return ((a[None, :] * (np.float32(1.0) - t)).astype(np.float32) + (b[None, :] * t).astype(np.float32)).astype(
        np.float32).reshape(-1, a.size)


def shortcut_pairs(path, valid_segment_length: float):
    """Skip edges (i, j >= i + 2) of a path with their interior states
    (segment_interior). Returns [(i, j, states)]."""
    p = np.asarray(path, np.float32)
    n = p.shape[0]
    return [(i, j, segment_interior(p[i], p[j], valid_segment_length)) for i in range(n - 2) for j in range(i + 2, n)]


def shortcut_greedy(n: int, free) -> list:
    """Waypoint indices kept by the greedy pass (shortcut.cuh:254-272): from
    i, the largest j > i + 1 with free[(i, j)], else i + 1."""
    keep = [0]
    i = 0
    while i < n - 1:
        far = i + 1
        for j in range(n - 1, i + 1, -1):
            if free.get((i, j), False):
                far = j
                break
        keep.append(far)
        i = far
    return keep


class Ambiguous(Exception):
    """An edge state lies within the ambiguous clearance band: the FP32 check
    may decide either way, so the sequential run cannot be reproduced."""


def _fma32(a, b, c):
    """FP32 fma through FP64: the product of two FP32 values is exact in FP64
    and the FP64 sum is rounded once more to FP32 (exact unless the FP64 sum
    lands on an FP32 halfway point)."""
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def _nearest(nodes, q):
    """Index and squared distance of the nearest row: per row, s = fma(t, t, s)
    over the dims in order, t = node - q in FP32; ties go to the smaller index
    (pRRTC.cu:1637-1666)."""
    t = (nodes - q[None, :]).astype(np.float32)
    s = np.zeros(len(nodes), np.float32)
    for k in range(nodes.shape[1]):
        s = (t[:, k].astype(np.float64) ** 2 + s.astype(np.float64)).astype(np.float32)
    i = int(np.argmin(s))
    return i, s[i]


def rrtc_single_workgroup(start, goals, lower, upper, collides, granularity=16, range_=0.5, balance=1,
                          tree_ratio=0.5, dynamic_domain=True, dd_alpha=1e-4, dd_radius=4.0, dd_min_radius=1.0,
                          max_samples=100000, max_iters=100000):
    """pRRTC with one workgroup (num_new_configs = 1), pRRTC.cu:1556-2010.
    Without concurrency the run is sequential, so the trees, the iteration
    that connects them and the path are fixed. Workgroup 0 samples the
    Halton sequence of the primes in order (pRRTC.cu:955-995). collides(states
    [k, D] FP32) -> bool says whether any state of one edge check collides.
    Returns dict(solved, iters, path start -> goal, reached_goal,
    start_tree_size, goal_tree_size)."""
    f32 = np.float32
    FLT_MAX = np.finfo(np.float32).max
    D = len(start)
    trees = [[np.asarray(start, f32)], [np.asarray(g, f32) for g in goals]]
    parents = [[0], list(range(len(goals)))]
    radii = [[FLT_MAX], [FLT_MAX] * len(goals)]
    b = np.float32(PRIMES[:D])
    hn, hd = np.zeros(D, f32), np.ones(D, f32)
    lo, hi = np.asarray(lower, f32), np.asarray(upper, f32)
    rng = (hi - lo).astype(f32)
    G = f32(granularity)
    one = f32(1.0)

    def halton():
        out = np.empty(D, f32)
        for i in range(D):
            x = f32(hd[i] - hn[i])
            if x == one:
                hd[i] = np.floor(f32(hd[i] * b[i]))
                hn[i] = one
            else:
                y = np.floor(f32(hd[i] / b[i]))
                while x <= y:
                    y = np.floor(f32(y / b[i]))
                hn[i] = f32(np.floor(f32((b[i] + one) * y)) - x)
            out[i] = f32(hn[i] / hd[i])
        return out

    def edge_free(base, step):
        states = np.stack([np.float32(np.float64(k + 1) * np.float64(step) + np.float64(base))
                           for k in range(granularity)])
        return not collides(states)

    def shrink(t, i):
        old = radii[t][i]
        radii[t][i] = f32(dd_radius) if old == FLT_MAX else max(f32(old * f32(one - f32(dd_alpha))), f32(dd_min_radius))

    def grow(t, i):
        if radii[t][i] != FLT_MAX:
            radii[t][i] = f32(radii[t][i] * f32(one + f32(dd_alpha)))

    def add(t, q, parent):
        if len(trees[t]) >= max_samples:
            return -1
        trees[t].append(q)
        parents[t].append(parent)
        radii[t].append(FLT_MAX)
        return len(trees[t]) - 1

    def result(solved, it, win=None):
        out = {"solved": solved, "iters": it, "start_tree_size": len(trees[0]), "goal_tree_size": len(trees[1])}
        if solved:
            wt, wlast, wo = win
            ends = {wt: wlast, 1 - wt: wo}
            walks = []
            for tree in (0, 1):
                cur, walk = ends[tree], []
                while parents[tree][cur] != cur:
                    walk.append(trees[tree][cur])
                    cur = parents[tree][cur]
                walks.append((walk, cur))
            rows = [trees[0][0]] + walks[0][0][::-1] + walks[1][0] + [trees[1][walks[1][1]]]
            out.update(path=np.stack(rows), reached_goal=walks[1][1])
        return out

    t = 0
    it = 0
    while True:
        it += 1
        if it > max_iters:
            return result(False, it - 1)
        f0, f1 = len(trees[0]), len(trees[1])
        if balance == 0 or it == 1:
            t = 1                                   # workgroup 0 of 1: 0 < 1 / 2 is false
        elif balance == 1:
            ratio = f32(f32(f0) / f32(f0 + f1))
            if f32(abs(f0 - f1)) < f32(1.5):
                t = 0 if f32(0.0) < f32(one - ratio) else 1
            else:
                t = 0 if ratio < f32(tree_ratio) else 1
        else:
            ft, fo = (f0, f1) if t == 0 else (f1, f0)
            t = 1 - t if f32(f32(abs(ft - fo)) / f32(ft)) < f32(tree_ratio) else t
        cfg = np.array([_fma32(h, r, l) for h, r, l in zip(halton(), rng, lo)], f32)
        o = 1 - t
        near, d2 = _nearest(np.stack(trees[t]), cfg)
        dist = np.sqrt(f32(d2))
        if dynamic_domain and radii[t][near] < dist:
            continue
        nq = trees[t][near]
        scale = min(one, f32(f32(range_) / dist)) if dist > 0 else one
        c = np.array([_fma32(f32(cfg[k] - nq[k]), scale, nq[k]) for k in range(D)], f32)
        step = ((c - nq).astype(f32) / G).astype(f32)
        if not edge_free(nq, step):
            if dynamic_domain:
                shrink(t, near)
            continue
        last = add(t, c, near)
        if last < 0:
            return result(False, it)
        if dynamic_domain:
            grow(t, near)
        onear, d2o = _nearest(np.stack(trees[o]), c)
        n_ext = int(np.ceil(f32(np.sqrt(f32(d2o)) / f32(range_))))
        cur = c
        if n_ext > 0:
            vec = ((trees[o][onear] - cur).astype(f32) / f32(n_ext)).astype(f32)
            stepc = (vec / G).astype(f32)
        e = 0
        while e < n_ext:
            if not edge_free(cur, stepc):
                break
            nxt = (cur + vec).astype(f32)
            idx = add(t, nxt, last)
            if idx < 0:
                return result(False, it)
            cur, last = nxt, idx
            e += 1
        if e == n_ext:
            return result(True, it, (t, last, onear))
