"""Restatement of pRRTC's deterministic parts (test infrastructure only).

Follows planning/pRRTC.cu (Halton sampling :955-995, edge states :1668-1677
and :1852-1866, the tree walk :1962-1988, the goal -> start assembly
:2167-2186) and planning/shortcut.cuh:16-274 (skip-edge states and the
greedy pass). The planner itself is concurrent and not reproducible state
for state; its outputs are checked against what these functions and the
FP64 collision oracle fix: bases, samples, edge states, path validity,
costs and the shortcut's choices.

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
