"""FP64 restatement of the STOMP optimiser loop (test infrastructure only).

Follows planning/stomp.cu:565-930 (solve_runtime_scene) and
planning/stomp_precompute.cpp:24-118 (the R, L and M matrices), with two
documented differences of the product that this restatement shares:

* Sampling. The reference draws z ~ N(0, 1) from cuRAND XORWOW streams
  (stomp.cu:48-53, 100-118). The product uses the counter-based
  Philox4x32-10 generator (Salmon, Moraes, Dror, Shaw, "Parallel random
  numbers: as easy as 1, 2, 3", SC'11) with the Box-Muller transform; the
  generator here is pinned by the Random123 known-answer vectors in the tests.
  A counter (t >> 2, r | d << 16, b, iteration) and key (seed low, seed high)
  name every draw, so results do not depend on launch order.
* Matrices. R^-1, L and M are formed in FP64 (the reference uses FP32 Eigen).

Everything else -- clamped samples around the optimum with pinned endpoints,
the reuse of the lowest-cost rollouts of the previous population, the
optimum as the last slot, per-(dof, waypoint) exponentiated-cost
probabilities, the M-projected update with pinned endpoints and no clamp,
and the best-valid / streak / polish bookkeeping -- is the reference's.
"""
from __future__ import annotations

import numpy as np

MASK = np.uint64(0xFFFFFFFF)
PHILOX_M0, PHILOX_M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
PHILOX_W0, PHILOX_W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)


def philox4x32_10(ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
    """ctr [..., 4], key [..., 2] (uint32 values) -> [..., 4] uint32.
    Ten rounds of (hi, lo) = M * c, c' = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0),
    the key advanced by the Weyl constants between rounds (Salmon et al., Fig. 2)."""
    c = np.asarray(ctr, dtype=np.uint64) & MASK
    k = np.asarray(key, dtype=np.uint64) & MASK
    c0, c1, c2, c3 = c[..., 0], c[..., 1], c[..., 2], c[..., 3]
    k0, k1 = k[..., 0], k[..., 1]
    for rnd in range(10):
        if rnd:
            k0 = (k0 + PHILOX_W0) & MASK
            k1 = (k1 + PHILOX_W1) & MASK
        p0 = PHILOX_M0 * c0
        p1 = PHILOX_M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def uniform_open(w: np.ndarray) -> np.ndarray:
    """(w >> 9 + 1/2) 2^-23: in (0, 1), exact in FP32 and FP64."""
    return ((w.astype(np.uint64) >> np.uint64(9)).astype(np.float64) + 0.5) * 2.0 ** -23


def normals(seed: int, it: int, b: int, r: int, d: int, T: int) -> np.ndarray:
    """z[t], t < T, of row (b, r, d) at iteration `it`: Box-Muller on the words
    (0, 1) for t % 4 in {0, 1} (cos, sin) and (2, 3) for {2, 3}."""
    t = np.arange(T)
    ctr = np.stack([t >> 2, np.full(T, r | (d << 16)), np.full(T, b), np.full(T, it)], axis=-1)
    key = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF])
    w = philox4x32_10(ctr, np.broadcast_to(key, (T, 2)))
    j = t & 3
    pair = j >> 1
    u1 = uniform_open(w[t, 2 * pair])
    u2 = uniform_open(w[t, 2 * pair + 1])
    rad = np.sqrt(-2.0 * np.log(u1))
    ang = 2.0 * np.pi * u2
    return np.where((j & 1) == 0, rad * np.cos(ang), rad * np.sin(ang))


def control_R(T: int, dt: float) -> np.ndarray:
    """dt A^T A, central T x T block, A the 5-point acceleration stencil on a
    grid padded by 6 each side (stomp_precompute.cpp:24-44)."""
    pad = 6
    N = T + 2 * pad
    c = np.array([-1.0, 16.0, -30.0, 16.0, -1.0]) / 12.0 / (dt * dt)
    A = np.zeros((N, N))
    for i in range(N):
        for j, v in zip(range(i - 2, i + 3), c):
            if 0 <= j < N:
                A[i, j] = v
    return (dt * A.T @ A)[pad:pad + T, pad:pad + T]


def stomp_matrices(T: int, dt: float, noise_scale: float = 1.0):
    """(R, L, M): L the Cholesky factor of R^-1 rescaled to max entry 1/T,
    symmetrised, times noise_scale (stomp_precompute.cpp:48-74); M = R^-1 with
    each column scaled so its maximum is 1/T (stomp_precompute.cpp:76-101)."""
    R = control_R(T, dt)
    Rinv = np.linalg.inv(R + 1e-8 * np.eye(T))
    S = Rinv * ((1.0 / T) / max(Rinv.max(), 1e-12))
    S = 0.5 * (S + S.T)
    L = np.linalg.cholesky(S) * (noise_scale if noise_scale > 0 else 1.0)
    M = Rinv * ((1.0 / T) / np.maximum(Rinv.max(axis=0), 1e-12))[None, :]
    return R, L, M


def weights(tot: np.ndarray, h: float) -> np.ndarray:
    """Per-(dof, waypoint) probabilities over the population axis 0
    (stomp.cu:310-350): exp(-h (c - min) / max(max - min, 1e-8)), normalised
    with the sum floored at 1e-12."""
    mn, mx = tot.min(axis=0), tot.max(axis=0)
    un = np.exp(-h * (tot - mn) / np.maximum(mx - mn, 1e-8))
    return un / np.maximum(un.sum(axis=0), 1e-12)


def kept_slots(totals: np.ndarray, n_old: int) -> np.ndarray:
    """The n_old cheapest slots, cheapest first, ties to the lower slot
    (the selection sort of stomp.cu:135-149)."""
    return np.argsort(totals, kind="stable")[:n_old]


class StompOracle:
    """One trajectory's STOMP state. `state_cost(x [n, D]) -> [n]` gives the
    weighted world + self hinge cost of configurations; `valid(theta [T, D])`
    the densified two-phase verdict."""

    def __init__(self, theta_dt: np.ndarray, lo, hi, state_cost, valid, *, Rn=30, Ro=20, dt=0.1, w_u=0.1, h=2.0,
                 noise_scale=1.0, seed=0x5BE5, b=0, polish=3, eps=1e-4):
        self.th = np.array(theta_dt, np.float64)      # [D, T]
        self.D, self.T = self.th.shape
        self.lo, self.hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
        self.cost_fn, self.valid_fn = state_cost, valid
        self.Rn, self.Ro, self.dt, self.w_u, self.h = Rn, Ro, dt, w_u, h
        self.seed, self.b, self.it = seed, b, 0
        self.Rm, self.L, self.M = stomp_matrices(self.T, dt, noise_scale)
        self.prev = None                                # (rollouts [R, D, T], totals [R])
        self.polish0, self.eps = polish, eps
        self.total = self.traj_total(self.th)
        ok = valid(self.th.T)
        self.best = self.th.copy() if ok else None
        self.best_cost = self.total if ok else np.inf
        self.prev_cost, self.streak, self.polish, self.done, self.iters = np.inf, 0, polish, False, 0
        if ok and polish <= 0:
            self.done = True

    def control_elem(self, x: np.ndarray) -> np.ndarray:
        """w x_t (R x)_t per (..., d, t) (stomp.cu:255-272)."""
        return self.w_u * x * (x @ self.Rm.T)

    def traj_total(self, x_dt: np.ndarray) -> float:
        return float(self.cost_fn(x_dt.T).sum() + self.control_elem(x_dt).sum())

    def population(self):
        D, T, Rn, Ro = self.D, self.T, self.Rn, self.Ro
        first = self.prev is None
        nf = Rn + Ro if first else Rn
        fresh = np.empty((nf, D, T))
        for r in range(nf):
            for d in range(D):
                z = normals(self.seed, self.it, self.b, r, d, T)
                e = self.L @ z
                e[0] = e[-1] = 0.0
                fresh[r, d] = np.clip(self.th[d] + e, self.lo[d], self.hi[d])
        parts = [fresh]
        if not first:
            rolls, totals = self.prev
            parts.append(rolls[kept_slots(totals, Ro)])
        parts.append(self.th[None])
        return np.concatenate(parts)

    def iterate(self):
        if self.done:
            return
        X = self.population()                              # [R, D, T]
        R_, D, T = X.shape
        s = self.cost_fn(X.transpose(0, 2, 1).reshape(-1, D)).reshape(R_, T)
        ctrl = self.control_elem(X)
        tot = s[:, None, :] + ctrl                         # [R, D, T] (stomp.cu:281-303)
        totals = s.sum(axis=1) + ctrl.sum(axis=(1, 2))
        self.prev = (X, totals)
        p = weights(tot, self.h)
        delta = (p * (X - self.th[None])).sum(axis=0)
        delta[:, 0] = delta[:, -1] = 0.0                   # stomp.cu:358-380
        step = delta @ self.M.T
        step[:, 0] = step[:, -1] = 0.0                     # stomp.cu:390-403
        self.th = self.th + step
        self.it += 1
        self.iters += 1
        self.total = self.traj_total(self.th)
        ok = self.valid_fn(self.th.T)
        if ok and self.total < self.best_cost:             # stomp.cu:868-875
            self.best, self.best_cost = self.th.copy(), self.total
        if self.best is not None:                          # stomp.cu:881-887
            impr = self.prev_cost - self.best_cost
            self.streak = self.streak + 1 if (0.0 <= impr < self.eps) else 0
            self.prev_cost = self.best_cost
            self.polish -= 1
            if self.streak >= 2 or self.polish <= 0:
                self.done = True
