"""The STOMP oracle (oracle/stomp_oracle.py) against known answers and the
mathematics that fixes its pieces: Random123's Philox4x32-10 known-answer
vectors, Box-Muller identities, the matrices' defining properties, closed
forms of the weighting, and the reference's selection order."""
import numpy as np

import pyoracle as PO
import stomp_oracle as SO
from prismplan_amd import workloads as W


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        got = SO.philox4x32_10(np.array(ctr), np.array(key))
        assert tuple(int(v) for v in got) == want


def test_box_muller_pairs_and_moments():
    T = 8
    z = SO.normals(0x5BE5, 3, 1, 2, 4, T)
    # t = 4k, 4k+1 share words (0, 1): z0^2 + z1^2 = -2 ln u1
    t = np.arange(T)
    w = SO.philox4x32_10(np.stack([t >> 2, np.full(T, 2 | (4 << 16)), np.full(T, 1), np.full(T, 3)], -1),
                         np.broadcast_to(np.array([0x5BE5, 0]), (T, 2)))
    for k in range(0, T, 4):
        u1 = SO.uniform_open(w[k, 0])
        assert abs(z[k] ** 2 + z[k + 1] ** 2 + 2.0 * np.log(u1)) < 1e-12
        u1b = SO.uniform_open(w[k, 2])
        assert abs(z[k + 2] ** 2 + z[k + 3] ** 2 + 2.0 * np.log(u1b)) < 1e-12
    zs = np.concatenate([SO.normals(11, i, b, r, d, 64) for i in range(4) for b in range(2) for r in range(8)
                         for d in range(7)])
    n = zs.size
    assert abs(zs.mean()) < 5.0 / np.sqrt(n)
    assert abs(zs.var() - 1.0) < 5.0 * np.sqrt(2.0 / n)
    # every counter word matters: neighbouring rows and iterations differ
    a = SO.normals(11, 0, 0, 0, 0, 16)
    for other in (SO.normals(11, 1, 0, 0, 0, 16), SO.normals(11, 0, 1, 0, 0, 16), SO.normals(11, 0, 0, 1, 0, 16),
                  SO.normals(11, 0, 0, 0, 1, 16), SO.normals(12, 0, 0, 0, 0, 16)):
        assert np.abs(a - other).max() > 0.1


def test_stomp_matrices():
    for T, dt, ns in ((12, 0.1, 1.0), (51, 0.1, 1.0), (30, 0.05, 1.7)):
        R, L, M = SO.stomp_matrices(T, dt, ns)
        # one metric: the pyoracle and workload restatements give the same R
        np.testing.assert_allclose(R, PO.control_R(T, dt), rtol=1e-12, atol=1e-9 * np.abs(R).max())
        np.testing.assert_allclose(R, W.control_R_closed_form(T, dt), rtol=1e-12, atol=1e-9 * np.abs(R).max())
        A = R + 1e-8 * np.eye(T)
        # L L^T is R^-1 scaled so the largest entry is 1/T (times noise_scale^2)
        S = L @ L.T / ns ** 2
        assert np.allclose(np.triu(L, 1), 0.0)
        assert abs(S.max() - 1.0 / T) < 1e-12
        P = A @ S
        c = P[0, 0]
        np.testing.assert_allclose(P, c * np.eye(T), atol=1e-6 * c)
        # M = R^-1 with column j scaled to max 1/T: A M is diagonal, max over rows is 1/T
        np.testing.assert_allclose(M.max(axis=0), 1.0 / T, rtol=1e-12)
        AM = A @ M
        np.testing.assert_allclose(AM - np.diag(np.diag(AM)), 0.0, atol=1e-6 * np.abs(np.diag(AM)).max())


def test_weights_closed_forms():
    h = 2.0
    # equal costs: uniform
    np.testing.assert_allclose(SO.weights(np.full((5, 3), 7.0), h), 0.2)
    # two rollouts: [1, e^-h] / (1 + e^-h)
    p = SO.weights(np.array([[0.0], [1.0]]), h)[:, 0]
    np.testing.assert_allclose(p, np.array([1.0, np.exp(-h)]) / (1.0 + np.exp(-h)), rtol=1e-15)
    # invariant under c -> a c + b, a > 0
    rng = np.random.default_rng(3)
    c = rng.normal(size=(9, 4))
    np.testing.assert_allclose(SO.weights(3.0 * c - 5.0, h), SO.weights(c, h), rtol=1e-12)
    # the cheapest rollout has the largest weight
    assert np.all(np.argmax(SO.weights(c, h), axis=0) == np.argmin(c, axis=0))


def test_kept_slots_order():
    tot = np.array([5.0, 1.0, 3.0, 1.0, 0.5, 3.0])
    # selection sort, strict <: cheapest first, equal costs by slot
    np.testing.assert_array_equal(SO.kept_slots(tot, 4), [4, 1, 3, 2])
    np.testing.assert_array_equal(SO.kept_slots(tot, 0), np.zeros(0, int))


def test_oracle_iteration_invariants():
    """One trajectory in free space (no obstacles, no self pairs): endpoints
    stay, the update is M applied to a convex combination of noise, and the
    run's counters follow the reference's loop."""
    T, D = 16, 3
    lo, hi = -np.ones(D) * 3, np.ones(D) * 3
    th0 = np.linspace(np.array([-1.0, 0.5, 0.0]), np.array([1.0, -0.5, 2.0]), T).T
    o = SO.StompOracle(th0, lo, hi, lambda x: np.zeros(len(x)), lambda th: True, Rn=6, Ro=3, polish=3)
    assert o.best is not None and not o.done
    o.iterate()
    np.testing.assert_array_equal(o.th[:, [0, -1]], th0[:, [0, -1]])
    assert o.iters == 1 and o.polish == 2
    while not o.done:
        o.iterate()
    assert o.iters <= 3
