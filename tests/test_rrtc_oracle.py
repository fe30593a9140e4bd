"""pRRTC's deterministic parts on the CPU: the restatement (oracle/
rrtc_oracle.py) against what mathematics and published vectors fix, the
product's host-side Halton base table against it, and the C ABI's argument
checks (no device work)."""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest

import rrtc_oracle as RO
from prismplan_amd import _ppl
from prismplan_amd import substrate as S


def test_splitmix64_published_vectors():
    # Vigna's splitmix64.c reference outputs
    x, out = 0, []
    for _ in range(4):
        x, v = RO.splitmix64(x)
        out.append(v)
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    x, out = 1234567, []
    for _ in range(3):
        x, v = RO.splitmix64(x)
        out.append(v)
    assert out == [0x599ED017FB08FC85, 0x2C73F08458540FA5, 0x883EBCE5A3F27C77]


def test_radical_inverse_closed_form():
    assert RO.radical_inverse(1, 2) == Fraction(1, 2)
    assert [RO.radical_inverse(k, 3) for k in range(1, 6)] == [Fraction(1, 3), Fraction(2, 3), Fraction(1, 9),
                                                                Fraction(4, 9), Fraction(7, 9)]
    assert RO.radical_inverse(11, 2) == Fraction(13, 16)   # 1011 -> 0.1101


def test_halton_generator_is_the_radical_inverse():
    """Every FP32 sample equals the correctly rounded radical inverse of k
    (the generator's numerator and denominator stay exact below 2^24)."""
    n = 1500
    seq = RO.halton_sequence(RO.PRIMES, n)
    for i, b in enumerate(RO.PRIMES):
        want = []
        for k in range(1, n + 1):
            f = RO.radical_inverse(k, b)
            want.append(np.float32(f.numerator) / np.float32(f.denominator))
        np.testing.assert_array_equal(seq[:, i], np.array(want, np.float32), err_msg=f"base {b}")
    assert np.all((seq > 0) & (seq < 1))


def test_halton_bases_table_matches_restatement():
    for seed in (1, 7, 2**63 + 5):
        got = S.rrtc_halton_bases(seed, 40, 7)
        np.testing.assert_array_equal(got, RO.halton_bases(seed, 40, 7))
    full = S.rrtc_halton_bases(3, 64, 16)
    np.testing.assert_array_equal(full[0], np.float32(RO.PRIMES))
    for row in full:
        assert sorted(row.tolist()) == list(map(float, RO.PRIMES))
    # different workgroups draw different permutations
    assert len({tuple(r) for r in full}) > 60


def test_scale_to_limits_and_segment_states():
    h = np.float32([0.0, 0.5, 0.25])
    q = RO.scale_to_limits(h, [-1.0, -2.0, 0.0], [1.0, 2.0, 4.0])
    np.testing.assert_array_equal(q, np.float32([-1.0, 0.0, 1.0]))
    st = RO.segment_states([0.0, 0.0], [1.0, -2.0], 4)
    np.testing.assert_allclose(st, [[0, 0], [0.25, -0.5], [0.5, -1.0], [0.75, -1.5], [1.0, -2.0]])


def test_shortcut_pairs_and_greedy():
    p = np.float32([[0, 0], [0.1, 0], [0.2, 0], [0.3, 0]])
    pairs = RO.shortcut_pairs(p, 0.05)
    assert [(i, j) for i, j, _ in pairs] == [(0, 2), (0, 3), (1, 3)]
    i, j, st = pairs[0]
    assert st.shape == (3, 2)   # |q2 - q0| = 0.2 -> 4 segments, 3 interior states
    np.testing.assert_allclose(st[:, 0], [0.05, 0.1, 0.15], rtol=1e-6)
    assert RO.shortcut_greedy(4, {(0, 2): True, (0, 3): True, (1, 3): True}) == [0, 3]
    assert RO.shortcut_greedy(4, {}) == [0, 1, 2, 3]
    assert RO.shortcut_greedy(5, {(0, 2): True, (2, 4): True, (0, 4): False, (0, 3): False}) == [0, 2, 4]
    assert RO.shortcut_greedy(2, {}) == [0, 1]


def test_rrtc_settings_defaults_and_arguments():
    s = S.rrtc_settings()
    assert (s.max_samples, s.max_iters, s.num_new_configs, s.granularity, s.balance, s.dynamic_domain) == \
        (1000000, 1000000, 600, 16, 1, 1)
    assert abs(s.range - 0.5) < 1e-7 and abs(s.time_limit_ms - 4000.0) < 1e-3
    assert abs(s.dd_alpha - 1e-4) < 1e-9 and s.dd_radius == 4.0 and s.dd_min_radius == 1.0
    assert abs(s.valid_segment_length - 0.025) < 1e-9 and abs(s.shortcut_collision_margin - 5e-4) < 1e-9
    with pytest.raises(AttributeError):
        S.rrtc_settings(no_such_field=1)
    L = _ppl.lib()
    h = C.c_void_p()
    assert L.ppl_rrtc_buffers_create(0, 100, 8, 16, C.byref(h)) == _ppl.PPL_ERR_INVALID
    assert L.ppl_rrtc_buffers_create(7, 100, 8, 65, C.byref(h)) == _ppl.PPL_ERR_INVALID
    assert L.ppl_rrtc_buffers_create(17, 100, 8, 16, C.byref(h)) == _ppl.PPL_ERR_INVALID
    assert L.ppl_rrtc_buffers_destroy(None) == _ppl.PPL_OK
    assert L.ppl_rrtc_solve(None, None, None, None, 1, None, None, None, 0, None, None) == _ppl.PPL_ERR_INVALID
    assert L.ppl_shortcut_path(None, None, None, 0, C.c_float(0.025), C.c_float(0), None, None, None) == \
        _ppl.PPL_ERR_INVALID
    assert L.ppl_rrtc_halton_bases(C.c_ulonglong(1), 2, 17, None) == _ppl.PPL_ERR_INVALID


def test_single_workgroup_run_in_free_space():
    """In free space the sequential run connects at iteration 1: workgroup 0
    of 1 grows the goal tree toward the first Halton sample (steered to at most
    `range`), then extends from that node to the start in ceil(d / range)
    equal steps (pRRTC.cu:1602-1627, 1668-1677, 1814-1960); the path is start,
    the extension nodes back to the steered node, goal (pRRTC.cu:2167-2186)."""
    lo = np.float32([-2.0, -1.0, -3.0])
    hi = np.float32([2.0, 1.0, 3.0])
    start = np.float32([0.5, 0.2, -1.0])
    goal = np.float32([-0.4, 0.1, 1.5])
    res = RO.rrtc_single_workgroup(start, goal[None], lo, hi, lambda s: False, range_=0.5)
    assert res["solved"] and res["iters"] == 1 and res["reached_goal"] == 0
    h = np.float32([1 / 3, 1 / 5, 1 / 7])                       # first samples of bases 3, 5, 7
    sample = lo + h * (hi - lo)
    d = np.linalg.norm(sample - goal)
    c = goal + min(1.0, 0.5 / d) * (sample - goal)
    n_ext = int(np.ceil(np.linalg.norm(start - c) / 0.5))
    path = res["path"]
    assert len(path) == n_ext + 3
    np.testing.assert_array_equal(path[0], start)
    np.testing.assert_array_equal(path[-1], goal)
    np.testing.assert_allclose(path[-2], c, atol=1e-6)
    np.testing.assert_allclose(path[1], start, atol=1e-5)          # the extension ends on the start
    seg = np.linalg.norm(np.diff(path.astype(np.float64), axis=0), axis=1)
    assert seg[1:].max() <= 0.5 + 1e-6
    assert res["start_tree_size"] == 1 and res["goal_tree_size"] == n_ext + 2


def test_single_workgroup_run_grows_around_a_wall():
    """A wall between start and goal (a 1-D slab the edge states may not
    enter): the run needs several iterations, and its path stays within
    `range` per segment and off the wall at every checked state."""
    lo, hi = np.float32([-1.0, -1.0]), np.float32([1.0, 1.0])
    start, goal = np.float32([-0.6, 0.0]), np.float32([0.6, 0.0])

    def wall(states):   # x in [-0.05, 0.05] and y < 0.5
        return bool(np.any((np.abs(states[:, 0]) <= 0.05) & (states[:, 1] < 0.5)))

    res = RO.rrtc_single_workgroup(start, goal[None], lo, hi, wall, range_=0.3, granularity=8, max_iters=5000)
    assert res["solved"] and res["iters"] > 1
    p = res["path"]
    states = np.concatenate([RO.segment_states(p[i], p[i + 1], 8) for i in range(len(p) - 1)])
    assert not wall(states)
    assert np.linalg.norm(np.diff(p.astype(np.float64), axis=0), axis=1).max() <= 0.3 + 1e-6
