"""Comparison helpers shared by the GPU parity tests."""
import numpy as np


# FP32 resolution of robot geometry, in metres. The FP32 path and the FP64
# oracle read the same FP32 inputs, yet an FP32 evaluation places a sphere
# away from its FP64 position by up to ~1.4e-7 m even when every operation
# after the FP32 sin/cos is exact and only the final position is rounded,
# and by up to ~4.6e-7 m with ordinary FP32 arithmetic (Panda, Fetch and
# Baxter, 4,000 random configurations; test_oracle_geometry.py pins both
# figures). A distance-like result therefore carries an absolute error of a
# few 1e-7 m whatever its size. The 1e-5 relative bar is applied with the
# denominator max(|ref|, TERM_FLOOR) for a clearance (1e-5 x 0.05 m =
# 5e-7 m), and a sum of hinge terms gets the sum of its terms' bars:
# max(|ref|, weight x TERM_FLOOR x (number of non-zero terms)), the term
# count taken from the FP64 oracle.
TERM_FLOOR = 0.05


def clearance_den(ref):
    """Denominator of the 1e-5 relative bar for a clearance (metres)."""
    return np.maximum(np.abs(np.asarray(ref, np.float64)), TERM_FLOOR)


def cost_den(ref, weight, n_terms):
    """Denominator of the 1e-5 relative bar for a weighted sum of hinge terms."""
    n = np.maximum(np.asarray(n_terms, np.float64), 1.0)
    return np.maximum(np.abs(np.asarray(ref, np.float64)), weight * TERM_FLOOR * n)


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def assert_values(got, ref, amb, den, what, tol=1e-5):
    """Relative error |got - ref| / den below tol at every result the oracle
    does not flag as value-ambiguous (an ESDF lookup at a voxel face or the
    grid's guard boundary, where the value jumps and FP32 may land on either
    side); flagged results are counted and printed, and must stay rare."""
    got = np.asarray(got, np.float64)
    err = np.abs(got - ref) / den
    amb = np.asarray(amb, bool)
    bad = (err >= tol) & ~amb
    worst = float(err[~amb].max()) if (~amb).any() else 0.0
    banner(f"{what}: n={err.size} max rel err {worst:.2e} (unflagged), flagged {int(amb.sum())}, "
           f"flagged and off {int(((err >= tol) & amb).sum())}")
    assert not bad.any(), f"{what}: {int(bad.sum())} results off by >= {tol} relative outside the flagged set"
    assert amb.mean() < 0.05, f"{what}: {amb.mean():.3f} of the results flagged"


def traj_rows(flat, off, D):
    """[waypoints, D] rows of a ragged dim-major trajectory array."""
    rows = []
    for b in range(len(off) - 1):
        T = off[b + 1] - off[b]
        rows.append(flat[off[b] * D:off[b + 1] * D].reshape(D, T).T)
    return np.concatenate(rows) if rows else np.zeros((0, D), flat.dtype)


def traj_terms(o, theta, off, D, params):
    """Non-zero hinge terms summed over each trajectory's waypoints."""
    nt = o.hinge_terms(traj_rows(theta, off, D), params)
    return np.add.reduceat(nt, np.asarray(off[:-1])) if len(off) > 1 else nt[:0]


def fd_bar(o, rows, params_unit, weight, eps):
    """Per-entry bar for an FP32 central-FD gradient kernel against the FP64
    FD: each entry is w (C+ - C-) / (2 eps) of two FP32 cost evaluations at
    q[d] += eps and then q[d] -= 2 eps (FP32), each within 1e-5 cost_den of
    FP64, so an entry may differ by w 1e-5 (den(C+) + den(C-)) / (2 eps).
    rows: [waypoints, D]; params_unit: the cost parameters at weight 1."""
    rows = np.asarray(rows, np.float32)
    n, D = rows.shape
    eps = np.float32(eps)
    qp = np.repeat(rows[:, None, :], D, axis=1)
    qm = qp.copy()
    idx = np.arange(D)
    plus = (rows + eps).astype(np.float32)
    qp[:, idx, idx] = plus
    qm[:, idx, idx] = (plus - np.float32(2 * eps)).astype(np.float32)
    cp, _, tp = o.hinge_cost_amb(qp.reshape(-1, D), params_unit, with_terms=True)
    cm, _, tm = o.hinge_cost_amb(qm.reshape(-1, D), params_unit, with_terms=True)
    den = cost_den(cp, 1.0, tp) + cost_den(cm, 1.0, tm)
    return (weight * 1e-5 * den / (2 * float(eps))).reshape(n, D)
