"""Comparison helpers shared by the GPU parity tests."""
import numpy as np


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
