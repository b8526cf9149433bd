"""Shared checks of HIP decode results against the CPU oracle (test helper).

Contract, per precision:
  * FP64 executors: every marginal within 1e-12 relative of the oracle's,
    structural zeros exactly zero, classes bit-exact on every syndrome that
    is not a tie (top-2 relative gap <= 1e-12), tie flags equal.
  * FP32 executors (power-of-two scaling, FP64 re-check of low margins):
    marginals within ``tol`` (2e-5 by default) of the oracle row's maximum,
    impossible syndromes (all-zero oracle rows) exactly zero, classes
    bit-exact on every non-tie syndrome.
"""
import numpy as np

from oracle import tn_oracle as O


def check_f64(mar, cls, tie, ref, what=""):
    ref = np.asarray(ref)
    nz = ref > 0
    assert np.all(mar[~nz] == 0.0), what
    err = float(np.max(np.abs(mar[nz] - ref[nz]) / ref[nz])) if nz.any() else 0.0
    assert err < 1e-12, (what, err)
    valid = ref.max(axis=1) > 0
    ref_tie = O.tie_flags(ref[valid], 1e-12)
    ok = ref_tie == 0
    assert np.array_equal(cls[valid][ok], O.findmax(ref[valid])[ok]), what
    assert np.array_equal(tie[valid], ref_tie), what
    return err


def check_f32(mar, cls, ref, what="", tol=2e-5):
    """Returns (max error relative to the row maximum, number of ties)."""
    ref = np.asarray(ref)
    valid = ref.max(axis=1) > 0
    assert np.all(mar[~valid] == 0.0), what
    err = float(np.max(np.abs(mar[valid] - ref[valid]) / ref[valid].max(axis=1, keepdims=True))) \
        if valid.any() else 0.0
    assert err < tol, (what, err)
    ok = valid & (O.tie_flags(ref, 1e-12) == 0)
    assert np.array_equal(cls[ok], O.findmax(ref)[ok]), what
    return err, int((valid & ~ok).sum())
