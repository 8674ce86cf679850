"""Verification of warm starts (CPU, the fit core's host build).

A warm start goes straight to the FP64 active-set phase with the predicted
passive set; the KKT test at its end verifies the result, and a prediction
that fails it is not discarded: the iterations go on from it.  That is
equivalent to a cold start because the normal matrix AtA0 is positive
definite (the problem strictly convex, its optimum unique) and the active-set
steps from any feasible point decrease the objective strictly at each added
index and never revisit a passive set, so they end -- in exact arithmetic --
at the unique KKT point whatever the start (Lawson and Hanson, Solving Least
Squares Problems, 1974, ch. 23; the FNNLS steps of Bro and De Jong, 1997, are
the same steps on the normal equations).  The FP64 KKT test decides the end
with the reference's eps.  As a fallback for what exact arithmetic does not
cover, a warm start whose iterations reach max_iter is dropped and the
channel refitted from x = 0 (MF_F_WARM_FALLBACK).

These tests start the fast path from deliberately wrong predictions and
check every channel against the oracle's cold fit."""
import numpy as np
import pytest

import parity
from oracle_binding import hostcore_fast_from, oracle_fit, oracle_opts


def _batch(n, seed=3):
    from test_regression_amd import workloads
    w = workloads.config("c2", "zero")
    w.seed += seed
    ct = w.chtypes()
    s, _, _ = w.generate(n)
    return ct, s.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("start", ["oracle", "random_sets", "all_passive", "empty"])
def test_wrong_predictions_converge_to_the_cold_optimum(pkg, start):
    from test_regression_amd import _abi
    ct, s = _batch(3000)
    ref = oracle_fit(ct, s, None, oracle_opts(), nthreads=4)
    xr = ref["x"].T
    rng = np.random.default_rng(11)
    if start == "oracle":
        xs = xr.copy()
    elif start == "random_sets":
        xs = np.where(rng.random(xr.shape) < 0.5, rng.uniform(0.1, 50, xr.shape), 0.0)
    elif start == "all_passive":
        xs = np.full(xr.shape, 5.0)
    else:
        xs = np.zeros(xr.shape)
    x, chi2, fl, it = hostcore_fast_from(ct, s, xs, True, _abi.default_opts(precision=1))
    rej = (fl & _abi.FLAGS["warm_reject"]) != 0
    print(f"{start}: rejected {rej.mean():.3f}")
    if start == "oracle":
        assert rej.mean() < 0.01      # the optimum's own set passes its KKT test
    elif start == "random_sets":
        assert rej.mean() > 0.3
    elif start == "all_passive":      # mostly steps back to the optimum's set
        assert rej.mean() > 0.1
    assert not np.any(fl & _abi.FLAGS["warm_fallback"])
    rep = parity.gate(x, chi2, ref, 1e-4, iters=it, compare_outer=False,
                      label=f"warm from {start}")
    # and the rejected channels alone
    parity.gate(x, chi2, ref, 1e-4, iters=it, compare_outer=False, mask=rej,
                label=f"warm from {start}, rejected channels")


def test_fallback_to_cold_start_at_max_iter(pkg):
    """With max_iter 4, a prediction of one wrong passive index makes most
    channels need more adds than that: they reach the cap, are refitted from
    x = 0 and flagged, and their results are those of the cold FP64
    iterations under the same cap (a start from the empty set), bit for bit."""
    from test_regression_amd import _abi
    ct, s = _batch(2000, seed=5)
    o = _abi.default_opts(precision=1, max_iter=4)
    n = s.shape[1]
    xs = np.zeros((10, n))
    xs[0] = 5.0          # BX -5 alone: far from any channel's optimum
    xw, cw, fw, iw = hostcore_fast_from(ct, s, xs, True, o)
    x0, c0, f0, i0 = hostcore_fast_from(ct, s, np.zeros_like(xs), True, o)
    fb = (fw & _abi.FLAGS["warm_fallback"]) != 0
    print(f"fallback {fb.mean():.3f}")
    assert fb.mean() > 0.5
    # an empty prediction is the cold start itself: it never falls back
    assert not np.any(f0 & _abi.FLAGS["warm_fallback"])
    assert np.any(f0 & _abi.FLAGS["fnnls_cap"])
    assert np.array_equal(xw[:, fb], x0[:, fb]) and np.array_equal(cw[fb], c0[fb])
    # without the cap nothing falls back, and every channel ends at the optimum
    xw, cw, fw, iw = hostcore_fast_from(ct, s, xs, True, _abi.default_opts(precision=1))
    assert not np.any(fw & _abi.FLAGS["warm_fallback"])
    ref = oracle_fit(ct, s, None, oracle_opts(), nthreads=4)
    parity.gate(xw, cw, ref, 1e-4, iters=iw, compare_outer=False, label="one wrong index")
