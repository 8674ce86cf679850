"""Lawson-Hanson NNLS with the semantics of the reference's Eigen::NNLS class
(nnls/interface/eigen_nnls.h:336-414): the oracle's restatement
(or_lh_nnls, Householder QR updated column by column) and the product's
(fit_core.hpp lh_nnls, masked Cholesky of the passive block) built for the
host.

Pins of the oracle:
  cold start       brute force over all passive sets (m = 6), scipy's
                   Lawson-Hanson and the KKT conditions (m = 10)
  identity A       closed forms: the k largest positive b enter one per solve,
                   max_iter (a count of solves) stops after k of them, and a
                   warm x whose gradient b - x is nowhere above eps is
                   returned unchanged after no solve (setX, :124, :361)
  fit (DoFit with the alternative of PulseChiSqSNNLS.cpp:253-257)
                   zero pulse covariance: the optimum every iteration, as the
                   pinned inplace_fnnls oracle; live covariance: each
                   iteration ends at the NNLS optimum of its problem or keeps
                   its warm start when no gradient is above eps
"""
import ctypes as C
import itertools

import numpy as np
import pytest
from scipy.optimize import nnls as scipy_nnls

from oracle_binding import (OR_SOLVER_INPLACE, OR_SOLVER_LH, OrStats, dptr, hostcore,
                            hostcore_fit, oracle, oracle_fit, oracle_opts)


def _lh(A, b, x0=None, max_iter=-1, eps=1e-11):
    m = A.shape[0]
    Ar = np.zeros((10, 10))
    Ar[:m, :m] = A
    bb = np.zeros(10)
    bb[:m] = b
    x = np.zeros(10)
    if x0 is not None:
        x[:m] = x0
    st = OrStats()
    ok = oracle().or_lh_nnls(dptr(Ar), dptr(bb), dptr(x), m, eps, max_iter, C.byref(st))
    return x[:m], ok, st


def _hc_lh(A, b, x0=None, max_iter=0, eps=1e-11):
    m = A.shape[0]
    Ac = np.zeros((10, 10))
    Ac[:m, :m] = A.T                    # column-major, leading dimension 10
    bb = np.zeros(10)
    bb[:m] = b
    x = np.zeros(10)
    if x0 is not None:
        x[:m] = x0
    ok = hostcore().hc_nnls(dptr(np.ascontiguousarray(Ac)), dptr(bb), dptr(x), m, 2, eps,
                            max_iter)
    return x[:m], ok


def _brute(A, b):
    n = A.shape[1]
    best, bx = np.sum(b * b), np.zeros(n)
    for k in range(1, n + 1):
        for S in itertools.combinations(range(n), k):
            xs, *_ = np.linalg.lstsq(A[:, S], b, rcond=None)
            if np.all(xs >= 0):
                r = A[:, S] @ xs - b
                if float(r @ r) < best - 1e-12:
                    best, bx = float(r @ r), np.zeros(n)
                    bx[list(S)] = xs
    return bx


def test_lh_cold_start_optimum(pkg):
    rng = np.random.default_rng(21)
    for trial in range(200):
        m = 6 if trial < 100 else 10
        A = rng.standard_normal((m, m))
        b = rng.standard_normal(m) * 5
        x, ok, st = _lh(A, b)
        assert ok
        ref = _brute(A, b) if m == 6 else scipy_nnls(A, b)[0]
        w = A.T @ (b - A @ x)
        scale = np.abs(A.T @ b).max() + 1
        assert np.all(x >= 0)
        assert np.all(w[x == 0] <= 1e-9 * scale) and np.all(np.abs(w[x > 0]) <= 1e-9 * scale)
        assert np.allclose(x, ref, rtol=1e-8, atol=1e-8), trial
        assert st.inner >= st.outer >= np.count_nonzero(x)
        xh, okh = _hc_lh(A, b)
        assert okh and np.abs(xh - x).max() <= 1e-10 * max(1.0, np.abs(x).max())


def test_lh_identity_closed_forms(pkg):
    b = np.array([3.0, -1.0, 7.0, 5.0, 0.5, -2.0, 4.0])
    A = np.eye(7)
    x, ok, st = _lh(A, b)
    assert ok and np.array_equal(x, np.maximum(b, 0.0))
    assert st.inner == st.outer == 5
    # max_iter counts solves: after 3 the three largest have entered
    x, ok, st = _lh(A, b, max_iter=3)
    assert not ok and st.inner == 3
    assert np.array_equal(x, np.where(b >= 4.0, b, 0.0))
    # warm start: the gradient b - x0 is nowhere above eps: x0 comes back
    x0 = np.array([3.5, 0.0, 7.5, 5.0, 0.5, 0.0, 4.2])
    x, ok, st = _lh(A, b, x0=x0)
    assert ok and st.inner == 0 and np.array_equal(x, x0)
    # warm start with one positive gradient: that index enters, the feasible
    # solve replaces all of x (the free entries become 0)
    x0 = np.array([3.5, 0.0, 6.0, 5.0, 0.5, 0.0, 4.2])
    x, ok, st = _lh(A, b, x0=x0, max_iter=1)
    assert not ok and st.inner == 1
    assert np.array_equal(x, np.where(np.arange(7) == 2, 7.0, 0.0))
    for x0, mi in ((None, -1), (None, 3), (np.array([3.5, 0.0, 6.0, 5.0, 0.5, 0.0, 4.2]), 1)):
        xo, oko, _ = _lh(A, b, x0=x0, max_iter=mi)
        xh, okh = _hc_lh(A, b, x0=x0, max_iter=max(mi, 0))
        assert oko == okh and np.array_equal(xo, xh)


def _whitened(t, x, s):
    from test_legacy import _whitened as w
    return w(t, x, s)


def test_lh_fit_zero_pi_equals_inplace(pkg):
    from test_regression_amd import workloads
    w = workloads.config("c2", "zero")
    ct = w.chtypes()
    s, t, _ = w.generate(3000)
    lh = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LH), nthreads=8)
    ip = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_INPLACE), nthreads=8)
    scale = np.maximum(np.abs(ip["x"]).max(axis=1), 1.0)
    assert (np.abs(lh["x"] - ip["x"]).max(axis=1) <= 1e-10 * scale).all()
    assert np.array_equal(lh["n_minimize"], ip["n_minimize"])


def test_lh_fit_live_pi_optimum_or_kept_warm_start(pkg):
    from test_regression_amd import workloads
    w = workloads.config("c2", "live")
    ct = w.chtypes()
    s, t, _ = w.generate(150)
    full = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LH), nthreads=8)
    nm = full["n_minimize"]
    prev = {k: oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LH, max_minimize=int(k) - 1),
                          nthreads=8)["x"] for k in np.unique(nm)}
    kept = optimal = 0
    for c in range(s.shape[1]):
        xp, x = prev[nm[c]][c], full["x"][c]
        A, b = _whitened(ct[0], xp, s[:, c])
        wv = A.T @ (b - A @ x)
        scale = np.abs(A.T @ b).max()
        if np.array_equal(x, xp):
            assert np.all(A.T @ (b - A @ xp) < 1e-11)    # no gradient above eps
            kept += 1
        else:
            assert np.all(x >= 0)
            assert np.all(np.where(x > 0, np.abs(wv), np.maximum(wv, 0)) <= 1e-9 * scale), c
            optimal += 1
    assert kept > 0 and optimal > 0


@pytest.mark.parametrize("cfg,pulse_cov", [("c2", "zero"), ("c2", "live"), ("c3", "live")])
def test_lh_core_matches_oracle(pkg, cfg, pulse_cov):
    from test_regression_amd import _abi, workloads
    w = workloads.config(cfg, pulse_cov)
    ct = w.chtypes()
    s, t, _ = w.generate(2000)
    ref = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LH, compute_errors=1), nthreads=8)
    x, chi2, fl, it, err = hostcore_fit(
        ct, s, t, _abi.default_opts(solver=_abi.MF_SOLVER_LAWSON_HANSON, compute_errors=1))
    xr = ref["x"].T
    scale = np.maximum(np.abs(xr).max(axis=0), 1.0)
    assert ((x > 0) == (xr > 0)).all()
    assert (np.abs(x - xr).max(axis=0) <= 1e-10 * scale).all()
    assert np.array_equal(it & 255, ref["n_minimize"])
    assert np.array_equal((it >> 8) & 255, ref["n_outer"])
    assert np.array_equal(it >> 16, ref["n_inner"])
    assert np.array_equal(fl, ref["flags"])
    e, er = err.T, ref["err"]
    assert np.array_equal(np.isnan(e), np.isnan(er))
    ok = ~np.isnan(er) & (er != 0)
    assert np.all(np.abs(e[ok] - er[ok]) <= 1e-9 * np.abs(er[ok]))
