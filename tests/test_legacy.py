"""The legacy NNLS (legacy_multifit_cpu/src/PulseChiSqSNNLS.cpp:268-363), whose
passive set and column order carry over from one Minimize iteration to the
next and into the MINOS refits (:151-157).

Pins of the oracle's restatement (OR_SOLVER_LEGACY), each chosen so that a
plausible mistake fails it:
  zero pulse covariance  the covariance does not depend on x, so every
                         Minimize iteration solves the same NNLS problem: the
                         amplitudes equal the pinned inplace_fnnls oracle's,
                         the carried-over passive set makes later iterations
                         cost exactly one solve each and no additions (a
                         reset passive set, or a KKT pass before the first
                         solve, breaks the counts)
  live pulse covariance  each Minimize iteration ends at the exact NNLS
                         optimum (KKT conditions) of the problem built from
                         the previous iteration's amplitudes; the same check
                         fails for inplace_fnnls, whose warm start keeps x but
                         resets the passive set, so it detects that mistake
  MINOS errors           with a zero pulse covariance the constrained refits
                         are NNLS problems with the in-time amplitude fixed:
                         brute force over the other pulses' passive sets gives
                         sigma+ / sigma- in closed form
Then the product's fit core (fit_core.hpp built for the host) against the
oracle: same amplitudes to rounding, identical iteration counts and flags.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

from oracle_binding import (OR_SOLVER_INPLACE, OR_SOLVER_LEGACY, dptr, hostcore_fit, oracle,
                            oracle_fit, oracle_opts)


def _workload(cfg, pulse_cov, n):
    from test_regression_amd import workloads
    w = workloads.config(cfg, pulse_cov)
    ct = w.chtypes()
    s, t, _ = w.generate(n)
    return ct, s, t


def _pulsemat(t):
    P = np.zeros((10, 10))
    oracle().or_build_pulsemat(C.byref(t), dptr(P))
    return P[:t.nsamples, :t.npulses]


def _whitened(t, x, s):
    """updateCov + LLT + the whitened (A, b) of NNLS (PulseChiSqSNNLS.cpp:
    230-258, 273-281) for amplitudes x, written with numpy."""
    ns = t.nsamples
    nc = np.array(t.noisecor).reshape(10, 10)[:ns, :ns]
    cov = np.array(t.fullpulsecov).reshape(19, 19)
    Cm = t.pederr ** 2 * nc
    for p in range(t.npulses):
        if x[p] == 0.0:
            continue
        bx = t.bx[p]
        f, o = max(0, bx + t.first_off), t.pulse_off - bx
        Cm[f:, f:] += x[p] ** 2 * cov[f + o:ns + o, f + o:ns + o]
    Cm = np.tril(Cm) + np.tril(Cm, -1).T
    L = np.linalg.cholesky(Cm)
    return np.linalg.solve(L, _pulsemat(t)), np.linalg.solve(L, s[:ns])


def _kkt_violation(A, b, x):
    """max over pulses of the NNLS optimality violation, relative to |A^T b|:
    passive pulses need w = 0, pulses at zero need w <= 0, w = A^T (b - A x)."""
    w = A.T @ (b - A @ x)
    scale = max(np.abs(A.T @ b).max(), 1e-300)
    v = np.where(x > 0, np.abs(w), np.maximum(w, 0.0))
    return v.max() / scale


@pytest.mark.parametrize("cfg,n", [("c1", 2000), ("c2", 3000)])
def test_legacy_zero_pi_equals_inplace_and_keeps_state(pkg, cfg, n):
    ct, s, t = _workload(cfg, "zero", n)
    new = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_INPLACE), nthreads=8)
    leg = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LEGACY), nthreads=8)
    one = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LEGACY, max_minimize=1), nthreads=8)
    scale = np.maximum(np.abs(new["x"]).max(axis=1), 1.0)
    assert (np.abs(leg["x"] - new["x"]).max(axis=1) <= 1e-12 * scale).all()
    assert np.array_equal(leg["n_minimize"], new["n_minimize"])
    # the first NNLS call starts from nothing in both: same additions
    assert np.array_equal(one["n_outer"], leg["n_outer"])
    assert np.array_equal(leg["n_outer"], new["n_outer"])
    # every later call: one solve of the carried-over passive set, no addition
    assert np.array_equal(leg["n_inner"], one["n_inner"] + leg["n_minimize"] - 1)
    assert (leg["n_minimize"] >= 2).mean() > 0.9


@pytest.mark.parametrize("cfg", ["c2", "c3"])
def test_legacy_minimize_iterations_reach_nnls_optimum(pkg, cfg):
    ct, s, t = _workload(cfg, "live", 150)
    viol = {}
    for solver in (OR_SOLVER_LEGACY, OR_SOLVER_INPLACE):
        full = oracle_fit(ct, s, t, oracle_opts(solver=solver), nthreads=8)
        nm = full["n_minimize"]
        # the amplitudes the last iteration's covariance was built from
        prev = {k: oracle_fit(ct, s, t, oracle_opts(solver=solver, max_minimize=int(k) - 1),
                              nthreads=8)["x"] for k in np.unique(nm)}
        v = []
        for c in range(s.shape[1]):
            ty = ct[t[c] if t is not None else 0]
            A, b = _whitened(ty, prev[nm[c]][c], s[:, c])
            v.append(_kkt_violation(A, b, full["x"][c][:ty.npulses]))
        viol[solver] = np.array(v)
    assert viol[OR_SOLVER_LEGACY].max() <= 1e-9
    # the check has teeth: the warm-start quirk of inplace_fnnls fails it
    assert (viol[OR_SOLVER_INPLACE] > 1e-6).mean() > 0.5


def _brute_nnls(A, b):
    """argmin |A x - b|^2 over x >= 0, by least squares on every support."""
    n = A.shape[1]
    best, bx = np.sum(b * b), np.zeros(n)
    for k in range(1, n + 1):
        for S in itertools.combinations(range(n), k):
            xs, *_ = np.linalg.lstsq(A[:, S], b, rcond=None)
            if np.all(xs >= 0):
                r = A[:, S] @ xs - b
                if float(r @ r) < best:
                    best, bx = float(r @ r), np.zeros(n)
                    bx[list(S)] = xs
    return bx


def _minimize_fixed(ty, s, x, ip, chi2, max_minimize=50, tol=1e-3):
    """Minimize (PulseChiSqSNNLS.cpp:208-224) with pulse ip's column zeroed and
    its amplitude fixed at x[ip] (the MINOS refits, :104-138), every NNLS
    solved by brute force.  Each NNLS optimum is unique, so the result does
    not depend on how an active-set method reaches it."""
    x = x.copy()
    others = [p for p in range(ty.npulses) if p != ip]
    for _ in range(max_minimize):
        A, b = _whitened(ty, x, s)
        x[others] = _brute_nnls(A[:, others], b)
        A[:, ip] = 0.0               # ComputeChiSq on the same covariance
        r = A[:, :ty.npulses] @ x[:ty.npulses] - b
        now = float(r @ r)
        d, chi2 = now - chi2, now
        if abs(d) < tol:
            break
    return x, chi2


@pytest.mark.parametrize("pulse_cov,n", [("zero", 30), ("live", 20)])
def test_legacy_minos_errors_brute_force(pkg, pulse_cov, n):
    """The in-time amplitude's MINOS error (PulseChiSqSNNLS.cpp:73-146) with
    every constrained refit solved by brute force.  With a live pulse
    covariance the in-time pulse must leave the carried-over passive set
    before the refits (:151-157): kept passive, its zeroed column would move
    its amplitude off the fixed value and out of the covariance."""
    ct, s, t = _workload("c2", pulse_cov, n)
    ty = ct[0]
    ip = list(ty.bx).index(0)
    res = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LEGACY, compute_errors=1), nthreads=8)
    nm = res["n_minimize"]
    prev = {k: oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LEGACY, max_minimize=int(k) - 1),
                          nthreads=8)["x"] for k in np.unique(nm)}
    for c in range(n):
        sc = s[:, c]
        xfit, chi2_0 = res["x"][c].copy(), res["chi2"][c]
        A, _ = _whitened(ty, prev[nm[c]][c], sc)     # covariance of the last iteration
        approxerr = 1.0 / np.linalg.norm(A[:, ip])   # ComputeApproxUncertainty
        x0 = xfit[ip]
        x = xfit.copy()
        x[ip] = x0 + approxerr
        sp_samples = sc - x[ip] * _pulsemat(ty)[:, ip]
        x, chip = _minimize_fixed(ty, sp_samples, x, ip, chi2_0)
        sp = approxerr / np.sqrt(chip - chi2_0)
        want = sp
        if x0 / sp > 0.5:
            x[ip] = max(0.0, x0 - approxerr)
            sm_samples = sc - x[ip] * _pulsemat(ty)[:, ip]
            xm = x[ip]
            x, chim = _minimize_fixed(ty, sm_samples, x, ip, chip)
            want = 0.5 * (sp + abs(xm - x0) / np.sqrt(chim - chi2_0))
        got = res["err"][c][ip]
        if np.isnan(want):
            assert np.isnan(got), c
        else:
            assert abs(got - want) <= 1e-7 * want, (c, got, want)


@pytest.mark.parametrize("cfg,pulse_cov,n", [
    ("c1", "zero", 2000), ("c2", "zero", 3000), ("c2", "live", 3000), ("c3", "live", 3000),
])
def test_legacy_core_matches_oracle(pkg, cfg, pulse_cov, n):
    from test_regression_amd import _abi
    ct, s, t = _workload(cfg, pulse_cov, n)
    ref = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LEGACY, compute_errors=1), nthreads=8)
    x, chi2, fl, it, err = hostcore_fit(
        ct, s, t, _abi.default_opts(solver=_abi.MF_SOLVER_LEGACY, compute_errors=1))
    xr = ref["x"].T
    scale = np.maximum(np.abs(xr).max(axis=0), 1.0)
    assert ((x > 0) == (xr > 0)).all()
    assert (np.abs(x - xr).max(axis=0) <= 1e-10 * scale).all()
    assert np.allclose(chi2, ref["chi2"], rtol=1e-10, atol=1e-10)
    assert np.array_equal(it & 255, ref["n_minimize"])
    assert np.array_equal((it >> 8) & 255, ref["n_outer"])
    assert np.array_equal(it >> 16, ref["n_inner"])
    assert np.array_equal(fl, ref["flags"])
    e, er = err.T, ref["err"]
    assert np.array_equal(np.isnan(e), np.isnan(er))
    ok = ~np.isnan(er) & (er != 0)
    assert np.all(np.abs(e[ok] - er[ok]) <= 1e-9 * np.abs(er[ok]))
