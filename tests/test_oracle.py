"""Pins the FP64 oracle (oracle/mf_oracle.c) before anything is compared with it.

Checks, per oracle function:
  dense primitives   golden vectors written by the reference's own
                     linalgebra headers (tests/golden/linalgebra_golden.npz),
                     numpy's Cholesky / solve, the 3x3 closed form
  inplace_fnnls      scipy's Lawson-Hanson NNLS, brute force over all passive
                     sets, KKT conditions (scipy 1.15's nnls returns
                     non-optimal points on some inputs, so brute force and
                     KKT decide where they disagree), and the reference's OpenCL
                     incremental-Cholesky FNNLS (tests/golden/fnnls_cl_golden.npz)
  fnnls              scipy, brute force
  DoFit              closed forms: noise-free recovery, identity covariance
                     with orthonormal pulses, single-pulse closed form, chi2
                     recomputed from the returned amplitudes
  planted mistakes   a copy of the oracle's FNNLS written here with one
                     deliberate error (a transposed operand, a dropped term)
                     must fail the same checks, so the checks have teeth
"""
import ctypes as C
import itertools
from pathlib import Path

import numpy as np
import pytest
from scipy.optimize import nnls as scipy_nnls

from oracle_binding import (OR_SOLVER_FNNLS, OR_SOLVER_INPLACE, OR_SOLVER_INPLACE_CHOL,
                            OrStats, dptr, oracle, oracle_fit, oracle_opts)

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def L():
    return oracle()


@pytest.fixture(scope="module")
def lin():
    return np.load(GOLD / "linalgebra_golden.npz")


@pytest.fixture(scope="module")
def clg():
    return np.load(GOLD / "fnnls_cl_golden.npz")


def _m(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- primitives
def test_cholesky_3x3_closed_form(L):
    M = _m([[4, -1, 2], [-1, 6, 0], [2, 0, 5]])
    Lc = np.zeros((3, 3))
    L.or_cholesky_decomp(dptr(M), dptr(Lc), 3, 3)
    l11 = np.sqrt(5.75)
    l21 = 0.5 / l11
    expect = np.array([[2, 0, 0], [-0.5, l11, 0], [1, l21, np.sqrt(4 - l21 ** 2)]])
    assert np.allclose(Lc, expect, rtol=0, atol=1e-15)


def test_linalgebra_golden(L, lin):
    # oracle restatement vs the reference's own headers, FP64, and vs numpy
    M3, b3 = lin["lit3_M_f64"], lin["lit3_b_f64"]
    Lc = np.zeros((3, 3))
    L.or_cholesky_decomp(dptr(_m(M3)), dptr(Lc), 3, 3)
    assert np.allclose(Lc, lin["lit3_L_f64"], rtol=1e-15, atol=1e-15)
    y, x = np.zeros(3), np.zeros(3)
    L.or_forward_subst(dptr(Lc), dptr(_m(b3)), dptr(y), 3, 3)
    L.or_backward_subst(dptr(Lc), dptr(y), dptr(x), 3, 3)
    assert np.allclose(x, lin["lit3_x_f64"], rtol=1e-14)
    assert np.allclose(x, np.linalg.solve(M3, b3), rtol=1e-14)
    Ms, bs, views = lin["spd_M_f64"], lin["spd_b_f64"], lin["spd_view"]
    for i in range(len(Ms)):
        v = int(views[i])
        Lc = np.zeros((10, 10))
        L.or_cholesky_decomp(dptr(_m(Ms[i])), dptr(Lc), 10, v)
        g = lin["spd_L_f64"][i]
        assert np.allclose(np.tril(Lc)[:v, :v], np.tril(g)[:v, :v], rtol=1e-13, atol=1e-13)
        assert np.allclose(np.tril(Lc)[:v, :v], np.linalg.cholesky(Ms[i][:v, :v]), rtol=1e-12,
                           atol=1e-12)
        y, x = np.zeros(10), np.zeros(10)
        L.or_forward_subst(dptr(Lc), dptr(_m(bs[i])), dptr(y), 10, v)
        L.or_backward_subst(dptr(Lc), dptr(y), dptr(x), 10, v)
        assert np.allclose(x[:v], lin["spd_x_f64"][i][:v], rtol=1e-11, atol=1e-11)
        assert np.allclose(x[:v], np.linalg.solve(Ms[i][:v, :v], bs[i][:v]), rtol=1e-9)


def test_swap_and_fused_updates_golden(L, lin):
    Ms, bs = lin["spd_M_f64"], lin["spd_b_f64"]
    for i in range(len(Ms)):
        Sw = _m(Ms[i]).copy()
        a, b = (int(v) for v in lin["swap_pairs"][i])
        L.or_swap_row_column(dptr(Sw), a, b, 10, 10)
        assert np.array_equal(Sw, lin["swap_M_f64"][i])
        # linalgebra's swap (common.hpp:11-42) runs its second loop through
        # elem == j after the diagonal swap, so it is the symmetric permutation
        # except at (i, j) and (j, j), which receive M[i][i] and M[i][j]
        p = list(range(10))
        p[a], p[b] = b, a
        want = Ms[i][np.ix_(p, p)].copy()
        want[a, b], want[b, b] = Ms[i][a, a], Ms[i][a, b]
        assert np.array_equal(Sw, want)
        # 9x9 factor, append row 9, compare with the reference and a full refactor
        Lf, yf = np.zeros((10, 10)), np.zeros(10)
        L.or_cholesky_decomp(dptr(_m(Ms[i])), dptr(Lf), 10, 9)
        L.or_forward_subst(dptr(Lf), dptr(_m(bs[i])), dptr(yf), 10, 9)
        L.or_fused_rcaddition(dptr(_m(Ms[i])), dptr(Lf), dptr(_m(bs[i])), dptr(yf), 10, 10)
        assert np.allclose(np.tril(Lf), np.tril(lin["fused_L_f64"][i]), rtol=1e-12, atol=1e-12)
        assert np.allclose(np.tril(Lf), np.linalg.cholesky(Ms[i]), rtol=1e-10, atol=1e-10)
        assert np.allclose(yf, lin["fused_y_f64"][i], rtol=1e-11, atol=1e-11)
        # remove row 6 (swap with the last of the 9-view, refactor from 6)
        Mr, br = _m(lin["removal_M_f64"][i]), _m(lin["removal_b_f64"][i])
        Lr, yr = Lf.copy(), yf.copy()
        L.or_fused_inbetween_removal(dptr(Mr), dptr(Lr), dptr(br), dptr(yr), 6, 10, 9)
        assert np.allclose(np.tril(Lr)[:9, :9], np.tril(lin["removal_L_f64"][i])[:9, :9],
                           rtol=1e-11, atol=1e-11)
        assert np.allclose(np.tril(Lr)[:9, :9], np.linalg.cholesky(Mr[:9, :9]), rtol=1e-10,
                           atol=1e-10)
        assert np.allclose(yr[:9], lin["removal_y_f64"][i][:9], rtol=1e-10, atol=1e-10)


def test_llt_failure_semantics(L):
    """Eigen LLT stops at the first non-positive pivot and leaves the rest."""
    M = _m([[1, 2, 0], [2, 1, 0], [0, 0, 3]])
    M10 = np.zeros((10, 10))
    M10[:3, :3] = M
    Lc = np.zeros((10, 10))
    k = L.or_llt(dptr(M10), dptr(Lc), 3)
    assert k == 1
    assert Lc[0, 0] == 1.0 and Lc[1, 0] == 2.0 and Lc[1, 1] == 1.0


# ---------------------------------------------------------------- NNLS
def _brute_nnls(A, b):
    """Exact NNLS optimum for full-column-rank A: best feasible LS over all supports."""
    n = A.shape[1]
    best, bx = np.inf, np.zeros(n)
    for k in range(n + 1):
        for S in itertools.combinations(range(n), k):
            x = np.zeros(n)
            if S:
                xs, *_ = np.linalg.lstsq(A[:, S], b, rcond=None)
                if np.any(xs < 0):
                    continue
                x[list(S)] = xs
            r = np.sum((A @ x - b) ** 2)
            if r < best - 1e-12:
                best, bx = r, x
    return bx


def _run(L, fn, A, b, m, x0=None, max_iter=1000):
    Ar = np.zeros((10, 10))
    Ar[:m, :m] = A
    bb = np.zeros(10)
    bb[:m] = b
    x = np.zeros(10) if x0 is None else x0.copy()
    st = OrStats()
    getattr(L, fn)(dptr(Ar), dptr(bb), dptr(x), m, 1e-11, max_iter, C.byref(st))
    return x[:m], st


@pytest.mark.parametrize("fn", ["or_inplace_fnnls", "or_inplace_fnnls_chol", "or_fnnls"])
def test_nnls_vs_scipy_and_brute_force(L, fn):
    rng = np.random.default_rng(11)
    for trial in range(300):
        m = 6 if trial < 100 else 10
        A = rng.standard_normal((m, m))
        b = rng.standard_normal(m) * 5
        x, st = _run(L, fn, A, b, m)
        xs, _ = scipy_nnls(A, b)
        if fn == "or_fnnls" and np.count_nonzero(x == 0) == 1 and not np.allclose(x, xs):
            # fnnls.cu:113 stops on !n_active, i.e. as soon as exactly one
            # index is still free, without testing it: the result is then the
            # optimum with that index held at 0, wherever the optimum is.
            z = np.nonzero(x == 0)[0]
            keep = [i for i in range(m) if i != z[0]]
            xr, _ = scipy_nnls(A[:, keep], b)
            assert np.allclose(x[keep], xr, rtol=1e-8, atol=1e-8)
            continue
        assert np.allclose(x, xs, rtol=1e-8, atol=1e-8), (trial, x, xs)
        if m == 6:
            assert np.allclose(x, _brute_nnls(A, b), rtol=1e-8, atol=1e-8)
        # KKT: x >= 0, w = A^T(b - Ax) <= tol on the zero set, ~0 on the support
        w = A.T @ (b - A @ x)
        scale = np.abs(A.T @ b).max() + 1
        assert np.all(x >= 0)
        assert np.all(w[x == 0] <= 1e-9 * scale)
        assert np.all(np.abs(w[x > 0]) <= 1e-9 * scale)


def test_inplace_fnnls_vs_reference_opencl_double(L, clg):
    """Incremental-Cholesky FNNLS, oracle vs the reference's .cl compiled as double."""
    A, b, x64 = clg["A"], clg["b"], clg["x_f64"]
    for i in range(len(A)):
        xo, _ = _run(L, "or_inplace_fnnls_chol", A[i], b[i], 10)
        scale = max(1.0, np.abs(x64[i]).max())
        assert np.all((xo > 0) == (x64[i] > 0)), i
        assert np.abs(xo - x64[i]).max() <= 1e-9 * scale, (i, xo, x64[i])
        # and the full-refactor variant reaches the same optimum
        xf, _ = _run(L, "or_inplace_fnnls", A[i], b[i], 10)
        assert np.abs(xf - x64[i]).max() <= 1e-8 * scale


def test_reference_fp32_criterion(L, clg):
    """The reference's only numeric pass/fail rule: FP32 FNNLS within 1e-4 of
    the CPU result (multifit_ocl/tests/inplace_fnnls_original_host.cpp:340-360),
    on its own kind of input (uniform random A, b).  On pulse-shaped problems
    the reference's FP32 solver itself breaks down (NaN or large errors), which
    is why the product's fast path finishes every fit in FP64."""
    A, b, x32, kind = clg["A"], clg["b"], clg["x_f32"], clg["kind"]
    for i in np.nonzero(kind == 0)[0]:
        xo, _ = _run(L, "or_inplace_fnnls", A[i], b[i], 10)
        assert np.abs(xo - x32[i]).max() <= 1e-4 * max(1.0, np.abs(xo).max()), i
    pulse = kind == 1
    bad = (~np.isfinite(x32[pulse]).all(axis=1)).sum() + \
        (np.nan_to_num(np.abs(x32[pulse] - clg["x_f64"][pulse]), nan=1).max(axis=1) > 1e-4).sum()
    assert bad > 0


def test_inplace_fnnls_warm_start_quirk(L):
    """Warm x is kept while the passive set restarts empty
    (nnls/src/inplace_fnnls.cu:97,137): started at the optimum of one problem,
    the solver's first step moves one coordinate to its single-variable least
    squares value and can stop there."""
    rng = np.random.default_rng(5)
    moved = 0
    for _ in range(200):
        A = np.abs(rng.standard_normal((10, 10))) + 2 * np.eye(10)
        b = A @ np.abs(rng.standard_normal(10))
        xopt, _ = _run(L, "or_inplace_fnnls", A, b, 10)
        b2 = b + 0.05 * rng.standard_normal(10)
        xw, st = _run(L, "or_inplace_fnnls", A, b2, 10, x0=np.concatenate([xopt, []]))
        xc, _ = _run(L, "or_inplace_fnnls", A, b2, 10)
        if st.outer > 0 and np.abs(xw - xc).max() > 1e-6:
            moved += 1
            # the coordinates never re-entered keep their stale warm value
            assert np.any(np.isclose(xw, xopt, rtol=0, atol=0) & (xopt > 0))
    assert moved > 0


def test_fnnls_max_iter_truncates(L):
    rng = np.random.default_rng(3)
    A = np.abs(rng.standard_normal((10, 10))) + np.eye(10)
    b = A @ np.ones(10)
    x, st = _run(L, "or_inplace_fnnls", A, b, 10, max_iter=3)
    assert st.hit_max == 1 and st.outer == 3 and np.count_nonzero(x) <= 3


# ---------------------------------------------------------------- DoFit
def _chtype_identity(pk):
    """A 10x10 type with identity noise and an orthonormal pulse matrix is out
    of reach of mf_chtype (pulses are shifted copies of one template); use a
    template with a single nonzero sample: P is then the shifted identity."""
    ct = pk._abi.ChType()
    ct.nsamples = ct.npulses = 10
    for p in range(10):
        ct.bx[p] = p - 5
    ct.first_off, ct.pulse_off, ct.pederr = 3, 4, 1.0
    for i in range(10):
        ct.noisecor[i * 10 + i] = 1.0
    ct.fullpulse[9] = 1.0
    return ct


def test_dofit_identity_closed_form(pkg):
    """Identity covariance and P = I: x = max(0, s), chi2 = sum of min(0, s)^2."""
    ct = _chtype_identity(pkg)
    arr = (pkg._abi.ChType * 1)(ct)
    rng = np.random.default_rng(9)
    s = rng.standard_normal((10, 200)) * 10
    r = oracle_fit(arr, s, None, oracle_opts())
    assert np.allclose(r["x"], np.maximum(s.T, 0), atol=1e-12)
    assert np.allclose(r["chi2"], (np.minimum(s, 0) ** 2).sum(axis=0), rtol=1e-12, atol=1e-12)


def test_dofit_noise_free_recovery_and_chi2(pkg):
    from test_regression_amd import workloads
    w = workloads.config("c1")
    ct = w.chtypes()
    P = np.zeros((10, 10))
    fp = np.array(ct[0].fullpulse[:])
    for j, bx in enumerate(ct[0].bx):
        f, o = max(0, bx + 3), 4 - bx
        P[f:, j] = fp[f + o:10 + o]
    rng = np.random.default_rng(2)
    amps = np.zeros((10, 300))
    amps[5] = rng.uniform(1, 100, 300)
    amps[[3, 6, 8]] = rng.uniform(0, 10, (3, 300)) * (rng.uniform(0, 1, (3, 300)) < 0.5)
    s = P @ amps
    r = oracle_fit(ct, s, None, oracle_opts())
    assert np.allclose(r["x"].T, amps, rtol=1e-6, atol=1e-6)
    assert np.all(r["chi2"] < 1e-10)
    # chi2 recomputed from returned amplitudes with noise
    s2 = s + rng.standard_normal(s.shape)
    r2 = oracle_fit(ct, s2, None, oracle_opts())
    R = np.array(ct[0].noisecor[:]).reshape(10, 10)
    Lr = np.linalg.cholesky(R)
    res = np.linalg.solve(Lr, P @ r2["x"].T - s2)
    assert np.allclose((res ** 2).sum(axis=0), r2["chi2"], rtol=1e-10)


def test_dofit_single_pulse_closed_form(pkg):
    """One pulse (bxs = {0}): x = max(0, p^T C^-1 s / p^T C^-1 p)."""
    from test_regression_amd import workloads
    ct = workloads.config("c1").chtypes()[0]
    one = pkg._abi.ChType()
    C.memmove(C.byref(one), C.byref(ct), C.sizeof(ct))
    one.npulses = 1
    one.bx[0] = 0
    arr = (pkg._abi.ChType * 1)(one)
    L = oracle()
    rng = np.random.default_rng(4)
    R = np.array(ct.noisecor[:]).reshape(10, 10)
    p = np.zeros(10)
    fp = np.array(ct.fullpulse[:])
    p[3:] = fp[7:14]
    for _ in range(100):
        s = p * rng.uniform(-20, 100) + rng.standard_normal(10)
        r = oracle_fit(arr, s[:, None].copy(), None, oracle_opts())
        Ci = np.linalg.inv(R)
        expect = max(0.0, p @ Ci @ s / (p @ Ci @ p))
        assert abs(r["x"][0][0] - expect) <= 1e-10 * max(1, abs(expect))


def test_dofit_pederr_scaling(pkg):
    """DoFit's pederr (PulseChiSqSNNLS.cpp:176, the covariance pederr^2 R):
    given per channel it equals a type of that pederr bit for bit; and with
    pederr 2 every whitened quantity is the pederr-1 one scaled by a power of
    two, so at zero pulse covariance the amplitudes are the same bit for bit
    and chi2 is a quarter of it, except where a decision (max w < eps,
    |dchi2| < 1e-3) falls between the two scales."""
    from boundary_cases import npulse_types
    from test_regression_amd import workloads
    s, _, _ = workloads.config("c2").generate(600)
    ct1 = npulse_types("zero", npulses=(10,))
    ct2 = npulse_types("zero", npulses=(10,), pederr=2.0)
    r1 = oracle_fit(ct1, s, None, oracle_opts())
    r2 = oracle_fit(ct1, s, None, oracle_opts(), pederr=np.full(s.shape[1], 2.0))
    r3 = oracle_fit(ct2, s, None, oracle_opts())
    assert np.array_equal(r2["x"], r3["x"]) and np.array_equal(r2["chi2"], r3["chi2"])
    same = (r1["x"] == r2["x"]).all(axis=1)
    assert same.mean() > 0.99
    assert np.array_equal(r2["chi2"][same] * 4.0, r1["chi2"][same])
    # a planted mistake, pederr entering once instead of squared, breaks it
    r4 = oracle_fit(ct1, s, None, oracle_opts(), pederr=np.full(s.shape[1], np.sqrt(2.0)))
    assert not np.allclose(r4["chi2"] * 4.0, r1["chi2"], rtol=1e-6)


# ---------------------------------------------------------------- planted mistakes
def _py_inplace_fnnls(A, b, eps=1e-11, max_iter=1000, mistake=None):
    """Python copy of the oracle's FNNLS (pure loops, small sizes) for planting
    mistakes; mistake=None must reproduce the oracle."""
    n = A.shape[1]
    AtA = A.T @ A if mistake != "transpose" else A @ A.T
    Atb = A.T @ b
    x = np.zeros(n)
    perm = list(range(n))
    P = 0
    for _ in range(max_iter):
        if P == n:
            break
        w = Atb[P:] - (AtA @ x)[P:]
        if mistake == "drop_term":
            w = Atb[P:]
        q = int(np.argmax(w)) + P
        if w[q - P] < eps:
            break
        for arr in (AtA,):
            arr[[P, q]] = arr[[q, P]]
            arr[:, [P, q]] = arr[:, [q, P]]
        Atb[[P, q]] = Atb[[q, P]]
        x[[P, q]] = x[[q, P]]
        perm[P], perm[q] = perm[q], perm[P]
        P += 1
        while P > 0:
            s = np.linalg.solve(AtA[:P, :P], Atb[:P])
            if s.min() > 0:
                x[:P] = s
                break
            ratios = [(x[i] / (x[i] - s[i]), i) for i in range(P) if s[i] <= 0]
            alpha, a = min(ratios)
            x[:P] += alpha * (s - x[:P])
            x[a] = 0
            P -= 1
            AtA[[P, a]] = AtA[[a, P]]
            AtA[:, [P, a]] = AtA[:, [a, P]]
            Atb[[P, a]] = Atb[[a, P]]
            x[[P, a]] = x[[a, P]]
            perm[P], perm[a] = perm[a], perm[P]
    out = np.zeros(n)
    out[perm] = x
    return out


@pytest.mark.parametrize("mistake", ["transpose", "drop_term"])
def test_planted_mistakes_are_caught(L, mistake):
    rng = np.random.default_rng(12)
    caught = 0
    for _ in range(40):
        A = rng.standard_normal((6, 6))
        b = rng.standard_normal(6) * 5
        good = _py_inplace_fnnls(A.copy(), b.copy())
        xo, _ = _run(L, "or_inplace_fnnls", A, b, 6)
        assert np.allclose(good, xo, rtol=1e-9, atol=1e-9)
        bad = _py_inplace_fnnls(A.copy(), b.copy(), mistake=mistake)
        if not np.allclose(bad, _brute_nnls(A, b), rtol=1e-6, atol=1e-6):
            caught += 1
    assert caught >= 20
