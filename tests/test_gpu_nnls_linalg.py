"""GPU parity of the batched solver and linalgebra entry points.

mf_inplace_fnnls_batch / mf_fnnls_batch against the oracle's inplace_fnnls /
fnnls, against exact brute force (m <= 7) and by the KKT conditions (a proof
of optimality for this convex problem; scipy 1.15's nnls is not used as the
reference: it returns non-optimal points on some of these inputs); mf_linalg_batch against the
oracle's restatement of linalgebra/interface/*.hpp (same operation order:
bit-identical in FP64 up to fma contraction, checked to 1e-13 relative).
"""
import numpy as np
import pytest
import torch
import itertools

from oracle_binding import OrStats, dptr, oracle

pytestmark = pytest.mark.gpu


def _brute_nnls(A, b):
    """Exact NNLS optimum: the best feasible least-squares fit over all supports."""
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


def _problems(n, m, seed, kind="random"):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, 10, 10))          # [c, col, row]
    b = np.zeros((n, 10))
    for c in range(n):
        M = rng.standard_normal((m, m))
        if kind == "pulse":
            M = np.abs(M) + np.eye(m) * 2
        A[c, :m, :m] = M.T             # column-major: A[c, col, row]
        b[c, :m] = rng.standard_normal(m) * 10
    return A, b


@pytest.mark.parametrize("solver", ["inplace_fnnls", "fnnls", "inplace_fnnls_chol"])
@pytest.mark.parametrize("m", [10, 7, 3, 1])
def test_nnls_batch_matches_oracle_and_scipy(pkg, gpu, solver, m):
    from test_regression_amd import engine
    n = 4000
    A, b = _problems(n, m, 100 + m)
    x0 = np.zeros((n, 10))
    x = torch.from_numpy(x0).cuda()
    print(f"[gpu-nnls] solver={solver} m={m} n={n}", flush=True)
    engine.nnls_batch(torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda(), x, m=m,
                      solver=solver)
    torch.cuda.synchronize()
    xg = x.cpu().numpy()
    L = oracle()
    f = {"inplace_fnnls": L.or_inplace_fnnls, "fnnls": L.or_fnnls,
         "inplace_fnnls_chol": L.or_inplace_fnnls_chol}[solver]
    worst = 0.0
    for c in range(n):
        Ar = np.zeros((10, 10))
        Ar[:m, :m] = A[c, :m, :m].T      # row-major for the oracle (ld 10)
        xr = np.zeros(10)
        st = OrStats()
        f(dptr(np.ascontiguousarray(Ar)), dptr(np.ascontiguousarray(b[c])), dptr(xr), m, 1e-11,
          1000, st)
        scale = max(np.abs(xr).max(), 1.0)
        # both solve the normal equations, in different elimination orders
        # (the oracle permutes as the reference does): they agree to about
        # cond(A)^2 eps, so the bound scales with it
        bound = max(1e-10, 1e-14 * np.linalg.cond(A[c, :m, :m]) ** 2)
        worst = max(worst, np.abs(xg[c, :m] - xr[:m]).max() / scale / bound)
        if c < 300:
            Am, bm, x = A[c, :m, :m].T, b[c, :m], xg[c, :m]
            if solver == "fnnls" and np.count_nonzero(x == 0) == 1:
                continue  # fnnls.cu:113 never tests the last free index (see test_oracle)
            # optimality proof: KKT of min |Am x - bm|^2, x >= 0 (convex)
            wv = Am.T @ (bm - Am @ x)
            tol = 1e-8 * (np.abs(Am.T @ bm).max() + 1) * np.linalg.cond(Am) ** 2
            assert np.all(x >= 0)
            assert np.all(wv[x == 0] <= tol) and np.all(np.abs(wv[x > 0]) <= tol), (c, x, wv)
            if m <= 7:
                # normal equations: error ~ cond(A)^2 eps against lstsq's cond(A) eps
                xb = _brute_nnls(Am, bm)
                rt = max(1e-7, 1e-14 * np.linalg.cond(Am) ** 2)
                assert np.allclose(x, xb, rtol=rt, atol=rt * max(1, np.abs(xb).max()))
    assert worst <= 1.0
    assert np.all(xg[:, m:] == 0)


def test_inplace_fnnls_warm_start_keeps_stale_values(pkg, gpu):
    """The reference keeps the warm x but resets the passive set
    (nnls/src/inplace_fnnls.cu:97,137): a warm start can end away from the
    cold optimum.  The batch entry reproduces that, as the oracle does."""
    from test_regression_amd import engine
    n, m = 2000, 10
    A, b = _problems(n, m, 7)
    rng = np.random.default_rng(8)
    x0 = np.abs(rng.standard_normal((n, 10))) * 3
    x = torch.from_numpy(x0.copy()).cuda()
    engine.nnls_batch(torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda(), x, m=m)
    torch.cuda.synchronize()
    xg = x.cpu().numpy()
    L = oracle()
    for c in range(n):
        Ar = np.ascontiguousarray(A[c].T)
        xr = x0[c].copy()
        L.or_inplace_fnnls(dptr(Ar), dptr(np.ascontiguousarray(b[c])), dptr(xr), m, 1e-11, 1000,
                           OrStats())
        assert np.abs(xg[c] - xr).max() <= 1e-10 * max(1, np.abs(xr).max())


def _spd(rng, n):
    M = rng.standard_normal((n, n))
    return M @ M.T + n * np.eye(n)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_linalg_batch(pkg, gpu, dtype):
    from test_regression_amd import engine, _abi
    rng = np.random.default_rng(3)
    n, full, view = 512, 10, 8
    Ms = np.zeros((n, full, full))
    bs = rng.standard_normal((n, full))
    for c in range(n):
        Ms[c, :view, :view] = _spd(rng, view)
    tol = 1e-12 if dtype == torch.float64 else 2e-4
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()  # noqa: E731
    M, L = T(Ms), T(np.zeros_like(Ms))
    print(f"[gpu-linalg] dtype={dtype}", flush=True)
    engine.linalg_batch(_abi.MF_LA_CHOLESKY, M=M, L=L, full=full, view=view)
    Lg = L.double().cpu().numpy()
    for c in range(0, n, 17):
        Lr = np.linalg.cholesky(Ms[c, :view, :view])
        assert np.allclose(np.tril(Lg[c, :view, :view]), Lr, rtol=tol, atol=tol)
    y = T(np.zeros((n, full)))
    engine.linalg_batch(_abi.MF_LA_FORWARD, L=L, b=T(bs), y=y, full=full, view=view)
    z = T(np.zeros((n, full)))
    engine.linalg_batch(_abi.MF_LA_BACKWARD, L=L, b=y, y=z, full=full, view=view)
    zg = z.double().cpu().numpy()
    for c in range(0, n, 17):
        xs = np.linalg.solve(Ms[c, :view, :view], bs[c, :view])
        assert np.allclose(zg[c, :view], xs, rtol=10 * tol, atol=10 * tol)
    # swap rows/cols 2 and 5 of M
    M2 = T(Ms)
    engine.linalg_batch(_abi.MF_LA_SWAP, M=M2, full=full, view=view, i=2, j=5)
    M2g = M2.double().cpu().numpy()
    # the reference's swap (linalgebra/interface/common.hpp:11-42) visits
    # elem == j after swapping the diagonal: (i, j) and (j, j) end up holding
    # M[i][i] and M[i][j]; everything else is the symmetric permutation
    p = list(range(view))
    p[2], p[5] = 5, 2
    for c in range(0, n, 37):
        want = Ms[c][:view, :view][np.ix_(p, p)].copy()
        want[2, 5], want[5, 5] = Ms[c][2, 2], Ms[c][2, 5]
        assert np.allclose(M2g[c, :view, :view], want, rtol=tol, atol=tol)
    # row/column addition: factor 7x7, then append row 7 -> 8x8
    L7 = T(np.zeros_like(Ms))
    engine.linalg_batch(_abi.MF_LA_CHOLESKY, M=M, L=L7, full=full, view=view - 1)
    y7 = T(np.zeros((n, full)))
    engine.linalg_batch(_abi.MF_LA_FORWARD, L=L7, b=T(bs), y=y7, full=full, view=view - 1)
    engine.linalg_batch(_abi.MF_LA_RCADDITION, M=M, L=L7, b=T(bs), y=y7, full=full, view=view)
    assert np.allclose(L7.double().cpu().numpy()[:, :view, :view], Lg[:, :view, :view],
                       rtol=10 * tol, atol=10 * tol)
    assert np.allclose(y7.double().cpu().numpy()[:, :view], y.double().cpu().numpy()[:, :view],
                       rtol=10 * tol, atol=10 * tol)


@pytest.mark.parametrize("sfx,tol", [("f64", 1e-13), ("f32", 2e-5)])
def test_linalg_batch_golden_every_op(pkg, gpu, sfx, tol):
    """Every mf_linalg_batch op against the golden vectors written by the
    reference's own linalgebra headers (tests/golden/make_golden.py): the 3x3
    literal of test_cholesky_with_solvers.cpp, seeded 10x10 and 7-of-10 SPD
    factors and solves, the symmetric swap, and the sequence of
    generate_and_test_fused_impl.cpp:66-119 -- factor 9x9, append row 9
    (rcaddition), swap rows 6 and 9 and refactor from position 6
    (inbetween_removal).  The same operation order; agreement to a few ulps
    (the GPU contracts multiply-adds into FMAs)."""
    from pathlib import Path
    from test_regression_amd import engine, _abi
    g = np.load(Path(__file__).resolve().parent / "golden" / "linalgebra_golden.npz")
    dt = torch.float64 if sfx == "f64" else torch.float32
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()  # noqa: E731
    H = lambda t: t.double().cpu().numpy()  # noqa: E731

    def close(got, want, what):
        sc = max(1.0, np.abs(want).max())
        err = np.abs(got - want).max() / sc
        assert err <= tol, (what, err)

    print(f"[gpu-linalg] golden {sfx}", flush=True)
    # 3x3 literal: factor, forward, backward
    M3, b3 = T(g[f"lit3_M_{sfx}"][None]), T(g[f"lit3_b_{sfx}"][None])
    L3, y3, x3 = T(np.zeros((1, 3, 3))), T(np.zeros((1, 3))), T(np.zeros((1, 3)))
    engine.linalg_batch(_abi.MF_LA_CHOLESKY, M=M3, L=L3, full=3, view=3)
    engine.linalg_batch(_abi.MF_LA_FORWARD, L=L3, b=b3, y=y3, full=3, view=3)
    engine.linalg_batch(_abi.MF_LA_BACKWARD, L=L3, b=y3, y=x3, full=3, view=3)
    close(np.tril(H(L3)[0]), np.tril(g[f"lit3_L_{sfx}"]), "lit3 L")
    close(H(y3)[0], g[f"lit3_y_{sfx}"], "lit3 y")
    close(H(x3)[0], g[f"lit3_x_{sfx}"], "lit3 x")
    # seeded SPD, views 10 and 7 (one call per view)
    Ms, bs, views = g[f"spd_M_{sfx}"], g[f"spd_b_{sfx}"], g["spd_view"]
    for v in (10, 7):
        sel = np.flatnonzero(views == v)
        M, b = T(Ms[sel]), T(bs[sel])
        L, y, x = T(np.zeros((sel.size, 10, 10))), T(np.zeros((sel.size, 10))), \
            T(np.zeros((sel.size, 10)))
        engine.linalg_batch(_abi.MF_LA_CHOLESKY, M=M, L=L, full=10, view=v)
        engine.linalg_batch(_abi.MF_LA_FORWARD, L=L, b=b, y=y, full=10, view=v)
        engine.linalg_batch(_abi.MF_LA_BACKWARD, L=L, b=y, y=x, full=10, view=v)
        close(np.tril(H(L))[:, :v, :v], np.tril(g[f"spd_L_{sfx}"][sel])[:, :v, :v], f"spd L {v}")
        close(H(y)[:, :v], g[f"spd_y_{sfx}"][sel][:, :v], f"spd y {v}")
        close(H(x)[:, :v], g[f"spd_x_{sfx}"][sel][:, :v], f"spd x {v}")
    # swap (i, j) per problem: one call per pair; bit for bit (moves only)
    pairs = g["swap_pairs"]
    for (i, j) in {tuple(p) for p in pairs}:
        sel = np.flatnonzero((pairs[:, 0] == i) & (pairs[:, 1] == j))
        M = T(Ms[sel])
        engine.linalg_batch(_abi.MF_LA_SWAP, M=M, full=10, view=10, i=int(i), j=int(j))
        assert np.array_equal(H(M), g[f"swap_M_{sfx}"][sel].astype(np.float64)), (i, j)
    # fused: factor 9, forward 9, append row 9
    n = Ms.shape[0]
    M, b = T(Ms), T(bs)
    L, y = T(np.zeros((n, 10, 10))), T(np.zeros((n, 10)))
    engine.linalg_batch(_abi.MF_LA_CHOLESKY, M=M, L=L, full=10, view=9)
    engine.linalg_batch(_abi.MF_LA_FORWARD, L=L, b=b, y=y, full=10, view=9)
    engine.linalg_batch(_abi.MF_LA_RCADDITION, M=M, L=L, b=b, y=y, full=10, view=10)
    close(np.tril(H(L)), np.tril(g[f"fused_L_{sfx}"]), "rcaddition L")
    close(H(y), g[f"fused_y_{sfx}"], "rcaddition y")
    # removal: swap rows/cols 6 and 9 of M (and of b), refactor from 6 in view 9
    Mr = T(Ms)
    engine.linalg_batch(_abi.MF_LA_SWAP, M=Mr, full=10, view=10, i=6, j=9)
    assert np.array_equal(H(Mr), g[f"removal_M_{sfx}"].astype(np.float64))
    br = bs.copy()
    br[:, [6, 9]] = br[:, [9, 6]]
    assert np.array_equal(br, g[f"removal_b_{sfx}"])
    engine.linalg_batch(_abi.MF_LA_INBETWEEN, M=Mr, L=L, b=T(br), y=y, full=10, view=9, i=6)
    close(np.tril(H(L))[:, :9, :9], np.tril(g[f"removal_L_{sfx}"])[:, :9, :9], "removal L")
    close(H(y)[:, :9], g[f"removal_y_{sfx}"][:, :9], "removal y")


def test_linalg_batch_many_problems_and_partial_blocks(pkg, gpu):
    """Problem counts around the 32-problem blocks (1, 31, 33, 1000): every
    problem's factor equals numpy's, none is skipped or mixed up."""
    from test_regression_amd import engine, _abi
    rng = np.random.default_rng(9)
    for n in (1, 31, 33, 1000):
        Ms = np.stack([_spd(rng, 10) for _ in range(n)])
        M = torch.from_numpy(Ms).cuda()
        L = torch.zeros_like(M)
        engine.linalg_batch(_abi.MF_LA_CHOLESKY, M=M, L=L, full=10, view=10)
        Lg = L.cpu().numpy()
        Lr = np.linalg.cholesky(Ms)
        assert np.allclose(np.tril(Lg), Lr, rtol=1e-12, atol=1e-12), n
