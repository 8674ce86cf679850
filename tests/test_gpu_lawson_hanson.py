"""GPU parity of the Lawson-Hanson solver (Eigen::NNLS semantics,
nnls/interface/eigen_nnls.h:336-414): the batched solver entry
(mf_lh_nnls_batch) and the fit mode (MF_SOLVER_LAWSON_HANSON, the
alternative of multifit_cpu/src/PulseChiSqSNNLS.cpp:253-257) against the
oracle.

Bars: batched solves within max(1e-10, 1e-14 cond(A)^2) of the oracle
relative to the largest amplitude (the oracle solves by Householder QR, the
kernel by Cholesky of the normal equations) with the same solved flag; fits as
the other FP64 parity tests (1e-10, identical iteration counts and flags on
>= 99.9% of channels).
"""
import ctypes as C

import numpy as np
import pytest

import parity
import torch

from oracle_binding import OR_SOLVER_LH, OrStats, dptr, oracle, oracle_fit, oracle_opts

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,kind,max_iter", [(10, "random", -1), (6, "random", -1),
                                             (10, "pulse", -1), (10, "random", 4)])
def test_lh_batch_matches_oracle(pkg, gpu, m, kind, max_iter):
    from test_regression_amd import engine
    rng = np.random.default_rng(m * 7 + (kind == "pulse"))
    n = 3000
    A = np.zeros((n, 10, 10))
    b = np.zeros((n, 10))
    for c in range(n):
        M = rng.standard_normal((m, m))
        if kind == "pulse":
            M = np.abs(M) + 2 * np.eye(m)
        A[c, :m, :m] = M.T
        b[c, :m] = rng.standard_normal(m) * 10
    x0 = np.zeros((n, 10))
    x0[n // 2:, :m] = np.abs(rng.standard_normal((n - n // 2, m)))   # setX warm starts
    x = torch.from_numpy(x0.copy()).cuda()
    print(f"[gpu-lh] m={m} kind={kind} max_iter={max_iter} n={n}", flush=True)
    _, solved = engine.lh_nnls_batch(torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda(), x,
                                     m=m, eps=1e-10, max_iter=max_iter)
    torch.cuda.synchronize()
    xg, sg = x.cpu().numpy(), solved.cpu().numpy()
    L = oracle()
    worst, same_ok = 0.0, 0
    for c in range(n):
        Ar = np.ascontiguousarray(np.pad(A[c, :m, :m].T, ((0, 10 - m), (0, 10 - m))))
        xr = x0[c].copy()
        ok = L.or_lh_nnls(dptr(Ar), dptr(np.ascontiguousarray(b[c])), dptr(xr), m, 1e-10,
                          max_iter, C.byref(OrStats()))
        same_ok += int(ok == sg[c])
        bound = max(1e-10, 1e-14 * np.linalg.cond(A[c, :m, :m]) ** 2)
        worst = max(worst, np.abs(xg[c, :m] - xr[:m]).max() / max(np.abs(xr).max(), 1.0) / bound)
    print(f"worst/bound {worst:.3e} same solved {same_ok / n:.6f} solved {sg.mean():.4f}")
    assert same_ok / n >= 0.999
    assert worst <= 1.0
    if max_iter < 0:
        assert sg.all()


@pytest.mark.parametrize("cfg,pulse_cov,n", [("c2", "zero", 20_000), ("c2", "live", 20_000),
                                             ("c3", "live", 20_000)])
def test_lh_fit_matches_oracle(pkg, gpu, cfg, pulse_cov, n):
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config(cfg, pulse_cov)
    ct = w.chtypes()
    s, t, _ = w.generate(n)
    eng = engine.Engine(list(ct))
    print(f"[gpu-lh-fit] cfg={cfg} pulse_cov={pulse_cov} n={n}", flush=True)
    r = eng.fit(torch.from_numpy(s).cuda(), torch.from_numpy(t).cuda() if t is not None else None,
                _abi.default_opts(solver=_abi.MF_SOLVER_LAWSON_HANSON))
    torch.cuda.synchronize()
    ref = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LH), nthreads=8)
    parity.gate(r.x.double().cpu().numpy(), r.chi2.double().cpu().numpy(), ref, 1e-10,
                iters=r.iters.cpu().numpy(), flags=r.flags.cpu().numpy(), compare_inner=True,
                label=f"lawson-hanson {cfg} {pulse_cov}")
