"""GPU parity of mf_fit_batch (plain FP64 path) against the FP64 oracle.

Bar (FP64), channel by channel (tests/parity.py): every amplitude within
1e-10 of the oracle's relative to itself (floor 1 ADC), chi2 within 1e-10
relative (floor 1), the same active set, Minimize / FNNLS iteration counts
and flags; a channel outside that must be explained by the oracle's own
record of its conditioning or of a decision taken within rounding, and at
most 0.1% of channels may be outside.  The oracle follows the reference's
operation order; the kernel solves the passive system in ascending index
order (masked Cholesky), so the two agree to rounding, not bit for bit.
"""
import numpy as np
import pytest
import torch

import parity
from oracle_binding import oracle_fit, oracle_opts

pytestmark = pytest.mark.gpu

TOL_X = 1e-10
TOL_CHI2 = 1e-10


def _run(pkg, cfg, pulse_cov, n, precision=0, seed_off=0, **kw):
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config(cfg, pulse_cov)
    w.seed += seed_off
    ct = w.chtypes()
    s, t, _ = w.generate(n)
    eng = engine.Engine(list(ct))
    opts = _abi.default_opts(precision=precision, **kw)
    ts = torch.from_numpy(s).cuda()
    tt = torch.from_numpy(t).cuda() if t is not None else None
    print(f"[gpu-fit] cfg={cfg} pulse_cov={pulse_cov} n={n} prec={precision}", flush=True)
    r = eng.fit(ts, tt, opts)
    torch.cuda.synchronize()
    return ct, s, t, r


def _gate(ct, s, t, r, oo=None, label="", **kw):
    ref = oracle_fit(ct, s, t, oo or oracle_opts(), nthreads=8)
    return parity.gate(r.x.double().cpu().numpy(), r.chi2.double().cpu().numpy(), ref, TOL_X,
                       tol_chi2=TOL_CHI2, iters=r.iters.cpu().numpy(),
                       flags=r.flags.cpu().numpy(), label=label, **kw), ref


@pytest.mark.parametrize("cfg,pulse_cov,n", [
    ("c1", "zero", 10_000),
    ("c2", "zero", 20_000),
    ("c2", "live", 20_000),
    ("c3", "live", 20_000),
    ("c5", "zero", 10_000),
])
def test_fp64_fit_matches_oracle(pkg, gpu, cfg, pulse_cov, n):
    ct, s, t, r = _run(pkg, cfg, pulse_cov, n)
    _gate(ct, s, t, r, label=f"fp64 {cfg} {pulse_cov}")


def test_fp64_fit_fnnls_solver(pkg, gpu):
    from test_regression_amd import _abi
    from oracle_binding import OR_SOLVER_FNNLS
    ct, s, t, r = _run(pkg, "c2", "zero", 5000, solver=_abi.MF_SOLVER_FNNLS)
    _gate(ct, s, t, r, oracle_opts(solver=OR_SOLVER_FNNLS), label="fp64 fnnls")


@pytest.mark.parametrize("cfg,pulse_cov,n", [("c1", "zero", 10_000), ("c2", "zero", 20_000),
                                             ("c2", "live", 20_000), ("c3", "live", 20_000),
                                             ("c5", "zero", 10_000), ("c5", "live", 10_000)])
def test_fp64_fit_incremental_cholesky_solver(pkg, gpu, cfg, pulse_cov, n):
    """MF_SOLVER_INPLACE_FNNLS_CHOL (inplace_fnnls_test_cholesky, the factor
    of the passive block updated from one solve to the next) against the
    oracle's restatement of the same function, and against the oracle's
    refactorising inplace_fnnls: the same fits to rounding."""
    from test_regression_amd import _abi
    from oracle_binding import OR_SOLVER_INPLACE_CHOL
    ct, s, t, r = _run(pkg, cfg, pulse_cov, n, solver=_abi.MF_SOLVER_INPLACE_FNNLS_CHOL)
    _gate(ct, s, t, r, oracle_opts(solver=OR_SOLVER_INPLACE_CHOL),
          label=f"fp64 incremental cholesky {cfg} {pulse_cov}")
    _gate(ct, s, t, r, label=f"fp64 incremental cholesky vs inplace_fnnls {cfg} {pulse_cov}")


def test_fp64_fit_cpu_max_iter(pkg, gpu):
    """multifit_cpu's FNNLS cap of 10 (PulseChiSqSNNLS.cpp:243) truncates some fits."""
    ct, s, t, r = _run(pkg, "c2", "live", 5000, max_iter=10)
    _gate(ct, s, t, r, oracle_opts(max_iter=10), label="fp64 max_iter 10")


def test_fp64_fit_edge_cases(pkg, gpu):
    """All-zero, NaN, huge and negative samples; empty batch."""
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config("c1")
    ct = w.chtypes()
    eng = engine.Engine(list(ct))
    s = np.zeros((10, 6))
    s[:, 1] = np.nan
    s[:, 2] = 1e6 * np.exp(-0.5 * (np.arange(10) - 5) ** 2)
    s[:, 3] = -5.0
    s[:, 4] = 1e-300
    s[3, 5] = np.inf
    r = eng.fit(torch.from_numpy(s).cuda(), None, _abi.default_opts())
    torch.cuda.synchronize()
    x = r.x.cpu().numpy()
    fl = r.flags.cpu().numpy()
    assert np.all(x[:, 0] == 0) and r.chi2[0].item() == 0.0
    assert fl[1] & _abi.FLAGS["nonfinite"] and fl[5] & _abi.FLAGS["nonfinite"]
    assert np.all(x[:, 1] == 0) and np.all(x[:, 5] == 0)
    ref = oracle_fit(ct, s[:, [0, 3, 4]].copy(), None, oracle_opts())
    xr = ref["x"].T
    got = x[:, [0, 3, 4]]
    scale = np.maximum(np.abs(xr).max(axis=0), 1.0)
    assert np.all(np.abs(got - xr) / scale <= 1e-10)
    # 1e6 ADC: FP64 rounding of w exceeds eps = 1e-11 and the reference's
    # decisions become rounding-dependent; the fit stays finite and feasible
    assert np.all(x[:, 2] >= 0) and np.isfinite(x[:, 2]).all() and x[5, 2] > 0
    # empty batch is a no-op
    eng.fit(torch.zeros((10, 0), dtype=torch.float64, device="cuda"), None, _abi.default_opts())


def test_fp64_fit_float_samples(pkg, gpu):
    """FP32 sample storage feeds the FP64 fit: same as FP64 samples rounded to FP32."""
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config("c2")
    ct = w.chtypes()
    s, t, _ = w.generate(3000)
    s32 = s.astype(np.float32)
    eng = engine.Engine(list(ct))
    r = eng.fit(torch.from_numpy(s32).cuda(), None, _abi.default_opts())
    torch.cuda.synchronize()
    ref = oracle_fit(ct, s32.astype(np.float64), None, oracle_opts(), nthreads=8)
    parity.gate(r.x.cpu().numpy(), r.chi2.cpu().numpy(), ref, TOL_X, iters=r.iters.cpu().numpy(),
                flags=r.flags.cpu().numpy(), label="fp64 float samples")


def test_fit_rejects_bad_input(pkg, gpu):
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config("c3")
    eng = engine.Engine(list(w.chtypes()))
    s = torch.zeros((10, 4), dtype=torch.float64, device="cuda")
    s[5] = 20.0
    bad_t = torch.tensor([0, 1, 2, 0], dtype=torch.uint8, device="cuda")
    # a type id that is not set: that channel is not fitted and is flagged
    # (the partition runs on the device; nothing is read back to the host)
    for prec in (_abi.MF_PREC_FP64, _abi.MF_PREC_FP32):
        r = eng.fit(s if prec == _abi.MF_PREC_FP64 else s.float(), bad_t,
                    _abi.default_opts(precision=prec))
        torch.cuda.synchronize()
        fl = r.flags.cpu().numpy()
        assert fl[2] == _abi.FLAGS["bad_type"] and torch.all(r.x[:, 2] == 0)
        assert not np.any(fl[[0, 1, 3]] & _abi.FLAGS["bad_type"])
        assert torch.all(r.x[5, [0, 3]] > 0)
    with pytest.raises(_abi.MFError):
        eng.fit(s, None, _abi.default_opts(precision=7))
    with pytest.raises(_abi.MFError):
        eng.fit(s, None, _abi.default_opts(block=100))


@pytest.mark.parametrize("mode", ["plain", "fast", "warm"])
def test_type_partition_equals_per_type_batches(pkg, gpu, mode):
    """A mixed batch is split by type on the device (per-chunk counts, scans,
    a stable scatter).  Batch sizes around the chunking (one wavefront, one
    block, several 256-channel passes per chunk), with unset type ids mixed
    in: every channel is fitted once with its own type, bit for bit as in a
    batch of that type alone, and a channel with an unset type is flagged."""
    from test_regression_amd import engine, workloads, warmstart, _abi
    w = workloads.config("c3")
    eng = engine.Engine(list(w.chtypes()))
    if mode == "warm":
        warmstart.ensure_trained(eng)
    prec = _abi.MF_PREC_FP64 if mode == "plain" else _abi.MF_PREC_FP32
    o = _abi.default_opts(precision=prec, warm_start=1 if mode == "warm" else 0)
    dt = torch.float64 if mode == "plain" else torch.float32
    rng = np.random.default_rng(31)
    for n in (1, 63, 257, 131_073, 300_001):
        s, _, _ = w.generate(n)
        t = rng.choice(np.array([0, 1, 5], dtype=np.uint8), size=n, p=[0.6, 0.38, 0.02])
        print(f"[gpu-fit] partition mode={mode} n={n}", flush=True)
        x = torch.from_numpy(s).to("cuda", dt)
        r = eng.fit(x, torch.from_numpy(t).cuda(), o)
        torch.cuda.synchronize()
        rx, rf, ri, rc = (r.x.cpu().numpy(), r.flags.cpu().numpy(), r.iters.cpu().numpy(),
                          r.chi2.cpu().numpy())
        bad = t == 5
        assert np.all(rf[bad] == _abi.FLAGS["bad_type"]) and np.all(rx[:, bad] == 0)
        assert not np.any(rf[~bad] & _abi.FLAGS["bad_type"])
        for ty in (0, 1):
            sel = np.flatnonzero(t == ty)
            if sel.size == 0:
                continue
            q = eng.fit(x[:, torch.from_numpy(sel).cuda()].contiguous(),
                        torch.full((sel.size,), ty, dtype=torch.uint8, device="cuda"), o)
            torch.cuda.synchronize()
            assert np.array_equal(q.x.cpu().numpy(), rx[:, sel]), (n, ty)
            assert np.array_equal(q.flags.cpu().numpy(), rf[sel])
            assert np.array_equal(q.iters.cpu().numpy(), ri[sel])
            assert np.array_equal(q.chi2.cpu().numpy(), rc[sel])


@pytest.mark.parametrize("pulse_cov", ["zero", "live"])
def test_fp64_fit_minos_errors(pkg, gpu, pulse_cov):
    """compute_errors: MINOS-like in-time uncertainty (PulseChiSqSNNLS.cpp:73-146).
    With a live pulse covariance the reference itself returns NaN for most
    channels (chi2 at x0 +- sigma below the fit's chi2 after its warm-start
    jump); the NaN pattern must match too."""
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config("c2", pulse_cov)
    ct = w.chtypes()
    s, _, _ = w.generate(5000)
    eng = engine.Engine(list(ct))
    print(f"[gpu-minos] pulse_cov={pulse_cov}", flush=True)
    r = eng.fit(torch.from_numpy(s).cuda(), None, _abi.default_opts(compute_errors=1))
    torch.cuda.synchronize()
    ref = oracle_fit(ct, s, None, oracle_opts(compute_errors=1), nthreads=8)
    parity.errors_gate(r.err.cpu().numpy()[5], ref["err"][:, 5], ref, 1e-8,
                       label=f"minos {pulse_cov}")
    assert np.all(r.err.cpu().numpy()[[0, 1, 2, 3, 4, 6, 7, 8, 9]] == 0)
    # the amplitudes are those of the fit without errors
    parity.gate(r.x.cpu().numpy(), r.chi2.cpu().numpy(), ref, TOL_X,
                label=f"minos fit {pulse_cov}")


@pytest.mark.parametrize("mode", ["plain", "fast", "warm"])
def test_repeated_fits_are_identical(pkg, gpu, mode):
    """The same call repeated on a mixed batch (type partition, per-type
    launches, the MLP inside the warm launch) gives the same bits every time:
    no result depends on the order in which blocks or atomics run."""
    import hashlib
    from test_regression_amd import engine, workloads, warmstart, _abi
    w = workloads.config("c3")
    s, t, _ = w.generate(100_000)
    eng = engine.Engine(list(w.chtypes()))
    if mode == "warm":
        eng.mlp_set_params(np.load(warmstart.weights_for("c3")))
    prec = _abi.MF_PREC_FP64 if mode == "plain" else _abi.MF_PREC_FP32
    o = _abi.default_opts(precision=prec, warm_start=1 if mode == "warm" else 0)
    x = torch.from_numpy(s).to("cuda", torch.float64 if mode == "plain" else torch.float32)
    tt = torch.from_numpy(t).cuda()
    out = eng.alloc_result(x.shape[1], prec)
    print(f"[gpu-fit] repeated fits mode={mode}", flush=True)

    def digest():
        h = hashlib.sha1()
        for a in (out.x, out.chi2, out.flags, out.iters):
            h.update(a.cpu().numpy().tobytes())
        return h.hexdigest()

    eng.fit(x, tt, o, out)
    torch.cuda.synchronize()
    ref = digest()
    for _ in range(10):
        out.x.zero_()
        out.flags.zero_()
        eng.fit(x, tt, o, out)
        torch.cuda.synchronize()
        assert digest() == ref
