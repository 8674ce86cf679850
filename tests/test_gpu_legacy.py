"""GPU parity of the legacy NNLS solver (MF_SOLVER_LEGACY,
legacy_multifit_cpu/src/PulseChiSqSNNLS.cpp:268-363) on the plain FP64 path
against the oracle's legacy restatement.

Bar (FP64): identical active sets, amplitudes within 1e-10 of the oracle
relative to the channel's largest amplitude (floor 1 ADC), chi2 within 1e-10
relative (floor 1), identical Minimize / FNNLS addition / solve counts and
flags on >= 99.9% of channels, MINOS errors within 1e-9 relative with the
same NaN pattern.
"""
import numpy as np
import pytest

import parity
import torch

from oracle_binding import OR_SOLVER_LEGACY, oracle_fit, oracle_opts

pytestmark = pytest.mark.gpu


def _fit(cfg, pulse_cov, n, **kw):
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config(cfg, pulse_cov)
    ct = w.chtypes()
    s, t, _ = w.generate(n)
    eng = engine.Engine(list(ct))
    opts = _abi.default_opts(solver=_abi.MF_SOLVER_LEGACY, **kw)
    print(f"[gpu-legacy] cfg={cfg} pulse_cov={pulse_cov} n={n} {kw}", flush=True)
    r = eng.fit(torch.from_numpy(s).cuda(),
                torch.from_numpy(t).cuda() if t is not None else None, opts)
    torch.cuda.synchronize()
    return ct, s, t, r


@pytest.mark.parametrize("cfg,pulse_cov,n", [
    ("c1", "zero", 10_000),
    ("c2", "zero", 20_000),
    ("c2", "live", 20_000),
    ("c3", "live", 20_000),
    ("c5", "zero", 10_000),
])
def test_legacy_fit_matches_oracle(pkg, gpu, cfg, pulse_cov, n):
    ct, s, t, r = _fit(cfg, pulse_cov, n)
    ref = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LEGACY), nthreads=8)
    parity.gate(r.x.double().cpu().numpy(), r.chi2.double().cpu().numpy(), ref, 1e-10,
                iters=r.iters.cpu().numpy(), flags=r.flags.cpu().numpy(), compare_inner=True,
                label=f"legacy {cfg} {pulse_cov}")


@pytest.mark.parametrize("pulse_cov", ["zero", "live"])
def test_legacy_minos_errors_match_oracle(pkg, gpu, pulse_cov):
    ct, s, t, r = _fit("c2", pulse_cov, 5000, compute_errors=1)
    ref = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LEGACY, compute_errors=1),
                     nthreads=8)
    parity.errors_gate(r.err.double().cpu().numpy()[5], ref["err"][:, 5], ref, 1e-9,
                       label=f"legacy minos {pulse_cov}")


def test_legacy_needs_fp64(pkg, gpu):
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config("c2", "zero")
    s, _, _ = w.generate(256)
    eng = engine.Engine(list(w.chtypes()))
    opts = _abi.default_opts(precision=_abi.MF_PREC_FP32, solver=_abi.MF_SOLVER_LEGACY)
    with pytest.raises(Exception, match="inplace_fnnls only"):
        eng.fit(torch.from_numpy(s.astype(np.float32)).cuda(), None, opts)
