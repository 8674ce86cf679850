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
    x = r.x.double().cpu().numpy()
    xr = ref["x"].T
    scale = np.maximum(np.abs(xr).max(axis=0), 1.0)
    err = (np.abs(x - xr) / scale).max(axis=0)
    aset = ((x > 0) == (xr > 0)).all(axis=0)
    chi2 = r.chi2.double().cpu().numpy()
    cerr = np.abs(chi2 - ref["chi2"]) / np.maximum(ref["chi2"], 1.0)
    it = r.iters.cpu().numpy()
    same_it = ((it & 255) == ref["n_minimize"]) & (((it >> 8) & 255) == ref["n_outer"]) & \
              ((it >> 16) == ref["n_inner"])
    fl = r.flags.cpu().numpy()
    print(f"max err {err.max():.3e} aset {aset.mean():.6f} chi2 {cerr.max():.3e} "
          f"same iters {same_it.mean():.6f} same flags {(fl == ref['flags']).mean():.6f}")
    assert aset.mean() >= 0.999
    assert np.quantile(err, 0.999) <= 1e-10
    assert np.quantile(cerr, 0.999) <= 1e-10
    assert same_it.mean() >= 0.999
    assert (fl == ref["flags"]).mean() >= 0.999


@pytest.mark.parametrize("pulse_cov", ["zero", "live"])
def test_legacy_minos_errors_match_oracle(pkg, gpu, pulse_cov):
    ct, s, t, r = _fit("c2", pulse_cov, 5000, compute_errors=1)
    ref = oracle_fit(ct, s, t, oracle_opts(solver=OR_SOLVER_LEGACY, compute_errors=1),
                     nthreads=8)
    e, er = r.err.double().cpu().numpy()[5], ref["err"][:, 5]
    nan_same = np.isnan(e) == np.isnan(er)
    ok = ~np.isnan(er) & ~np.isnan(e)
    rel = np.abs(e[ok] - er[ok]) / np.abs(er[ok])
    print(f"nan same {nan_same.mean():.6f} max rel {rel.max():.3e}")
    assert nan_same.mean() >= 0.999
    assert np.quantile(rel, 0.999) <= 1e-9


def test_legacy_needs_fp64(pkg, gpu):
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config("c2", "zero")
    s, _, _ = w.generate(256)
    eng = engine.Engine(list(w.chtypes()))
    opts = _abi.default_opts(precision=_abi.MF_PREC_FP32, solver=_abi.MF_SOLVER_LEGACY)
    with pytest.raises(Exception, match="inplace_fnnls only"):
        eng.fit(torch.from_numpy(s.astype(np.float32)).cuda(), None, opts)
