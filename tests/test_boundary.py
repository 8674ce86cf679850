"""The fit core's host build (the kernels' source) against the oracle on the
boundary the reference's callers use: BX lists of 1, 5 and 10 pulses in one
batch and a pederr per channel.  The oracle fits each channel with its own
npulses and pederr; the engine's types carry pederr 1 and the per-channel
pederr scales their covariance."""
import numpy as np
import pytest

import parity

from boundary_cases import mixed_batch, npulse_types
from oracle_binding import hostcore_fit, oracle_fit, oracle_opts


def _check(x, chi2, fl, it, ref, t, tol, npulses=(1, 5, 10)):
    xr = ref["x"].T
    scale = np.maximum(np.abs(xr).max(axis=0), 1.0)
    err = (np.abs(x - xr) / scale).max(axis=0)
    aset = ((x > 0) == (xr > 0)).all(axis=0)
    cerr = np.abs(chi2 - ref["chi2"]) / np.maximum(ref["chi2"], 1.0)
    # slots past a channel's pulses stay 0
    npc = np.asarray(npulses)[t]
    pad = np.arange(10)[:, None] >= npc[None, :]
    assert np.all(x[pad] == 0)
    return err, aset, cerr


@pytest.mark.parametrize("pulse_cov", ["zero", "live"])
def test_plain_core_pederr_and_npulses(pkg, pulse_cov):
    from test_regression_amd import _abi
    ct = npulse_types(pulse_cov)
    s, t, ped = mixed_batch(3000, pulse_cov=pulse_cov)
    ref = oracle_fit(ct, s, t, oracle_opts(), nthreads=4, pederr=ped)
    x, chi2, fl, it = hostcore_fit(ct, s, t, _abi.default_opts(), pederr=ped)
    _check(x, chi2, fl, it, ref, t, 1e-10)
    parity.gate(x, chi2, ref, 1e-10, iters=it, flags=fl, label=f"plain core boundary {pulse_cov}")


@pytest.mark.parametrize("pulse_cov", ["zero", "live"])
def test_fast_core_pederr_and_npulses(pkg, pulse_cov):
    from test_regression_amd import _abi
    ct = npulse_types(pulse_cov)
    s, t, ped = mixed_batch(3000, pulse_cov=pulse_cov)
    s = s.astype(np.float32).astype(np.float64)
    ref = oracle_fit(ct, s, t, oracle_opts(), nthreads=4, pederr=ped)
    x, chi2, fl, it = hostcore_fit(ct, s, t, _abi.default_opts(precision=_abi.MF_PREC_FP32),
                                   pederr=ped)
    _check(x, chi2, fl, it, ref, t, 1e-4)
    parity.gate(x, chi2, ref, 1e-4, iters=it, compare_outer=False,
                label=f"fast core boundary {pulse_cov}")


def test_pederr_scaling_invariants(pkg):
    """A per-channel pederr p on a type of pederr 1 equals a type of pederr p
    (FP64, the same arithmetic: the covariance is p^2 noisecor either way)."""
    from test_regression_amd import _abi
    ct1 = npulse_types("live", npulses=(10,))
    s, _, ped = mixed_batch(500, pulse_cov="live")
    ped[:] = 1.7
    ct2 = npulse_types("live", npulses=(10,), pederr=1.7)
    a = hostcore_fit(ct1, s, None, _abi.default_opts(), pederr=ped)
    b = hostcore_fit(ct2, s, None, _abi.default_opts())
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # pederr enters squared, as in the reference (updateCov): -1 fits as 1;
    # a zero or non-finite pederr is not fitted
    ped2 = np.full(5, 1.0)
    ped2[1], ped2[2], ped2[3], ped2[4] = 0.0, -1.0, np.nan, np.inf
    x, chi2, fl, it = hostcore_fit(ct1, s[:, :5].copy(), None, _abi.default_opts(), pederr=ped2)
    bad = np.array([False, True, False, True, True])
    assert np.all(fl[bad] & _abi.FLAGS["nonfinite"]) and np.all(x[:, bad] == 0)
    assert not np.any(fl[~bad] & _abi.FLAGS["nonfinite"])
    x1, *_ = hostcore_fit(ct1, s[:, [2]].copy(), None, _abi.default_opts(), pederr=np.ones(1))
    assert np.array_equal(x[:, 2], x1[:, 0])
