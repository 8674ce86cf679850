"""The per-channel parity accounting (tests/parity.py) itself: it passes the
fit core's host build against the oracle, and it fails -- names the channel
unexplained -- for mistakes planted in a well-conditioned channel whose
decisions were far from any threshold: an amplitude off by 1e-9, a passive
amplitude dropped to zero, a Minimize count off by one."""
import numpy as np
import pytest

import parity
from oracle_binding import hostcore_fit, oracle_fit, oracle_opts


@pytest.fixture(scope="module")
def fits(pkg):
    from test_regression_amd import workloads, _abi
    w = workloads.config("c2", "live")
    ct = w.chtypes()
    s, t, _ = w.generate(3000)
    ref = oracle_fit(ct, s, t, oracle_opts(), nthreads=4)
    x, chi2, fl, it = hostcore_fit(ct, s, t, _abi.default_opts())
    return ref, x, chi2, fl, it


def _clean_channel(ref):
    dec = np.minimum.reduce([ref["m_eps"], ref["m_arg"], ref["m_pos"], ref["m_alpha"],
                             ref["m_chi2"]])
    ok = (ref["pivot"] > 1e-3) & (ref["pivot"] < np.inf) & (dec > 1e4) & \
         ((ref["x"] > 1.0).sum(axis=1) >= 3)
    assert ok.any()
    return int(np.flatnonzero(ok)[0])


def test_accounting_passes_the_host_build(fits):
    ref, x, chi2, fl, it = fits
    rep = parity.gate(x, chi2, ref, 1e-10, iters=it, flags=fl, label="host build")
    assert rep["unexplained"] == 0


@pytest.mark.parametrize("mistake", ["amplitude", "dropped", "iterations"])
def test_accounting_catches_planted_mistakes(fits, mistake):
    ref, x, chi2, fl, it = fits
    c = _clean_channel(ref)
    x, it = x.copy(), it.copy()
    p = int(np.argmax(ref["x"][c]))
    if mistake == "amplitude":
        x[p, c] *= 1 + 1e-9
    elif mistake == "dropped":
        q = int(np.flatnonzero(ref["x"][c] > 1.0)[-1])
        x[q, c] = 0.0
    else:
        it[c] += 1
    rep = parity.account(x, chi2, ref, 1e-10, iters=it, flags=fl)
    assert rep["unexplained_channels"] == [c]
    with pytest.raises(AssertionError):
        parity.check(rep)
