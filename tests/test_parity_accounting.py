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


@pytest.fixture(scope="module")
def near_exact(pkg):
    """Two c2 channels (zero pulse covariance) fitted almost exactly (chi2 ~
    1e-11) by nine pulses, where the sets with BX -5 and with BX -4 both pass
    the reference's stop test: the oracle's cold start ends at one, a warm
    start from a prediction of the other ends there."""
    from test_regression_amd import workloads
    w = workloads.config("c2", "zero")
    ct = w.chtypes()
    s, _, _ = w.generate(720_000)
    s = np.ascontiguousarray(s[:, [119053, 719443]].astype(np.float32).astype(np.float64))
    ref = oracle_fit(ct, s, None, oracle_opts())
    return ct, s, ref


def _alt(ref, c):
    pm = int(sum(1 << i for i in range(10) if ref["x"][c][i] > 0))
    return pm ^ 0x3   # BX -5 and BX -4 exchanged


def test_another_end_point_of_the_stop_test_is_explained(near_exact):
    from oracle_binding import oracle_kkt_sets
    ct, s, ref = near_exact
    assert (ref["chi2"] < 1e-9).all()
    pm = [_alt(ref, c) for c in range(2)]
    k = oracle_kkt_sets(ct, s, None, [0, 1], pm)
    assert k["feasible"].all() and (k["margin"] < 0).all()
    # the fit continued from the other set keeps it
    assert ((k["fit"]["x"] > 0) == ((np.array(pm)[:, None] >> np.arange(10)) & 1)).all()
    x = k["fit"]["x"].T.astype(np.float32).astype(np.float64)
    rep = parity.account(x, k["fit"]["chi2"], ref, 1e-4, sets=(ct, s, None))
    assert rep["outside_tolerance"] == 2 and rep["by_alternative_end_point"] == 2
    assert rep["unexplained"] == 0
    # without the samples the accounting cannot tell: both unexplained
    assert parity.account(x, k["fit"]["chi2"], ref, 1e-4)["unexplained"] == 2


@pytest.mark.parametrize("mistake", ["off_by_1e-3", "in_time_dropped"])
def test_planted_mistakes_are_not_another_end_point(near_exact, mistake):
    from oracle_binding import oracle_kkt_sets
    ct, s, ref = near_exact
    pm = [_alt(ref, c) for c in range(2)]
    if mistake == "in_time_dropped":
        pm = [p & ~(1 << 5) for p in pm]   # the in-time pulse left out: w_5 far above eps
    k = oracle_kkt_sets(ct, s, None, [0, 1], pm)
    x = k["fit"]["x"].T.copy()
    if mistake == "off_by_1e-3":
        x[5] *= 1 + 1e-3
    else:
        assert (k["margin"] > 1e6).all()
    rep = parity.account(x, k["fit"]["chi2"], ref, 1e-4, sets=(ct, s, None))
    assert rep["unexplained"] == 2 and rep["by_alternative_end_point"] == 0
