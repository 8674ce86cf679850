"""Per-channel parity accounting against the FP64 oracle.

Every channel is checked; none is hidden by a quantile.  A channel is
*outside tolerance* when any of these holds:
  - an amplitude differs by more than tol x max(|x_oracle|, floor), the
    per-amplitude relative error with an absolute floor (1 ADC);
  - its passive set (x > 0) differs from the oracle's;
  - chi2 differs by more than tol x max(chi2_oracle, 1);
  - (where compared) its Minimize / FNNLS iteration counts or flags differ.
Each such channel must be *explained* by the oracle's own record of how
close its fit came to a different outcome (or_result.m_*, pivot):
  - conditioning: with the same passive set, counts and flags, every
    amplitude difference is within BAND_X rounding units of the channel's
    scale (its largest amplitude, floor 1), a unit being u (n + 1/rcond)
    with u = 2^-53 and rcond the reciprocal condition number of the passive
    blocks whose solves produced the final amplitudes: two backward-stable
    solves of a block that ill-conditioned legitimately differ by that much;
  - a decision within rounding: one of the oracle's decisions (stop test,
    argmax of w, sign of a passive solution, step ratio, Minimize's |dchi2|
    test) was within BAND_D rounding units of going the other way (the
    oracle expresses each margin in units of the rounding error of the
    quantity decided on);
  - another end point of the same algorithm (when the samples are given,
    sets=(chtypes, samples, type_of[, pederr])): the stop test max w < eps
    leaves more than one passive set acceptable where the fit is nearly
    exact (chi2 ~ 1e-11: every gradient is tiny), and a warm start can end
    the first NNLS call at another one.  The code's passive set, solved
    exactly on the first call's FP64 normal equations, has positive
    amplitudes and passes the stop test (within BAND_D rounding units of
    w); the oracle's fit continued from it (or_dofit_from_set) gives the
    code's chi2 and Minimize count to the tolerance, and its amplitudes to
    the tolerance or, where that set's passive block is ill-conditioned
    (early-BX pulses that are nearly the same template tail), to BAND_X of
    that block's rounding units: the optimum's chi2 is unique, its
    minimizer is fixed only to the block's conditioning.
A channel with none of these is a failure, and the numbers are capped: at
most 0.5% of the batch outside tolerance, at most 0.1% explained only by a
decision margin or another end point (at least 8 channels either way).
"""
from __future__ import annotations

import numpy as np

U = 2.0 ** -53
BAND_X = 16.0
BAND_D = 8.0


def _units(ref):
    """u (n + 1/rcond) behind each channel's final amplitudes (10 u where the
    oracle solved no passive block; unbounded where one was not positive
    definite)."""
    rc = ref["pivot"]
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(np.isinf(rc) & (rc > 0), 10 * U,
                        np.where(rc > 0, U * (10 + 1.0 / rc), np.inf))


def account(x, chi2, ref, tol, *, tol_chi2=None, floor=1.0, iters=None, flags=None,
            compare_outer=True, compare_inner=False, mask=None, sets=None):
    """x [10, n] and chi2 [n] of the code under test against the oracle's
    records ref (oracle_fit with margins).  iters / flags: compared when
    given (FP64 paths; the fast path takes its own route, so only the
    Minimize count is compared there, compare_outer=False)."""
    tol_chi2 = tol if tol_chi2 is None else tol_chi2
    x = np.asarray(x, dtype=np.float64)
    xr = ref["x"].T
    n = x.shape[1]
    sel = np.ones(n, dtype=bool) if mask is None else np.asarray(mask, dtype=bool)
    d = np.abs(x - xr)
    allowed = tol * np.maximum(np.abs(xr), floor)
    amp_bad = (d > allowed).any(axis=0)
    scale = np.maximum(np.abs(xr).max(axis=0), floor)
    err_ch = d.max(axis=0) / scale
    amp_rel = (d / np.maximum(np.abs(xr), floor)).max(axis=0)
    aset = ((x > 0) == (xr > 0)).all(axis=0)
    cerr = np.abs(np.asarray(chi2, dtype=np.float64) - ref["chi2"]) / np.maximum(ref["chi2"], 1.0)
    chi_bad = ~(cerr <= tol_chi2)
    it_bad = np.zeros(n, dtype=bool)
    fnnls_bad = np.zeros(n, dtype=bool)   # FNNLS counts, where compared
    if iters is not None:
        it = np.asarray(iters)
        it_bad |= (it & 255) != ref["n_minimize"]
        if compare_outer:
            fnnls_bad |= ((it >> 8) & 255) != ref["n_outer"]
        if compare_inner:
            fnnls_bad |= (it >> 16) != ref["n_inner"]
        it_bad |= fnnls_bad
    fl_bad = np.zeros(n, dtype=bool) if flags is None else (np.asarray(flags) != ref["flags"])
    bad = sel & (amp_bad | ~aset | chi_bad | it_bad | fl_bad)
    cond = _units(ref)
    dec = np.minimum.reduce([ref["m_eps"], ref["m_arg"], ref["m_pos"], ref["m_alpha"],
                             ref["m_chi2"]])
    why_cond = aset & ~it_bad & ~fl_bad & (d.max(axis=0) <= BAND_X * cond * scale) & \
        (cerr <= np.maximum(tol_chi2, BAND_X * cond))
    why_dec = dec <= BAND_D
    why_alt = np.zeros(n, dtype=bool)
    cand = np.where(bad & ~why_cond & ~why_dec & ~fl_bad & ~fnnls_bad)[0]
    if sets is not None and cand.size:
        from oracle_binding import oracle_kkt_sets
        ct, s, type_of = sets[:3]
        pm = ((x[:, cand] > 0) * (1 << np.arange(10))[:, None]).sum(axis=0)
        k = oracle_kkt_sets(ct, s, type_of, cand, pm, pederr=sets[3] if len(sets) > 3 else None)
        z = k["fit"]["x"].T
        # amplitudes to the tolerance, or to the conditioning band of the
        # continued fit's own passive blocks (as for the oracle's end point);
        # chi2 always to the tolerance: the optimum's value is well defined
        # however ill-conditioned its minimizer
        band = BAND_X * _units(k["fit"]) * scale[cand]
        close = (np.abs(x[:, cand] - z) <=
                 np.maximum(tol * np.maximum(np.abs(z), floor), band)).all(axis=0) & \
            (np.abs(np.asarray(chi2, dtype=np.float64)[cand] - k["fit"]["chi2"]) <=
             tol_chi2 * np.maximum(k["fit"]["chi2"], 1.0))
        if iters is not None:
            close &= (np.asarray(iters)[cand] & 255) == k["fit"]["n_minimize"]
        why_alt[cand] = k["feasible"] & (k["margin"] <= BAND_D) & close
    explained = why_cond | why_dec | why_alt
    unexpl = bad & ~explained
    return {
        "n": int(sel.sum()),
        "outside_tolerance": int(bad.sum()),
        "unexplained": int(unexpl.sum()),
        "by_conditioning": int((bad & why_cond).sum()),
        "by_decision_margin": int((bad & ~why_cond & why_dec).sum()),
        "by_alternative_end_point": int((bad & why_alt).sum()),
        "different_active_sets": int((sel & ~aset).sum()),
        "err_channel_max": float(err_ch[sel].max()) if sel.any() else 0.0,
        "err_amplitude_max": float(amp_rel[sel].max()) if sel.any() else 0.0,
        "err_amplitude_max_explained_excluded": float(amp_rel[sel & ~bad].max())
        if (sel & ~bad).any() else 0.0,
        "chi2_rel_max": float(cerr[sel].max()) if sel.any() else 0.0,
        "unexplained_channels": [int(c) for c in np.where(unexpl)[0][:20]],
        "_detail": {int(c): {"err": float(err_ch[c]), "amp": float(amp_rel[c]),
                             "aset": bool(aset[c]), "cerr": float(cerr[c]),
                             "it_bad": bool(it_bad[c]), "fl_bad": bool(fl_bad[c]),
                             "units": float(cond[c]), "dec": float(dec[c])}
                    for c in np.where(unexpl)[0][:10]},
    }


def check(rep, max_frac=5e-3, dec_frac=1e-3, label=""):
    """Every channel outside tolerance explained; their number capped: at
    most max_frac of the channels outside tolerance, at most dec_frac of
    them explained only by a decision margin (at least 8 either way)."""
    print(f"parity {label}: {rep['outside_tolerance']} of {rep['n']} outside tolerance "
          f"({rep['by_conditioning']} by conditioning, {rep['by_decision_margin']} by a decision "
          f"margin, {rep['by_alternative_end_point']} another end point, {rep['unexplained']} "
          f"unexplained); amplitude err max {rep['err_amplitude_max']:.2e}"
          f" (within tolerance: {rep['err_amplitude_max_explained_excluded']:.2e}), chi2 "
          f"{rep['chi2_rel_max']:.2e}", flush=True)
    assert rep["unexplained"] == 0, (label, rep["_detail"])
    assert rep["outside_tolerance"] <= max(8, int(max_frac * rep["n"])), (label, rep)
    assert rep["by_decision_margin"] + rep["by_alternative_end_point"] + rep["unexplained"] <= \
        max(8, int(dec_frac * rep["n"])), (label, rep)
    return rep


def gate(x, chi2, ref, tol, max_frac=5e-3, dec_frac=1e-3, label="", **kw):
    """account() + check() in one call; returns the report."""
    return check(account(x, chi2, ref, tol, **kw), max_frac=max_frac, dec_frac=dec_frac,
                 label=label)


def errors_gate(e, er, ref, tol, max_frac=1e-3, label=""):
    """MINOS in-time errors e (code under test) against er (oracle): the NaN
    pattern and the relative error, channel by channel.  A channel outside
    tolerance is explained by its fit's conditioning (the refits at x0 +- sigma
    inherit it: |de|/e within BAND_X u / rcond) or by a decision margin of
    the fit; the rest fail."""
    e = np.asarray(e, dtype=np.float64)
    er = np.asarray(er, dtype=np.float64)
    n = e.size
    nan_bad = np.isnan(e) != np.isnan(er)
    ok = ~np.isnan(e) & ~np.isnan(er)
    rel = np.zeros(n)
    rel[ok] = np.abs(e[ok] - er[ok]) / np.maximum(np.abs(er[ok]), 1e-300)
    bad = nan_bad | (rel > tol)
    cond = _units(ref)
    dec = np.minimum.reduce([ref["m_eps"], ref["m_arg"], ref["m_pos"], ref["m_alpha"],
                             ref["m_chi2"]])
    why = (~nan_bad & (rel <= BAND_X * cond)) | (dec <= BAND_D)
    unexpl = bad & ~why
    print(f"parity {label}: {int(bad.sum())} of {n} errors outside tolerance, "
          f"{int(unexpl.sum())} unexplained; rel max within tolerance "
          f"{rel[ok & ~bad].max() if (ok & ~bad).any() else 0.0:.2e}", flush=True)
    assert not unexpl.any(), (label, np.where(unexpl)[0][:10], rel[unexpl][:10])
    assert bad.sum() <= max(8, int(max_frac * n)), (label, int(bad.sum()))
