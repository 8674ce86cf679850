"""GPU verification of warm starts: channels whose prediction the KKT test
rejects, and starts from deliberately wrong predictions, end at the oracle's
cold optimum; a warm start whose iterations reach max_iter falls back to a
cold start (MF_F_WARM_FALLBACK).

A wrong prediction is planted through the network itself: all weights 0 and
the output biases chosen, so that every channel is predicted the same passive
set (logit > 0) with amplitude softplus(bias) times its sample scale.  Both
warm paths run: the single launch (prediction and FP64 phase in one kernel)
and the persistent kernels (prediction through memory, lanes refilled from a
queue).  Bar of the parity checks: tests/parity.py's channel accounting,
1e-4 per amplitude and chi2 (the fast path's FP32 bar)."""
import numpy as np
import pytest
import torch

import parity
from oracle_binding import oracle_fit, oracle_opts

pytestmark = pytest.mark.gpu

NPARAM_B3 = 20   # the last 20 parameters: output biases (10 logits, 10 amplitudes)


def _constant_network(eng, passive, amp_bias=2.0):
    from test_regression_amd import _abi
    p = np.zeros(_abi.MF_MLP_NPARAM, dtype=np.float32)
    b3 = p[-NPARAM_B3:]
    b3[:10] = -1.0
    for i in passive:
        b3[i] = 1.0
    b3[10:] = amp_bias
    eng.mlp_set_params(p)


def _setup(cfg, pulse_cov, n, seed=0):
    from test_regression_amd import engine, workloads
    w = workloads.config(cfg, pulse_cov)
    w.seed += seed
    ct = w.chtypes()
    s, t, _ = w.generate(n)
    eng = engine.Engine(list(ct))
    x = torch.from_numpy(s.astype(np.float32)).cuda()
    tt = torch.from_numpy(t).cuda() if t is not None else None
    s_ref = s.astype(np.float32).astype(np.float64)
    return eng, ct, x, tt, s_ref, t


def _fit(eng, x, tt, **kw):
    from test_regression_amd import _abi
    r = eng.fit(x, tt, _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=1, **kw))
    torch.cuda.synchronize()
    return r


def _np(r):
    return (r.x.double().cpu().numpy(), r.chi2.double().cpu().numpy(),
            r.flags.cpu().numpy(), r.iters.cpu().numpy())


@pytest.mark.parametrize("cfg,pulse_cov,n", [("c2", "zero", 50_000), ("c2", "live", 20_000),
                                             ("c3", "zero", 20_000)])
@pytest.mark.parametrize("persistent", [0, 1])
def test_rejected_warm_starts_match_the_oracle(pkg, gpu, cfg, pulse_cov, n, persistent):
    """The committed network's predictions: the channels whose predicted set
    failed its KKT test, gated alone against the oracle; at the default
    max_iter no channel falls back."""
    from test_regression_amd import _abi, warmstart
    eng, ct, x, tt, s_ref, t = _setup(cfg, pulse_cov, n)
    warmstart.ensure_trained(eng, path=warmstart.weights_for(cfg))
    print(f"[gpu-warm] rejected cfg={cfg} {pulse_cov} n={n} persistent={persistent}", flush=True)
    xw, cw, fw, iw = _np(_fit(eng, x, tt, persistent=persistent))
    ref = oracle_fit(ct, s_ref, t, oracle_opts(), nthreads=8)
    rej = (fw & _abi.FLAGS["warm_reject"]) != 0
    print(f"rejected {rej.mean():.4f}", flush=True)
    assert rej.sum() > 50
    assert not np.any(fw & _abi.FLAGS["warm_fallback"])
    parity.gate(xw, cw, ref, 1e-4, iters=iw, compare_outer=False, mask=rej,
                label=f"rejected warm starts {cfg} {pulse_cov}", sets=(ct, s_ref, t))


@pytest.mark.parametrize("passive", [(0,), (0, 1, 2, 3, 4, 6, 7, 8, 9), tuple(range(10))])
@pytest.mark.parametrize("persistent", [0, 1])
def test_wrong_predictions_end_at_the_oracle_optimum(pkg, gpu, passive, persistent):
    """Every channel started from the same wrong set (one early pulse alone,
    all but the in-time pulse, all ten) ends at the oracle's cold optimum."""
    from test_regression_amd import _abi
    eng, ct, x, tt, s_ref, t = _setup("c2", "zero", 30_000, seed=7)
    _constant_network(eng, passive)
    print(f"[gpu-warm] constant prediction {passive} persistent={persistent}", flush=True)
    xw, cw, fw, iw = _np(_fit(eng, x, tt, persistent=persistent))
    rej = (fw & _abi.FLAGS["warm_reject"]) != 0
    print(f"rejected {rej.mean():.4f}", flush=True)
    # from all ten, most channels only step back (removals), which the
    # KKT test at the end accepts
    assert rej.mean() > (0.1 if len(passive) == 10 else 0.5)
    assert not np.any(fw & _abi.FLAGS["warm_fallback"])
    ref = oracle_fit(ct, s_ref, t, oracle_opts(), nthreads=8)
    parity.gate(xw, cw, ref, 1e-4, iters=iw, compare_outer=False,
                label=f"warm from {passive}", sets=(ct, s_ref, t))


@pytest.mark.parametrize("persistent", [0, 1])
def test_fallback_at_max_iter_is_the_cold_start(pkg, gpu, persistent):
    """With max_iter 4 and every channel predicted BX -5 alone, most channels
    need more adds than that from the prediction: they reach the cap, are
    refitted from x = 0 and flagged.  Their amplitudes and chi2 are, bit for
    bit, those of a start from the empty set (a network predicting nothing
    passive) under the same cap."""
    from test_regression_amd import _abi
    eng, ct, x, tt, s_ref, t = _setup("c2", "zero", 30_000, seed=9)
    print(f"[gpu-warm] fallback persistent={persistent}", flush=True)
    _constant_network(eng, (0,))
    xw, cw, fw, iw = _np(_fit(eng, x, tt, persistent=persistent, max_iter=4))
    _constant_network(eng, ())
    x0, c0, f0, i0 = _np(_fit(eng, x, tt, persistent=persistent, max_iter=4))
    fb = (fw & _abi.FLAGS["warm_fallback"]) != 0
    print(f"fallback {fb.mean():.4f}", flush=True)
    assert fb.mean() > 0.5
    assert not np.any(f0 & _abi.FLAGS["warm_fallback"])
    assert np.array_equal(xw[:, fb], x0[:, fb]) and np.array_equal(cw[fb], c0[fb])
    # the fallback keeps the cap's flag when the cold iterations reach it too
    cap = _abi.FLAGS["fnnls_cap"]
    assert np.array_equal(fw[fb] & cap, f0[fb] & cap)
