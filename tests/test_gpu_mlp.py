"""GPU tests of the warm-start MLP kernels against the CPU oracle
(tests/mlp_ref.py), and of the warm-started fit.

  gradient   mf_mlp_grad vs the oracle's analytic gradient (itself pinned by
             finite differences in test_mlp_cpu.py), same bfloat16 weights
  Adam       one mf_mlp_train_step vs the oracle's Adam on the GPU gradient
  inference  mf_mlp_predict (MFMA bf16) vs the oracle's forward with the same
             bfloat16 rounding of weights and of each layer's input
  training   a short training run on plain-FP64 labels learns the passive set
  warm start the warm-started fast fit equals the cold one and the oracle (the
             result is KKT-verified in FP64), with fewer active-set solves
"""
import numpy as np
import pytest

import parity
import torch

import mlp_ref as M
from oracle_binding import oracle_fit, oracle_opts

pytestmark = pytest.mark.gpu


def _engine_and_data(n, seed_off=0, cfg="c2"):
    from test_regression_amd import engine, workloads
    w = workloads.config(cfg)
    w.seed += seed_off
    ct = w.chtypes()
    s, t, _ = w.generate(n)
    return engine.Engine(list(ct)), ct, s


def _params(seed=3):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed)
    p[640:704] = rng.normal(0, 0.1, 64)
    p[4800:4864] = rng.normal(0, 0.1, 64)
    p[-20:] = rng.normal(0, 0.5, 20)
    return p.astype(np.float32)


def test_mlp_gradient_matches_oracle(pkg, gpu):
    eng, ct, s = _engine_and_data(512)
    ref = oracle_fit(ct, s, None, oracle_opts(), nthreads=8)
    xref = np.ascontiguousarray(ref["x"].T, dtype=np.float32)
    s32 = s.astype(np.float32)
    p = _params()
    eng.mlp_set_params(p)
    print("[gpu-mlp] gradient B=512", flush=True)
    g, loss = eng.mlp_grad(torch.from_numpy(s32).cuda(), torch.from_numpy(xref).cuda())
    torch.cuda.synchronize()
    lr, gr = M.loss_and_grad(p.astype(np.float64), s32.astype(np.float64),
                             xref.astype(np.float64))
    gg = g.cpu().numpy().astype(np.float64)
    assert abs(loss.item() - lr) <= 1e-5 * abs(lr)
    scale = np.abs(gr).max()
    assert np.all(np.abs(gg - gr) <= 1e-4 * scale + 1e-3 * np.abs(gr))
    assert np.dot(gg, gr) / (np.linalg.norm(gg) * np.linalg.norm(gr)) > 0.999999


def test_mlp_adam_step_matches_oracle(pkg, gpu):
    eng, ct, s = _engine_and_data(256, 1)
    ref = oracle_fit(ct, s, None, oracle_opts(), nthreads=8)
    xref = torch.from_numpy(np.ascontiguousarray(ref["x"].T, dtype=np.float32)).cuda()
    s32 = torch.from_numpy(s.astype(np.float32)).cuda()
    p = _params(5)
    eng.mlp_set_params(p)
    g, _ = eng.mlp_grad(s32, xref)
    eng.mlp_train_step(s32, xref, 1e-3)
    torch.cuda.synchronize()
    q = eng.mlp_get_params()
    qr, _, _ = M.adam(p.astype(np.float64), g.cpu().numpy().astype(np.float64),
                      np.zeros(M.NPARAM), np.zeros(M.NPARAM), 1, 1e-3)
    assert np.allclose(q, qr, rtol=1e-6, atol=1e-7)


def test_mlp_predict_mfma_matches_oracle(pkg, gpu):
    eng, ct, s = _engine_and_data(3000, 2)
    p = _params(7)
    eng.mlp_set_params(p)
    s32 = s.astype(np.float32)
    print("[gpu-mlp] predict n=3000", flush=True)
    out = eng.mlp_predict(torch.from_numpy(s32).cuda()).cpu().numpy().astype(np.float64)
    # oracle forward with the kernel's roundings: bf16 weights, bf16 layer inputs
    X, m = M.features(s32.astype(np.float64))
    W1, b1, W2, b2, W3, b3 = M.unflatten(M.bf16_round(p).astype(np.float64))
    q = lambda a: M.bf16_round(a.astype(np.float32)).astype(np.float64)  # noqa: E731
    h1 = M.gelu(q(X) @ W1 + b1)
    h2 = M.gelu(q(h1) @ W2 + b2)
    z3 = q(h2) @ W3 + b3
    logits, amps = z3[:, :10].T, (M.softplus(z3[:, 10:]) * m[:, None]).T
    # float32 vs float64 rounding of a layer input can land on the other side
    # of a bfloat16 rounding boundary (one bf16 ulp, 2^-8): rare, bounded
    el = np.abs(out[:10] - logits) / (1 + np.abs(logits))
    ea = np.abs(out[10:] - amps) / (1 + np.abs(amps))
    assert np.median(el) < 1e-6 and np.median(ea) < 1e-6       # the fragment layout is exact
    assert (el <= 2e-4).mean() >= 0.999 and (ea <= 2e-4).mean() >= 0.999
    assert el.max() <= 3e-2 and ea.max() <= 3e-2


def test_mlp_training_and_warm_start(pkg, gpu):
    from test_regression_amd import _abi, warmstart
    eng, ct, s = _engine_and_data(262_144, 10)
    s64 = torch.from_numpy(s).cuda()
    print("[gpu-mlp] label + train 262144", flush=True)
    xref = warmstart.label(eng, s64)
    rep = warmstart.train(eng, s64.float(), xref, epochs=4, batch=4096, lr=3e-3)
    assert rep.steps == 4 * 64 and np.isfinite(rep.final_loss)
    # held-out channels
    from test_regression_amd import workloads
    w = workloads.config("c2")
    w.seed += 99
    sh, _, _ = w.generate(40_000)
    sh32 = torch.from_numpy(sh.astype(np.float32)).cuda()
    xh = warmstart.label(eng, torch.from_numpy(sh).cuda())
    ev = warmstart.evaluate(eng, sh32, xh)
    print(f"[gpu-mlp] held-out {ev}", flush=True)
    assert ev["bit_accuracy"] > 0.85
    cold = eng.fit(sh32, None, _abi.default_opts(precision=_abi.MF_PREC_FP32))
    warm = eng.fit(sh32, None, _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=1))
    torch.cuda.synchronize()
    ref = oracle_fit(ct, sh.astype(np.float32).astype(np.float64), None, oracle_opts(),
                     nthreads=8)
    parity.gate(warm.x.double().cpu().numpy(), warm.chi2.double().cpu().numpy(), ref, 1e-4,
                iters=warm.iters.cpu().numpy(), compare_outer=False, label="freshly trained warm")
    # ablation: the prediction replaces most of the active-set solves
    inner_c = cold.n_inner().float().mean().item()
    inner_w = warm.n_inner().float().mean().item()
    print(f"[gpu-mlp] solves per channel: cold {inner_c:.2f} warm {inner_w:.2f}", flush=True)
    assert inner_w < 0.6 * inner_c


def test_committed_networks_on_held_out_channels(pkg, gpu):
    """The committed parameters on channels neither network was trained on
    (another seed): the c2 network on the c2 distribution; on the c3 mixture
    (ECAL and HCAL-like types), the network trained on that mixture predicts
    the exact passive set far more often than the c2 one, and the warm path
    then takes fewer solves."""
    from test_regression_amd import _abi, engine, warmstart, workloads
    res = {}
    for cfg, wts in (("c2", "c2"), ("c3", "c2"), ("c3", "c3")):
        w = workloads.config(cfg)
        w.seed += 2000
        s, t, _ = w.generate(100_000)
        eng = engine.Engine(list(w.chtypes()))
        eng.mlp_set_params(np.load(warmstart.weights_for(wts)))
        tt = torch.from_numpy(t).cuda() if t is not None else None
        xref = warmstart.label(eng, torch.from_numpy(s).cuda(), tt)
        s32 = torch.from_numpy(s.astype(np.float32)).cuda()
        ev = warmstart.evaluate(eng, s32, xref)
        r = eng.fit(s32, tt, _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=1))
        torch.cuda.synchronize()
        res[(cfg, wts)] = (ev["exact_active_set"], r.n_inner().float().mean().item())
        print(f"[gpu-mlp] {cfg} with the {wts} network: exact {ev['exact_active_set']:.3f}, "
              f"solves {res[(cfg, wts)][1]:.2f}", flush=True)
        ref = oracle_fit(w.chtypes(), s.astype(np.float32).astype(np.float64), t, oracle_opts(),
                         nthreads=8)
        parity.gate(r.x.double().cpu().numpy(), r.chi2.double().cpu().numpy(), ref, 1e-4,
                    iters=r.iters.cpu().numpy(), compare_outer=False,
                    label=f"{cfg} warm with the {wts} network")
    assert res[("c2", "c2")][0] > 0.7 and res[("c3", "c3")][0] > 0.7
    assert res[("c3", "c3")][0] > res[("c3", "c2")][0] + 0.15
    assert res[("c3", "c3")][1] < res[("c3", "c2")][1]
