"""mf_autotune on a mixed, heavy-tailed batch (c3: 70% ECAL-like, 30%
HCAL-like, pileup U[0,80]): it returns a valid launch shape and its time,
keeps it in the context, and the fit under the tuned shape (block 0,
persistent = auto) is bit for bit the fit under the default shape -- the
launch shape changes where channels run, not what they compute."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_autotune_keeps_results(pkg, gpu, prec):
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config("c3")
    s, t, _ = w.generate(200_000)
    eng = engine.Engine(list(w.chtypes()))
    p = _abi.MF_PREC_FP32 if prec == "fp32" else _abi.MF_PREC_FP64
    x = torch.from_numpy(s).cuda()
    if prec == "fp32":
        x = x.float()
    tt = torch.from_numpy(t).cuda()
    print(f"[gpu-autotune] prec={prec}", flush=True)
    ref = eng.fit(x, tt, _abi.default_opts(precision=p))
    torch.cuda.synchronize()
    tune, ms = eng.autotune(x, tt, _abi.default_opts(precision=p))
    print(f"[gpu-autotune] chose {tune} at {ms:.3f} ms", flush=True)
    assert ms > 0 and tune["fit_block"] in (64, 128, 256)
    assert eng.get_tune() == tune
    got = eng.fit(x, tt, _abi.default_opts(precision=p, block=0,
                                           persistent=_abi.MF_PERSISTENT_AUTO))
    torch.cuda.synchronize()
    assert torch.equal(got.x, ref.x) and torch.equal(got.chi2, ref.chi2)
    assert torch.equal(got.flags, ref.flags) and torch.equal(got.iters, ref.iters)
