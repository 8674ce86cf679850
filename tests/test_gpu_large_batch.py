"""Batches whose arrays hold more than 2^31 elements: 249,995,750 channels,
so sample row 8 crosses element 2^31 and row 9 lies above it (and the
live-covariance hand-over, 11 values per channel, crosses it from row 2 on).

The batch is one generated block of M channels repeated K times on the
device; every channel's fit depends only on its own samples, so each copy's
outputs must equal, bit for bit, those of the block fitted on its own.  That
checks every index formed from the channel count (sample loads, result
stores, the hand-over, the type partition's lists) at the largest batches
the boundary accepts in this memory (about 45 GB at most here).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

M, K = 999_983, 250
N = M * K


def _same_as_block(big, small, label):
    for name in ("x", "chi2", "flags", "iters"):
        a, b = getattr(big, name), getattr(small, name)
        if a is None:
            continue
        if a.dim() == 2:
            ok = torch.equal(a.view(a.shape[0], K, M), b.view(b.shape[0], 1, M).expand(-1, K, -1))
        else:
            ok = torch.equal(a.view(K, M), b.view(1, M).expand(K, -1))
        assert ok, f"{label}: {name} of a repeated block differs from the block's own fit"


@pytest.mark.parametrize("path", ["warm", "warm_type_of", "cold_live", "plain_fp64"])
def test_large_batch_indexing(pkg, path):
    from test_regression_amd import _abi, engine, warmstart, workloads
    assert N * 10 > 2 ** 31 and N * 8 < 2 ** 31 < N * 9
    live = path == "cold_live"
    w = workloads.config("c2", "live" if live else "zero")
    s, _, _ = w.generate(M, nthreads=8)
    eng = engine.Engine(list(w.chtypes()))
    try:
        if path.startswith("warm"):
            warmstart.ensure_trained(eng, path=warmstart.weights_for("c2"))
            o = _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=1)
        elif live:
            o = _abi.default_opts(precision=_abi.MF_PREC_FP32)
        else:
            o = _abi.default_opts(precision=_abi.MF_PREC_FP64)
        dt = torch.float64 if path == "plain_fp64" else torch.float32
        block = torch.from_numpy(s).to("cuda", dt)
        print(f"[gpu-large] {path}: block of {M} channels, batch of {N}", flush=True)
        small = eng.fit(block, None, o)
        big_in = block.repeat(1, K)
        t_small = t_big = None
        if path == "warm_type_of":
            t_small = torch.zeros(M, dtype=torch.uint8, device="cuda")
            t_big = torch.zeros(N, dtype=torch.uint8, device="cuda")
            small = eng.fit(block, t_small, o)
        big = eng.fit(big_in, t_big, o)
        torch.cuda.synchronize()
        print(f"[gpu-large] {path}: fitted, comparing", flush=True)
        _same_as_block(big, small, path)
        del big, big_in
    finally:
        eng.close()
        torch.cuda.empty_cache()
