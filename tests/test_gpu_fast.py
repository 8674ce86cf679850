"""GPU parity of the fast path (MF_PREC_FP32) against the FP64 oracle.

Bar (the FP32 fast path's accuracy target), channel by channel (tests/parity.py):
every amplitude within 1e-4 of the oracle's relative to itself (floor 1 ADC),
chi2 within 1e-4 relative (floor 1), the same active set and Minimize
iteration count; a channel outside that must be explained by the oracle's
record of its conditioning or of a decision taken within rounding, and at
most 0.1% of channels may be outside.  (The fast path reaches the optimum by
its own route, so its FNNLS counts are not compared.)
"""
from pathlib import Path

import numpy as np
import pytest
import torch

import parity
from oracle_binding import oracle_fit, oracle_opts

pytestmark = pytest.mark.gpu

TOL_X = 1e-4
TOL_CHI2 = 1e-4
# the samples behind each oracle record, for the accounting's end-point test
ref_sets = {}


def _fit_fast(cfg, pulse_cov, n, sample_dtype=np.float32, seed=0, **kw):
    from test_regression_amd import engine, workloads, warmstart, _abi
    w = workloads.config(cfg, pulse_cov)
    w.seed += seed
    ct = w.chtypes()
    s, t, _ = w.generate(n)
    s_in = s.astype(sample_dtype)
    eng = engine.Engine(list(ct))
    if kw.get("warm_start"):
        warmstart.ensure_trained(eng, path=warmstart.weights_for(cfg))
    print(f"[gpu-fast] cfg={cfg} pulse_cov={pulse_cov} n={n} dtype={sample_dtype.__name__} "
          f"seed+{seed} {kw}", flush=True)
    r = eng.fit(torch.from_numpy(s_in).cuda(),
                torch.from_numpy(t).cuda() if t is not None else None,
                _abi.default_opts(precision=_abi.MF_PREC_FP32, **kw))
    torch.cuda.synchronize()
    # the oracle sees exactly the samples the fast path fitted (FP32-rounded)
    s_ref = s_in.astype(np.float32).astype(np.float64)
    ref = oracle_fit(ct, s_ref, t, oracle_opts(), nthreads=8)
    ref_sets[id(ref)] = (ct, s_ref, t)
    return r, ref


def _check(r, ref, label=""):
    return parity.gate(r.x.double().cpu().numpy(), r.chi2.double().cpu().numpy(), ref, TOL_X,
                       tol_chi2=TOL_CHI2, iters=r.iters.cpu().numpy(), compare_outer=False,
                       label=label, sets=ref_sets.get(id(ref)))


@pytest.mark.parametrize("cfg,pulse_cov,n", [
    ("c1", "zero", 10_000),
    ("c2", "zero", 50_000),
    ("c2", "live", 20_000),
    ("c3", "zero", 20_000),
    ("c3", "live", 20_000),
    ("c5", "zero", 20_000),
    ("c5", "live", 10_000),
])
@pytest.mark.parametrize("warm", [0, 1])
def test_fast_fit_matches_oracle(pkg, gpu, cfg, pulse_cov, n, warm):
    """Cold and warm-started (the workload's committed network) fast path
    against the oracle on the same FP32-rounded samples."""
    r, ref = _fit_fast(cfg, pulse_cov, n, warm_start=warm)
    _check(r, ref, f"fast {cfg} {pulse_cov} warm={warm}")
    fl = r.flags.cpu().numpy()
    from test_regression_amd import _abi
    assert not np.any(fl & _abi.FLAGS["nonfinite"])


@pytest.mark.parametrize("seed", [11, 12, 13])
@pytest.mark.parametrize("cfg,pulse_cov,n", [("c2", "zero", 200_000), ("c2", "live", 100_000),
                                             ("c3", "zero", 200_000)])
def test_fast_fit_seed_sweep(pkg, gpu, cfg, pulse_cov, n, seed):
    """The warm-started path on batches generated under other seeds than the
    benchmark's and the training's: every channel accounted for against the
    oracle, as above."""
    r, ref = _fit_fast(cfg, pulse_cov, n, seed=seed, warm_start=1)
    _check(r, ref, f"fast {cfg} {pulse_cov} warm seed+{seed}")


@pytest.mark.parametrize("persistent", [0, 1])
def test_warm_fit_ill_conditioned_blocks(pkg, gpu, persistent):
    """The three c2 channels of tests/golden/degenerate_channels.npz, whose
    warm starts (the committed network's predictions) put BX -5 to BX -1
    together in the passive set (condition numbers up to 1.3e14), through
    the one-launch kernel and the persistent one: every chi2 is
    the oracle's optimum (to the FP32 output's rounding), and every channel
    is accounted for (the amplitudes of such a block are fixed only to its
    conditioning; the oracle continued from the same set agrees to that)."""
    from test_regression_amd import _abi, engine, warmstart, workloads
    z = np.load(Path(__file__).resolve().parent / "golden" / "degenerate_channels.npz")
    w = workloads.config("c2")
    ct = w.chtypes()
    s32 = z["samples"].astype(np.float32)
    eng = engine.Engine(list(ct))
    warmstart.ensure_trained(eng, path=warmstart.weights_for("c2"))
    print(f"[gpu-fast] ill-conditioned warm starts, persistent={persistent}", flush=True)
    r = eng.fit(torch.from_numpy(s32).cuda(), None,
                _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=1,
                                  persistent=persistent))
    torch.cuda.synchronize()
    s = s32.astype(np.float64)
    ref = oracle_fit(ct, s, None, oracle_opts())
    ref_sets[id(ref)] = (ct, s, None)
    chi2 = r.chi2.double().cpu().numpy()
    assert np.all(np.abs(chi2 - ref["chi2"]) <= 1e-6 * np.maximum(ref["chi2"], 1.0)), chi2
    rep = parity.account(r.x.double().cpu().numpy(), chi2, ref, TOL_X, tol_chi2=TOL_CHI2,
                         iters=r.iters.cpu().numpy(), compare_outer=False, sets=(ct, s, None))
    assert rep["unexplained"] == 0, rep


def test_fast_fit_fp64_samples(pkg, gpu):
    """FP64 sample storage is rounded to FP32 on load."""
    r, ref = _fit_fast("c2", "zero", 5000, sample_dtype=np.float64)
    _check(r, ref)


@pytest.mark.parametrize("warm", [0, 1])
def test_fast_fit_edge_cases(pkg, gpu, warm):
    """Zero, NaN, negative and infinite samples and a 1e5 ADC pulse, from a
    cold start and from the MLP's prediction (which sees the same samples)."""
    from test_regression_amd import engine, workloads, warmstart, _abi
    w = workloads.config("c1")
    ct = w.chtypes()
    eng = engine.Engine(list(ct))
    if warm:
        warmstart.ensure_trained(eng)
    s = np.zeros((10, 5), dtype=np.float32)
    s[:, 1] = np.nan
    s[:, 2] = 1e5 * np.exp(-0.5 * (np.arange(10) - 5) ** 2)
    s[:, 3] = -5.0
    s[4, 4] = np.inf
    print(f"[gpu-fast] edge cases warm={warm}", flush=True)
    r = eng.fit(torch.from_numpy(s).cuda(), None,
                _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=warm))
    torch.cuda.synchronize()
    x = r.x.cpu().numpy()
    fl = r.flags.cpu().numpy()
    assert np.all(x[:, 0] == 0) and r.chi2[0].item() == 0.0
    assert fl[1] & _abi.FLAGS["nonfinite"] and fl[4] & _abi.FLAGS["nonfinite"]
    ref = oracle_fit(ct, s[:, [0, 3]].astype(np.float64), None, oracle_opts())
    xr = ref["x"].T
    scale = np.maximum(np.abs(xr).max(axis=0), 1.0)
    assert np.all(np.abs(x[:, [0, 3]] - xr) / scale <= TOL_X)
    assert np.all((x[:, [0, 3]] > 0) == (xr > 0))
    # A 1e5 ADC pulse: FP64 rounding of w (~1e-16 |Atb|) exceeds the
    # reference's eps 1e-11, so the reference's second FNNLS call decides its
    # warm-start jump on rounding noise and may end away from the optimum.
    # The fast path returns the NNLS optimum: its chi2 is never worse.
    big = oracle_fit(ct, s[:, [2]].astype(np.float64), None, oracle_opts())
    assert r.chi2[2].item() <= big["chi2"][0] * (1 + 1e-6)
    assert np.all(x[:, 2] >= 0) and np.isfinite(x[:, 2]).all()


@pytest.mark.parametrize("cfg,n", [("c1", 10_000), ("c2", 50_000), ("c3", 20_000),
                                   ("c5", 20_000)])
def test_persistent_fast_fit_matches_oracle(pkg, gpu, cfg, n):
    """The persistent kernels (lanes refill from a device work queue) against
    the oracle, same bar."""
    r, ref = _fit_fast(cfg, "zero", n, persistent=1)
    _check(r, ref)


@pytest.mark.parametrize("warm", [0, 1])
def test_persistent_equals_per_channel_launch(pkg, gpu, warm):
    """Each lane of the persistent kernels runs the same step sequence on its
    channel as the per-channel kernels: the outputs are identical, bit for
    bit, cold and warm-started."""
    from test_regression_amd import engine, workloads, _abi, warmstart
    w = workloads.config("c2")
    ct = w.chtypes()
    s, _, _ = w.generate(100_000)
    eng = engine.Engine(list(ct))
    if warm:
        warmstart.ensure_trained(eng)
    x = torch.from_numpy(s.astype(np.float32)).cuda()
    print(f"[gpu-fast] persistent vs per-channel warm={warm}", flush=True)
    a = eng.fit(x, None, _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=warm))
    b = eng.fit(x, None, _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=warm,
                                           persistent=1))
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.chi2, b.chi2)
    assert torch.equal(a.flags, b.flags) and torch.equal(a.iters, b.iters)


@pytest.mark.parametrize("cfg,n,layout", [("c2", 100_003, "type_of"), ("c3", 50_000, "type_of"),
                                          ("c5", 75_848, "ranges")])
def test_warm_single_launch_equals_persistent(pkg, gpu, cfg, n, layout):
    """The warm single launch (the MLP's prediction and the FP64 phase in one
    kernel; blocks of c5's type ranges hold channels of both types) against
    the persistent kernels (the prediction through memory, lanes refilled
    from a queue): each channel runs the same steps, so the outputs are
    identical, bit for bit, including a last block that is only partly
    filled."""
    from test_regression_amd import engine, workloads, _abi, warmstart
    w = workloads.config(cfg)
    eng = engine.Engine(list(w.chtypes()))
    warmstart.ensure_trained(eng)
    kw = {}
    if layout == "ranges":
        s, t = w.generate(events=1)[:2]
        kw["type_ranges"] = [0, w.event[0], sum(w.event)]
        t = None
    else:
        s, t, _ = w.generate(n)
        t = torch.from_numpy(t).cuda() if t is not None else None
    x = torch.from_numpy(s.astype(np.float32)).cuda()
    print(f"[gpu-fast] warm single launch vs persistent cfg={cfg} n={x.shape[1]} {layout}",
          flush=True)
    o = dict(precision=_abi.MF_PREC_FP32, warm_start=1)
    a = eng.fit(x, t, _abi.default_opts(**o), **kw)
    b = eng.fit(x, t, _abi.default_opts(persistent=1, **o), **kw)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.chi2, b.chi2)
    assert torch.equal(a.flags, b.flags) and torch.equal(a.iters, b.iters)
    # the warm starts were used: most channels took one solve, some more
    inner = a.iters.cpu().numpy() >> 16
    assert (inner == 1).mean() > 0.2 and (inner > 1).mean() > 0.05


def test_persistent_edge_cases(pkg, gpu):
    """Non-finite and empty channels, and batches smaller than one wavefront
    or not a multiple of it, through the queue."""
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config("c1")
    ct = w.chtypes()
    eng = engine.Engine(list(ct))
    for n in (1, 5, 63, 65, 1000):
        s = np.abs(np.random.default_rng(n).normal(0, 1, (10, n))).astype(np.float32)
        s[:, 0] += 50 * np.exp(-0.5 * (np.arange(10) - 5) ** 2)
        if n >= 5:
            s[:, 1] = np.nan
            s[:, 3] = 0.0
        r = eng.fit(torch.from_numpy(s).cuda(), None,
                    _abi.default_opts(precision=_abi.MF_PREC_FP32, persistent=1))
        q = eng.fit(torch.from_numpy(s).cuda(), None, _abi.default_opts(precision=_abi.MF_PREC_FP32))
        torch.cuda.synchronize()
        assert torch.equal(r.x, q.x) and torch.equal(r.flags, q.flags)
        assert torch.equal(r.chi2, q.chi2) and torch.equal(r.iters, q.iters)
        if n >= 5:
            assert r.flags[1].item() & _abi.FLAGS["nonfinite"]


@pytest.mark.parametrize("cfg,warm,persistent,max_iter", [
    ("c2", 0, 0, 1000), ("c2", 1, 0, 1000), ("c3", 1, 0, 1000), ("c3", 1, 1, 1000),
    ("c5", 1, 0, 1000), ("c2", 1, 0, 2), ("c3", 0, 0, 1), ("c3", 1, 0, 3), ("c5", 1, 0, 2)])
def test_live_split_equals_single_kernel(pkg, gpu, cfg, warm, persistent, max_iter):
    """With a live pulse covariance mf_fit_batch runs the first iteration in
    a light kernel and hands the FP64 state over to the later iterations'
    kernels: k_fast_p3 runs the iterations that stop at their start or add a
    single index, k_fast_p4 finishes the channels that need the normal
    equations; mf_fit_batch_noalloc keeps all of it in k_fast_p2.  The same
    arithmetic: outputs identical bit for bit (warm starts through the
    in-kernel MLP, or through the prediction kernel when persistent).  A
    small max_iter sends channels to k_fast_p4 (the single-index step needs
    two solves' allowance): from a list, type-partitioned lists and type
    ranges."""
    from test_regression_amd import engine, workloads, _abi, warmstart
    w = workloads.config(cfg, "live")
    eng = engine.Engine(list(w.chtypes()))
    if warm:
        warmstart.ensure_trained(eng, path=warmstart.weights_for(cfg))
    kw = {}
    if w.event:
        s, _, _ = w.generate(events=1)
        kw["type_ranges"] = [0, w.event[0], sum(w.event)]
        t = None
    else:
        s, t, _ = w.generate(60_000)
        t = torch.from_numpy(t).cuda() if t is not None else None
    x = torch.from_numpy(s.astype(np.float32)).cuda()
    print(f"[gpu-fast] live split cfg={cfg} warm={warm} persistent={persistent} "
          f"max_iter={max_iter}", flush=True)
    o = _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=warm, persistent=persistent,
                          max_iter=max_iter)
    a = eng.fit(x, t, o, **kw)
    b = eng.fit_noalloc(x, t, o, **kw)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.chi2, b.chi2)
    assert torch.equal(a.flags, b.flags) and torch.equal(a.iters, b.iters)
    assert ((a.iters.cpu().numpy() & 255) > 1).mean() > 0.5   # later iterations ran
