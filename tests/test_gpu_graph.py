"""Graph capture of a fit call (mf_graph_capture / mf_graph_launch) on a
mixed-type event (two channel types, partitioned on the device): every replay
equals the eager call bit for bit, with eager work on the same context and
host synchronizations between replays, and a replay after new data is copied
into the captured input equals an eager fit of that data."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _event(w, k, seed_off):
    w.seed += seed_off
    s, t, _ = w.generate(events=k)
    return s, t


@pytest.mark.parametrize("mode,layout", [("plain", "type_of"), ("fast", "type_of"),
                                         ("warm", "type_of"), ("persistent", "type_of"),
                                         ("plain", "ranges"), ("warm", "ranges"),
                                         ("warm_live", "type_of"), ("fast_live", "ranges"),
                                         ("plain_live", "ranges")])
def test_graph_replays_equal_eager(pkg, gpu, mode, layout):
    """_live: a live pulse covariance (the fast path's hand-over scratch and
    pending lists, cleared inside the captured kernels)."""
    from test_regression_amd import engine, workloads, warmstart, _abi
    live = mode.endswith("_live")
    mode = mode.removesuffix("_live")
    w = workloads.config("c5", "live" if live else "zero")
    eng = engine.Engine(list(w.chtypes()))
    if mode == "warm":
        warmstart.ensure_trained(eng)
    prec = _abi.MF_PREC_FP64 if mode == "plain" else _abi.MF_PREC_FP32
    o = _abi.default_opts(precision=prec, warm_start=1 if mode == "warm" else 0,
                          persistent=1 if mode == "persistent" else 0)
    dt = torch.float64 if mode == "plain" else torch.float32
    s1, t1 = _event(w, 1, 0)
    s2, t2 = _event(w, 1, 17)
    x = torch.from_numpy(s1).to("cuda", dt)
    tt = torch.from_numpy(t1).cuda()
    kw = {}
    if layout == "ranges":
        tt = None
        kw["type_ranges"] = [0, w.event[0], sum(w.event)]
    print(f"[gpu-graph] mode={mode} live={live} layout={layout} nch={x.shape[1]}", flush=True)
    ref = eng.fit(x, tt, o, **kw)
    torch.cuda.synchronize()
    g = eng.capture(x, tt, o, **kw)
    other = torch.from_numpy(s2).to("cuda", dt)
    for rep in range(3):
        # eager work on the same context between replays, then the replay
        eng.fit(other, torch.from_numpy(t2).cuda(), o)
        torch.cuda.synchronize()
        g.result.x.zero_()
        g.result.flags.zero_()
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(g.result.x, ref.x), rep
        assert torch.equal(g.result.chi2, ref.chi2), rep
        assert torch.equal(g.result.flags, ref.flags), rep
        assert torch.equal(g.result.iters, ref.iters), rep
    # new data in the captured input buffer
    x.copy_(other)
    g.launch()
    torch.cuda.synchronize()
    ref2 = eng.fit(other, torch.from_numpy(t2).cuda(), o)
    if layout == "ranges":   # the same fit through the other layout: identical
        ref3 = eng.fit(other, None, o, type_ranges=kw["type_ranges"])
        torch.cuda.synchronize()
        assert torch.equal(ref3.x, ref2.x) and torch.equal(ref3.flags, ref2.flags)
    torch.cuda.synchronize()
    assert torch.equal(g.result.x, ref2.x) and torch.equal(g.result.flags, ref2.flags)
    # both types were fitted (no bad-type flags; in-time amplitudes found)
    fl = g.result.flags.cpu().numpy()
    assert not np.any(fl & _abi.FLAGS["bad_type"])
    assert (g.result.x[5].double().cpu().numpy() > 0).mean() > 0.99
    g.close()


def test_graph_guards_type_tables(pkg, gpu):
    """A captured graph's kernels read the context's type tables: mf_set_types
    is refused while the graph exists (the graph still replays the fit it
    captured), allowed again once it is destroyed; a graph whose context is
    destroyed refuses to launch."""
    from test_regression_amd import engine, workloads, _abi
    w = workloads.config("c5")
    eng = engine.Engine(list(w.chtypes()))
    s, t = _event(w, 1, 3)
    x = torch.from_numpy(s).cuda()
    tt = torch.from_numpy(t).cuda()
    o = _abi.default_opts()
    print("[gpu-graph] set_types under a live graph", flush=True)
    ref = eng.fit(x, tt, o)
    torch.cuda.synchronize()
    g = eng.capture(x, tt, o)
    with pytest.raises(_abi.MFError, match="graph"):
        eng.set_types(list(w.chtypes())[::-1])
    g.launch()
    torch.cuda.synchronize()
    assert torch.equal(g.result.x, ref.x) and torch.equal(g.result.flags, ref.flags)
    g.close()
    eng.set_types(list(w.chtypes()))
    # a graph outliving its context: the C ABI detaches it at mf_context_destroy
    g2 = eng.capture(x, tt, o)
    h = g2._h
    lib = eng._lib
    eng._graphs.discard(g2)          # keep it open across the context's destruction
    lib.mf_context_destroy(eng._ctx)
    eng._ctx = type(eng._ctx)()
    import ctypes as C
    assert lib.mf_graph_launch(h, C.c_void_p(None)) == _abi.MF_ERR_INVALID
    assert b"destroyed" in lib.mf_last_error()
    assert lib.mf_graph_destroy(h) == _abi.MF_OK
    g2._h = None
