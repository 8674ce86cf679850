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
                                         ("plain", "ranges"), ("warm", "ranges")])
def test_graph_replays_equal_eager(pkg, gpu, mode, layout):
    from test_regression_amd import engine, workloads, warmstart, _abi
    w = workloads.config("c5")
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
    print(f"[gpu-graph] mode={mode} layout={layout} nch={x.shape[1]}", flush=True)
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
