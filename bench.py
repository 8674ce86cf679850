#!/usr/bin/env python3
"""bench.py -- fitted channels/s of the multifit engine on one MI355X.

A step is one mf_fit_batch over one batch of synthetic channels already
resident in HBM (structure of arrays).  Default workload (c2): 1,000,000
ECAL-like channels, 10 samples x 10 BX, in-time amplitude U[1,500], per-BX
Poisson(40/10) pileup of Exponential(2) ADC, N(0, 1 ns) jitter, zero pulse
covariance (as the reference's drivers pass it).  Paths timed on the same
GPU and the same batch:

  plain   plain cold-start FP64 fit, the reference's semantics
          (multifit_gpu: inplace_fnnls with max_iter 1000)
  fast    fast path, cold start: FP32 active-set exploration on the type's
          precomputed normal equations, FP64 polish and KKT verification
  warm    fast path started from the MLP's predicted passive set and
          amplitudes (inference included in the step) -- the reported value

Reported next to the value: the plain path's rate and the speedups, the
error of every fitted amplitude and chi2 of the warm path against the plain
FP64 fit over the whole batch (and against the FP64 oracle on a sample), the
roofline of the dominant kernel (the fit kernel of the warm path: algorithmic
flops per launch -- the plain method's operation count, counted by the oracle
on a sample of the batch -- over the kernel's launch time from HIP events on
its stream), and the oracle's throughput on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tests"))

PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}

# HBM bytes per launch of a warm step's kernel (k_fast_warm)
# on the default workload, from two rocprofv3 --pmc runs of
# `bench.py --traffic-probe`: FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, the
# factors that make the counters read a 256 MiB device copy as 256 MiB each
# way.  Other workloads report null.
MEASURED_TRAFFIC = {("c2", "zero", 1_000_000): {
    "bytes": 94.5e6, "how": "rocprofv3 --pmc FETCH_SIZE (x2.00) and WRITE_SIZE (x1.00), "
                             "bench.py --traffic-probe, copy-calibrated"}}


# What "achieved" counts: the plain method's operations, counted by the
# oracle on a sample (the fast path reaches the same optimum with fewer).
ACHIEVED_COUNTS = ("algorithmic flops of the plain method per channel (about 10.6 solves and "
                   "two Minimize iterations at zero pulse covariance, counted by the oracle) x "
                   "channels / kernel time: a rate of useful fit work. The warm kernel reaches "
                   "the same optimum in about 1.4 solves and needs no second iteration, so it "
                   "issues far fewer flops (issued-instruction counts are under profiles/)")


# Forward pass of the warm-start MLP per channel (10 -> 64 -> 64 -> 20): MFMA
# (bfloat16) work inside the warm fit kernel, not counted in the VALU
# roofline's achieved figure.
MLP_FLOPS_PER_CHANNEL = 2 * (10 * 64 + 64 * 64 + 64 * 20)


def fit_kernel_name(pulse_cov: str, warm: bool = True) -> str:
    """The fast path's fit kernels for a pulse-covariance mode."""
    if pulse_cov == "live":
        return (("k_fast_warm (MLP and first iteration)" if warm else
                 "k_fast_fused (first iteration)") +
                " + k_fast_p3 (later iterations) + k_fast_p4 (channels needing the normal "
                "equations): the step's kernels")
    if warm:
        return "k_fast_warm (MLP prediction and FP64 active-set phase in one launch)"
    return "k_fast_fused (both phases in one launch)"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# Cycles of the spin kernel (torch.cuda._sleep) that holds the launch stream
# before a timed region: 0.8 ms or more at any shader clock, longer than the
# host takes to queue the first step.
GATE_CYCLES = 2_000_000


def gate_stream():
    """Holds the current stream for GATE_CYCLES (torch's spin kernel, a
    private helper: where it is missing the region starts ungated and also
    holds the host's latency for its first call)."""
    import torch
    sleep = getattr(torch.cuda, "_sleep", None)
    if sleep is not None:
        sleep(GATE_CYCLES)


def time_steps(eng, fn, steps, warmup, stream, kernel_steps=5):
    """K steps back to back between two events on the launch stream, with a
    synchronize on both sides (ms per step).  A spin kernel holds the stream
    while the host queues the first step, so the events time the device's K
    steps rather than the host's latency for its first call (about 0.1 ms).
    Also: the host clock over K more steps, from the first call to the end of
    the last (ms per step), and, outside both, a few steps with the library's
    kernel events (fit and MLP kernel time)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    gate_stream()
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    eng.set_timing(True)
    kern = []
    for _ in range(kernel_steps):
        fn()
        kern.append(eng.timings())   # waits for this step's kernel events
    eng.set_timing(False)
    return a.elapsed_time(b) / steps, wall, kern


def host_cpu():
    """The host's CPU model (/proc/cpuinfo) and its logical CPU count."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu": model, "nproc": os.cpu_count()}


def cpu_baseline(ct, s, t, seconds, threads):
    """Oracle throughput on host cores over a bounded slice of the batch:
    5 runs of about seconds/6 each with `threads` threads (their median is
    the baseline), then one thread on a smaller slice for about seconds/6."""
    from oracle_binding import oracle_fit, oracle_opts
    n = min(s.shape[1], 200_000)

    def rate(nsl, th, budget):
        a = np.ascontiguousarray(s[:, :nsl])
        b = t[:nsl] if t is not None else None
        done, t0 = 0, time.perf_counter()
        while True:
            oracle_fit(ct, a, b, oracle_opts(record_margins=0), nthreads=th)
            done += nsl
            dt = time.perf_counter() - t0
            if dt >= budget:
                return done / dt, done, dt

    runs = [rate(n, threads, seconds / 6) for _ in range(5)]
    n1 = min(n, 20_000)
    single = rate(n1, 1, seconds / 6)
    return {"value": float(np.median([r[0] for r in runs])), "unit": "channels/s",
            "cores": threads, "kind": "port",
            "runs": [r[0] for r in runs], "single_thread": single[0],
            "fits": sum(r[1] for r in runs) + single[1],
            "seconds": sum(r[2] for r in runs) + single[2], "slice": n, "slice_1t": n1,
            "host": host_cpu()}


def oracle_sample(ct, s, t, n):
    from oracle_binding import oracle_fit, oracle_opts
    return oracle_fit(ct, np.ascontiguousarray(s[:, :n]), t[:n] if t is not None else None,
                      oracle_opts(count_flops=1), nthreads=min(16, os.cpu_count() or 1))


def errors(x, chi2, xr, chi2r, floor=1.0):
    """One path against another over the whole batch: the per-amplitude
    relative error |x - xr| / max(|xr|, floor) (floor 1 ADC), the channel
    error relative to the channel's largest amplitude (same floor), chi2's
    relative error (floor 1), and the counts above 1e-4 and of channels whose
    active sets differ."""
    d = np.abs(x - xr)
    a = (d / np.maximum(np.abs(xr), floor)).max(axis=0)
    e = d.max(axis=0) / np.maximum(np.abs(xr).max(axis=0), floor)
    c = np.abs(chi2 - chi2r) / np.maximum(chi2r, 1.0)
    aset = ((x > 0) == (xr > 0)).all(axis=0)
    return {"amp_rel_err_max": float(a.max()), "amp_rel_err_p999": float(np.quantile(a, 0.999)),
            "amp_above_1e-4": int((a > 1e-4).sum()),
            "channel_rel_err_max": float(e.max()), "channel_above_1e-4": int((e > 1e-4).sum()),
            "chi2_rel_err_max": float(c.max()), "chi2_above_1e-4": int((c > 1e-4).sum()),
            "different_active_sets": int((~aset).sum()),
            "identical_active_sets": float(aset.mean()), "channels": int(x.shape[1]),
            "amplitude_floor_adc": floor}


def event_stream(args, eng, w, log, with_cpu=True):
    """Event stream: `events` events back to back, each one fit call on one
    event's channels (both types, given as channel ranges), replayed from a
    captured graph; per-event latency from HIP events around each replay.
    A pool of `event_pool` distinct generated events stays resident in HBM
    and is cycled (event i uses pool slot i mod pool)."""
    import torch
    from test_regression_amd import _abi, engine, warmstart
    P, E = args.event_pool, args.events
    nev = sum(w.event)
    t0 = time.perf_counter()
    s, _, _ = w.generate(events=P, nthreads=min(16, os.cpu_count() or 1))
    log(f"generated {P} events of {nev} channels in {time.perf_counter() - t0:.1f}s")
    pool = torch.from_numpy(np.ascontiguousarray(
        s.reshape(10, P, nev).transpose(1, 0, 2)).astype(np.float32)).cuda()
    ranges = np.concatenate([[0], np.cumsum(w.event)]).tolist()   # barrel, then endcap
    wpath = Path(args.mlp_weights) if args.mlp_weights else warmstart.weights_for(args.config)
    mlp_info = warmstart.ensure_trained(eng, warmstart.training_workload(args.config),
                                        path=wpath if wpath.exists() else None, log=log)
    if args.save_mlp:
        np.save(args.save_mlp, eng.mlp_get_params())
    o = _abi.default_opts(precision=_abi.MF_PREC_FP32, warm_start=1)
    outs = [eng.alloc_result(nev, _abi.MF_PREC_FP32) for _ in range(P)]
    graphs = [eng.capture(pool[e], None, o, outs[e], type_ranges=ranges) for e in range(P)]
    stream = torch.cuda.current_stream()

    def run(launch):
        for i in range(min(P, E)):
            launch(i)
        torch.cuda.synchronize()
        st = [torch.cuda.Event(enable_timing=True) for _ in range(E)]
        en = [torch.cuda.Event(enable_timing=True) for _ in range(E)]
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for i in range(E):
            st[i].record(stream)
            launch(i)
            en[i].record(stream)
        b.record(stream)
        torch.cuda.synchronize()
        lat = np.array([x.elapsed_time(y) for x, y in zip(st, en)])
        return a.elapsed_time(b), lat

    def run_inflight(k, gk):
        """Event i replayed on stream i mod k (default priority), so that k
        events are in flight.  The graphs gk of stream j were captured on a
        context of their own (a context's scratch serves one call at a
        time), and k divides the pool size, so a graph and its output buffers
        are only ever replayed on one stream."""
        streams = [torch.cuda.Stream() for _ in range(k)]
        for i in range(min(P, E)):
            gk[i % P].launch(streams[i % k])
        torch.cuda.synchronize()
        st = [torch.cuda.Event(enable_timing=True) for _ in range(E)]
        en = [torch.cuda.Event(enable_timing=True) for _ in range(E)]
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for sm in streams:
            sm.wait_event(a)
        for i in range(E):
            sm = streams[i % k]
            st[i].record(sm)
            gk[i % P].launch(sm)
            en[i].record(sm)
        for sm in streams:
            d = torch.cuda.Event()
            d.record(sm)
            stream.wait_event(d)
        b.record(stream)
        torch.cuda.synchronize()
        lat = np.array([x.elapsed_time(y) for x, y in zip(st, en)])
        return a.elapsed_time(b), lat

    total, lat = run(lambda i: graphs[i % P].launch())
    total_e, lat_e = run(lambda i: eng.fit(pool[i % P], None, o, outs[i % P], type_ranges=ranges))
    log(f"graph: {total:.1f} ms for {E} events, p50 {np.median(lat):.4f} ms, "
        f"p99 {np.quantile(lat, 0.99):.4f} ms; eager: {total_e:.1f} ms")
    K_INFLIGHT = 4 if P % 4 == 0 else 1
    # one context per stream, same types, network and launch shapes
    engs = [eng] + [engine.Engine(list(w.chtypes())) for _ in range(K_INFLIGHT - 1)]
    for e in engs[1:]:
        e.mlp_set_params(eng.mlp_get_params())
        e.set_tune(**eng.get_tune())
    outs_k = [eng.alloc_result(nev, _abi.MF_PREC_FP32) for _ in range(P)]
    graphs_k = [engs[e % K_INFLIGHT].capture(pool[e], None, o, outs_k[e], type_ranges=ranges)
                for e in range(P)]
    total_k, lat_k = run_inflight(K_INFLIGHT, graphs_k)
    # every event's in-flight outputs against its one-at-a-time outputs
    inflight_equal = all(torch.equal(outs_k[e].x, outs[e].x) and
                         torch.equal(outs_k[e].chi2, outs[e].chi2) and
                         torch.equal(outs_k[e].flags, outs[e].flags) for e in range(P))
    log(f"graph, {K_INFLIGHT} events in flight: {total_k:.1f} ms for {E} events, "
        f"p50 {np.median(lat_k):.4f} ms, p99 {np.quantile(lat_k, 0.99):.4f} ms; outputs "
        f"{'identical' if inflight_equal else 'DIFFERENT'} to one at a time")
    if not inflight_equal:
        raise SystemExit("in-flight outputs differ from one-at-a-time outputs")
    for g in graphs_k:
        g.close()
    # accuracy of one event against the plain FP64 fit
    ref = eng.fit(pool[0].double(), None, _abi.default_opts(precision=_abi.MF_PREC_FP64),
                  type_ranges=ranges)
    graphs[0].launch()
    torch.cuda.synchronize()
    acc = errors(outs[0].x.double().cpu().numpy(), outs[0].chi2.double().cpu().numpy(),
                 ref.x.cpu().numpy(), ref.chi2.cpu().numpy())
    # kernel time of an eager event (library events), for the roofline
    eng.set_timing(True)
    kern = []
    for i in range(5):
        eng.fit(pool[i % P], None, o, outs[i % P], type_ranges=ranges)
        kern.append(eng.timings())
    eng.set_timing(False)
    for g in graphs:
        g.close()
    fit_ms = float(np.median([k["fit_ms"] for k in kern]))
    ct = w.chtypes()
    sample = np.ascontiguousarray(s[:, :nev]).astype(np.float32).astype(np.float64)
    orc = oracle_sample(ct, sample, w.event_types(), nev)   # the whole event, both types
    af = float(orc["aflops"].mean())
    achieved = af * nev / (fit_ms / 1e3) / 1e12
    value = E * nev / (total / 1e3)
    line = {
        "metric": "fitted channels/sec (10 samples x 10 BX), warm-start FP32 vs plain FP64, same GPU",
        "value": value, "unit": "channels/s", "n_gpus": 1, "steps": E, "warmup": min(P, E),
        "ms_per_step": total / E, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"c5: {E} events x {nev} channels ({w.event[0]} barrel PU40, "
                               f"{w.event[1]} endcap PU80), warm-start FP32 path, one "
                               f"graph-captured fit call per event",
                   "events": E, "channels_per_event": nev, "event_pool": P},
        "latency_ms": {"p50": float(np.median(lat)), "p99": float(np.quantile(lat, 0.99)),
                       "mean": float(lat.mean()), "max": float(lat.max())},
        "inflight": {"events_in_flight": K_INFLIGHT, "contexts": K_INFLIGHT,
                     "outputs_identical_to_one_at_a_time": inflight_equal,
                     "channels_per_s": E * nev / (total_k / 1e3),
                     "latency_ms_p50": float(np.median(lat_k)),
                     "latency_ms_p99": float(np.quantile(lat_k, 0.99))},
        "eager": {"channels_per_s": E * nev / (total_e / 1e3),
                  "latency_ms_p50": float(np.median(lat_e)),
                  "latency_ms_p99": float(np.quantile(lat_e, 0.99))},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["f32"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["f32"], "traffic": None,
                     "kernel": fit_kernel_name("zero") + ", both types", "kernel_ms": fit_ms,
                     "aflops_per_channel": af,
                     "mlp_mfma_flops_per_channel_not_counted": MLP_FLOPS_PER_CHANNEL},
        "accuracy_vs_plain_fp64": acc,
        "mlp": {k: v for k, v in mlp_info.items() if k != "report"},
    }
    if not args.no_cpu and with_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cb = cpu_baseline(ct, sample, w.event_types(), args.cpu_seconds, threads)
        cb["sample"] = (f"one event ({cb['slice']} channels) fitted by the oracle: median of "
                        f"5 runs of {args.cpu_seconds / 6:.0f} s on {threads} threads; "
                        f"single_thread: its first {cb['slice_1t']} channels on one thread")
        line["cpu_baseline"] = cb
    return line


def batch_block(args, cfg, stream, threads):
    """Another batch workload (zero pulse covariance) timed as the headline:
    warm, cold and plain FP64 paths, autotuned, over the workload's whole
    batch; the warm path's errors against the plain fit and its channel
    accounting against the oracle over the whole batch; the warm kernels'
    roofline."""
    import torch
    from oracle_binding import oracle_fit, oracle_opts
    from test_regression_amd import _abi, engine, warmstart, workloads
    w = workloads.config(cfg, "zero")
    nch = w.nch
    ct = w.chtypes()
    s, t, _ = w.generate(nch, nthreads=threads)
    eng = engine.Engine(list(ct))
    mlp_info = warmstart.ensure_trained(eng, warmstart.training_workload(cfg),
                                        path=warmstart.weights_for(cfg), log=log)
    s64 = torch.from_numpy(s).cuda()
    s32 = s64.float()
    tt = torch.from_numpy(t).cuda() if t is not None else None
    O = _abi.default_opts
    auto = dict(block=0, persistent=_abi.MF_PERSISTENT_AUTO)
    paths = {
        "plain": (s64, O(precision=_abi.MF_PREC_FP64, block=0), _abi.MF_PREC_FP64),
        "fast": (s32, O(precision=_abi.MF_PREC_FP32, **auto), _abi.MF_PREC_FP32),
        "warm": (s32, O(precision=_abi.MF_PREC_FP32, warm_start=1, **auto), _abi.MF_PREC_FP32),
    }
    log(f"{cfg}: {nch} channels")
    res, outs, tuned, graph_ms = time_paths(eng, paths, tt, nch, args, stream)
    o_np = {k: np_out(v) for k, v in outs.items()}
    del outs
    s_r = np.ascontiguousarray(s.astype(np.float32).astype(np.float64))
    ref = oracle_fit(ct, s_r, t, oracle_opts(record_margins=1), nthreads=threads)
    orc = oracle_sample(ct, s_r, t, args.oracle_sample)
    af = float(orc["aflops"].mean())
    k = res["warm"]["fit_ms"] + res["warm"]["mlp_ms"]
    a = af * nch / (k / 1e3) / 1e12
    blk = {
        "workload": f"{cfg}: {w.description}",
        "channels": nch,
        "ms_per_step": {p: v["ms"] for p, v in res.items()},
        "channels_per_s": {p: nch / (v["ms"] / 1e3) for p, v in res.items()},
        "kernel_ms": {p: v["fit_ms"] for p, v in res.items()},
        "warm_graph_ms_per_step": graph_ms,
        "speedup_warm_vs_plain_fp64": res["plain"]["ms"] / res["warm"]["ms"],
        "speedup_cold_vs_plain_fp64": res["plain"]["ms"] / res["fast"]["ms"],
        "roofline": {"warm": {"bound": "valu", "achieved": a, "peak": PEAK_TFLOPS["f32"],
                              "unit": "TFLOP/s", "frac": a / PEAK_TFLOPS["f32"], "kernel_ms": k,
                              "kernel": fit_kernel_name("zero") + ", per type"},
                     "aflops_per_channel": af, "achieved_counts": ACHIEVED_COUNTS},
        "accuracy": {"warm_vs_plain_fp64": errors(*o_np["warm"][:2], *o_np["plain"][:2]),
                     "warm_vs_oracle": accounting(o_np["warm"], ref, 1e-4, False,
                                                  sets=(ct, s_r, t))},
        "solves_per_channel": {p: float(((o_np[p][3] >> 16) & 0xFFFF).mean())
                               for p in ("warm", "fast")},
        "autotune": tuned or None,
        "mlp": {k2: v for k2, v in mlp_info.items() if k2 != "report"},
    }
    log(f"{cfg}: plain {res['plain']['ms']:.3f} ms, cold {res['fast']['ms']:.3f} ms, warm "
        f"{res['warm']['ms']:.3f} ms ({blk['speedup_warm_vs_plain_fp64']:.2f}x)")
    del eng, ref
    torch.cuda.empty_cache()
    return blk


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--pulse-cov", default="zero", choices=["zero", "live"])
    ap.add_argument("--nch", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--oracle-sample", type=int, default=20_000)
    ap.add_argument("--mlp-weights", default="", help="parameters to load (default: the "
                    "committed weights for the workload: c3's own, else those of the c2 "
                    "distribution; missing ones are trained on 4M channels)")
    ap.add_argument("--save-mlp", default="", help="write the trained parameters here")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip mf_autotune (default launch shapes)")
    ap.add_argument("--events", type=int, default=2000, help="event stream (c5): events")
    ap.add_argument("--event-pool", type=int, default=64,
                    help="event stream (c5): distinct events resident in HBM, cycled")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for counter runs (rocprofv3 --pmc): a 256 MiB device copy, the "
                         "calibration of the byte counters, then --steps steps of each path "
                         "of --probe-paths; no timing, no JSON")
    ap.add_argument("--no-live", dest="live", action="store_false",
                    help="skip the live-pulse-covariance block of the default workload")
    ap.add_argument("--no-workloads", dest="workloads", action="store_false",
                    help="skip the c3 (mixed batch) and c5 (event stream) blocks of the default "
                         "run")
    ap.add_argument("--probe-paths", default="warm,fast",
                    help="paths of --traffic-probe, in order: any of plain, plain_chol, fast, "
                         "warm")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the multifit path has no cross-GPU exchange "
              "(channels shard by range); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from oracle_binding import load_pkg
    load_pkg()
    from test_regression_amd import _abi, engine, warmstart, workloads

    w = workloads.config(args.config, args.pulse_cov)
    if w.event:
        print(json.dumps(event_stream(args, engine.Engine(list(w.chtypes())), w, log)),
              flush=True)
        return
    nch = args.nch or w.nch
    ct = w.chtypes()
    t0 = time.perf_counter()
    s, t, _ = w.generate(nch, nthreads=min(16, os.cpu_count() or 1))
    log(f"generated {nch} channels of {args.config} in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda", 0)
    eng = engine.Engine(list(ct))
    s64 = torch.from_numpy(s).to(dev)
    s32 = s64.float()
    tt = torch.from_numpy(t).to(dev) if t is not None else None
    stream = torch.cuda.current_stream()

    # learned warm start: load or train (not timed)
    wpath = Path(args.mlp_weights) if args.mlp_weights else warmstart.weights_for(args.config)
    t0 = time.perf_counter()
    mlp_info = warmstart.ensure_trained(eng, warmstart.training_workload(args.config),
                                        path=wpath if wpath.exists() else None, log=log)
    mlp_info["seconds"] = time.perf_counter() - t0
    if args.save_mlp:
        np.save(args.save_mlp, eng.mlp_get_params())
    log(f"mlp: {mlp_info}")

    O = _abi.default_opts
    if args.traffic_probe:
        a = torch.ones(64 << 20, dtype=torch.float32, device=dev)   # 256 MiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        probe = {"warm": (s32, O(precision=_abi.MF_PREC_FP32, warm_start=1), _abi.MF_PREC_FP32),
                 "fast": (s32, O(precision=_abi.MF_PREC_FP32), _abi.MF_PREC_FP32),
                 "plain": (s64, O(precision=_abi.MF_PREC_FP64), _abi.MF_PREC_FP64),
                 "plain_chol": (s64, O(precision=_abi.MF_PREC_FP64,
                                       solver=_abi.MF_SOLVER_INPLACE_FNNLS_CHOL),
                                _abi.MF_PREC_FP64)}
        for name in args.probe_paths.split(","):
            x_in, o, prec = probe[name]
            out = eng.alloc_result(nch, prec)
            for _ in range(args.steps):
                eng.fit(x_in, tt, o, out)
            torch.cuda.synchronize()
            log(f"traffic probe: {args.steps} steps of {name}")
        return
    auto = dict(block=0, persistent=_abi.MF_PERSISTENT_AUTO)
    paths = {
        "plain": (s64, O(precision=_abi.MF_PREC_FP64, block=0), _abi.MF_PREC_FP64),
        "plain_chol": (s64, O(precision=_abi.MF_PREC_FP64, block=0,
                              solver=_abi.MF_SOLVER_INPLACE_FNNLS_CHOL), _abi.MF_PREC_FP64),
        "fast": (s32, O(precision=_abi.MF_PREC_FP32, **auto), _abi.MF_PREC_FP32),
        "warm": (s32, O(precision=_abi.MF_PREC_FP32, warm_start=1, **auto), _abi.MF_PREC_FP32),
    }
    res, outs, tuned, graph_ms = time_paths(eng, paths, tt, nch, args, stream)

    # accuracy, channel by channel, over the whole batch
    threads = min(16, os.cpu_count() or 1)
    o_np = {k: np_out(v) for k, v in outs.items()}
    acc = {k: errors(*o_np[k][:2], *o_np["plain"][:2]) for k in ("warm", "fast")}
    acc.update(oracle_accounting(ct, s, t, o_np, threads, eng, s64, tt, paths))
    # the same sample gives the flop count
    orc = oracle_sample(ct, s.astype(np.float32).astype(np.float64), t, args.oracle_sample)
    af = float(orc["aflops"].mean())
    it_w = o_np["warm"][3]
    it_f = o_np["fast"][3]
    mlp_eval = warmstart.evaluate(eng, s32, outs["plain"].x.float())
    fl_w = o_np["warm"][2]
    o_chol, o_plain = o_np["plain_chol"], o_np["plain"]
    del outs, o_np
    torch.cuda.empty_cache()

    ms = res["warm"]["ms"]
    value = nch / (ms / 1e3)
    kern_ms = res["warm"]["fit_ms"]
    achieved = af * nch / (kern_ms / 1e3) / 1e12
    peak = PEAK_TFLOPS["f32"]
    line = {
        "metric": "fitted channels/sec (10 samples x 10 BX), warm-start FP32 vs plain FP64, same GPU",
        "value": value,
        "unit": "channels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "ms_per_step_host_clock": res["warm"]["wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {w.description}", "channels": nch,
                   "samples": 10, "pulses": 10, "pulse_cov": args.pulse_cov,
                   "path": "warm-start fast path (MLP prediction on MFMA, FP64 active-set finish)"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": (MEASURED_TRAFFIC.get((args.config, args.pulse_cov, nch)) or
                                 {}).get("bytes"),
                     "traffic_how": (MEASURED_TRAFFIC.get((args.config, args.pulse_cov, nch)) or
                                     {}).get("how"),
                     "kernel": fit_kernel_name(args.pulse_cov),
                     "kernel_ms": kern_ms,
                     "aflops_per_channel": af,
                     "achieved_counts": ACHIEVED_COUNTS,
                     "mlp_mfma_flops_per_channel_not_counted": MLP_FLOPS_PER_CHANNEL},
        "plain_fp64": {"channels_per_s": nch / (res["plain"]["ms"] / 1e3),
                       "ms_per_step": res["plain"]["ms"], "kernel_ms": res["plain"]["fit_ms"],
                       "roofline_frac_fp64": af * nch / (res["plain"]["fit_ms"] / 1e3) / 1e12 /
                       PEAK_TFLOPS["f64"]},
        "plain_fp64_incremental_cholesky": {
            "solver": "MF_SOLVER_INPLACE_FNNLS_CHOL (passive factor kept across solves)",
            "ms_per_step": res["plain_chol"]["ms"], "kernel_ms": res["plain_chol"]["fit_ms"],
            "vs_plain": errors(o_chol[0], o_chol[1], o_plain[0], o_plain[1])},
        "fast_cold_fp32": {"channels_per_s": nch / (res["fast"]["ms"] / 1e3),
                           "ms_per_step": res["fast"]["ms"], "kernel_ms": res["fast"]["fit_ms"]},
        "warm_graph_ms_per_step": graph_ms,
        "speedup_warm_vs_plain_fp64": res["plain"]["ms"] / res["warm"]["ms"],
        "speedup_cold_vs_plain_fp64": res["plain"]["ms"] / res["fast"]["ms"],
        "mlp_ms_per_step": res["warm"]["mlp_ms"],
        "autotune": tuned or None,
        "accuracy": acc,
        "solves_per_channel": {"warm": float(((it_w >> 16) & 0xFFFF).mean()),
                               "cold": float(((it_f >> 16) & 0xFFFF).mean())},
        "warm_start_rejected": float(((fl_w & _abi.FLAGS["warm_reject"]) != 0).mean()),
        "mlp": {"eval_on_batch": mlp_eval, **{k: v for k, v in mlp_info.items() if k != "report"}},
    }
    if "report" in mlp_info:
        line["mlp"]["train"] = mlp_info["report"]
    if args.live and args.pulse_cov == "zero" and args.config == "c2":
        line["live"] = live_block(args, stream, threads)
    if args.workloads and args.pulse_cov == "zero" and args.config == "c2" and not args.nch:
        # the other configurations, timed in the same run: the uneven
        # mixed-type batch and the event stream
        line["workloads"] = {"c3": batch_block(args, "c3", stream, threads)}
        w5 = workloads.config("c5", "zero")
        line["workloads"]["c5"] = event_stream(args, engine.Engine(list(w5.chtypes())), w5, log,
                                               with_cpu=False)
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cb = cpu_baseline(ct, s, t, args.cpu_seconds, threads)
        cb["sample"] = (f"first {cb['slice']} channels of the same batch fitted by the oracle "
                        f"(plain FP64 method, FNNLS max_iter 1000): median of 5 runs of "
                        f"{args.cpu_seconds / 6:.0f} s on {threads} threads; single_thread: "
                        f"the first {cb['slice_1t']} channels on one thread")
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


def time_paths(eng, paths, tt, nch, args, stream, graph_path="warm"):
    """Each path of `paths` (name -> samples, opts, output precision):
    launch shape chosen by mf_autotune for it on this batch (unless
    disabled), then timed (time_steps); the graph_path call is also captured
    as a graph (mf_graph_capture runs it once eagerly first, into the same
    outputs) and replayed, timed the same way.  Returns the timings, the
    outputs, the tunes and the graph's ms per step."""
    import torch
    res, outs, tuned = {}, {}, {}
    default_tune = eng.get_tune()
    nsteps = len(paths)
    for name, (x_in, o, prec) in paths.items():
        out = eng.alloc_result(nch, prec)
        eng.set_tune(**default_tune)
        if args.autotune:
            t0 = time.perf_counter()
            tune, tms = eng.autotune(x_in, tt, o, out)
            tuned[name] = {**tune, "ms": tms, "seconds": time.perf_counter() - t0}
            log(f"autotune {name}: {tuned[name]}")
        ms, wall, kern = time_steps(eng, lambda: eng.fit(x_in, tt, o, out), args.steps,
                                    args.warmup, stream)
        res[name] = {"ms": ms, "wall_ms": wall,
                     "fit_ms": float(np.mean([k["fit_ms"] for k in kern])),
                     "mlp_ms": float(np.mean([k["mlp_ms"] for k in kern]))}
        outs[name] = out
        log(f"{name}: {res[name]['ms']:.3f} ms/step (host clock {wall:.3f}; fit kernels "
            f"{res[name]['fit_ms']:.3f} ms, mlp {res[name]['mlp_ms']:.3f} ms) -> "
            f"{nch / (res[name]['ms'] / 1e3):.3e} ch/s")
    torch.cuda.synchronize()
    graph_ms = None
    if graph_path in paths:
        x_in, o, _ = paths[graph_path]
        g = eng.capture(x_in, tt, o, outs[graph_path])
        for _ in range(args.warmup):
            g.launch()
        torch.cuda.synchronize()
        ga, gb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        gate_stream()
        ga.record(stream)
        for _ in range(args.steps):
            g.launch()
        gb.record(stream)
        torch.cuda.synchronize()
        graph_ms = ga.elapsed_time(gb) / args.steps
        g.close()
        log(f"{graph_path}, graph replays: {graph_ms:.3f} ms/step")
    return res, outs, tuned, graph_ms


def np_out(out):
    """x [10, n], chi2, flags, iters of a result, on the host (x, chi2 in FP64)."""
    return (out.x.double().cpu().numpy(), out.chi2.double().cpu().numpy(),
            out.flags.cpu().numpy(), out.iters.cpu().numpy())


def accounting(o, ref, tol, fp64, mask=None, sets=None):
    """Per-channel parity accounting (tests/parity.py) of one path's outputs
    against the oracle's records: counts of channels outside tolerance and
    how each is explained.  FP64 paths also compare FNNLS counts and flags."""
    from parity import account
    x, c, f, it = o
    rep = account(x, c, ref, tol, iters=it, flags=f if fp64 else None, compare_outer=fp64,
                  mask=mask, sets=sets)
    rep = {k: v for k, v in rep.items() if not k.startswith("_")}
    rep["tol"] = tol
    rep["amplitude_floor_adc"] = 1.0
    return rep


def oracle_accounting(ct, s, t, o_np, threads, eng, s64, tt, paths):
    """The oracle (FP64, with its decision margins) over the whole batch,
    twice: on the FP32-rounded samples the fast path fits and on the FP64
    samples.  Returns, per path, the channel accounting (fast paths at 1e-4,
    the plain fit at 1e-10); for the warm path's rejected channels, their
    accounting alone and chi2 against the plain (cold-start) FP64 fit; and
    the warm path fed the FP64 samples (rounded to FP32 at load) against the
    oracle on the unrounded samples, the error the input rounding adds."""
    from oracle_binding import oracle_fit, oracle_opts
    from test_regression_amd import _abi
    t0 = time.perf_counter()

    def oracle_batch(samples, label):
        # in slices of 4M channels, so that a large batch reports its progress
        n, step = samples.shape[1], 4_000_000
        parts = []
        for a in range(0, n, step):
            b = min(n, a + step)
            parts.append(oracle_fit(ct, np.ascontiguousarray(samples[:, a:b]),
                                    None if t is None else np.ascontiguousarray(t[a:b]),
                                    oracle_opts(record_margins=1), nthreads=threads))
            if n > step:
                log(f"oracle on {label}: {b} of {n} channels ({time.perf_counter() - t0:.0f}s)")
        return parts[0] if len(parts) == 1 else np.concatenate(parts)

    s_r = np.ascontiguousarray(s.astype(np.float32).astype(np.float64))
    ref32 = oracle_batch(s_r, "FP32-rounded samples")
    sets = (ct, s_r, t)
    out = {"vs_oracle": {"warm": accounting(o_np["warm"], ref32, 1e-4, False, sets=sets)}}
    log(f"warm path accounted ({time.perf_counter() - t0:.0f}s)")
    out["vs_oracle"]["fast"] = accounting(o_np["fast"], ref32, 1e-4, False, sets=sets)
    log(f"cold fast path accounted ({time.perf_counter() - t0:.0f}s)")
    fw = o_np["warm"][2]
    rej = (fw & _abi.FLAGS["warm_reject"]) != 0
    cw, cp = o_np["warm"][1], o_np["plain"][1]
    d = np.abs(cw - cp) / np.maximum(cp, 1.0)
    out["warm_rejected"] = {
        "channels": int(rej.sum()), "fraction": float(rej.mean()),
        "fallback_to_cold_start": int(((fw & _abi.FLAGS["warm_fallback"]) != 0).sum()),
        "chi2_rel_vs_plain_fp64_max": float(d[rej].max()) if rej.any() else 0.0,
        "chi2_rel_vs_plain_fp64_above_1e-4": int((d[rej] > 1e-4).sum()),
        "vs_oracle": accounting(o_np["warm"], ref32, 1e-4, False, mask=rej, sets=sets)}
    ref64 = oracle_batch(s, "FP64 samples")
    out["vs_oracle"]["plain"] = accounting(o_np["plain"], ref64, 1e-10, True)
    # FP64 samples into the FP32 path: rounded to FP32 at load, so the
    # outputs are those of the FP32 samples; their distance from the oracle
    # on the unrounded samples is the input rounding's, which the oracle
    # itself shows between its two runs
    x_in, o, prec = paths["warm"]
    r = np_out(eng.fit(s64, tt, o))
    out["warm_fp64_input"] = {
        "outputs_equal_fp32_input": bool(np.array_equal(r[0], o_np["warm"][0]) and
                                         np.array_equal(r[1], o_np["warm"][1])),
        "vs_oracle_unrounded": errors(r[0], r[1], ref64["x"].T, ref64["chi2"]),
        "oracle_rounded_vs_unrounded": errors(ref32["x"].T, ref32["chi2"], ref64["x"].T,
                                              ref64["chi2"])}
    del ref32, ref64
    out["oracle_seconds"] = time.perf_counter() - t0
    out["oracle_threads"] = threads
    log(f"oracle accounting: {out['oracle_seconds']:.1f}s; warm "
        f"{out['vs_oracle']['warm']['outside_tolerance']} outside tolerance "
        f"({out['vs_oracle']['warm']['unexplained']} unexplained), plain "
        f"{out['vs_oracle']['plain']['outside_tolerance']} "
        f"({out['vs_oracle']['plain']['unexplained']} unexplained)")
    return out


def live_block(args, stream, threads):
    """The same workload with a live pulse covariance (the noise covariance
    plus the amplitude-weighted pulse covariance, rebuilt per channel at each
    Minimize iteration): plain, cold and warm paths timed as the headline,
    with their kernels' rooflines and the channel accounting against the
    oracle over the whole batch."""
    import torch
    from test_regression_amd import _abi, engine, warmstart, workloads
    w = workloads.config(args.config, "live")
    nch = args.nch or w.nch
    ct = w.chtypes()
    s, t, _ = w.generate(nch, nthreads=threads)
    eng = engine.Engine(list(ct))
    warmstart.ensure_trained(eng, warmstart.training_workload(args.config),
                             path=warmstart.weights_for(args.config), log=log)
    s64 = torch.from_numpy(s).cuda()
    s32 = s64.float()
    tt = torch.from_numpy(t).cuda() if t is not None else None
    O = _abi.default_opts
    auto = dict(block=0, persistent=_abi.MF_PERSISTENT_AUTO)
    paths = {
        "plain": (s64, O(precision=_abi.MF_PREC_FP64, block=0), _abi.MF_PREC_FP64),
        "fast": (s32, O(precision=_abi.MF_PREC_FP32, **auto), _abi.MF_PREC_FP32),
        "warm": (s32, O(precision=_abi.MF_PREC_FP32, warm_start=1, **auto), _abi.MF_PREC_FP32),
    }
    log(f"live pulse covariance: {nch} channels")
    res, outs, tuned, _ = time_paths(eng, paths, tt, nch, args, stream, graph_path=None)
    o_np = {k: np_out(v) for k, v in outs.items()}
    del outs
    acc = {k: errors(*o_np[k][:2], *o_np["plain"][:2]) for k in ("warm", "fast")}
    acc.update(oracle_accounting(ct, s, t, o_np, threads, eng, s64, tt, paths))
    orc = oracle_sample(ct, s.astype(np.float32).astype(np.float64), t, args.oracle_sample)
    af = float(orc["aflops"].mean())

    def roof(name, peak_key):
        k = res[name]["fit_ms"] + res[name]["mlp_ms"]
        a = af * nch / (k / 1e3) / 1e12
        return {"bound": "valu", "achieved": a, "peak": PEAK_TFLOPS[peak_key],
                "unit": "TFLOP/s", "frac": a / PEAK_TFLOPS[peak_key], "kernel_ms": k,
                "kernel": fit_kernel_name("live", name == "warm") if name != "plain"
                else "k_fit_plain"}

    fl = o_np["warm"][2]
    blk = {
        "workload": f"{args.config} with a live pulse covariance: {w.description}",
        "channels": nch,
        "ms_per_step": {k: v["ms"] for k, v in res.items()},
        "channels_per_s": {k: nch / (v["ms"] / 1e3) for k, v in res.items()},
        "kernel_ms": {k: v["fit_ms"] for k, v in res.items()},
        "mlp_ms": res["warm"]["mlp_ms"],
        "speedup_warm_vs_plain_fp64": res["plain"]["ms"] / res["warm"]["ms"],
        "speedup_cold_vs_plain_fp64": res["plain"]["ms"] / res["fast"]["ms"],
        "roofline": {"warm": roof("warm", "f32"), "fast": roof("fast", "f32"),
                     "plain": roof("plain", "f64"), "aflops_per_channel": af,
                     "achieved_counts": ACHIEVED_COUNTS},
        "accuracy": acc,
        "solves_per_channel": {k: float(((o_np[k][3] >> 16) & 0xFFFF).mean())
                               for k in ("warm", "fast", "plain")},
        "minimize_iterations": {k: float((o_np[k][3] & 255).mean())
                                for k in ("warm", "fast", "plain")},
        "warm_start_rejected": float(((fl & _abi.FLAGS["warm_reject"]) != 0).mean()),
        "autotune": tuned or None,
    }
    log(f"live: plain {res['plain']['ms']:.3f} ms, cold {res['fast']['ms']:.3f} ms, warm "
        f"{res['warm']['ms']:.3f} ms ({blk['speedup_warm_vs_plain_fp64']:.2f}x)")
    del eng
    torch.cuda.empty_cache()
    return blk


if __name__ == "__main__":
    main()
