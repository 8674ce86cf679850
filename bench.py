#!/usr/bin/env python3
"""bench.py -- fitted channels/s of the multifit engine on one MI355X.

A step is one mf_fit_batch over one batch of synthetic channels already
resident in HBM (structure of arrays).  Default workload: 1,000,000 ECAL-like
channels, 10 samples x 10 BX, pileup 40, in-time amplitude U[1,500], pulse
jitter N(0, 1 ns).  Reported:

  value        channels/s of the fast path (FP32 bulk, warm active set,
               FP64 refinement) -- or of the plain FP64 path while the fast
               path is not built
  plain_fp64   channels/s of the plain cold-start FP64 fit (reference
               semantics, FNNLS max_iter 1000)
  roofline     dominant kernel: algorithmic flops per launch (the plain
               method's operation count, counted by the oracle on a sample of
               the same workload) / its mean launch time from HIP events on
               the launch stream, against the MI355X vector peak
  cpu_baseline the oracle (a C restatement of multifit_cpu) on the host's
               cores, on a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tests"))

PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [a.elapsed_time(b) for a, b in evs]
    return wall, per


def cpu_baseline(ct, s, t, seconds, threads):
    """Oracle throughput on host cores over a bounded prefix of the batch."""
    from oracle_binding import oracle_fit, oracle_opts
    n_probe = 2000
    t0 = time.perf_counter()
    oracle_fit(ct, s[:, :n_probe].copy(), t[:n_probe] if t is not None else None,
               oracle_opts(), nthreads=threads)
    rate = n_probe / max(time.perf_counter() - t0, 1e-6)
    n = int(min(s.shape[1], max(n_probe, rate * seconds)))
    t0 = time.perf_counter()
    oracle_fit(ct, np.ascontiguousarray(s[:, :n]), t[:n] if t is not None else None,
               oracle_opts(), nthreads=threads)
    dt = time.perf_counter() - t0
    return n / dt, n, dt


def algorithmic_flops(ct, s, t, n=20000):
    from oracle_binding import oracle_fit, oracle_opts
    r = oracle_fit(ct, np.ascontiguousarray(s[:, :n]), t[:n] if t is not None else None,
                   oracle_opts(count_flops=1), nthreads=min(16, os.cpu_count() or 1))
    return float(r["aflops"].mean())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--pulse-cov", default="zero", choices=["zero", "live"])
    ap.add_argument("--nch", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="", help="plain|fast: time one path only")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the multifit path has no cross-GPU exchange "
              "(channels shard by range); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from oracle_binding import load_pkg
    load_pkg()
    from test_regression_amd import _abi, engine, workloads

    w = workloads.config(args.config, args.pulse_cov)
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

    res = {}
    o64 = _abi.default_opts(precision=_abi.MF_PREC_FP64)
    out64 = eng.alloc_result(nch, _abi.MF_PREC_FP64)
    if args.only in ("", "plain"):
        wall, per = time_steps(lambda: eng.fit(s64, tt, o64, out64), args.steps, args.warmup,
                               stream)
        res["plain"] = dict(wall=wall, per=per)
        log(f"plain fp64: {np.mean(per):.3f} ms/step -> {nch / (np.mean(per) / 1e3):.3e} ch/s")

    fast_ok = False
    if args.only in ("", "fast"):
        try:
            o32 = _abi.default_opts(precision=_abi.MF_PREC_FP32)
            out32 = eng.alloc_result(nch, _abi.MF_PREC_FP32)
            eng.fit(s32, tt, o32, out32)
            torch.cuda.synchronize()
            fast_ok = True
            if fast_ok:
                wall, per = time_steps(lambda: eng.fit(s32, tt, o32, out32), args.steps,
                                       args.warmup, stream)
                res["fast"] = dict(wall=wall, per=per)
                log(f"fast fp32: {np.mean(per):.3f} ms/step -> "
                    f"{nch / (np.mean(per) / 1e3):.3e} ch/s")
        except _abi.MFError as e:
            log(f"fast path unavailable: {e}")

    head = "fast" if fast_ok else "plain"
    per = res[head]["per"]
    ms = float(np.mean(per))
    value = nch / (ms / 1e3)
    af = algorithmic_flops(ct, s, t)
    achieved = af * nch / (ms / 1e3) / 1e12
    peak = PEAK_TFLOPS["fp32" if head == "fast" else "fp64"]
    line = {
        "metric": "fitted channels/sec (10 samples x 10 BX), warm-start FP32 vs plain FP64, same GPU",
        "value": value,
        "unit": "channels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fast" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {w.description}", "channels": nch,
                   "samples": 10, "pulses": 10, "pulse_cov": args.pulse_cov,
                   "path": head},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "aflops_per_channel": af},
    }
    if "plain" in res:
        line["plain_fp64"] = {"channels_per_s": nch / (np.mean(res["plain"]["per"]) / 1e3),
                              "ms_per_step": float(np.mean(res["plain"]["per"]))}
    if "fast" in res and "plain" in res:
        line["speedup_vs_plain_fp64"] = (np.mean(res["plain"]["per"]) / np.mean(res["fast"]["per"]))
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        rate, n, dt = cpu_baseline(ct, s, t, args.cpu_seconds, threads)
        line["cpu_baseline"] = {"value": rate, "unit": "channels/s", "cores": threads,
                                "kind": "port",
                                "sample": f"first {n} channels of the same batch, {dt:.1f} s, "
                                          f"FNNLS max_iter 1000, {threads} threads"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
