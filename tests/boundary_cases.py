"""Inputs of the boundary tests: channel types of 1, 5 and 10 pulses over the
c2 template and noise (the reference's dynamic BX list, bxs.rows() <= 10,
multifit_cpu/src/PulseChiSqSNNLS.cpp:32-33,50-62), and per-channel pederr
(DoFit's pederr argument, multifit_gpu/interface/DeviceData.h:6-26)."""
from __future__ import annotations

import numpy as np

BX_LISTS = {1: [0], 5: [-2, -1, 0, 1, 2], 10: list(range(-5, 5))}


def npulse_types(pulse_cov="zero", npulses=(1, 5, 10), pederr=1.0):
    """ChType array: the c2 type with each BX list of npulses."""
    from test_regression_amd import _abi, workloads
    base = workloads.config("c2", pulse_cov).chtypes()[0]
    out = (_abi.ChType * len(npulses))()
    for k, n in enumerate(npulses):
        t = out[k]
        C = type(base)
        C.from_buffer(t)  # noqa: B018 (same layout)
        import ctypes
        ctypes.memmove(ctypes.addressof(t), ctypes.addressof(base), ctypes.sizeof(base))
        t.npulses = n
        bx = BX_LISTS[n]
        for p in range(10):
            t.bx[p] = bx[p] if p < n else 0
        t.pederr = pederr
    return out


def mixed_batch(n, seed=11, pulse_cov="zero"):
    """c2 samples, a type id per channel (0, 1, 2 -> 1, 5, 10 pulses) and a
    pederr per channel drawn from U[0.5, 3]."""
    from test_regression_amd import workloads
    w = workloads.config("c2", pulse_cov)
    w.seed += seed
    s, _, _ = w.generate(n)
    rng = np.random.default_rng(seed)
    t = rng.integers(0, 3, size=n).astype(np.uint8)
    ped = rng.uniform(0.5, 3.0, size=n)
    return s, t, ped
