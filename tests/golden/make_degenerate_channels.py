#!/usr/bin/env python3
"""Writes tests/golden/degenerate_channels.npz: three channels of the c2
workload (seed 2; channel indices 11962371, 22806810 and 48707875 of its
stream, made by the library's own generator) whose warm-started fits reach
passive sets holding BX -5 to BX -1 together.  Those pulses are nearly the
same template tail, so the block is ill-conditioned (condition numbers up to
1.3e14) and a product with its rounded inverse alone misses the block's
least-squares optimum (chi2 0.0017 instead of 4.5e-12 on the first channel).
The fixture is data only (10 samples per channel); making it generates the
48.7M channels up to the last one (about two minutes, 3.9 GB of memory).

    python tests/golden/make_degenerate_channels.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
IDS = [11962371, 22806810, 48707875]


def main() -> None:
    sys.path.insert(0, str(HERE.parent))
    from oracle_binding import load_pkg
    load_pkg()
    from test_regression_amd import workloads
    w = workloads.config("c2")
    s, _, _ = w.generate(max(IDS) + 1, nthreads=8)
    np.savez(HERE / "degenerate_channels.npz", samples=np.ascontiguousarray(s[:, IDS]),
             ids=np.array(IDS), seed=np.array(w.seed))


if __name__ == "__main__":
    main()
