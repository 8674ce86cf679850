#!/usr/bin/env python3
"""Writes tests/golden/reference_samples.npz from the reference's own data files,
read with the engine's ROOT-free reader (include/mf_io.h).

Run in a container where the reference tree is present:

  * samples, amplitudeTruth -- the 100 events of the `Samples` tree of
    data/mysample_100_0.000_0.000_10_25.00_10.00_0.00_1.000_1.00_CRRC43_0.00_slew_1.00.root
    (gen/src/gen_data.cpp: signal amplitude 10, no pileup, noise sigma 0.044
    with the CRRC correlation of tau 43 ns);
  * graph_x, graph_y, tail -- the pulse shape of data/EmptyFile.root
    (TGraph PulseShape/grPulseShape; timeMin, expAmplitude, expTime of the
    tree PulseShape/Tail), from which the reference's CPU driver builds its
    template (multifit_cpu/interface/Pulse.h:69-131, multifit.cpp:44-83).

The fixture is data only: the reference's generated inputs and its pulse
shape.  GPU tests fit these events without the reference tree present.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = Path("/root/reference/data")
SAMPLES = "mysample_100_0.000_0.000_10_25.00_10.00_0.00_1.000_1.00_CRRC43_0.00_slew_1.00.root"
SHAPE = "EmptyFile.root"


def main() -> None:
    sys.path.insert(0, str(HERE.parent))
    from oracle_binding import load_pkg
    load_pkg()
    from test_regression_amd import refdata

    s, a = refdata.read_samples(str(REF / SAMPLES))
    ps = refdata.PulseShape.read(str(REF / SHAPE))
    out = HERE / "reference_samples.npz"
    np.savez_compressed(out, samples=s, amplitudeTruth=a, graph_x=ps.x, graph_y=ps.y,
                        tail=np.array([ps.t_min, ps.par0, ps.par1]))
    print(f"wrote {out}: {s.shape[1]} events, {ps.x.size} pulse-shape points")


if __name__ == "__main__":
    main()
