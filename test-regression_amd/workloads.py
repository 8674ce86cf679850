"""Synthetic channel populations of the benchmark configurations.

Each configuration is a set of generator specs (mf_gen_spec), a channel
count, a type mix and a seed; libmf_amd.so's host generator (Philox4x32-10
counter-based, so the output is independent of the thread count) turns them
into structure-of-arrays samples.

  ecal      10 samples / 10 BX (-5..+4), alpha 1.138, beta 1.655*25 ns,
            peak at sample 5, sigma 1 ADC, rho 0.5
  hcal      8 samples / 8 BX (-4..+3), alpha 1.138, beta 1.2*25 ns, peak at
            sample 4, sigma 1.5 ADC, rho 0.3
  endcap    ecal shape with sigma 2 ADC, rho 0.6

  c1  10,000 ecal channels, in-time amplitude U[1,100], no pileup, no jitter
  c2  1,000,000 ecal channels, in-time U[1,500], per BX Poisson(40/10) pileup
      pulses of Exponential(mean 2) amplitude, N(0, 1 ns) jitter per pulse
  c3  2,000,000 channels, 70% ecal / 30% hcal, pileup mean ~ U[0,80] per
      channel, otherwise as c2
  c5  events of 61,200 barrel (ecal, pileup 40) + 14,648 endcap (pileup 80)
      channels, otherwise as c2

pulse_cov: "zero" leaves fullpulsecov at 0 (every reference driver passes a
zero matrix, multifit_cpu/src/multifit.cpp:27,193); "live" sets it to
sigma_t^2 A'(t) A'(t)^T with sigma_t = 1 ns, the covariance of a pulse whose
time jitters by N(0, sigma_t).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi

ECAL = dict(nsamples=10, peak_sample=5, alpha=1.138, beta_ns=1.655 * 25, sigma=1.0, rho=0.5)
HCAL = dict(nsamples=8, peak_sample=4, alpha=1.138, beta_ns=1.2 * 25, sigma=1.5, rho=0.3)
ENDCAP = dict(ECAL, sigma=2.0, rho=0.6)


def spec(shape: dict, a0=(1.0, 500.0), pu=(0.0, 0.0), pu_amp=2.0, jitter=0.0,
         sigma_t=0.0) -> _abi.GenSpec:
    s = _abi.GenSpec()
    s.nsamples = shape["nsamples"]
    s.peak_sample = shape["peak_sample"]
    s.alpha = shape["alpha"]
    s.beta_ns = shape["beta_ns"]
    s.sigma = shape["sigma"]
    s.rho = shape["rho"]
    s.a0_lo, s.a0_hi = a0
    s.pu_lo, s.pu_hi = pu
    s.pu_amp_mean = pu_amp
    s.jitter_ns = jitter
    s.sigma_t_ns = sigma_t
    return s


@dataclass
class Workload:
    name: str
    specs: list
    nch: int
    seed: int
    mix: float = 0.0          # fraction of type-1 channels (0: all type 0)
    description: str = ""
    types: list = field(default_factory=list)
    event: tuple = ()         # channels per type in one event, in order (event streams)

    def chtypes(self):
        out = (_abi.ChType * len(self.specs))()
        for i, s in enumerate(self.specs):
            _abi.check(_abi.lib().mf_gen_type(C.byref(s), C.byref(out[i])), "mf_gen_type")
        return out

    def event_types(self) -> np.ndarray:
        """Type of each channel of one event: event[0] of type 0, then
        event[1] of type 1, ..."""
        return np.concatenate([np.full(k, t, dtype=np.uint8) for t, k in enumerate(self.event)])

    def generate(self, nch: int | None = None, nthreads: int = 0, with_truth: bool = False,
                 events: int = 0):
        """Returns (samples [10, n] float64, type_of [n] uint8 or None, truth or None).
        events > 0 (workloads with an event layout): n = events x the event
        size, channels event after event, each event laid out by event_types()."""
        n = self.nch if nch is None else int(nch)
        L = _abi.lib()
        type_of = None
        if events > 0:
            if not self.event:
                raise ValueError(f"{self.name} has no event layout")
            type_of = np.tile(self.event_types(), events)
            n = type_of.size
        elif self.mix > 0:
            type_of = np.empty(n, dtype=np.uint8)
            _abi.check(L.mf_gen_mix(n, self.mix, self.seed + 7919,
                                    type_of.ctypes.data_as(C.c_void_p)), "mf_gen_mix")
        samples = np.empty((10, n), dtype=np.float64)
        truth = np.empty((10, n), dtype=np.float64) if with_truth else None
        arr = (_abi.GenSpec * len(self.specs))(*self.specs)
        _abi.check(L.mf_gen_samples(
            arr, len(self.specs),
            type_of.ctypes.data_as(C.c_void_p) if type_of is not None else None,
            n, self.seed, samples.ctypes.data_as(C.c_void_p),
            truth.ctypes.data_as(C.c_void_p) if truth is not None else None,
            nthreads), "mf_gen_samples")
        return samples, type_of, truth


def config(name: str, pulse_cov: str = "zero") -> Workload:
    st = 1.0 if pulse_cov == "live" else 0.0
    if pulse_cov not in ("zero", "live"):
        raise ValueError("pulse_cov must be 'zero' or 'live'")
    if name == "c1":
        return Workload("c1", [spec(ECAL, a0=(1, 100), sigma_t=st)], 10_000, 1,
                        description="10k ECAL-like, in-time U[1,100], no pileup")
    if name == "c2":
        return Workload("c2", [spec(ECAL, a0=(1, 500), pu=(40, 40), jitter=1.0, sigma_t=st)],
                        1_000_000, 2,
                        description="1M ECAL-like, in-time U[1,500], pileup 40, jitter 1 ns")
    if name == "c3":
        return Workload("c3", [spec(ECAL, a0=(1, 500), pu=(0, 80), jitter=1.0, sigma_t=st),
                               spec(HCAL, a0=(1, 500), pu=(0, 80), jitter=1.0, sigma_t=st)],
                        2_000_000, 3, mix=0.3,
                        description="2M, 70% ECAL-like / 30% HCAL-like, pileup U[0,80]")
    if name == "c4":
        return Workload("c4", [spec(ECAL, a0=(1, 500), pu=(40, 40), jitter=1.0, sigma_t=st)],
                        4_000_000, 4, description="MLP training set, c2 distribution")
    if name == "c5":
        return Workload("c5", [spec(ECAL, a0=(1, 500), pu=(40, 40), jitter=1.0, sigma_t=st),
                               spec(ENDCAP, a0=(1, 500), pu=(80, 80), jitter=1.0, sigma_t=st)],
                        75_848, 5, mix=14_648 / 75_848,
                        description="event: 61,200 barrel PU40 + 14,648 endcap PU80",
                        event=(61_200, 14_648))
    raise KeyError(name)
