"""The reference's own data files, read with the ROOT-free reader (include/mf_io.h),
and its CPU driver's fit set-up on top of the engine.

- read_branch / read_graph / tree_entries: the reader's C ABI on numpy arrays.
- PulseShape / fshape: the pulse shape of multifit_cpu/interface/Pulse.h:69-131
  (TGraph PulseShape/grPulseShape and tree PulseShape/Tail of
  data/EmptyFile.root).
- noisecor: Pulse::NoiseInit (Pulse.h:134-141).
- driver_chtype: the DoFit constants multifit_cpu/src/multifit.cpp:44-83 sets
  up (template, noise correlation, active BX -5..4) with pedrms 1 (:194-196)
  and a zero pulse covariance.
- read_samples / write_samples: the `Samples` tree of
  gen/src/gen_data.cpp:224-248 (what multifit.cpp:88-94 reads).
- fit_samples / write_reco: the driver's per-entry outputs
  (multifit.cpp:196-227): chi2, the in-time amplitude and samplesReco (the
  amplitudes by BX + 5), fitted on the GPU.  The reference stores them as the
  `RecoAndSim` ROOT tree (:108-129, 266-279); write_reco writes that tree to a
  ROOT file (mf_io_write_tree) with the same branches; write_tree writes any
  tree of /F, /D, /I and vector branches.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _abi

# multifit_cpu/interface/Pulse.h:19-33
NSAMPLES, NFREQ, IDSTART, TAU = 10, 25, 150 + 13, 43.0
ACTIVE_BX = (-5, -4, -3, -2, -1, 0, 1, 2, 3, 4)


class RootFile:
    """An open ROOT file (mf_io_open); use as a context manager."""

    def __init__(self, path: str):
        self.path = str(path)
        h = C.c_void_p()
        _abi.check(_abi.lib().mf_io_open(self.path.encode(), C.byref(h)), f"mf_io_open({path})")
        self._h = h

    def close(self):
        if self._h:
            _abi.lib().mf_io_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def tree_entries(self, tree: str) -> int:
        n = C.c_int64()
        _abi.check(_abi.lib().mf_io_tree_entries(self._h, tree.encode(), C.byref(n)),
                   "mf_io_tree_entries")
        return n.value

    def read_branch(self, tree: str, branch: str, kind: int):
        """(values, counts): every value of the branch in entry order, and the
        number of values of each entry."""
        L = _abi.lib()
        nv, ne = C.c_int64(), C.c_int64()
        _abi.check(L.mf_io_read_branch(self._h, tree.encode(), branch.encode(), kind, None, 0,
                                       C.byref(nv), None, 0, C.byref(ne)), "mf_io_read_branch")
        vals = np.zeros(nv.value, dtype=np.float64)
        counts = np.zeros(ne.value, dtype=np.int64)
        _abi.check(L.mf_io_read_branch(self._h, tree.encode(), branch.encode(), kind,
                                       vals.ctypes.data_as(C.c_void_p), vals.size, C.byref(nv),
                                       counts.ctypes.data_as(C.c_void_p), counts.size,
                                       C.byref(ne)), "mf_io_read_branch")
        return vals, counts

    def read_vectors(self, tree: str, branch: str, kind: int = _abi.MF_IO_VEC_F64) -> np.ndarray:
        """A vector branch whose entries all have the same length, as [entries, length]."""
        vals, counts = self.read_branch(tree, branch, kind)
        if counts.size and not np.all(counts == counts[0]):
            raise _abi.MFError(f"{branch}: entries of different lengths")
        return vals.reshape(counts.size, -1 if counts.size else 0)

    def read_graph(self, name: str):
        L = _abi.lib()
        n = C.c_int64()
        _abi.check(L.mf_io_read_graph(self._h, name.encode(), None, None, 0, C.byref(n)),
                   "mf_io_read_graph")
        x = np.zeros(n.value)
        y = np.zeros(n.value)
        _abi.check(L.mf_io_read_graph(self._h, name.encode(), x.ctypes.data_as(C.c_void_p),
                                      y.ctypes.data_as(C.c_void_p), n.value, C.byref(n)),
                   "mf_io_read_graph")
        return x, y


@dataclass
class PulseShape:
    """Pulse shape file content: graph points and the exponential tail."""
    x: np.ndarray
    y: np.ndarray
    t_min: float
    par0: float
    par1: float

    @classmethod
    def read(cls, path: str) -> "PulseShape":
        """PulseShape/grPulseShape and the first entry of PulseShape/Tail
        (Pulse.h:69-89); the tail parameters are float branches."""
        with RootFile(path) as f:
            x, y = f.read_graph("PulseShape/grPulseShape")
            tail = {b: f.read_branch("PulseShape/Tail", b, _abi.MF_IO_F32)[0][0]
                    for b in ("timeMin", "expAmplitude", "expTime")}
        return cls(x, y, tail["timeMin"], tail["expAmplitude"], tail["expTime"])

    def eval_graph(self, t: float) -> float:
        """Linear interpolation between the graph's two points around t (the
        graph's points are sorted); beyond the ends, the line through the two
        end points; exactly the point's value on a point."""
        x, y = self.x, self.y
        k = int(np.searchsorted(x, t))
        if k < x.size and x[k] == t:
            return float(y[k])
        lo = min(max(k - 1, 0), x.size - 2)
        hi = lo + 1
        return float(y[hi] + (t - x[hi]) * (y[lo] - y[hi]) / (x[lo] - x[hi]))


def fshape(ps: PulseShape, t: float, clamp: bool = True) -> float:
    """Pulse::fShape (multifit_cpu/interface/Pulse.h:116-131): the graph for
    0 < t < 800 ns (negative values clamped to 0), fPar0 exp(-t fPar1)
    beyond, 0 for t <= 0.  clamp=False: the generator's version
    (gen/interface/Pulse.h fShape), which does not clamp."""
    if t <= 0.0:
        return 0.0
    if t < 800.0:
        v = ps.eval_graph(t)
        return max(v, 0.0) if clamp else v
    return ps.par0 * math.exp(-t * ps.par1)


def noisecor(nsamples: int = NSAMPLES, tau: float = TAU) -> np.ndarray:
    """Pulse::NoiseInit (Pulse.h:134-141): R[i][j] = c(|i - j|) with
    c(k) = 1 - (1 - exp(-25 k / (sqrt(2) tau)))^2."""
    c = np.empty(nsamples)
    for i in range(nsamples):
        y = 1.0 - math.exp(-float(NFREQ * i) / (math.sqrt(2.0) * tau))
        c[i] = 1.0 - y * y
    idx = np.abs(np.arange(nsamples)[:, None] - np.arange(nsamples)[None, :])
    return c[idx]


def driver_fullpulse(ps: PulseShape, idstart: float = IDSTART) -> np.ndarray:
    """fullpulse of multifit.cpp:48-68: entries 7..18 = fShape(IDSTART +
    25 (i + 3) + 25 - 250) for i < 12, zero before.  The driver's IDSTART is
    163 ns; the generator places its pulses with IDSTART 150
    (gen/src/gen_data.cpp:115, 328), so the driver's template peaks 13 ns
    later than the generated pulses (idstart=150 aligns them)."""
    fp = np.zeros(19)
    for i in range(NSAMPLES + 2):
        fp[i + 7] = fshape(ps, float(idstart + NFREQ * (i + 3) + NFREQ - 500 / 2))
    return fp


def driver_chtype(ps: PulseShape, pederr: float = 1.0, idstart: float = IDSTART) -> _abi.ChType:
    """The mf_chtype of the reference's CPU driver: its template and noise
    correlation, BX -5..4 (multifit.cpp:77-82), DoFit's ECAL layout
    (firstsamplet = max(0, bx + 3), offset = 4 - bx), zero pulse covariance."""
    t = _abi.ChType()
    t.nsamples = NSAMPLES
    t.npulses = len(ACTIVE_BX)
    for i, b in enumerate(ACTIVE_BX):
        t.bx[i] = b
    t.first_off, t.pulse_off, t.pederr = 3, 4, pederr
    for i, v in enumerate(noisecor().ravel()):
        t.noisecor[i] = v
    for i, v in enumerate(driver_fullpulse(ps, idstart)):
        t.fullpulse[i] = v
    return t


def read_samples(path: str):
    """(samples [10, n], amplitudeTruth [n] or None) of a `Samples` tree.  A
    file whose writer stopped early can hold fewer entries of one branch than
    of another, or none: n is what both hold, and amplitudeTruth is None when
    the file holds none of it."""
    with RootFile(path) as f:
        s = f.read_vectors("Samples", "samples")
        try:
            a, _ = f.read_branch("Samples", "amplitudeTruth", _abi.MF_IO_F64)
        except _abi.MFError as e:
            if "no basket" not in str(e):
                raise
            a = None
    if a is None:
        return np.ascontiguousarray(s.T), None
    n = min(s.shape[0], a.size)
    return np.ascontiguousarray(s[:n].T), a[:n]


def fit_samples(samples: np.ndarray, chtype: _abi.ChType, opts=None, device="cuda"):
    """DoFit of every entry on the GPU, as the driver's kernel lambda does
    (multifit.cpp:196-227, errors off): returns chi2, ampl (X at the first
    BX == 0) and samplesReco [n, 10] (X by BX + 5)."""
    import torch
    from . import engine
    eng = engine.Engine([chtype])
    dt = np.float32 if samples.dtype == np.float32 else np.float64
    r = eng.fit(torch.from_numpy(np.ascontiguousarray(samples, dtype=dt)).to(device),
                None, opts or _abi.default_opts())
    x = r.x.double().cpu().numpy()
    bx = np.array(ACTIVE_BX[:chtype.npulses])
    it = int(np.flatnonzero(bx == 0)[0])
    reco = np.zeros((x.shape[1], 10))
    reco[:, bx + 5] = x[:chtype.npulses].T
    return {"chi2": r.chi2.double().cpu().numpy(), "ampl": x[it].copy(), "samplesReco": reco,
            "flags": r.flags.cpu().numpy()}


def write_tree(path: str, tree: str, branches: dict, title: str = "") -> None:
    """Writes a ROOT file holding one tree (mf_io_write_tree).  branches:
    name -> (kind, values) for /F, /D, /I (one value per entry) or (kind,
    list of per-entry sequences) for the vector kinds, in branch order."""
    cols, keep, n = [], [], None
    for name, (kind, vals) in branches.items():
        if kind in (_abi.MF_IO_VEC_F64, _abi.MF_IO_VEC_I32):
            counts = np.array([len(v) for v in vals], dtype=np.int64)
            flat = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64).ravel()
                                                        for v in vals]) if len(vals) else
                                        np.zeros(0), dtype=np.float64)
            m = counts.size
        else:
            counts = None
            flat = np.ascontiguousarray(np.asarray(vals, dtype=np.float64).ravel())
            m = flat.size
        if n is None:
            n = m
        if m != n:
            raise _abi.MFError(f"write_tree: branch {name} has {m} entries, not {n}")
        keep += [flat, counts, name.encode()]
        cols.append(_abi.IoColumn(keep[-1], kind, 0, flat.ctypes.data_as(C.c_void_p),
                                  counts.ctypes.data_as(C.c_void_p) if counts is not None
                                  else None))
    arr = (_abi.IoColumn * len(cols))(*cols)
    _abi.check(_abi.lib().mf_io_write_tree(str(path).encode(), tree.encode(), title.encode(),
                                           C.cast(arr, C.c_void_p), len(cols), n or 0),
               "mf_io_write_tree")


# The generator's Samples branches in its order (gen/src/gen_data.cpp:225-248)
# with their kinds; wf_name (/C, a string) is not written.
SAMPLES_BRANCHES = (
    ("pulse_shift", _abi.MF_IO_F32), ("pileup_shift", _abi.MF_IO_F32),
    ("nSmpl", _abi.MF_IO_I32), ("nFreq", _abi.MF_IO_F32),
    ("amplitudeTruth", _abi.MF_IO_F64), ("samples", _abi.MF_IO_VEC_F64),
    ("samples_noise", _abi.MF_IO_VEC_F64), ("nPU", _abi.MF_IO_F32), ("BX0", _abi.MF_IO_I32),
    ("nBX", _abi.MF_IO_I32), ("nWF", _abi.MF_IO_I32), ("nMinBias", _abi.MF_IO_VEC_F64),
    ("energyPU", _abi.MF_IO_VEC_F64), ("pulse_signal", _abi.MF_IO_VEC_F64),
    ("pileup_signal", _abi.MF_IO_VEC_F64), ("signalTruth", _abi.MF_IO_F64),
    ("sigmaNoise", _abi.MF_IO_F32), ("puFactor", _abi.MF_IO_F32),
    ("pulse_tau", _abi.MF_IO_F32), ("input_pedestal", _abi.MF_IO_F32),
    ("distortion_sample_4", _abi.MF_IO_F32))


def write_samples(path: str, samples: np.ndarray, amplitude_truth: np.ndarray,
                  extra: dict | None = None) -> None:
    """Writes a `Samples` tree as the generator lays it out
    (gen/src/gen_data.cpp:224-248), readable by read_samples and by the
    reference's driver (multifit.cpp:88-94): samples [nsamples, n] (one
    vector per entry), amplitudeTruth [n], nSmpl (the sample count), and any
    other generator branch given in extra (name -> per-entry values, or a
    list of per-entry sequences for the vector branches), all in the
    generator's branch order."""
    s = np.asarray(samples, dtype=np.float64)
    if s.ndim != 2:
        raise _abi.MFError("write_samples: samples must be [nsamples, n]")
    n = s.shape[1]
    vals = dict(extra or {})
    unknown = set(vals) - {name for name, _ in SAMPLES_BRANCHES}
    if unknown:
        raise _abi.MFError(f"write_samples: not a generator branch: {sorted(unknown)}")
    vals["samples"] = list(np.ascontiguousarray(s.T))
    vals["amplitudeTruth"] = np.asarray(amplitude_truth, dtype=np.float64)
    vals.setdefault("nSmpl", np.full(n, s.shape[0]))
    write_tree(path, "Samples", {name: (kind, vals[name]) for name, kind in SAMPLES_BRANCHES
                                 if name in vals})


def write_reco(path: str, fit: dict, ps: PulseShape, source: dict | None = None) -> None:
    """The CPU driver's output file (multifit.cpp:108-129, 247-262, 266-279):
    a ROOT file with the tree RecoAndSim, one entry per fitted input entry.
    Its branches are the input tree's (CloneTree(0)): source, name -> (kind,
    values) as write_tree takes them, e.g. the Samples branches this reader
    covers; then the driver's own, filled as the driver fills them: chi2
    (/F; the driver never assigns it, so it keeps its initial -99),
    samplesReco (the amplitudes by BX + 5), ipulseintime (/I, 0),
    complete_chi2 and complete_pedestal (empty vectors), best_pedestal and
    best_chi2 (/F, 0), pulseShapeTemplate (fShape(IDSTART + 25 i + 75 - 250),
    i < 17) and activeBXs (-5..4).  The driver's complete_samplesReco
    (std::vector<std::vector<double>>, empty) is not written."""
    reco = np.asarray(fit["samplesReco"], dtype=np.float64)
    n = reco.shape[0]
    template = np.array([fshape(ps, float(IDSTART + NFREQ * i + 3 * 25.0 - 500 / 2.0))
                         for i in range(NSAMPLES + 7 * (25 // NFREQ))])
    br = dict(source or {})
    br["chi2"] = (_abi.MF_IO_F32, np.full(n, -99.0))
    br["samplesReco"] = (_abi.MF_IO_VEC_F64, list(reco))
    br["ipulseintime"] = (_abi.MF_IO_I32, np.zeros(n))
    br["complete_chi2"] = (_abi.MF_IO_VEC_F64, [np.zeros(0)] * n)
    br["complete_pedestal"] = (_abi.MF_IO_VEC_F64, [np.zeros(0)] * n)
    br["best_pedestal"] = (_abi.MF_IO_F32, np.zeros(n))
    br["best_chi2"] = (_abi.MF_IO_F32, np.zeros(n))
    br["pulseShapeTemplate"] = (_abi.MF_IO_VEC_F64, [template] * n)
    br["activeBXs"] = (_abi.MF_IO_VEC_I32, [np.array(ACTIVE_BX, dtype=np.float64)] * n)
    write_tree(path, "RecoAndSim", br)
