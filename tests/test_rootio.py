"""The ROOT-free reader (include/mf_io.h, test-regression_amd/csrc/rootio.cpp) on the
reference's own data files, and the reference CPU driver's fit set-up
(test-regression_amd/refdata.py).

The reader is pinned by the generator's arithmetic (gen/src/gen_data.cpp),
which ties branches to each other bit for bit, so a wrong offset, byte order,
entry boundary or basket order in any of them breaks an identity:
  samples[i] = ((samples_noise[i] + pulse_signal[p_i]) + pileup_signal[p_i])
               + input_pedestal,  p_i = 4 (150 + 25 i)           (:318-349)
  amplitudeTruth = signalTruth + energyPU[BX0]                    (:352)
  pulse_signal[w] = signalTruth * fShape(w / 4 - 225)             (:303-305)
The last one also ties the TGraph reader and the linear interpolation of
fShape to the generator's own output (ROOT's TGraph::Eval), bit for bit.
The events come from the reference tree when it is present; the fixture
tests/golden/reference_samples.npz holds the fitted inputs for GPU runs.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle_binding import oracle_fit, oracle_opts

REF = Path("/root/reference/data")
F100 = REF / "mysample_100_0.000_0.000_10_25.00_10.00_0.00_1.000_1.00_CRRC43_0.00_slew_1.00.root"
F1000 = REF / "mysample_1000_0.000_0.000_10_25.00_10.00_0.00_1.000_1.00_CRRC43_0.00_slew_1.00.root"
SHAPE = REF / "EmptyFile.root"
FIXTURE = Path(__file__).resolve().parent / "golden" / "reference_samples.npz"

needs_ref = pytest.mark.skipif(not F100.exists(), reason="reference data files not present")


def _pulse_shape_from_fixture():
    from test_regression_amd import refdata
    z = np.load(FIXTURE)
    t = z["tail"]
    return refdata.PulseShape(z["graph_x"], z["graph_y"], t[0], t[1], t[2])


@needs_ref
@pytest.mark.parametrize("path,entries", [(F100, 100), (F1000, -1)])
def test_samples_tree_identities(pkg, path, entries):
    """mysample_1000 was never closed by its writer: its header stops at the
    first record, the tree record is missing and its branches were flushed
    unevenly; the reader recovers the records by scanning."""
    from test_regression_amd import _abi, refdata
    ps = refdata.PulseShape.read(str(SHAPE))
    with refdata.RootFile(str(path)) as f:
        assert f.tree_entries("Samples") == entries
        s = f.read_vectors("Samples", "samples")
        sn = f.read_vectors("Samples", "samples_noise")
        sig = f.read_vectors("Samples", "pulse_signal")
        pu = f.read_vectors("Samples", "pileup_signal")
        n = min(len(s), len(sn), len(sig), len(pu))
        assert n >= (100 if entries > 0 else 300)
        ped = np.zeros(n)
        if entries > 0:
            ped = f.read_branch("Samples", "input_pedestal", _abi.MF_IO_F32)[0]
            nsmpl = f.read_branch("Samples", "nSmpl", _abi.MF_IO_I32)[0]
            assert np.all(nsmpl == 10)
            truth = f.read_branch("Samples", "amplitudeTruth", _abi.MF_IO_F64)[0]
            st = f.read_branch("Samples", "signalTruth", _abi.MF_IO_F64)[0]
            bx0 = f.read_branch("Samples", "BX0", _abi.MF_IO_I32)[0].astype(int)
            epu = f.read_vectors("Samples", "energyPU")
            nmb = f.read_vectors("Samples", "nMinBias", _abi.MF_IO_VEC_I32)
            assert epu.shape == nmb.shape == (100, 2800)
            assert np.array_equal(truth, st + epu[np.arange(100), bx0])
            assert np.all((epu == 0) == (nmb == 0))
            assert np.all((bx0 >= 30) & (bx0 <= 2800 - 30))
        p = 4 * (150 + 25 * np.arange(10))
        assert np.array_equal(s[:n], ((sn[:n] + sig[:n][:, p]) + pu[:n][:, p]) + ped[:n, None])
        shape = np.array([10.0 * refdata.fshape(ps, w / 4.0 - 225.0, clamp=False)
                          for w in range(sig.shape[1])])
        assert np.array_equal(sig[:n], np.broadcast_to(shape, sig[:n].shape))


@needs_ref
def test_pulse_shape_file(pkg):
    from test_regression_amd import refdata
    ps = refdata.PulseShape.read(str(SHAPE))
    assert np.array_equal(ps.x, np.arange(1000.0))
    assert 0.999 < ps.y.max() <= 1.0 and ps.x[np.argmax(ps.y)] == 62.0
    assert ps.t_min == 800.0
    assert ps.par1 == np.float32(0.01)
    # the tail continues the graph at its end (fShape switches at 800 ns)
    assert abs(ps.par0 * np.exp(-799.0 * ps.par1) - ps.y[799]) < 1e-3 * ps.y.max()


@needs_ref
def test_fixture_matches_reference_files(pkg):
    from test_regression_amd import refdata
    s, a = refdata.read_samples(str(F100))
    ps = refdata.PulseShape.read(str(SHAPE))
    z = np.load(FIXTURE)
    assert np.array_equal(z["samples"], s) and np.array_equal(z["amplitudeTruth"], a)
    assert np.array_equal(z["graph_x"], ps.x) and np.array_equal(z["graph_y"], ps.y)
    assert np.array_equal(z["tail"], [ps.t_min, ps.par0, ps.par1])


@needs_ref
def test_reader_errors(pkg, tmp_path):
    from test_regression_amd import _abi, refdata
    with refdata.RootFile(str(F100)) as f:
        with pytest.raises(_abi.MFError, match="not the kind asked for"):
            f.read_branch("Samples", "samples", _abi.MF_IO_F64)
        with pytest.raises(_abi.MFError, match="entry offsets"):
            f.read_branch("Samples", "amplitudeTruth", _abi.MF_IO_VEC_F64)
        with pytest.raises(_abi.MFError, match="no basket"):
            f.read_branch("Samples", "no_such_branch", _abi.MF_IO_F64)
        with pytest.raises(_abi.MFError, match="no TGraph"):
            f.read_graph("grPulseShape")
    bad = tmp_path / "not_root.root"
    bad.write_bytes(b"this is not a ROOT file" * 10)
    with pytest.raises(_abi.MFError, match="not a ROOT file"):
        refdata.RootFile(str(bad))
    # a file cut short: the records before the cut are still read
    cut = tmp_path / "cut.root"
    data = F100.read_bytes()
    cut.write_bytes(data[:len(data) // 2])
    with refdata.RootFile(str(cut)) as f:
        assert f.tree_entries("Samples") == -1
        e, c = f.read_branch("Samples", "energyPU", _abi.MF_IO_VEC_F64)
        assert 0 < c.size < 100 and np.all(c == 2800)


@needs_ref
def test_reader_survives_corrupted_files(pkg, tmp_path):
    """Corrupted copies of the reference's files (bytes flipped, cut short,
    length fields made huge or negative): every call returns a value or an
    MFError, none crashes or reads outside the file."""
    from test_regression_amd import _abi, refdata
    rng = np.random.default_rng(5)
    branches = [("samples", _abi.MF_IO_VEC_F64), ("amplitudeTruth", _abi.MF_IO_F64),
                ("energyPU", _abi.MF_IO_VEC_F64), ("nMinBias", _abi.MF_IO_VEC_I32),
                ("BX0", _abi.MF_IO_I32), ("input_pedestal", _abi.MF_IO_F32)]
    for src in (F100, SHAPE):
        data = np.frombuffer(src.read_bytes(), dtype=np.uint8)
        for trial in range(150):
            m = data.copy()
            mode = trial % 4
            if mode == 0:
                i = rng.integers(0, m.size, 8)
                m[i] ^= rng.integers(1, 256, 8).astype(np.uint8)
            elif mode == 1:
                m = m[:rng.integers(64, m.size)]
            elif mode == 2:
                i = int(rng.integers(0, m.size - 4))
                m[i:i + 4] = np.frombuffer(np.array([0xFFFFFFF0 if trial % 8 < 4 else 0x7FFFFF00],
                                                    dtype=">u4").tobytes(), dtype=np.uint8)
            else:
                i = rng.integers(0, min(2048, m.size), 4)
                m[i] ^= rng.integers(1, 256, 4).astype(np.uint8)
            f = tmp_path / f"m{trial}.root"
            f.write_bytes(m.tobytes())
            try:
                with refdata.RootFile(str(f)) as r:
                    for b, kind in branches:
                        try:
                            r.read_branch("Samples", b, kind)
                        except _abi.MFError:
                            pass
                    try:
                        r.read_graph("PulseShape/grPulseShape")
                    except _abi.MFError:
                        pass
            except _abi.MFError:
                pass
            f.unlink()


def test_driver_constants(pkg):
    """Pulse::NoiseInit (Pulse.h:134-141) and the driver's template
    (multifit.cpp:48-68) from the pulse shape."""
    from test_regression_amd import refdata
    R = refdata.noisecor()
    assert np.allclose(np.diag(R), 1.0) and np.allclose(R, R.T)
    assert np.isclose(R[0, 1], 1 - (1 - np.exp(-25 / (np.sqrt(2) * 43))) ** 2, rtol=0, atol=1e-15)
    assert np.all(np.linalg.eigvalsh(R) > 0)
    ps = _pulse_shape_from_fixture()
    fp = refdata.driver_fullpulse(ps)
    assert np.all(fp[:7] == 0)
    # grid points: x = 13 + 25 i falls on the graph's 1 ns points
    assert np.array_equal(fp[7:], ps.y[13 + 25 * np.arange(12)])


def test_driver_fit_of_reference_events(pkg):
    """The reference driver's fit of its own generated events (oracle): with
    the driver's template the in-time amplitude is biased low by the 13 ns
    offset between the driver's and the generator's IDSTART; with the
    generator's timing it recovers amplitudeTruth within the noise."""
    from test_regression_amd import _abi, refdata
    z = np.load(FIXTURE)
    s, a = z["samples"], z["amplitudeTruth"]
    ps = _pulse_shape_from_fixture()
    for idstart, lo, hi in ((refdata.IDSTART, -2.9, -2.6), (150.0, -0.02, 0.02)):
        ct = (_abi.ChType * 1)(refdata.driver_chtype(ps, idstart=idstart))
        r = oracle_fit(ct, s, None, oracle_opts(), nthreads=4)
        dA = r["x"][:, 5] - a
        assert lo < dA.mean() < hi, (idstart, dA.mean())
        assert dA.std() < 0.05
        assert np.all(r["flags"] == 0)
