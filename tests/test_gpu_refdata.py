"""The reference's own generated events (tests/golden/reference_samples.npz, read
from its data files by the ROOT-free reader) fitted on the GPU with the
reference CPU driver's set-up (multifit_cpu/src/multifit.cpp:44-83, 196-227),
against the oracle on the same inputs.  Bars as the other parity tests:
FP64 amplitudes within 1e-10 of the oracle (relative to the channel's
largest amplitude, floor 1), FP32 fast path within 1e-4.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle_binding import OR_SOLVER_LEGACY, oracle_fit, oracle_opts

pytestmark = pytest.mark.gpu
FIXTURE = Path(__file__).resolve().parent / "golden" / "reference_samples.npz"


def _setup(idstart=None):
    from test_regression_amd import _abi, refdata
    z = np.load(FIXTURE)
    t = z["tail"]
    ps = refdata.PulseShape(z["graph_x"], z["graph_y"], t[0], t[1], t[2])
    ct = refdata.driver_chtype(ps, idstart=idstart if idstart is not None else refdata.IDSTART)
    return z["samples"], z["amplitudeTruth"], ps, ct, (_abi.ChType * 1)(ct)


@pytest.mark.parametrize("mode", ["fp64", "legacy", "fp32"])
def test_reference_events_match_oracle(pkg, gpu, mode):
    from test_regression_amd import _abi, refdata
    s, a, ps, ct, cts = _setup()
    kw, oo, tol = {}, oracle_opts(), 1e-10
    if mode == "legacy":
        kw, oo = {"solver": _abi.MF_SOLVER_LEGACY}, oracle_opts(solver=OR_SOLVER_LEGACY)
    if mode == "fp32":
        kw, tol = {"precision": _abi.MF_PREC_FP32}, 1e-4
        s = s.astype(np.float32).astype(np.float64)
    print(f"[gpu-refdata] mode={mode} events={s.shape[1]}", flush=True)
    fit = refdata.fit_samples(s if mode != "fp32" else s.astype(np.float32), ct,
                              _abi.default_opts(**kw))
    ref = oracle_fit(cts, s, None, oo, nthreads=4)
    scale = np.maximum(np.abs(ref["x"]).max(axis=1), 1.0)
    err = np.abs(fit["samplesReco"] - ref["x"]).max(axis=1) / scale
    assert err.max() <= tol, err.max()
    assert np.allclose(fit["chi2"], ref["chi2"], rtol=max(tol, 1e-10), atol=1e-10)
    assert np.array_equal(fit["ampl"], fit["samplesReco"][:, 5])
    # the fast path's own status bits (FP64 pivot, warm reject, refined) have
    # no oracle counterpart
    assert np.all((fit["flags"] & 15) == ref["flags"])


def test_reference_events_recover_truth(pkg, gpu, tmp_path):
    """With the generator's timing (IDSTART 150) the fitted in-time amplitude
    matches amplitudeTruth within the noise; the RecoAndSim file the driver
    writes is read back with the reader: the cloned input branches and the
    fit's samplesReco, bit for bit."""
    from test_regression_amd import _abi, refdata
    s, a, ps, ct, _ = _setup(idstart=150.0)
    fit = refdata.fit_samples(s, ct, _abi.default_opts())
    dA = fit["ampl"] - a
    print(f"dA mean {dA.mean():.4f} std {dA.std():.4f}")
    assert abs(dA.mean()) < 0.02 and dA.std() < 0.05
    out = tmp_path / "reco.root"
    src = {"samples": (_abi.MF_IO_VEC_F64, list(s.T)),
           "amplitudeTruth": (_abi.MF_IO_F64, a)}
    refdata.write_reco(str(out), fit, ps, source=src)
    with refdata.RootFile(str(out)) as f:
        assert f.tree_entries("RecoAndSim") == s.shape[1]
        assert np.array_equal(f.read_vectors("RecoAndSim", "samplesReco"), fit["samplesReco"])
        assert np.array_equal(f.read_vectors("RecoAndSim", "samples"), s.T)
        assert np.array_equal(f.read_branch("RecoAndSim", "amplitudeTruth", _abi.MF_IO_F64)[0], a)
        assert f.read_vectors("RecoAndSim", "pulseShapeTemplate").shape == (s.shape[1], 17)


def test_cpp_driver_round_trip(pkg, gpu, tmp_path):
    """The CPU driver's loop in C++ on the library (tests/cpp/host_api_check
    --driver): the events written as a Samples tree (refdata.write_samples),
    read back with mf_io_*, fitted entry by entry by PulseChiSqSNNLS (errors
    off, as the driver runs them), and written as RecoAndSim with
    mf_io_write_tree.  The file read back holds the input branches bit for
    bit and the amplitudes of the batched fit of the same events; both are
    within the FP64 bar of the oracle."""
    import subprocess
    from test_regression_amd import _abi, refdata
    prog = Path(__file__).resolve().parent / "cpp" / "host_api_check"
    if not prog.exists():
        pytest.fail("tests/cpp/host_api_check is not built (make -C test-regression_amd)")
    s, a, ps, ct, cts = _setup()
    samples = tmp_path / "samples.root"
    refdata.write_samples(str(samples), s, a)
    nc = np.ctypeslib.as_array(ct.noisecor)
    tb = np.concatenate([nc.ravel(), np.ctypeslib.as_array(ct.fullpulse).ravel(),
                         np.ctypeslib.as_array(ct.fullpulsecov).ravel(), [ct.pederr, ct.npulses],
                         np.ctypeslib.as_array(ct.bx).astype(np.float64)[:10]])
    typ = tmp_path / "type.bin"
    tb.astype(np.float64).tofile(typ)
    out = tmp_path / "reco.root"
    print(f"[gpu-refdata] C++ driver on {s.shape[1]} events", flush=True)
    r = subprocess.run([str(prog), "--driver", str(samples), str(typ), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    fit = refdata.fit_samples(s, ct, _abi.default_opts())
    ref = oracle_fit(cts, s, None, oracle_opts(), nthreads=4)
    with refdata.RootFile(str(out)) as f:
        assert f.tree_entries("RecoAndSim") == s.shape[1]
        assert np.array_equal(f.read_vectors("RecoAndSim", "samples"), s.T)
        assert np.array_equal(f.read_branch("RecoAndSim", "amplitudeTruth", _abi.MF_IO_F64)[0], a)
        assert np.all(f.read_branch("RecoAndSim", "chi2", _abi.MF_IO_F32)[0] == -99.0)
        reco = f.read_vectors("RecoAndSim", "samplesReco")
        bxs = f.read_vectors("RecoAndSim", "activeBXs", _abi.MF_IO_VEC_I32)
    assert np.array_equal(bxs, np.tile(np.array(refdata.ACTIVE_BX, dtype=np.float64), (s.shape[1], 1)))
    scale = np.maximum(np.abs(ref["x"]).max(axis=1), 1.0)
    assert (np.abs(reco - ref["x"]).max(axis=1) / scale).max() <= 1e-10
    assert (np.abs(reco - fit["samplesReco"]).max(axis=1) / scale).max() <= 1e-12
