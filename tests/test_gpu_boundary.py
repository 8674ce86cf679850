"""GPU parity on the boundary the reference's callers use: one batch mixing
BX lists of 1, 5 and 10 pulses (the reference's dynamic BXVector,
multifit_cpu/src/PulseChiSqSNNLS.cpp:32-33,50-62) with a pederr per channel
(DoFitArgs::pederr, multifit_gpu/interface/DeviceData.h:6-26), through
mf_fit_batch on every path and through the C++ doFitWrapper on a full c5
event (75,848 channels, every pederr distinct).  Against the FP64 oracle,
which fits each channel with its own BX list and pederr."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import parity
import torch

from boundary_cases import BX_LISTS, mixed_batch, npulse_types
from oracle_binding import oracle_fit, oracle_opts

pytestmark = pytest.mark.gpu
PROG = Path(__file__).resolve().parent / "cpp" / "host_api_check"
NPULSES = (1, 5, 10)


def _errors(x, chi2, ref, t):
    xr = ref["x"].T
    scale = np.maximum(np.abs(xr).max(axis=0), 1.0)
    err = (np.abs(x - xr) / scale).max(axis=0)
    aset = ((x > 0) == (xr > 0)).all(axis=0)
    cerr = np.abs(chi2 - ref["chi2"]) / np.maximum(ref["chi2"], 1.0)
    pad = np.arange(10)[:, None] >= np.asarray(NPULSES)[t][None, :]
    assert np.all(x[pad] == 0), "slots past a channel's pulses must stay 0"
    return err, aset, cerr


@pytest.mark.parametrize("pulse_cov", ["zero", "live"])
@pytest.mark.parametrize("mode", ["plain", "cold", "warm", "persistent"])
def test_pederr_and_npulses_every_path(pkg, gpu, pulse_cov, mode):
    from test_regression_amd import engine, warmstart, _abi
    ct = npulse_types(pulse_cov)
    n = 20_000
    s, t, ped = mixed_batch(n, pulse_cov=pulse_cov)
    eng = engine.Engine(list(ct))
    fp64 = mode == "plain"
    if mode == "warm":
        warmstart.ensure_trained(eng)
    o = _abi.default_opts(precision=_abi.MF_PREC_FP64 if fp64 else _abi.MF_PREC_FP32,
                          warm_start=1 if mode == "warm" else 0,
                          persistent=1 if mode == "persistent" else 0)
    s_in = s if fp64 else s.astype(np.float32)
    print(f"[gpu-boundary] mode={mode} pulse_cov={pulse_cov} n={n}", flush=True)
    r = eng.fit(torch.from_numpy(s_in).cuda(), torch.from_numpy(t).cuda(), o,
                pederr=torch.from_numpy(ped).cuda())
    torch.cuda.synchronize()
    s_ref = s_in.astype(np.float64)
    ref = oracle_fit(ct, s_ref, t, oracle_opts(), nthreads=8, pederr=ped)
    x = r.x.double().cpu().numpy()
    chi2 = r.chi2.double().cpu().numpy()
    _errors(x, chi2, ref, t)
    tol = 1e-10 if fp64 else 1e-4
    parity.gate(x, chi2, ref, tol, iters=r.iters.cpu().numpy(),
                flags=r.flags.cpu().numpy() if fp64 else None, compare_outer=fp64,
                label=f"boundary {mode} {pulse_cov}")
    for k, it in enumerate((0, 2, 5)):   # every pulse count fits its in-time pulse (BX 0)
        assert (x[it, t == k] > 0).mean() > 0.95


def test_do_fit_wrapper_event_distinct_pederr(pkg, gpu, tmp_path):
    """doFitWrapper on one c5-sized event (75,848 channels), BX lists of 1, 5
    and 10 pulses, a distinct pederr on every channel."""
    from test_regression_amd import workloads
    assert PROG.exists(), "build the test program: make -C test-regression_amd"
    n = sum(workloads.config("c5").event)
    ct = npulse_types("zero")
    s, t, ped = mixed_batch(n, seed=23)
    ped = np.sort(ped)[np.random.default_rng(5).permutation(n)]
    assert np.unique(ped).size == n
    inp, outp = tmp_path / "w.bin", tmp_path / "o.bin"
    with open(inp, "wb") as f:
        f.write(np.int32(n).tobytes())
        f.write(np.int32(len(ct)).tobytes())
        for k, ty in enumerate(ct):
            rec = list(ty.noisecor[:]) + list(ty.fullpulse[:]) + list(ty.fullpulsecov[:])
            bx = BX_LISTS[NPULSES[k]]
            rec += [len(bx)] + bx + [0] * (10 - len(bx))
            f.write(np.asarray(rec, dtype=np.float64).tobytes())
        ch = np.concatenate([t[None, :].astype(np.float64), ped[None, :], s], axis=0)
        f.write(np.ascontiguousarray(ch.T).tobytes())
    print(f"[gpu-boundary] doFitWrapper n={n}", flush=True)
    r = subprocess.run([str(PROG), "--wrapper", str(inp), str(outp)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    o = np.fromfile(outp, dtype=np.float64).reshape(n, 13)
    ref = oracle_fit(ct, s, t, oracle_opts(), nthreads=8, pederr=ped)
    _errors(o[:, :10].T.copy(), o[:, 10], ref, t)
    parity.gate(o[:, :10].T.copy(), o[:, 10], ref, 1e-10, label="doFitWrapper c5 event")
    # Output::ampl is X at the first BX 0: pulse 0, 2 and 5 of the three lists
    it = np.array([0, 2, 5])[t]
    assert np.array_equal(o[:, 11], o[np.arange(n), it])
    assert np.all(o[:, 12] == 1)
