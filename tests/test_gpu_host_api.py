"""The C++ host library (include/mf_host.hpp) through its test program
tests/cpp/host_api_check: PulseChiSqSNNLS::DoFit (errors on, as the
reference's default), doFitWrapper, inplace_fnnls / inplace_fnnls_wrapper and
the linalgebra templates, each against the oracle or a closed form."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle_binding import OR_SOLVER_LEGACY, OrStats, dptr, oracle, oracle_fit, oracle_opts

pytestmark = pytest.mark.gpu
PROG = Path(__file__).resolve().parent / "cpp" / "host_api_check"


def test_cpp_host_library(pkg, gpu, tmp_path):
    from test_regression_amd import workloads
    assert PROG.exists(), "build the test program: make -C test-regression_amd"
    n = 64
    for pulse_cov in ("zero", "live"):
        w = workloads.config("c2", pulse_cov)
        ct = w.chtypes()
        s, _, _ = w.generate(n)
        t = ct[0]
        rec = []
        for c in range(n):
            rec += list(s[:, c]) + list(t.noisecor[:]) + [t.pederr] + list(t.fullpulse[:])
            rec += list(t.fullpulsecov[:]) + list(t.bx[:])
        inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
        with open(inp, "wb") as f:
            f.write(np.int32(n).tobytes())
            f.write(np.asarray(rec, dtype=np.float64).tobytes())
        print(f"[gpu-host-api] pulse_cov={pulse_cov} n={n}", flush=True)
        r = subprocess.run([str(PROG), str(inp), str(outp)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        o = np.fromfile(outp, dtype=np.float64)
        k = 0
        dofit = o[k:k + n * 21].reshape(n, 21)
        k += n * 21
        wrap = o[k:k + n * 12].reshape(n, 12)
        k += n * 12
        nn = o[k:k + n * 20].reshape(n, 20)
        k += n * 20
        la = o[k:k + 30]
        k += 30
        leg = o[k:].reshape(n, 21)
        ref = oracle_fit(ct, s, None, oracle_opts(compute_errors=1), nthreads=4)
        ref0 = oracle_fit(ct, s, None, oracle_opts(), nthreads=4)
        scale = np.maximum(np.abs(ref["x"]).max(axis=1), 1.0)[:, None]
        assert np.all(np.abs(dofit[:, :10] - ref["x"]) <= 1e-10 * scale)
        e, er = dofit[:, 15], ref["err"][:, 5]
        assert np.array_equal(np.isnan(e), np.isnan(er))
        ok = ~np.isnan(er)
        assert np.all(np.abs(e[ok] - er[ok]) <= 1e-9 * np.abs(er[ok]))
        assert np.allclose(dofit[:, 20], ref["chi2"], rtol=1e-10, atol=1e-10)
        assert np.all(np.abs(wrap[:, :10] - ref0["x"]) <= 1e-10 * scale)
        assert np.allclose(wrap[:, 10], ref0["chi2"], rtol=1e-10, atol=1e-10)
        assert np.allclose(wrap[:, 11], ref0["x"][:, 5], rtol=1e-10, atol=1e-10)  # BX 0
        # NNLS on A = I + pulsemat, b = samples
        L = oracle()
        P = np.zeros((10, 10))
        fp = np.array(t.fullpulse[:])
        for j, bx in enumerate(t.bx):
            f0, off = max(0, bx + 3), 4 - bx
            P[f0:, j] = fp[f0 + off:10 + off]
        A = P + np.eye(10)
        for c in range(n):
            xr = np.zeros(10)
            L.or_inplace_fnnls(dptr(np.ascontiguousarray(A)), dptr(np.ascontiguousarray(s[:, c])),
                               dptr(xr), 10, 1e-11, 1000, OrStats())
            sc = max(1.0, np.abs(xr).max())
            assert np.abs(nn[c, :10] - xr).max() <= 1e-10 * sc
            assert np.abs(nn[c, 10:] - xr).max() <= 1e-10 * sc
        # linalgebra 3x3 literal: closed form
        l11 = np.sqrt(5.75)
        Lx = np.array([[2, 0, 0], [-0.5, l11, 0], [1, 0.5 / l11, np.sqrt(4 - 0.25 / 5.75)]])
        M3 = np.array([[4, -1, 2], [-1, 6, 0], [2, 0, 5]], dtype=float)
        x3 = np.linalg.solve(M3, [3, 3, 4])
        assert np.allclose(la[:9].reshape(3, 3), Lx, atol=1e-14)
        assert np.allclose(la[12:15], x3, atol=1e-14)
        # legacy_multifit_cpu's class: the legacy solver, errors on
        refl = oracle_fit(ct, s, None, oracle_opts(compute_errors=1, solver=OR_SOLVER_LEGACY),
                          nthreads=4)
        scl = np.maximum(np.abs(refl["x"]).max(axis=1), 1.0)[:, None]
        assert np.all(np.abs(leg[:, :10] - refl["x"]) <= 1e-10 * scl)
        e, er = leg[:, 15], refl["err"][:, 5]
        assert np.array_equal(np.isnan(e), np.isnan(er))
        ok = ~np.isnan(er)
        assert np.all(np.abs(e[ok] - er[ok]) <= 1e-9 * np.abs(er[ok]))
        assert np.allclose(leg[:, 20], refl["chi2"], rtol=1e-10, atol=1e-10)
        assert np.allclose(la[15:24].reshape(3, 3), Lx, atol=1e-6)
        assert np.allclose(la[27:30], x3, atol=1e-5)
