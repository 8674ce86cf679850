"""ctypes bindings of the test infrastructure: the FP64 oracle
(oracle/liboracle.so), the reference's own Eigen-free sources compiled by
oracle/Makefile (oracle/_ref/*.so, only where the reference tree was present
at build time), and the host build of the product's fit core
(tests/hostcore/libhostcore.so).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
HOSTCORE_DIR = ROOT / "tests" / "hostcore"


def load_pkg():
    """Imports test-regression_amd/ as module `test_regression_amd`."""
    name = "test_regression_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg = ROOT / "test-regression_amd"
    spec = importlib.util.spec_from_file_location(name, pkg / "__init__.py",
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


class OrOpts(C.Structure):
    _fields_ = [("eps", C.c_double), ("max_iter", C.c_uint), ("chi2_tol", C.c_double),
                ("max_minimize", C.c_int), ("compute_errors", C.c_int), ("solver", C.c_int),
                ("count_flops", C.c_int), ("record_margins", C.c_int)]


class OrResult(C.Structure):
    _fields_ = [("x", C.c_double * 10), ("err", C.c_double * 10), ("chi2", C.c_double),
                ("status", C.c_int), ("flags", C.c_int), ("n_minimize", C.c_int),
                ("n_outer", C.c_int), ("n_inner", C.c_int), ("ksum", C.c_longlong),
                ("k2sum", C.c_longlong), ("k3sum", C.c_longlong),
                ("m_eps", C.c_double), ("m_arg", C.c_double), ("m_pos", C.c_double),
                ("m_alpha", C.c_double), ("pivot", C.c_double), ("m_chi2", C.c_double),
                ("xerr", C.c_double), ("aflops", C.c_double)]


class OrStats(C.Structure):
    _fields_ = [("outer", C.c_int), ("inner", C.c_int), ("hit_max", C.c_int),
                ("llt_fail", C.c_int), ("ksum", C.c_longlong), ("k2sum", C.c_longlong),
                ("k3sum", C.c_longlong), ("aflops", C.c_double)]


OR_SOLVER_INPLACE, OR_SOLVER_INPLACE_CHOL, OR_SOLVER_FNNLS, OR_SOLVER_LEGACY, OR_SOLVER_LH = \
    0, 1, 2, 3, 4

_cache: dict = {}


def _build(dirpath: Path, target: str):
    subprocess.run(["make", "-s", "-C", str(dirpath), target], check=True)


def oracle():
    if "oracle" in _cache:
        return _cache["oracle"]
    so = ORACLE_DIR / "liboracle.so"
    if not so.exists():
        _build(ORACLE_DIR, "oracle")
    L = C.CDLL(str(so))
    dp = C.POINTER(C.c_double)
    L.or_llt.argtypes = [dp, dp, C.c_int]
    L.or_llt.restype = C.c_int
    for f in ("or_cholesky_decomp",):
        getattr(L, f).argtypes = [dp, dp, C.c_int, C.c_int]
    for f in ("or_forward_subst", "or_backward_subst"):
        getattr(L, f).argtypes = [dp, dp, dp, C.c_int, C.c_int]
    L.or_swap_row_column.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.or_fused_rcaddition.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int]
    L.or_fused_inbetween_removal.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int, C.c_int]
    L.or_transpose_multiply.argtypes = [dp, dp, C.c_int]
    for f in ("or_inplace_fnnls", "or_inplace_fnnls_chol", "or_fnnls"):
        getattr(L, f).argtypes = [dp, dp, dp, C.c_int, C.c_double, C.c_uint,
                                  C.POINTER(OrStats)]
    L.or_lh_nnls.argtypes = [dp, dp, dp, C.c_int, C.c_double, C.c_int, C.POINTER(OrStats)]
    L.or_lh_nnls.restype = C.c_int
    L.or_build_pulsemat.argtypes = [C.c_void_p, dp]
    L.or_dofit.argtypes = [C.c_void_p, dp, C.POINTER(OrOpts), C.POINTER(OrResult)]
    L.or_dofit.restype = C.c_int
    L.or_dofit_batch.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_longlong,
                                 C.POINTER(OrOpts), C.POINTER(OrResult), C.c_int]
    L.or_dofit_batch_pederr.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_void_p, C.c_longlong,
                                        C.POINTER(OrOpts), C.POINTER(OrResult), C.c_int]
    L.or_dofit_from_set.argtypes = [C.c_void_p, dp, C.POINTER(OrOpts), C.c_uint,
                                    C.POINTER(OrResult), C.POINTER(C.c_int)]
    L.or_dofit_from_set.restype = C.c_double
    L.or_kkt_batch.argtypes = [C.c_void_p, C.c_void_p, dp, C.c_void_p, C.c_longlong, C.c_void_p,
                               C.c_longlong, C.c_void_p, C.POINTER(OrOpts), C.c_void_p, dp,
                               C.c_void_p]
    _cache["oracle"] = L
    return L


def ref_lib(name: str):
    """oracle/_ref/lib<name>.so or None when it was not built here."""
    so = ORACLE_DIR / "_ref" / f"lib{name}.so"
    if not so.exists():
        return None
    return C.CDLL(str(so))


def hostcore():
    if "hostcore" in _cache:
        return _cache["hostcore"]
    so = HOSTCORE_DIR / "libhostcore.so"
    _build(HOSTCORE_DIR, "libhostcore.so")
    L = C.CDLL(str(so))
    dp = C.POINTER(C.c_double)
    L.hc_fit_batch.argtypes = [C.c_void_p, C.c_int, C.c_void_p, dp, C.c_int64, C.c_void_p,
                               dp, dp, C.c_void_p, C.c_void_p, dp]
    L.hc_fit_batch.restype = C.c_int
    L.hc_fit_batch2.argtypes = [C.c_void_p, C.c_int, C.c_void_p, dp, C.c_void_p, C.c_int64,
                                C.c_void_p, dp, dp, C.c_void_p, C.c_void_p, dp]
    L.hc_fit_batch2.restype = C.c_int
    L.hc_nnls.argtypes = [dp, dp, dp, C.c_int, C.c_int, C.c_double, C.c_uint]
    L.hc_nnls.restype = C.c_int
    L.hc_fast_from.argtypes = [C.c_void_p, C.c_int, C.c_void_p, dp, dp, C.c_int64, C.c_int,
                               C.c_void_p, dp, dp, C.c_void_p, C.c_void_p]
    L.hc_fast_from.restype = C.c_int
    L.hc_mask_inverse.argtypes = [C.c_void_p, C.c_int, dp, dp]
    L.hc_mask_inverse.restype = C.c_int
    L.hc_set_use_tables.argtypes = [C.c_int]
    L.hc_set_layout.argtypes = [C.c_int]
    L.hc_set_split.argtypes = [C.c_int]
    L.hc_pending_count.restype = C.c_longlong
    _cache["hostcore"] = L
    return L


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def oracle_opts(eps=1e-11, max_iter=1000, chi2_tol=1e-3, max_minimize=50,
                compute_errors=0, solver=OR_SOLVER_INPLACE, count_flops=0,
                record_margins=1) -> OrOpts:
    """Oracle options; decision margins are recorded by default (the parity
    tests use them), the timed CPU baseline turns them off."""
    return OrOpts(eps=eps, max_iter=max_iter, chi2_tol=chi2_tol, max_minimize=max_minimize,
                  compute_errors=compute_errors, solver=solver, count_flops=count_flops,
                  record_margins=record_margins)


RESULT_DTYPE = np.dtype([
    ("x", "f8", 10), ("err", "f8", 10), ("chi2", "f8"), ("status", "i4"),
    ("flags", "i4"), ("n_minimize", "i4"), ("n_outer", "i4"), ("n_inner", "i4"),
    ("ksum", "i8"), ("k2sum", "i8"), ("k3sum", "i8"), ("m_eps", "f8"), ("m_arg", "f8"),
    ("m_pos", "f8"), ("m_alpha", "f8"), ("pivot", "f8"), ("m_chi2", "f8"), ("xerr", "f8"),
    ("aflops", "f8")], align=True)
assert RESULT_DTYPE.itemsize == C.sizeof(OrResult)


def oracle_fit(chtypes, samples: np.ndarray, type_of=None, opts: OrOpts | None = None,
               nthreads: int = 1, pederr: np.ndarray | None = None):
    """Runs the oracle's DoFit on every channel.  samples [10, n] float64;
    pederr (optional, [n] float64): DoFit's pederr per channel."""
    L = oracle()
    n = samples.shape[1]
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    res = (OrResult * n)()
    if pederr is not None:
        pederr = np.ascontiguousarray(pederr, dtype=np.float64)
        assert pederr.shape == (n,)
    L.or_dofit_batch_pederr(C.cast(chtypes, C.c_void_p),
                            type_of.ctypes.data_as(C.c_void_p) if type_of is not None else None,
                            dptr(samples),
                            pederr.ctypes.data_as(C.c_void_p) if pederr is not None else None,
                            n, C.byref(opts or oracle_opts()), res, nthreads)
    return np.frombuffer(res, dtype=RESULT_DTYPE)


def oracle_kkt_sets(chtypes, samples: np.ndarray, type_of, idx, pm, opts: OrOpts | None = None,
                    pederr: np.ndarray | None = None):
    """The oracle's DoFit of channel idx[a] continued from passive set pm[a]
    (bit mask) as the end of its first NNLS call (or_kkt_batch).  Returns the
    continued fits (an oracle_fit record array) with two more fields:
    margin, the first call's stop-test margin in rounding units of w (<= 0
    passes outright), and feasible (the set's amplitudes all > 0)."""
    L = oracle()
    n = samples.shape[1]
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    pm = np.ascontiguousarray(pm, dtype=np.uint32)
    k = idx.size
    res = (OrResult * max(k, 1))()
    margin = np.zeros(max(k, 1))
    feas = np.zeros(max(k, 1), dtype=np.int32)
    if pederr is not None:
        pederr = np.ascontiguousarray(pederr, dtype=np.float64)
    if k:
        L.or_kkt_batch(C.cast(chtypes, C.c_void_p),
                       type_of.ctypes.data_as(C.c_void_p) if type_of is not None else None,
                       dptr(samples),
                       pederr.ctypes.data_as(C.c_void_p) if pederr is not None else None,
                       n, idx.ctypes.data_as(C.c_void_p), k, pm.ctypes.data_as(C.c_void_p),
                       C.byref(opts or oracle_opts()), res, dptr(margin),
                       feas.ctypes.data_as(C.c_void_p))
    r = np.frombuffer(res, dtype=RESULT_DTYPE)[:k]
    return {"fit": r, "margin": margin[:k], "feasible": feas[:k] != 0}


def hostcore_fit(chtypes, samples: np.ndarray, type_of, opts, pederr=None):
    """Runs fit_core.hpp's host build.  opts is the product's mf_opts;
    pederr (optional) as mf_batch.pederr."""
    L = hostcore()
    n = samples.shape[1]
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    x = np.zeros((10, n))
    chi2 = np.zeros(n)
    flags = np.zeros(n, dtype=np.int32)
    iters = np.zeros(n, dtype=np.int32)
    err = np.zeros((10, n))
    if pederr is not None:
        pederr = np.ascontiguousarray(pederr, dtype=np.float64)
    rc = L.hc_fit_batch2(C.cast(chtypes, C.c_void_p), len(chtypes),
                         type_of.ctypes.data_as(C.c_void_p) if type_of is not None else None,
                         dptr(samples),
                         pederr.ctypes.data_as(C.c_void_p) if pederr is not None else None,
                         n, C.cast(C.byref(opts), C.c_void_p), dptr(x),
                         dptr(chi2), flags.ctypes.data_as(C.c_void_p),
                         iters.ctypes.data_as(C.c_void_p), dptr(err))
    assert rc == 0
    if opts.compute_errors:
        return x, chi2, flags, iters, err
    return x, chi2, flags, iters


def hostcore_fast_from(chtypes, samples: np.ndarray, xstart: np.ndarray, warm: bool, opts):
    """The fast path of fit_core.hpp's host build from given starting
    amplitudes (warm start: passive set = xstart > 0)."""
    L = hostcore()
    n = samples.shape[1]
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    xs = np.ascontiguousarray(xstart, dtype=np.float64)
    x = np.zeros((10, n))
    chi2 = np.zeros(n)
    flags = np.zeros(n, dtype=np.int32)
    iters = np.zeros(n, dtype=np.int32)
    rc = L.hc_fast_from(C.cast(chtypes, C.c_void_p), len(chtypes), None, dptr(samples), dptr(xs),
                        n, 1 if warm else 0, C.cast(C.byref(opts), C.c_void_p), dptr(x),
                        dptr(chi2), flags.ctypes.data_as(C.c_void_p),
                        iters.ctypes.data_as(C.c_void_p))
    assert rc == 0, rc
    return x, chi2, flags, iters
