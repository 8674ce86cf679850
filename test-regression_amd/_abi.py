"""ctypes view of libmf_amd.so (include/mf_abi.h, include/mf_gen.h).

The library is built in place by ``make -C test-regression_amd``.  Loading
fails loudly when it is missing: there is no Python or CPU fallback for any
entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libmf_amd.so"

MF_OK, MF_ERR_INVALID, MF_ERR_HIP, MF_ERR_NOMEM, MF_ERR_UNSUPPORTED = range(5)
MF_PREC_FP64, MF_PREC_FP32 = 0, 1
(MF_SOLVER_INPLACE_FNNLS, MF_SOLVER_FNNLS, MF_SOLVER_LEGACY, MF_SOLVER_LAWSON_HANSON,
 MF_SOLVER_INPLACE_FNNLS_CHOL) = 0, 1, 2, 3, 4
MF_DTYPE_F64, MF_DTYPE_F32 = 0, 1
MF_MLP_NPARAM = 6164
MF_NOALLOC_SCRATCH_BYTES = 64 << 20
(MF_LA_CHOLESKY, MF_LA_FORWARD, MF_LA_BACKWARD, MF_LA_SWAP, MF_LA_RCADDITION,
 MF_LA_INBETWEEN) = range(6)

FLAGS = {
    "minimize_cap": 1, "llt_fail": 2, "fnnls_cap": 4, "nonfinite": 8,
    "fp64_pivot": 16, "warm_reject": 32, "refined": 64, "bad_type": 128, "warm_fallback": 256,
}


class ChType(C.Structure):
    """mf_chtype: one channel type (DoFit's per-channel constants)."""
    _fields_ = [
        ("nsamples", C.c_int32), ("npulses", C.c_int32), ("bx", C.c_int32 * 10),
        ("first_off", C.c_int32), ("pulse_off", C.c_int32), ("pederr", C.c_double),
        ("noisecor", C.c_double * 100), ("fullpulse", C.c_double * 19),
        ("fullpulsecov", C.c_double * 361),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("solver", C.c_int32), ("max_iter", C.c_int32),
        ("max_minimize", C.c_int32), ("eps", C.c_double), ("chi2_tol", C.c_double),
        ("use_pulse_cov", C.c_int32), ("warm_start", C.c_int32),
        ("persistent", C.c_int32), ("block", C.c_int32),
        ("compute_errors", C.c_int32), ("pad", C.c_int32),
    ]


class Batch(C.Structure):
    _fields_ = [
        ("nch", C.c_int64), ("sample_dtype", C.c_int32), ("nranges", C.c_int32),
        ("samples", C.c_void_p), ("type_of", C.c_void_p), ("ranges", C.c_void_p),
        ("pederr", C.c_void_p),
    ]


class Result(C.Structure):
    _fields_ = [
        ("x", C.c_void_p), ("chi2", C.c_void_p), ("flags", C.c_void_p),
        ("iters", C.c_void_p), ("err", C.c_void_p),
    ]


class GenSpec(C.Structure):
    _fields_ = [
        ("nsamples", C.c_int32), ("peak_sample", C.c_int32), ("alpha", C.c_double),
        ("beta_ns", C.c_double), ("sigma", C.c_double), ("rho", C.c_double),
        ("a0_lo", C.c_double), ("a0_hi", C.c_double), ("pu_lo", C.c_double),
        ("pu_hi", C.c_double), ("pu_amp_mean", C.c_double), ("jitter_ns", C.c_double),
        ("sigma_t_ns", C.c_double),
    ]


class Timings(C.Structure):
    _fields_ = [("mlp_ms", C.c_float), ("fit_ms", C.c_float), ("fit_launches", C.c_int32),
                ("pad", C.c_int32)]


class Tune(C.Structure):
    _fields_ = [("block", C.c_int32), ("p1_wpc", C.c_int32), ("p2_wpc", C.c_int32),
                ("p1_refill", C.c_int32), ("p2_refill", C.c_int32), ("fit_block", C.c_int32),
                ("use_persistent", C.c_int32), ("pad", C.c_int32)]


MF_PERSISTENT_AUTO = 2
MF_IO_F32, MF_IO_F64, MF_IO_I32, MF_IO_VEC_F64, MF_IO_VEC_I32 = 0, 1, 2, 3, 4


class IoColumn(C.Structure):
    """mf_io_column (include/mf_io.h)."""
    _fields_ = [("name", C.c_char_p), ("kind", C.c_int32), ("pad", C.c_int32),
                ("values", C.c_void_p), ("counts", C.c_void_p)]


class MFError(RuntimeError):
    pass


_lib = None

# (name, restype, argtypes): every symbol include/*.h declares
SIGNATURES = [
    ("mf_last_error", C.c_char_p, []),
    ("mf_abi_version", C.c_int, []),
    ("mf_context_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("mf_context_destroy", C.c_int, [C.c_void_p]),
    ("mf_set_types", C.c_int, [C.c_void_p, C.POINTER(ChType), C.c_int32]),
    ("mf_default_opts", C.c_int, [C.POINTER(Opts)]),
    ("mf_fit_batch", C.c_int, [C.c_void_p, C.POINTER(Batch), C.POINTER(Result),
                               C.POINTER(Opts), C.c_void_p]),
    ("mf_fit_batch_noalloc", C.c_int, [C.c_void_p, C.POINTER(Batch), C.POINTER(Result),
                                       C.POINTER(Opts), C.c_void_p]),
    ("mf_inplace_fnnls_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                         C.c_int32, C.c_double, C.c_uint32, C.c_void_p]),
    ("mf_fnnls_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                 C.c_int32, C.c_double, C.c_uint32, C.c_void_p]),
    ("mf_inplace_fnnls_chol_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                              C.c_int32, C.c_double, C.c_uint32, C.c_void_p]),
    ("mf_lh_nnls_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                   C.c_int32, C.c_double, C.c_int32, C.c_void_p, C.c_void_p]),
    ("mf_linalg_batch", C.c_int, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                  C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    ("mf_mlp_init", C.c_int, [C.c_void_p, C.c_uint64]),
    ("mf_mlp_set_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    ("mf_mlp_get_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    ("mf_mlp_grad", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                              C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mf_mlp_train_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                    C.c_int64, C.c_float, C.c_void_p, C.c_void_p]),
    ("mf_mlp_predict", C.c_int, [C.c_void_p, C.POINTER(Batch), C.c_void_p, C.c_void_p]),
    ("mf_set_timing", C.c_int, [C.c_void_p, C.c_int32]),
    ("mf_get_timings", C.c_int, [C.c_void_p, C.POINTER(Timings)]),
    ("mf_get_tune", C.c_int, [C.c_void_p, C.POINTER(Tune)]),
    ("mf_graph_capture", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_void_p)]),
    ("mf_graph_launch", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mf_graph_destroy", C.c_int, [C.c_void_p]),
    ("mf_set_tune", C.c_int, [C.c_void_p, C.POINTER(Tune)]),
    ("mf_autotune", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.POINTER(Tune), C.POINTER(C.c_float)]),
    ("mf_gen_type", C.c_int, [C.POINTER(GenSpec), C.POINTER(ChType)]),
    ("mf_gen_samples", C.c_int, [C.POINTER(GenSpec), C.c_int32, C.c_void_p, C.c_int64,
                                 C.c_uint64, C.c_void_p, C.c_void_p, C.c_int32]),
    ("mf_gen_mix", C.c_int, [C.c_int64, C.c_double, C.c_uint64, C.c_void_p]),
    ("mf_io_open", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("mf_io_close", C.c_int, [C.c_void_p]),
    ("mf_io_tree_entries", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64)]),
    ("mf_io_read_branch", C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int32, C.c_void_p,
                                    C.c_int64, C.POINTER(C.c_int64), C.c_void_p, C.c_int64,
                                    C.POINTER(C.c_int64)]),
    ("mf_io_read_graph", C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_int64,
                                   C.POINTER(C.c_int64)]),
    ("mf_io_write_tree", C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_void_p, C.c_int32,
                                   C.c_int64]),
]


def lib():
    """Loads libmf_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise MFError(f"{LIB_PATH} is not built: run `make -C {PKG_DIR}` "
                      "(or __graft_entry__.build())")
    L = C.CDLL(str(LIB_PATH))
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status: int, what: str = "") -> None:
    if status != MF_OK:
        msg = lib().mf_last_error().decode(errors="replace")
        raise MFError(f"{what or 'libmf_amd'} failed (status {status}): {msg}")


def default_opts(**kw) -> Opts:
    o = Opts()
    check(lib().mf_default_opts(C.byref(o)), "mf_default_opts")
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def env_flag(name: str, default: str = "") -> str:
    return os.environ.get(name, default)
