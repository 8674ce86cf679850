"""MI355X multifit engine: Python view of libmf_amd.so.

The directory name is not a Python identifier; callers load it under the
module name ``test_regression_amd`` (see ``load()`` in bench.py / tests).

    _abi       ctypes declarations of include/mf_abi.h and include/mf_gen.h
    engine     device batches through mf_fit_batch / mf_*_fnnls_batch / mf_linalg_batch
    workloads  synthetic channel populations of the benchmark configurations
"""
from . import _abi  # noqa: F401
from ._abi import MFError, default_opts  # noqa: F401
