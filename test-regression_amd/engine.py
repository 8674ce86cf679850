"""Batched multifit on the GPU through libmf_amd.so.

torch is used only as device-memory and stream plumbing: tensors hold the
structure-of-arrays batch and results, the fit itself runs in the library's
gfx950 kernels (mf_fit_batch).  Nothing here computes a fit on the CPU.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass

import torch

from . import _abi


@dataclass
class FitResult:
    x: torch.Tensor        # [npulses_max=10, nch] amplitudes (pulse order of the type)
    chi2: torch.Tensor     # [nch]
    flags: torch.Tensor    # [nch] int32 MF_F_* bits
    iters: torch.Tensor    # [nch] int32: n_minimize | outer << 8 | inner << 16
    err: torch.Tensor | None = None  # [10, nch] in-time uncertainty (compute_errors)

    def n_minimize(self):
        return self.iters & 0xFF

    def n_outer(self):
        return (self.iters >> 8) & 0xFF

    def n_inner(self):
        return (self.iters >> 16) & 0xFFFF


def _stream_ptr(stream) -> int | None:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


class FitGraph:
    """A captured fit call (mf_graph_*): launch() repeats it on the same
    buffers; result holds the outputs.  The engine keeps track of its graphs:
    Engine.close() destroys them first, and Engine.set_types() is refused by
    the library while any of them is open (their kernels read the tables)."""

    def __init__(self, eng, handle, keep):
        self._eng, self._h, self._keep = eng, handle, keep
        self.result = keep[2]
        eng._graphs.add(self)

    def launch(self, stream=None):
        _abi.check(self._eng._lib.mf_graph_launch(self._h, C.c_void_p(_stream_ptr(stream))),
                   "mf_graph_launch")

    def close(self):
        if self._h:
            self._eng._lib.mf_graph_destroy(self._h)
            self._h = None
        self._eng._graphs.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One device context holding the channel-type table (mf_context)."""

    def __init__(self, types, device: int = 0):
        self._lib = _abi.lib()
        if not torch.cuda.is_available():
            raise _abi.MFError("no GPU visible: the multifit engine runs on the GPU only")
        self.device = device
        self._graphs = weakref.WeakSet()
        self._ctx = C.c_void_p()
        _abi.check(self._lib.mf_context_create(C.byref(self._ctx), device), "mf_context_create")
        self.set_types(types)

    def set_types(self, types):
        if isinstance(types, _abi.ChType):
            types = [types]
        arr = (_abi.ChType * len(types))(*types)
        _abi.check(self._lib.mf_set_types(self._ctx, arr, len(types)), "mf_set_types")
        self.types = list(arr)

    def close(self):
        for g in list(self._graphs):
            g.close()
        if self._ctx:
            self._lib.mf_context_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc_result(self, nch: int, precision: int = _abi.MF_PREC_FP64) -> FitResult:
        dt = torch.float64 if precision == _abi.MF_PREC_FP64 else torch.float32
        dev = torch.device("cuda", self.device)
        return FitResult(
            x=torch.zeros((10, nch), dtype=dt, device=dev),
            chi2=torch.zeros(nch, dtype=dt, device=dev),
            flags=torch.zeros(nch, dtype=torch.int32, device=dev),
            iters=torch.zeros(nch, dtype=torch.int32, device=dev),
            err=torch.zeros((10, nch), dtype=dt, device=dev),
        )

    def _batch_result(self, samples, type_of, opts, out, type_ranges, pederr=None):
        """ctypes mf_batch / mf_result for a call (and the objects to keep alive)."""
        if samples.dim() != 2 or not samples.is_cuda:
            raise ValueError("samples must be a 2-D device tensor [samples, nch]")
        samples = samples.contiguous()
        nch = samples.shape[1]
        need = max(t.nsamples for t in self.types)
        if samples.shape[0] < need:
            raise ValueError(f"samples has {samples.shape[0]} rows, the types need {need}")
        if type_of is not None and (type_of.dtype != torch.uint8 or type_of.numel() != nch
                                    or not type_of.is_cuda):
            raise ValueError("type_of must be a uint8 device tensor with one entry per channel")
        if pederr is not None and (pederr.dtype != torch.float64 or pederr.numel() != nch
                                   or not pederr.is_cuda or not pederr.is_contiguous()):
            raise ValueError("pederr must be a contiguous float64 device tensor with one entry "
                             "per channel")
        if samples.dtype == torch.float64:
            sdt = _abi.MF_DTYPE_F64
        elif samples.dtype == torch.float32:
            sdt = _abi.MF_DTYPE_F32
        else:
            raise ValueError("samples must be float64 or float32")
        if out is None:
            out = self.alloc_result(nch, opts.precision)
        rng = None
        if type_ranges is not None:
            # channels [type_ranges[t], type_ranges[t+1]) are of type t
            rng = (C.c_int64 * len(type_ranges))(*[int(v) for v in type_ranges])
        b = _abi.Batch(nch=nch, sample_dtype=sdt,
                       nranges=len(type_ranges) - 1 if type_ranges is not None else 0,
                       samples=samples.data_ptr(),
                       type_of=type_of.data_ptr() if type_of is not None else None,
                       ranges=C.cast(rng, C.c_void_p) if rng is not None else None,
                       pederr=pederr.data_ptr() if pederr is not None else None)
        r = _abi.Result(x=out.x.data_ptr(), chi2=out.chi2.data_ptr(),
                        flags=out.flags.data_ptr(), iters=out.iters.data_ptr(),
                        err=out.err.data_ptr() if out.err is not None else None)
        return b, r, (samples, type_of, out, rng, pederr)

    def fit(self, samples: torch.Tensor, type_of: torch.Tensor | None = None, opts=None,
            out: FitResult | None = None, stream=None, capture: bool = False,
            type_ranges=None, pederr: torch.Tensor | None = None):
        """samples: [>=nsamples, nch] float64/float32 on the device (row i = sample i);
        pederr: optional [nch] float64 device tensor, DoFit's pederr per channel."""
        if opts is None:
            opts = _abi.default_opts()
        b, r, keep = self._batch_result(samples, type_of, opts, out, type_ranges, pederr)
        if capture:
            g = C.c_void_p()
            _abi.check(self._lib.mf_graph_capture(self._ctx, C.byref(b), C.byref(r),
                                                  C.byref(opts), C.c_void_p(_stream_ptr(stream)),
                                                  C.byref(g)), "mf_graph_capture")
            return FitGraph(self, g, keep)
        _abi.check(self._lib.mf_fit_batch(self._ctx, C.byref(b), C.byref(r), C.byref(opts),
                                          C.c_void_p(_stream_ptr(stream))), "mf_fit_batch")
        return keep[2]

    def fit_noalloc(self, samples: torch.Tensor, type_of: torch.Tensor | None = None,
                    opts=None, out: FitResult | None = None, stream=None, type_ranges=None,
                    pederr: torch.Tensor | None = None):
        """mf_fit_batch_noalloc: the same fit with the library's fixed per-GPU
        scratch (no device allocation in the call; out must then be given by
        a caller that cannot allocate)."""
        if opts is None:
            opts = _abi.default_opts()
        b, r, keep = self._batch_result(samples, type_of, opts, out, type_ranges, pederr)
        _abi.check(self._lib.mf_fit_batch_noalloc(self._ctx, C.byref(b), C.byref(r),
                                                  C.byref(opts), C.c_void_p(_stream_ptr(stream))),
                   "mf_fit_batch_noalloc")
        return keep[2]

    def capture(self, samples: torch.Tensor, type_of: torch.Tensor | None = None, opts=None,
                out: FitResult | None = None, stream=None, type_ranges=None,
                pederr: torch.Tensor | None = None) -> "FitGraph":
        """mf_graph_capture: the fit of these buffers as a replayable graph
        (runs it once eagerly first).  The graph keeps the tensors alive."""
        return self.fit(samples, type_of, opts, out, stream, capture=True,
                        type_ranges=type_ranges, pederr=pederr)


    # ---- kernel timing (HIP events on the launch stream) ----
    def set_timing(self, enable: bool = True):
        _abi.check(self._lib.mf_set_timing(self._ctx, 1 if enable else 0), "mf_set_timing")

    def timings(self):
        t = _abi.Timings()
        _abi.check(self._lib.mf_get_timings(self._ctx, C.byref(t)), "mf_get_timings")
        return {"mlp_ms": t.mlp_ms, "fit_ms": t.fit_ms, "fit_launches": t.fit_launches}

    TUNE_FIELDS = ("block", "p1_wpc", "p2_wpc", "p1_refill", "p2_refill", "fit_block",
                   "use_persistent")

    def get_tune(self) -> dict:
        """Launch shapes (mf_tune)."""
        t = _abi.Tune()
        _abi.check(self._lib.mf_get_tune(self._ctx, C.byref(t)), "mf_get_tune")
        return {k: getattr(t, k) for k in self.TUNE_FIELDS}

    def autotune(self, samples, type_of=None, opts=None, out=None, stream=None,
                 type_ranges=None):
        """mf_autotune on this batch: (chosen tune, ms per call)."""
        if opts is None:
            opts = _abi.default_opts()
        b, r, keep = self._batch_result(samples, type_of, opts, out, type_ranges)
        t, ms = _abi.Tune(), C.c_float()
        _abi.check(self._lib.mf_autotune(self._ctx, C.byref(b), C.byref(r), C.byref(opts),
                                         C.c_void_p(_stream_ptr(stream)), C.byref(t),
                                         C.byref(ms)), "mf_autotune")
        return {k: getattr(t, k) for k in self.TUNE_FIELDS}, ms.value

    def set_tune(self, **kw):
        t = _abi.Tune(**{**self.get_tune(), **kw})
        _abi.check(self._lib.mf_set_tune(self._ctx, C.byref(t)), "mf_set_tune")

    # ---- learned warm start (mf_mlp_*) ----
    def mlp_init(self, seed: int = 0):
        _abi.check(self._lib.mf_mlp_init(self._ctx, seed), "mf_mlp_init")

    def mlp_set_params(self, params):
        import numpy as np
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != _abi.MF_MLP_NPARAM:
            raise ValueError(f"expected {_abi.MF_MLP_NPARAM} parameters")
        _abi.check(self._lib.mf_mlp_set_params(self._ctx, p.ctypes.data, p.size),
                   "mf_mlp_set_params")

    def mlp_get_params(self):
        import numpy as np
        p = np.empty(_abi.MF_MLP_NPARAM, dtype=np.float32)
        _abi.check(self._lib.mf_mlp_get_params(self._ctx, p.ctypes.data, p.size),
                   "mf_mlp_get_params")
        return p

    @staticmethod
    def _f32_2d(t, name):
        if not (t.is_cuda and t.dtype == torch.float32 and t.dim() == 2 and t.shape[0] == 10
                and t.stride(1) == 1):
            raise ValueError(f"{name} must be a float32 device tensor [10, B] with unit column stride")

    def mlp_grad(self, samples, xref, stream=None):
        """(gradient [6164], loss) of the mean loss over the batch's columns."""
        self._f32_2d(samples, "samples")
        self._f32_2d(xref, "xref")
        B = samples.shape[1]
        g = torch.empty(_abi.MF_MLP_NPARAM, dtype=torch.float32, device=samples.device)
        loss = torch.empty(1, dtype=torch.float32, device=samples.device)
        _abi.check(self._lib.mf_mlp_grad(self._ctx, samples.data_ptr(), samples.stride(0),
                                         xref.data_ptr(), xref.stride(0), B, g.data_ptr(),
                                         loss.data_ptr(), C.c_void_p(_stream_ptr(stream))),
                   "mf_mlp_grad")
        return g, loss

    def mlp_train_step(self, samples, xref, lr: float, loss=None, stream=None):
        self._f32_2d(samples, "samples")
        self._f32_2d(xref, "xref")
        _abi.check(self._lib.mf_mlp_train_step(
            self._ctx, samples.data_ptr(), samples.stride(0), xref.data_ptr(), xref.stride(0),
            samples.shape[1], lr, loss.data_ptr() if loss is not None else None,
            C.c_void_p(_stream_ptr(stream))), "mf_mlp_train_step")

    def mlp_predict(self, samples, stream=None):
        """[20, nch] float32: logits (passive if > 0), then amplitudes."""
        samples = samples.contiguous()
        nch = samples.shape[1]
        out = torch.empty((20, nch), dtype=torch.float32, device=samples.device)
        sdt = _abi.MF_DTYPE_F64 if samples.dtype == torch.float64 else _abi.MF_DTYPE_F32
        b = _abi.Batch(nch=nch, sample_dtype=sdt, nranges=0, samples=samples.data_ptr(),
                       type_of=None, ranges=None)
        _abi.check(self._lib.mf_mlp_predict(self._ctx, C.byref(b), out.data_ptr(),
                                            C.c_void_p(_stream_ptr(stream))), "mf_mlp_predict")
        return out


def nnls_batch(A: torch.Tensor, b: torch.Tensor, x: torch.Tensor, m: int = 10,
               eps: float = 1e-11, max_iter: int = 1000, solver: str = "inplace_fnnls",
               stream=None) -> torch.Tensor:
    """Batched inplace_fnnls / inplace_fnnls_test_cholesky ("inplace_fnnls_chol")
    / fnnls.  A [n,10,10] column-major per problem (A[c, col, row]), b [n,10],
    x [n,10] in/out (warm start)."""
    L = _abi.lib()
    for t in (A, b, x):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError("A, b, x must be contiguous float64 device tensors")
    f = {"inplace_fnnls": L.mf_inplace_fnnls_batch, "fnnls": L.mf_fnnls_batch,
         "inplace_fnnls_chol": L.mf_inplace_fnnls_chol_batch}[solver]
    _abi.check(f(A.data_ptr(), b.data_ptr(), x.data_ptr(), A.shape[0], m, eps, max_iter,
                 C.c_void_p(_stream_ptr(stream))), solver)
    return x


def lh_nnls_batch(A: torch.Tensor, b: torch.Tensor, x: torch.Tensor, m: int = 10,
                  eps: float = 1e-10, max_iter: int = -1, stream=None):
    """Batched Lawson-Hanson NNLS (Eigen::NNLS semantics, mf_lh_nnls_batch);
    layout as nnls_batch, x in/out (the setX starting point; eps default as
    Eigen::NNLS's constructor).  Returns (x, solved) with solved [n] int32,
    1 where the solve converged before max_iter least-squares solves."""
    L = _abi.lib()
    for t in (A, b, x):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError("A, b, x must be contiguous float64 device tensors")
    solved = torch.zeros(A.shape[0], dtype=torch.int32, device=A.device)
    _abi.check(L.mf_lh_nnls_batch(A.data_ptr(), b.data_ptr(), x.data_ptr(), A.shape[0], m, eps,
                                  max_iter, C.c_void_p(solved.data_ptr()),
                                  C.c_void_p(_stream_ptr(stream))), "mf_lh_nnls_batch")
    return x, solved


def linalg_batch(op: int, M=None, L=None, b=None, y=None, full: int = 10, view: int = 10,
                 i: int = 0, j: int = 0, stream=None):
    """Batched linalgebra primitive (see mf_linalg_batch)."""
    ts = [t for t in (M, L, b, y) if t is not None]
    dt = ts[0].dtype
    if dt not in (torch.float64, torch.float32) or any(t.dtype != dt for t in ts):
        raise ValueError("operands must share one dtype, float64 or float32")
    n = (M if M is not None else L).shape[0]
    ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
    _abi.check(_abi.lib().mf_linalg_batch(
        op, _abi.MF_DTYPE_F64 if dt == torch.float64 else _abi.MF_DTYPE_F32,
        ptr(M), ptr(L), ptr(b), ptr(y), n, full, view, i, j,
        C.c_void_p(_stream_ptr(stream))), "mf_linalg_batch")
