#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ from the reference's own code.

Run in a container where the reference tree is present and `make -C oracle
ref` has built oracle/_ref/ from it:

  * linalgebra_golden.npz -- linalgebra/interface/{cholesky,substitution,
    common,fused_cholesky_substitution}.hpp instantiated for double and float
    (oracle/_ref/libref_linalg.so) on the 3x3 literal of
    linalgebra/tests/test_cholesky.cpp:25-27 and b of
    test_cholesky_with_solvers.cpp:33, on seeded 10x10 SPD matrices (as
    generate_and_test.cpp:31-66 builds them, AtA of a random A), and on the
    9x9 -> 10x10 append and row-6 removal of
    generate_and_test_fused_impl.cpp:66-119.
  * fnnls_cl_golden.npz -- inplace_fnnls of
    multifit_ocl/device/regression_inplace_fnnls.cl:354-587 (incremental
    Cholesky FNNLS), as shipped (float) and as a double instance
    (oracle/_ref/libref_fnnls_cl{32,64}.so), on seeded problems: random A, b
    (the harness's own kind, multifit_ocl/tests/inplace_fnnls_original_host.cpp:181-215)
    and pulse-like A (shifted template columns whitened by a correlated
    noise factor, b = A a + noise).

The fixtures are data only (inputs and the reference's outputs).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = HERE.parents[1] / "oracle" / "_ref"


def _lib(name):
    p = REF / f"lib{name}.so"
    if not p.exists():
        raise SystemExit(f"{p} missing: run `make -C oracle ref` where the reference tree exists")
    return C.CDLL(str(p))


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def linalg(out):
    L = _lib("ref_linalg")
    data = {}
    rng = np.random.default_rng(20180101)
    for dt, sfx, ct in ((np.float64, "f64", C.c_double), (np.float32, "f32", C.c_float)):
        def call(fn, *arrs_and_ints):
            args = []
            for a in arrs_and_ints:
                args.append(_ptr(a, ct) if isinstance(a, np.ndarray) else C.c_int(a))
            getattr(L, f"ref_{fn}_{sfx}")(*args)

        # 3x3 literal (test_cholesky.cpp:25-27, test_cholesky_with_solvers.cpp:33)
        M3 = np.array([[4, -1, 2], [-1, 6, 0], [2, 0, 5]], dtype=dt)
        b3 = np.array([3, 3, 4], dtype=dt)
        L3 = np.zeros((3, 3), dtype=dt)
        call("cholesky_decomp", M3, L3, 3, 3)
        y3 = np.zeros(3, dtype=dt)
        x3 = np.zeros(3, dtype=dt)
        call("forward", L3, b3, y3, 3, 3)
        call("backward", L3, y3, x3, 3, 3)
        data[f"lit3_M_{sfx}"], data[f"lit3_b_{sfx}"] = M3, b3
        data[f"lit3_L_{sfx}"], data[f"lit3_y_{sfx}"], data[f"lit3_x_{sfx}"] = L3, y3, x3

        # seeded 10x10 AtA, full and view 7 of full 10
        n = 32
        Ms = np.zeros((n, 10, 10), dtype=dt)
        bs = rng.standard_normal((n, 10)).astype(dt)
        Ls = np.zeros((n, 10, 10), dtype=dt)
        ys = np.zeros((n, 10), dtype=dt)
        xs = np.zeros((n, 10), dtype=dt)
        views = np.array([10 if i % 2 == 0 else 7 for i in range(n)], dtype=np.int32)
        for i in range(n):
            A = rng.uniform(-1, 1, (10, 10))
            Ms[i] = (A.T @ A + 0.1 * np.eye(10)).astype(dt)
            call("cholesky_decomp", Ms[i], Ls[i], 10, int(views[i]))
            call("forward", Ls[i], bs[i], ys[i], 10, int(views[i]))
            call("backward", Ls[i], ys[i], xs[i], 10, int(views[i]))
        data[f"spd_M_{sfx}"], data[f"spd_b_{sfx}"], data[f"spd_view"] = Ms, bs, views
        data[f"spd_L_{sfx}"], data[f"spd_y_{sfx}"], data[f"spd_x_{sfx}"] = Ls, ys, xs

        # swap rows/cols (i, j) with j > i inside view 10
        Sw = Ms.copy()
        pairs = np.array([(i % 5, 5 + i % 5) for i in range(n)], dtype=np.int32)
        for i in range(n):
            call("swap_row_column", Sw[i], int(pairs[i, 0]), int(pairs[i, 1]), 10, 10)
        data[f"swap_pairs"], data[f"swap_M_{sfx}"] = pairs, Sw

        # fused: factor 9x9, append row 9 (rcaddition), then remove row 6:
        # swap 6 <-> 8 inside the 9-view and refactor from position 6
        Lf = np.zeros((n, 10, 10), dtype=dt)
        yf = np.zeros((n, 10), dtype=dt)
        Lr = np.zeros((n, 10, 10), dtype=dt)
        yr = np.zeros((n, 10), dtype=dt)
        Mr = Ms.copy()
        br = bs.copy()
        for i in range(n):
            call("cholesky_decomp", Ms[i], Lf[i], 10, 9)
            call("forward", Lf[i], bs[i], yf[i], 10, 9)
            call("rcaddition", Ms[i], Lf[i], bs[i], yf[i], 10, 10)
            Lr[i] = Lf[i]
            yr[i] = yf[i]
            call("swap_row_column", Mr[i], 6, 9, 10, 10)
            br[i, [6, 9]] = br[i, [9, 6]]
            call("inbetween_removal", Mr[i], Lr[i], br[i], yr[i], 6, 10, 9)
        data[f"fused_L_{sfx}"], data[f"fused_y_{sfx}"] = Lf, yf
        data[f"removal_M_{sfx}"], data[f"removal_b_{sfx}"] = Mr, br
        data[f"removal_L_{sfx}"], data[f"removal_y_{sfx}"] = Lr, yr
    np.savez_compressed(out, **data)


def _pulse_matrix(shift_rng, n=10):
    """Columns: the alpha-beta template (alpha 1.138, beta 41.375 ns) shifted
    by 25 ns per column, whitened by the Cholesky factor of 0.5^|i-j|."""
    a, b = 1.138, 1.655 * 25
    t = 25.0 * (np.arange(n)[:, None] - np.arange(n)[None, :]) + shift_rng
    with np.errstate(invalid="ignore"):
        P = np.where(t > -a * b, np.power(np.maximum(1 + t / (a * b), 0), a) * np.exp(-t / b), 0.0)
    R = 0.5 ** np.abs(np.arange(n)[:, None] - np.arange(n)[None, :])
    return np.linalg.solve(np.linalg.cholesky(R), P)


def fnnls_cl(out):
    L32, L64 = _lib("ref_fnnls_cl32"), _lib("ref_fnnls_cl64")
    rng = np.random.default_rng(20180202)
    n = 160
    A = np.zeros((n, 10, 10))
    b = np.zeros((n, 10))
    kind = np.zeros(n, dtype=np.int32)
    for i in range(n):
        if i < n // 2:
            A[i] = rng.uniform(0, 1, (10, 10))
            b[i] = rng.uniform(0, 1, 10)
        else:
            kind[i] = 1
            A[i] = _pulse_matrix(rng.normal(0, 2))
            amp = rng.exponential(8.0, 10) * (rng.uniform(0, 1, 10) < 0.6)
            b[i] = A[i] @ amp + rng.standard_normal(10)
    x32 = np.zeros((n, 10), dtype=np.float32)
    x64 = np.zeros((n, 10))
    for fn, x, dt, ct in ((L32.ref_inplace_fnnls_cl_f32, x32, np.float32, C.c_float),
                          (L64.ref_inplace_fnnls_cl_f64, x64, np.float64, C.c_double)):
        fn.argtypes = [C.POINTER(ct), C.POINTER(ct), C.POINTER(ct), C.c_double, C.c_uint]
        for i in range(n):
            Ai = np.ascontiguousarray(A[i], dtype=dt)   # row-major, as the .cl reads it
            bi = np.ascontiguousarray(b[i], dtype=dt)
            xi = np.zeros(10, dtype=dt)                  # the facade zeroes x (:603-604)
            fn(_ptr(Ai, ct), _ptr(bi, ct), _ptr(xi, ct), 1e-11, 1000)
            x[i] = xi
    np.savez_compressed(out, A=A, b=b, kind=kind, x_f32=x32, x_f64=x64, eps=1e-11,
                        max_iter=1000)


if __name__ == "__main__":
    linalg(HERE / "linalgebra_golden.npz")
    fnnls_cl(HERE / "fnnls_cl_golden.npz")
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)
