"""Multi-RHS product Y = K V (no reference counterpart) against the CPU oracle,
column by column on the same seeded inputs.

Bars (normwise relative error per column, stated per precision):
  FP64            1e-12, and bitwise equal to the single-vector FP64 product
                  only up to summation order, so 1e-13 against it
  FP32 (VALU/MFMA) 5e-6   (measured ~3.5e-7 at N = 2048)
  BF16x3          5e-5   (measured ~4.4e-6)
  BF16            1e-2   (measured ~2.2e-3)
"""
import ctypes as C
import math

import numpy as np
import pytest

from mpla_amd import _native
from tests.gpu_util import dev, empty, host, point_table, require_gpu, stream, workspace

pytestmark = pytest.mark.gpu

TOL = {0: 1e-12, 1: 5e-6, 2: 1e-2, 3: 5e-5}


@pytest.fixture(scope="module")
def lib():
    require_gpu()
    return _native.mplak()


def desc_for(T, dim, n, kind, ell, nugget, row_off=0, rows=None, col_off=0, cols=None):
    d = _native.mplak_kernel_desc()
    d.points, d.dim, d.kind, d.length_scale, d.nugget, d.coincident_shift = T.data_ptr(), dim, kind, ell, nugget, 0.0
    d.row_offset, d.row_count = row_off, n if rows is None else rows
    d.col_offset, d.col_count = col_off, n if cols is None else cols
    return d


def mrhs(lib, d, prec, V):
    cols, R = V.shape
    Vd = dev(np.asfortranarray(V).ravel(order="F"))
    Y = empty(d.row_count * R)
    wb = lib.mplak_kernel_mrhs_workspace_bytes(C.byref(d), prec, R)
    w = workspace(wb)
    _native.check(lib.mplak_kernel_mrhs(C.byref(d), prec, R, Vd.data_ptr(), cols, Y.data_ptr(), d.row_count,
                                        w.data_ptr(), wb, None, stream()))
    return host(Y).reshape((d.row_count, R), order="F")


def case(oracle, n, dim, R, seed):
    pts = oracle.uniform(dim * n, seed, 0).reshape(dim, n)
    V = oracle.normal(n * R, seed, 1).reshape((n, R), order="F")
    return pts, V


def col_errors(Y, ref):
    return [np.linalg.norm(Y[:, r] - ref[:, r]) / np.linalg.norm(ref[:, r]) for r in range(Y.shape[1])]


@pytest.mark.parametrize("prec,R", [(0, 1), (0, 5), (0, 16), (0, 20), (1, 3), (1, 8), (1, 16), (1, 64), (1, 65), (2, 32), (2, 64),
                                    (3, 40), (3, 64), (3, 70)])
@pytest.mark.parametrize("kind,ell", [(0, 0.05), (1, 0.1)])
def test_mrhs_vs_oracle(lib, oracle, prec, R, kind, ell):
    n = 1500
    pts, V = case(oracle, n, 3, R, 1235 + R)
    P = dev(pts)
    T = point_table(P)
    Y = mrhs(lib, desc_for(T, 3, n, kind, ell, 0.5), prec, V)
    k = oracle.Kernel(pts, kind, ell, nugget=0.5)
    cols = range(R) if R <= 8 else [0, R // 3, R - 1]
    ref = np.stack([k.matvec(V[:, r], mode=1) for r in cols], axis=1)
    errs = col_errors(Y[:, list(cols)], ref)
    assert max(errs) <= TOL[prec], errs


@pytest.mark.parametrize("dim", [1, 2, 4])
@pytest.mark.parametrize("prec,R", [(0, 3), (1, 16), (2, 32), (3, 20)])
def test_mrhs_dims(lib, oracle, dim, prec, R):
    """Every dimension the MFMA kernels specialise (and FP64) against the oracle."""
    n = 1100
    pts, V = case(oracle, n, dim, R, 300 + dim + R)
    P = dev(pts)
    T = point_table(P)
    ell = 0.05 if dim > 1 else 0.005
    Y = mrhs(lib, desc_for(T, dim, n, 0, ell, 0.5), prec, V)
    k = oracle.Kernel(pts, 0, ell, nugget=0.5)
    cols = [0, R // 2, R - 1]
    ref = np.stack([k.matvec(V[:, r], mode=1) for r in cols], axis=1)
    assert max(col_errors(Y[:, cols], ref)) <= TOL[prec]


def test_mrhs_f64_matches_single_vector_product(lib, oracle):
    n, R = 2000, 6
    pts, V = case(oracle, n, 3, R, 7)
    P = dev(pts)
    T = point_table(P)
    d = desc_for(T, 3, n, 1, 0.1, 0.5)
    Y = mrhs(lib, d, 0, V)
    wb = lib.mplak_kernel_matvec_workspace_bytes(C.byref(d))
    w = workspace(wb)
    for r in range(R):
        x = dev(V[:, r].copy())
        y = empty(n)
        _native.check(lib.mplak_kernel_matvec(C.byref(d), 0, x.data_ptr(), y.data_ptr(), w.data_ptr(), wb, None,
                                              stream()))
        y = host(y)
        assert np.linalg.norm(Y[:, r] - y) <= 1e-13 * np.linalg.norm(y)


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_mrhs_offsets_and_ragged(lib, oracle, prec):
    n, R = 3001, 33
    pts, V = case(oracle, n, 3, R, 11)
    P = dev(pts)
    T = point_table(P)
    d = desc_for(T, 3, n, 0, 0.05, 1.0, row_off=200, rows=1777, col_off=100, cols=2901)
    Vs = np.asfortranarray(V[100:3001])
    Y = mrhs(lib, d, prec, Vs)
    k = oracle.Kernel(pts, 0, 0.05, nugget=1.0)
    for r in (0, 17, 32):
        ref = k.matvec(Vs[:, r], mode=1, row_off=200, rows=1777, col_off=100, cols=2901)
        assert np.linalg.norm(Y[:, r] - ref) / np.linalg.norm(ref) <= TOL[prec]


@pytest.mark.parametrize("prec", [0, 2, 3])
def test_mrhs_deterministic(lib, oracle, prec):
    n, R = 4096, 64
    pts, V = case(oracle, n, 3, R, 5)
    P = dev(pts)
    T = point_table(P)
    d = desc_for(T, 3, n, 1, 0.1, 0.5)
    assert np.array_equal(mrhs(lib, d, prec, V), mrhs(lib, d, prec, V))


def test_mrhs_full_size_c3_linearity(lib, oracle):
    """Matern-3/2 operator at full size (N = 131072, Matern-3/2 l = 0.1, 64 RHS) in
    BF16x3: linearity in V and spot rows against the oracle."""
    n, R = 131072, 64
    pts = oracle.uniform(3 * n, 1235, 0).reshape(3, n)
    P = dev(pts)
    T = point_table(P)
    d = desc_for(T, 3, n, 1, 0.1, 0.5)
    V = oracle.normal(n * R, 1235, 1).reshape((n, R), order="F")
    Y = mrhs(lib, d, 3, V)
    Y2 = mrhs(lib, d, 3, 2.0 * V)  # scaling by 2 is exact in BF16
    assert np.array_equal(2.0 * Y, Y2)
    k = oracle.Kernel(pts, 1, 0.1, nugget=0.5)
    for i in (0, 77777, n - 1):
        for r in (0, 63):
            ref = k.matvec(V[:, r], mode=1, row_off=i, rows=1)[0]
            bound = k.matvec(np.abs(V[:, r]), mode=1, row_off=i, rows=1)[0]
            assert abs(Y[i, r] - ref) <= 5e-5 * bound
