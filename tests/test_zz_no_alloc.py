"""The allocation-free kernel-matrix product (mplak_kernel_matvec_noalloc,
no reference counterpart): for callers that cannot allocate device memory.

Bar: bitwise equal to mplak_kernel_matvec (the same kernels on the library's
fixed device scratch instead of a caller workspace), on two calls in a row,
the second with a smaller input, so stale scratch contents from the first
call cannot leak into the second.  A product whose workspace exceeds the
fixed scratch is refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from mpla_amd import _native
from tests.gpu_util import dev, empty, host, point_table, require_gpu, stream, workspace

pytestmark = pytest.mark.gpu


def _desc(table, dim, n, kind, ell, nugget, shift=0.0):
    d = _native.mplak_kernel_desc()
    d.points, d.dim, d.kind = table.data_ptr(), dim, kind
    d.length_scale, d.nugget, d.coincident_shift = ell, nugget, shift
    d.row_offset, d.row_count, d.col_offset, d.col_count = 0, n, 0, n
    return d


def _both(lib, oracle, n, dim, kind, ell, nugget, precision, seed):
    pts = oracle.uniform(dim * n, seed, 0).reshape(dim, n)
    xh = oracle.uniform(n, seed, 1, -1.0, 1.0)
    P = dev(pts)
    table = point_table(P)
    d = _desc(table, dim, n, kind, ell, nugget)
    x = dev(xh)
    y_ref, y_na = empty(n), empty(n)
    wb = lib.mplak_kernel_matvec_workspace_bytes(C.byref(d))
    w = workspace(wb)
    _native.check(lib.mplak_kernel_matvec(C.byref(d), precision, x.data_ptr(), y_ref.data_ptr(), w.data_ptr(), wb,
                                          None, stream()))
    _native.check(lib.mplak_kernel_matvec_noalloc(C.byref(d), precision, x.data_ptr(), y_na.data_ptr(), None,
                                                  stream()))
    return host(y_ref), host(y_na), pts, xh


@pytest.mark.parametrize("precision", [_native.MPLAK_PREC_F64, _native.MPLAK_PREC_F32])
def test_noalloc_matches_workspace_product_twice(oracle, precision):
    require_gpu()
    lib = _native.mplak()
    print(f"noalloc: precision {precision}, n 20000 (symmetric) then n 3000 (plain)", flush=True)
    # first call: 20000 points, the symmetric tile-pair product (largest use of the scratch)
    a_ref, a_na, pts, xh = _both(lib, oracle, 20000, 3, _native.MPLAK_KERNEL_GAUSSIAN, 0.05, 1.0, precision, 71)
    assert np.array_equal(a_ref, a_na)
    if precision == _native.MPLAK_PREC_F64:
        ref = oracle.Kernel(pts, oracle.GAUSSIAN, 0.05, nugget=1.0).matvec(xh)
        assert np.linalg.norm(a_na - ref) <= 1e-12 * np.linalg.norm(ref)
    # second call, smaller: 3000 points in 2-D, Matern, the plain slab product
    b_ref, b_na, pts, xh = _both(lib, oracle, 3000, 2, _native.MPLAK_KERNEL_MATERN32, 0.1, 0.5, precision, 72)
    assert np.array_equal(b_ref, b_na)
    if precision == _native.MPLAK_PREC_F64:
        ref = oracle.Kernel(pts, oracle.MATERN32, 0.1, nugget=0.5).matvec(xh)
        assert np.linalg.norm(b_na - ref) <= 1e-12 * np.linalg.norm(ref)


def test_noalloc_refuses_beyond_its_scratch(oracle):
    require_gpu()
    lib = _native.mplak()
    cap = lib.mplak_kernel_matvec_noalloc_max_bytes()
    assert cap == 80 << 20
    # n = 65536 fits (the symmetric product's 128 x 129 / 2 tile pairs of partials: 67.6 MB)
    d = _native.mplak_kernel_desc()
    d.points, d.dim, d.kind, d.length_scale, d.nugget = 1, 3, _native.MPLAK_KERNEL_GAUSSIAN, 0.05, 1.0
    d.row_offset = d.col_offset = 0
    d.row_count = d.col_count = 65536
    assert lib.mplak_kernel_matvec_workspace_bytes(C.byref(d)) <= cap
    n = 65536
    while lib.mplak_kernel_matvec_workspace_bytes(C.byref(d)) <= cap:
        n += 512
        d.row_count = d.col_count = n
    # real inputs of that size, so that nothing out of bounds could be read
    # even if the size check failed; the call must be refused before any launch
    P = dev(oracle.uniform(3 * n, 73, 0).reshape(3, n))
    table = point_table(P)
    x, y = dev(oracle.uniform(n, 73, 1)), empty(n)
    d = _desc(table, 3, n, _native.MPLAK_KERNEL_GAUSSIAN, 0.05, 1.0)
    st = lib.mplak_kernel_matvec_noalloc(C.byref(d), _native.MPLAK_PREC_F64, x.data_ptr(), y.data_ptr(), None,
                                         stream())
    assert st == _native.MPLAK_ERR_TOO_LARGE
