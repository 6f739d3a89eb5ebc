"""Parity of the gfx950 kernel library (C ABI of include/mplak.h) against the
CPU oracle on identical seeded inputs.

Bars: bit-exact where the arithmetic is exact (integer closed forms of
src/test_library.cpp, generated inputs); FP64 products within 1e-12 relative
(the bar of src/test_library.cpp:135-151) normwise and component-wise against
sum_j |A_ij x_j|; FP32 kernel entries within the stated 2e-5 normwise.
"""
import ctypes as C
import math

import numpy as np
import pytest

from mpla_amd import _native
from tests.gpu_util import dev, empty, host, point_table, require_gpu, stream, workspace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    require_gpu()
    return _native.mplak()


def ck(status):
    _native.check(status)


# --------------------------------------------------------------------------
# inputs


@pytest.mark.parametrize("n,lo,hi", [(1, 0.0, 1.0), (1000, -1.0, 1.0), (123457, 0.0, 1.0)])
def test_fill_uniform_bitwise_equals_host(lib, oracle, n, lo, hi):
    out = empty(n)
    ck(lib.mplak_fill_uniform_f64(out.data_ptr(), n, 1234, 3, 17, lo, hi, stream()))
    ref = oracle.uniform(n, 1234, 3, lo, hi, first=17)
    assert np.array_equal(host(out), ref)


# --------------------------------------------------------------------------
# dgemv (src/mpla.cpp:729, src/test_library.cpp)


def dgemv_gpu(lib, A_dev, lda, m, n, x_dev, skip=None):
    y = empty(m)
    wb = lib.mplak_dgemv_workspace_bytes(m, n)
    w = workspace(wb)
    ck(lib.mplak_dgemv_f64(m, n, A_dev.data_ptr(), lda, x_dev.data_ptr(), y.data_ptr(), w.data_ptr(),
                           wb, skip, stream()))
    return y


# (6000, 6000) and (6001, 6000) hold more than 256 MiB: the kernel then reads A
# with nontemporal loads, 16 bytes per lane for the even leading dimension and
# 8 bytes for the odd one.  Every partial sum is an integer below 2^53.
@pytest.mark.parametrize("m,n", [(1, 1), (3, 2), (127, 129), (128, 64), (1000, 1000), (4096, 300),
                                 (333, 5000), (6000, 6000), (6001, 6000)])
def test_dgemv_test_library_pattern_exact(lib, m, n):
    A = np.arange(m * n, dtype=np.float64)  # column-major A[idx] = idx
    x = np.arange(n, dtype=np.float64)
    y = host(dgemv_gpu(lib, dev(A), m, m, n, dev(x)))
    i = np.arange(m, dtype=np.float64)
    expect = m * (n - 1) * n * (2 * n - 1) / 6 + i * n * (n - 1) / 2
    assert np.array_equal(y, expect)


# Above 256 MiB (nontemporal loads): square, odd leading dimension, a wide
# block (one row panel, 256 column slabs) and a tall one (one slab, 8-byte
# loads).
@pytest.mark.parametrize("m,n,lda", [(513, 700, 513), (512, 700, 515), (4096, 4096, 4096),
                                     (1, 9000, 1), (9000, 1, 9000), (6000, 6000, 6000),
                                     (6000, 6000, 6001), (100, 700000, 100), (300000, 150, 300001)])
def test_dgemv_random_vs_oracle(lib, oracle, m, n, lda):
    A = oracle.uniform(lda * n, 1234, 1, -1.0, 1.0)
    x = oracle.uniform(n, 1234, 2, -1.0, 1.0)
    y = host(dgemv_gpu(lib, dev(A), lda, m, n, dev(x)))
    Af = A.reshape((lda, n), order="F")[:m]
    ref = oracle.dgemv(np.asfortranarray(Af), x, mode=1)
    bound = np.abs(Af) @ np.abs(x)
    assert np.all(np.abs(y - ref) <= 1e-12 * bound)
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


def test_dgemv_deterministic_and_skip(lib, oracle):
    m = n = 2048
    A = dev(oracle.uniform(m * n, 9, 0, -1, 1))
    x = dev(oracle.uniform(n, 9, 1, -1, 1))
    y1 = host(dgemv_gpu(lib, A, m, m, n, x))
    y2 = host(dgemv_gpu(lib, A, m, m, n, x))
    assert np.array_equal(y1, y2)
    flag = dev(np.array([1], dtype=np.int32), dtype=__import__("torch").int32)
    y = empty(m)
    y.fill_(7.0)
    wb = lib.mplak_dgemv_workspace_bytes(m, n)
    w = workspace(wb)
    ck(lib.mplak_dgemv_f64(m, n, A.data_ptr(), m, x.data_ptr(), y.data_ptr(), w.data_ptr(), wb,
                           flag.data_ptr(), stream()))
    assert np.all(host(y) == 7.0)


# --------------------------------------------------------------------------
# level 1 (src/mpla.cpp:753-815)


@pytest.mark.parametrize("n", [1, 255, 10000, 1 << 20])
def test_ddot_closed_form_exact(lib, n):
    x = np.arange(n, dtype=np.float64)
    out = empty(1)
    wb = lib.mplak_dot_workspace_bytes(n)
    w = workspace(wb)
    xd, yd = dev(x), dev(x + 1)  # keep both alive: temporaries would share one block
    ck(lib.mplak_ddot_f64(n, xd.data_ptr(), yd.data_ptr(), out.data_ptr(), w.data_ptr(), wb, stream()))
    assert host(out)[0] == (n - 1) * n * (n + 1) / 3


def test_ddot_random_vs_oracle(lib, oracle):
    n = 300001
    x = oracle.uniform(n, 5, 0, -1, 1)
    y = oracle.uniform(n, 5, 1, -1, 1)
    out = empty(1)
    wb = lib.mplak_dot_workspace_bytes(n)
    w = workspace(wb)
    xd, yd = dev(x), dev(y)
    ck(lib.mplak_ddot_f64(n, xd.data_ptr(), yd.data_ptr(), out.data_ptr(), w.data_ptr(), wb, stream()))
    ref = oracle.ddot(x, y, 1)
    assert abs(host(out)[0] - ref) <= 1e-12 * np.dot(np.abs(x), np.abs(y))


def test_axpy_waxpby_fill_copy(lib, oracle):
    n = 100003
    x = np.arange(n, dtype=np.float64)
    y = 3.0 * np.ones(n)
    xd, yd, y0 = dev(x), dev(y), dev(y)
    ck(lib.mplak_daxpy_f64(n, -2.0, xd.data_ptr(), yd.data_ptr(), stream()))
    assert np.array_equal(host(yd), 3.0 - 2.0 * x)
    w = empty(n)
    ck(lib.mplak_dwaxpby_f64(n, 0.5, xd.data_ptr(), 4.0, y0.data_ptr(), w.data_ptr(), stream()))
    assert np.array_equal(host(w), 0.5 * x + 12.0)
    ck(lib.mplak_fill_f64(w.data_ptr(), n, -1.25, stream()))
    assert np.all(host(w) == -1.25)
    ck(lib.mplak_copy_f64(w.data_ptr(), xd.data_ptr(), n, stream()))
    assert np.array_equal(host(w), x)


# --------------------------------------------------------------------------
# matrix-free kernel matvec (src/kernel.cu:17-171)


def kernel_desc(P_dev, table, kind, l, nugget=0.0, shift=0.0, row_off=0, rows=None, col_off=0,
                cols=None):
    dim, n = P_dev.shape
    d = _native.mplak_kernel_desc()
    d.points = table.data_ptr()
    d.dim = dim
    d.kind = kind
    d.length_scale = l
    d.nugget = nugget
    d.coincident_shift = shift
    d.row_offset = row_off
    d.row_count = n if rows is None else rows
    d.col_offset = col_off
    d.col_count = n if cols is None else cols
    return d


def kmv_gpu(lib, desc, x_dev, precision=0):
    y = empty(desc.row_count)
    wb = lib.mplak_kernel_matvec_workspace_bytes(C.byref(desc))
    w = workspace(wb)
    ck(lib.mplak_kernel_matvec(C.byref(desc), precision, x_dev.data_ptr(), y.data_ptr(), w.data_ptr(),
                               wb, None, stream()))
    return y


def kmv_case(oracle, n, dim, kind, l, nugget, shift, seed):
    pts = oracle.uniform(dim * n, seed, 0).reshape(dim, n)
    x = oracle.uniform(n, seed, 1, -1, 1)
    return pts, x


def rel_err(y, ref):
    return np.linalg.norm(y - ref) / np.linalg.norm(ref)


@pytest.mark.parametrize("kind,l", [(0, 1 / math.sqrt(2.0)), (0, 0.05), (1, 0.1), (1, math.sqrt(3.0))])
@pytest.mark.parametrize("dim", [1, 2, 3, 4, 6])
def test_kernel_matvec_f64_vs_oracle(lib, oracle, kind, l, dim):
    n = 1537  # ragged: 3 target tiles, partial last chunk
    pts, x = kmv_case(oracle, n, dim, kind, l, 0.0, 0.0, 1235 + dim)
    P = dev(pts)
    T = point_table(P)
    y = host(kmv_gpu(lib, kernel_desc(P, T, kind, l, nugget=0.5), dev(x)))
    k = oracle.Kernel(pts, kind, l, nugget=0.5)
    ref = k.matvec(x, mode=1)
    assert rel_err(y, ref) <= 1e-12
    # component-wise: entries are positive, so sum_j |K_ij x_j| = (K|x|)_i
    bound = k.matvec(np.abs(x), mode=1)
    assert np.all(np.abs(y - ref) <= 1e-12 * bound)


def test_kernel_matvec_reference_plugin_semantics(lib, oracle):
    # src/kernel.cu: exp(-r^2) (l = 1/sqrt(2)), +1e-8 where r < 1e-14, incl.
    # duplicate points; no nugget.
    n = 700
    pts = oracle.uniform(3 * n, 77, 0).reshape(3, n)
    pts[:, 10] = pts[:, 20]  # duplicate point
    x = oracle.uniform(n, 77, 1, -1, 1)
    P = dev(pts)
    T = point_table(P)
    y = host(kmv_gpu(lib, kernel_desc(P, T, 0, 1 / math.sqrt(2.0), shift=1e-8), dev(x)))
    ref = oracle.Kernel(pts, 0, 1 / math.sqrt(2.0), coincident_shift=1e-8).matvec(x, mode=1)
    assert rel_err(y, ref) <= 1e-12
    # the shift is visible: removing it changes rows 10 and 20 by 1e-8 * (x_i + x_other)
    ref0 = oracle.Kernel(pts, 0, 1 / math.sqrt(2.0)).matvec(x, mode=1)
    assert abs((ref - ref0)[10] - 1e-8 * (x[10] + x[20])) < 1e-15


def test_kernel_matvec_offsets(lib, oracle):
    n = 2000
    pts, x = kmv_case(oracle, n, 3, 1, 0.1, 0, 0, 4)
    P = dev(pts)
    T = point_table(P)
    desc = kernel_desc(P, T, 1, 0.1, nugget=0.75, row_off=300, rows=900, col_off=100, cols=1500)
    y = host(kmv_gpu(lib, desc, dev(x[100:1600])))
    ref = oracle.Kernel(pts, 1, 0.1, nugget=0.75).matvec(x[100:1600], row_off=300, rows=900, col_off=100,
                                                          cols=1500)
    assert rel_err(y, ref) <= 1e-12


@pytest.mark.parametrize("n", [1, 2, 255, 257, 513])
def test_kernel_matvec_small_and_ragged(lib, oracle, n):
    pts, x = kmv_case(oracle, n, 3, 0, 0.05, 1.0, 0, 11)
    P = dev(pts)
    T = point_table(P)
    y = host(kmv_gpu(lib, kernel_desc(P, T, 0, 0.05, nugget=1.0), dev(x)))
    ref = oracle.Kernel(pts, 0, 0.05, nugget=1.0).matvec(x, mode=1)
    assert rel_err(y, ref) <= 1e-12


@pytest.mark.parametrize("kind,l", [(0, 0.05), (1, 0.1)])
def test_kernel_matvec_f32_tolerance(lib, oracle, kind, l):
    n = 4096
    pts, x = kmv_case(oracle, n, 3, kind, l, 0, 0, 21)
    P = dev(pts)
    T = point_table(P)
    y = host(kmv_gpu(lib, kernel_desc(P, T, kind, l, nugget=0.5), dev(x), precision=1))
    ref = oracle.Kernel(pts, kind, l, nugget=0.5).matvec(x, mode=1)
    assert rel_err(y, ref) <= 2e-5


def test_kernel_matvec_full_size_linearity(lib, oracle):
    # Gaussian operator, N = 65536, l = 0.05, nugget 1 (symmetric path by size):
    # linearity K(a x + b y) = a Kx + b Ky
    n = 65536
    pts = oracle.uniform(3 * n, 1235, 0).reshape(3, n)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, 0, 0.05, nugget=1.0)
    x = oracle.uniform(n, 1235, 1, -1, 1)
    z = oracle.uniform(n, 1235, 2, -1, 1)
    kx = host(kmv_gpu(lib, d, dev(x)))
    kz = host(kmv_gpu(lib, d, dev(z)))
    kxz = host(kmv_gpu(lib, d, dev(0.5 * x + 2.0 * z)))
    assert rel_err(kxz, 0.5 * kx + 2.0 * kz) <= 1e-13
    # spot rows against the oracle
    rows = np.array([0, 1, 4097, 65535])
    k = oracle.Kernel(pts, 0, 0.05, nugget=1.0)
    for i in rows:
        ref = k.matvec(x, mode=1, row_off=int(i), rows=1)[0]
        bound = k.matvec(np.abs(x), mode=1, row_off=int(i), rows=1)[0]
        assert abs(kx[i] - ref) <= 1e-12 * bound


# --------------------------------------------------------------------------
# symmetric product (square block, rows == columns): each entry evaluated once
# and used for row i and row j.  MPLAK_KMV_SYMMETRIC=1 forces it at sizes the
# oracle checks quickly; unset it is chosen from 12288 points up.


@pytest.mark.parametrize("kind,l", [(0, 1 / math.sqrt(2.0)), (0, 0.05), (1, 0.1)])
@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("n", [1, 2, 513, 1537])
def test_kernel_matvec_symmetric_f64_vs_oracle(lib, oracle, monkeypatch, kind, l, dim, n):
    monkeypatch.setenv("MPLAK_KMV_SYMMETRIC", "1")
    pts, x = kmv_case(oracle, n, dim, kind, l, 0.0, 0.0, 4000 + 7 * dim + n)
    P = dev(pts)
    T = point_table(P)
    y = host(kmv_gpu(lib, kernel_desc(P, T, kind, l, nugget=0.5), dev(x)))
    k = oracle.Kernel(pts, kind, l, nugget=0.5)
    ref = k.matvec(x, mode=1)
    assert rel_err(y, ref) <= 1e-12
    bound = k.matvec(np.abs(x), mode=1)
    assert np.all(np.abs(y - ref) <= 1e-12 * bound)


@pytest.mark.parametrize("kind,l", [(0, 0.05), (1, 0.1)])
@pytest.mark.parametrize("dim", [1, 3])
def test_kernel_matvec_symmetric_f32_tolerance(lib, oracle, monkeypatch, kind, l, dim):
    monkeypatch.setenv("MPLAK_KMV_SYMMETRIC", "1")
    n = 3000
    pts, x = kmv_case(oracle, n, dim, kind, l, 0, 0, 31 + dim)
    P = dev(pts)
    T = point_table(P)
    y = host(kmv_gpu(lib, kernel_desc(P, T, kind, l, nugget=0.5), dev(x), precision=1))
    ref = oracle.Kernel(pts, kind, l, nugget=0.5).matvec(x, mode=1)
    assert rel_err(y, ref) <= 2e-5


def test_kernel_matvec_symmetric_coincident_points(lib, oracle, monkeypatch):
    # duplicates inside one tile, across the two tiles of a pair, and in the
    # padded last tile: the 1e-8 shift lands on exactly those entries
    monkeypatch.setenv("MPLAK_KMV_SYMMETRIC", "1")
    n = 1100
    pts = oracle.uniform(3 * n, 78, 0).reshape(3, n)
    for a, b in [(10, 20), (30, 700), (600, 1099)]:
        pts[:, a] = pts[:, b]
    x = oracle.uniform(n, 78, 1, -1, 1)
    P = dev(pts)
    T = point_table(P)
    for prec, tol in [(0, 1e-12), (1, 2e-5)]:
        y = host(kmv_gpu(lib, kernel_desc(P, T, 0, 1 / math.sqrt(2.0), shift=1e-8), dev(x), precision=prec))
        ref = oracle.Kernel(pts, 0, 1 / math.sqrt(2.0), coincident_shift=1e-8).matvec(x, mode=1)
        assert rel_err(y, ref) <= tol
    y = host(kmv_gpu(lib, kernel_desc(P, T, 0, 1 / math.sqrt(2.0), shift=1e-8), dev(x)))
    y0 = host(kmv_gpu(lib, kernel_desc(P, T, 0, 1 / math.sqrt(2.0)), dev(x)))
    for a, b in [(10, 20), (30, 700), (600, 1099)]:
        assert abs((y - y0)[a] - 1e-8 * (x[a] + x[b])) < 1e-14


def test_kernel_matvec_symmetric_equals_plain_and_is_deterministic(lib, oracle, monkeypatch):
    n = 20000  # 40 tiles, last one ragged; above the automatic threshold
    pts, x = kmv_case(oracle, n, 3, 0, 0.05, 1.0, 0, 91)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, 0, 0.05, nugget=1.0)
    X = dev(x)
    monkeypatch.setenv("MPLAK_KMV_SYMMETRIC", "0")
    plain = {p: host(kmv_gpu(lib, d, X, precision=p)) for p in (0, 1)}
    monkeypatch.delenv("MPLAK_KMV_SYMMETRIC")
    assert lib.mplak_kernel_matvec_workspace_bytes(C.byref(d)) == 2 * (40 * 41 // 2) * 512 * 8
    for p, tol in [(0, 1e-13), (1, 2e-5)]:
        a = host(kmv_gpu(lib, d, X, precision=p))
        b = host(kmv_gpu(lib, d, X, precision=p))
        assert np.array_equal(a, b)
        assert rel_err(a, plain[p]) <= tol
    ref = oracle.Kernel(pts, 0, 0.05, nugget=1.0).matvec(x, mode=1, row_off=19000, rows=1000)
    assert rel_err(a[19000:], ref) <= 2e-5


def test_kernel_matvec_symmetric_skip_flag(lib, oracle, monkeypatch):
    import torch
    monkeypatch.setenv("MPLAK_KMV_SYMMETRIC", "1")
    n = 1500
    pts, x = kmv_case(oracle, n, 2, 1, 0.2, 0, 0, 5)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, 1, 0.2)
    y = dev(np.full(n, 7.0))
    wb = lib.mplak_kernel_matvec_workspace_bytes(C.byref(d))
    w = workspace(wb)
    flag = torch.ones(1, dtype=torch.int32, device="cuda")
    X = dev(x)
    ck(lib.mplak_kernel_matvec(C.byref(d), 0, X.data_ptr(), y.data_ptr(), w.data_ptr(), wb,
                               flag.data_ptr(), stream()))
    assert np.all(host(y) == 7.0)
    flag.zero_()
    ck(lib.mplak_kernel_matvec(C.byref(d), 0, X.data_ptr(), y.data_ptr(), w.data_ptr(), wb, flag.data_ptr(),
                               stream()))
    ref = oracle.Kernel(pts, 1, 0.2).matvec(x, mode=1)
    assert rel_err(host(y), ref) <= 1e-12


def test_gaussian_rhs_vs_oracle(lib, oracle):
    n = 5000
    pts = oracle.uniform(3 * n, 3, 0).reshape(3, n)
    P = dev(pts)
    T = point_table(P)
    b = empty(n)
    ck(lib.mplak_gaussian_kernel_rhs(T.data_ptr(), 3, 0, n, b.data_ptr(), stream()))
    ref = oracle.Kernel(pts, 0, 1 / math.sqrt(2)).gaussian_rhs()
    assert np.allclose(host(b), ref, rtol=1e-14, atol=0)


# --------------------------------------------------------------------------
# fused CG steps (src/mpla.cpp:817-1001)


def run_fused_cg(lib, A, b, x0, iter_max, eps, rule):
    n = len(b)
    import torch
    Ad = dev(np.asfortranarray(A).ravel(order="F"))
    bd, xd = dev(b), dev(x0)
    r, d, z = empty(n), empty(n), empty(n)
    st = torch.zeros(2 * C.sizeof(_native.mplak_cg_state) // 8, dtype=torch.float64, device="cuda")
    s0 = st.data_ptr()
    s1 = s0 + C.sizeof(_native.mplak_cg_state)
    gw = lib.mplak_dgemv_workspace_bytes(n, n)
    gwork = workspace(gw)
    cw = lib.mplak_cg_workspace_bytes(n)
    cwork = workspace(cw)
    s = stream()
    ck(lib.mplak_dgemv_f64(n, n, Ad.data_ptr(), n, xd.data_ptr(), z.data_ptr(), gwork.data_ptr(), gw, None, s))
    ck(lib.mplak_cg_start(n, bd.data_ptr(), z.data_ptr(), r.data_ptr(), d.data_ptr(), s0, rule, eps, iter_max,
                          float(n), cwork.data_ptr(), cw, s))
    done_off = _native.mplak_cg_state.done.offset
    for j in range(iter_max + 1):
        sin, sout = (s0, s1) if j % 2 == 0 else (s1, s0)
        ck(lib.mplak_dgemv_f64(n, n, Ad.data_ptr(), n, d.data_ptr(), z.data_ptr(), gwork.data_ptr(), gw,
                               sin + done_off, s))
        ck(lib.mplak_cg_iterate(n, xd.data_ptr(), r.data_ptr(), d.data_ptr(), z.data_ptr(), sin, sout,
                                cwork.data_ptr(), cw, s))
    torch.cuda.synchronize()
    final = s0 if (iter_max + 1) % 2 == 0 else s1
    raw = st.cpu().numpy().tobytes()
    state = _native.mplak_cg_state.from_buffer_copy(raw[final - s0: final - s0 + C.sizeof(_native.mplak_cg_state)])
    return host(xd), state


def test_fused_cg_manufactured_laplacian(lib, oracle):
    n = 300
    A = (np.diag(2.0 * np.ones(n)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1))
    xt = oracle.uniform(n, 1235, 0)
    b = A @ xt
    x0 = oracle.uniform(n, 1235, 1)
    x, st = run_fused_cg(lib, A, b, x0, 2000, 1e-10, _native.MPLAK_STOP_RMS)
    xo, info = oracle.cg_rms(oracle.dense_op(A), b, x0, 2000, 1e-10)
    assert st.done == _native.MPLAK_CONVERGED and info.converged == 1
    assert abs(st.k - info.iterations) <= max(2, info.iterations // 50)
    assert np.sqrt(np.mean((b - A @ x) ** 2)) < 1e-10
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_fused_cg_relative_rule_and_limit(lib, oracle):
    n = 400
    A = np.diag(np.linspace(1.0, 50.0, n))
    b = oracle.uniform(n, 8, 0, -1, 1)
    x, st = run_fused_cg(lib, A, b, np.zeros(n), 2001, 1e-10, _native.MPLAK_STOP_RELATIVE)
    xo, info = oracle.cg_relative(oracle.dense_op(A), b, np.zeros(n), 2001, 1e-10)
    assert st.done == _native.MPLAK_CONVERGED and st.k == info.iterations
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10
    # iteration limit: k runs 1 .. iter_max-1
    x, st = run_fused_cg(lib, A, b, np.zeros(n), 5, 1e-14, _native.MPLAK_STOP_RELATIVE)
    assert st.done == _native.MPLAK_ITER_LIMIT and st.k == 4
