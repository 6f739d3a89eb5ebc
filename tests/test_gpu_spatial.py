"""Spatial ordering and the screened product (no reference counterpart).

Bars: the permutation equals an independent numpy computation of the same
Morton keys (exact integer work: bitwise); gather/scatter are exact inverses;
the screened product with nothing to drop is bitwise the exact symmetric
product; with a drop level of 2^-30 on ordered points every row differs from
the oracle's exact FP64 product by at most the dropped mass bound
n * 2^-30 * max|x| plus the FP32 product tolerance; the planned product
(screening plan built once, all-dropped tile pairs never launched) is, in
FP64, bitwise the screened product, also when the plan is reused for new
vectors.  In FP32 the planned product takes the squared distances of compact
tiles from the matrix cores (tile-centred form), so it agrees with the
screened product row by row within 2e-5 of sum_j |K_ij x_j| and with the FP64
oracle within the same bar, and it is bitwise reproducible.
"""
import ctypes as C
import math

import numpy as np
import pytest

from mpla_amd import _native
from tests.gpu_util import dev, empty, host, point_table, require_gpu, stream, workspace
from tests.test_gpu_mplak import kernel_desc, kmv_gpu, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    require_gpu()
    return _native.mplak()


def ck(status):
    _native.check(status)


def morton_perm_numpy(pts):
    """Z-order over 2^(30/dim) cells per axis of the bounding box, ties by index."""
    dim, n = pts.shape
    bits = 30 // dim
    cells = float(1 << bits)
    code = np.zeros(n, dtype=np.uint64)
    for d in range(dim):
        lo, hi = pts[d].min(), pts[d].max()
        ext = hi - lo
        c = np.floor((pts[d] - lo) / ext * cells) if ext > 0 else np.zeros(n)
        c = np.clip(c, 0.0, cells - 1.0).astype(np.uint64)
        for b in range(bits):
            code |= ((c >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * dim + d)
    return np.argsort(code, kind="stable")


def spatial_order(lib, T, dim, n, offset=0):
    import torch
    perm = torch.empty(n, dtype=torch.int64, device="cuda")
    wb = lib.mplak_spatial_order_workspace_bytes(n, dim)
    assert wb > 0
    w = workspace(wb)
    ck(lib.mplak_spatial_order(T.data_ptr(), dim, offset, n, perm.data_ptr(), w.data_ptr(), wb, stream()))
    return perm


@pytest.mark.parametrize("dim,n", [(1, 1), (1, 1000), (2, 4097), (3, 20000), (4, 777)])
def test_spatial_order_matches_numpy_morton(lib, oracle, dim, n):
    pts = oracle.uniform(dim * n, 300 + dim, 0, -2.0, 3.0).reshape(dim, n)
    pts[:, n // 2] = pts[:, 0]  # duplicate point: tie broken by index
    P = dev(pts)
    T = point_table(P)
    perm = host(spatial_order(lib, T, dim, n))
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.array_equal(perm, morton_perm_numpy(pts))


def test_spatial_order_offset_and_degenerate_axis(lib, oracle):
    n, off = 3000, 500
    pts = oracle.uniform(3 * (n + off), 9, 0).reshape(3, n + off)
    pts[2, :] = 0.25  # flat axis: every point in cell 0 along z
    P = dev(pts)
    T = point_table(P)
    perm = host(spatial_order(lib, T, 3, n, offset=off))
    assert np.array_equal(perm, morton_perm_numpy(pts[:, off:]))


def test_gather_scatter_points_round_trip(lib, oracle):
    import torch
    n, dim = 5000, 3
    pts = oracle.uniform(dim * n, 4, 0).reshape(dim, n)
    P = dev(pts)
    T = point_table(P)
    perm = spatial_order(lib, T, dim, n)
    ph = host(perm)
    out = torch.empty(dim * (n + 3), dtype=torch.float64, device="cuda")
    ck(lib.mplak_gather_points(T.data_ptr(), dim, 0, perm.data_ptr(), n, out.data_ptr(), n + 3, stream()))
    o = host(out).reshape(dim, n + 3)[:, :n]
    assert np.array_equal(o, pts[:, ph])
    x = oracle.uniform(n, 4, 1, -1, 1)
    X = dev(x)
    g = empty(n)
    ck(lib.mplak_gather_f64(g.data_ptr(), X.data_ptr(), perm.data_ptr(), n, stream()))
    assert np.array_equal(host(g), x[ph])
    back = empty(n)
    ck(lib.mplak_scatter_f64(back.data_ptr(), g.data_ptr(), perm.data_ptr(), n, stream()))
    assert np.array_equal(host(back), x)
    assert lib.mplak_gather_f64(g.data_ptr(), g.data_ptr(), perm.data_ptr(), n, stream()) == _native.MPLAK_ERR_ARG


def screened(lib, desc, X, prec, drop):
    y = empty(desc.row_count)
    wb = lib.mplak_kernel_matvec_screened_workspace_bytes(C.byref(desc))
    w = workspace(wb)
    ck(lib.mplak_kernel_matvec_screened(C.byref(desc), prec, drop, X.data_ptr(), y.data_ptr(), w.data_ptr(), wb,
                                        None, stream()))
    return host(y)


def ordered_points(lib, oracle, n, dim, seed):
    pts = oracle.uniform(dim * n, seed, 0).reshape(dim, n)
    return pts[:, morton_perm_numpy(pts)].copy()


@pytest.mark.parametrize("kind,l", [(0, 0.05), (1, 0.1)])
@pytest.mark.parametrize("ordered", [True, False])
def test_screened_without_drops_is_the_exact_symmetric_product(lib, oracle, monkeypatch, kind, l, ordered):
    """Screened with nothing to drop: bitwise the exact symmetric product in
    both precisions.  Planned FP32 with nothing to drop (matrix-core form on
    the compact tiles of ordered points): within the FP32 bar of the FP64
    oracle, row by row."""
    monkeypatch.setenv("MPLAK_KMV_SYMMETRIC", "1")
    n = 3000
    pts = ordered_points(lib, oracle, n, 3, 51) if ordered else oracle.uniform(3 * n, 51, 0).reshape(3, n)
    x = oracle.uniform(n, 51, 1, -1, 1)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, kind, l, nugget=0.5)
    X = dev(x)
    assert np.array_equal(screened(lib, d, X, 0, 2000.0), host(kmv_gpu(lib, d, X, precision=0)))
    exact32 = host(kmv_gpu(lib, d, X, precision=1))
    y32 = screened(lib, d, X, 1, 2000.0)
    assert np.array_equal(y32, exact32)
    if ordered:
        plan, wb = build_plan(lib, d, 1, 2000.0)
        yp = planned(lib, d, 1, 2000.0, plan, wb, X)
        k = oracle.Kernel(pts, kind, l, nugget=0.5)
        ref = k.matvec(x, mode=1)
        assert rel_err(yp, ref) <= 2e-5
        assert np.all(np.abs(yp - ref) <= 2e-5 * k.matvec(np.abs(x), mode=1))


@pytest.mark.parametrize("kind,l,dim", [(0, 0.05, 3), (0, 0.02, 2), (1, 0.05, 3), (0, 0.01, 1)])
def test_screened_error_within_dropped_mass_bound(lib, oracle, kind, l, dim):
    n = 20000
    pts = ordered_points(lib, oracle, n, dim, 60 + dim)
    x = oracle.uniform(n, 61, 1, -1, 1)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, kind, l, nugget=1.0)
    X = dev(x)
    k = oracle.Kernel(pts, kind, l, nugget=1.0)
    rows = np.arange(0, n, 97)
    ref = np.array([k.matvec(x, mode=1, row_off=int(i), rows=1)[0] for i in rows])
    absref = np.array([k.matvec(np.abs(x), mode=1, row_off=int(i), rows=1)[0] for i in rows])
    drop_bound = n * 2.0 ** -30 * np.abs(x).max()
    for prec, tol in [(0, 1e-12), (1, 2e-5)]:
        y = screened(lib, d, X, prec, 30.0)[rows]
        assert np.all(np.abs(y - ref) <= drop_bound + tol * absref)
    # something was dropped: the FP64 screened result is not the exact one
    y64 = screened(lib, d, X, 0, 30.0)
    exact = host(kmv_gpu(lib, d, X, precision=0))
    assert not np.array_equal(y64, exact)
    assert rel_err(y64, exact) <= 1e-7


def test_screened_falls_back_to_exact_for_rectangular_blocks(lib, oracle):
    n = 2000
    pts, x = oracle.uniform(3 * n, 8, 0).reshape(3, n), oracle.uniform(1500, 8, 1, -1, 1)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, 1, 0.1, nugget=0.75, row_off=300, rows=900, col_off=100, cols=1500)
    X = dev(x)
    assert np.array_equal(screened(lib, d, X, 0, 30.0), host(kmv_gpu(lib, d, X)))
    y = empty(900)
    assert lib.mplak_kernel_matvec_screened(C.byref(d), 0, 0.0, X.data_ptr(), y.data_ptr(), None, 0, None,
                                            stream()) == _native.MPLAK_ERR_ARG


def build_plan(lib, desc, prec, drop):
    wb = lib.mplak_screen_plan_bytes(C.byref(desc))
    assert wb > 0
    plan = workspace(wb)
    ck(lib.mplak_screen_plan_build(C.byref(desc), prec, drop, plan.data_ptr(), wb, stream()))
    return plan, wb


def planned(lib, desc, prec, drop, plan, wb, X):
    y = empty(desc.row_count)
    ck(lib.mplak_kernel_matvec_planned(C.byref(desc), prec, drop, plan.data_ptr(), wb, X.data_ptr(), y.data_ptr(),
                                       None, stream()))
    return host(y)


@pytest.mark.parametrize("kind,l,dim,n", [(0, 0.05, 3, 20000), (1, 0.1, 3, 7000), (0, 0.02, 2, 9000),
                                          (0, 0.01, 1, 5000), (0, 0.05, 3, 511)])
@pytest.mark.parametrize("drop", [20.0, 30.0])
def test_planned_product_matches_screened(lib, oracle, kind, l, dim, n, drop):
    pts = ordered_points(lib, oracle, n, dim, 70 + dim)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, kind, l, nugget=1.0)
    for prec in (1, 0):
        plan, wb = build_plan(lib, d, prec, drop)
        # the plan is reused across products with different vectors
        for seed in (1, 2):
            x = oracle.uniform(n, 71, seed, -1, 1)
            X = dev(x)
            yp = planned(lib, d, prec, drop, plan, wb, X)
            ys = screened(lib, d, X, prec, drop)
            if prec == 0:
                assert np.array_equal(yp, ys)
            else:
                mass = screened(lib, d, dev(np.abs(x)), 0, drop)
                assert np.all(np.abs(yp - ys) <= 2e-5 * mass)
                assert np.array_equal(planned(lib, d, prec, drop, plan, wb, X), yp)


@pytest.mark.parametrize("kind,l", [(0, 0.05), (1, 0.1)])
def test_planned_fp32_matrix_core_form_against_oracle(lib, oracle, kind, l):
    """20000 ordered points: every target tile is compact, so the planned FP32
    product takes all its tile pairs on the matrix cores.  With nothing
    dropped, sampled rows are within 2e-5 of sum_j |K_ij x_j| of the FP64
    oracle; the result is not bitwise the vector-unit screened product (the
    other arithmetic path)."""
    n = 20000
    pts = ordered_points(lib, oracle, n, 3, 80 + kind)
    x = oracle.uniform(n, 81, 1, -1, 1)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, kind, l, nugget=0.5)
    X = dev(x)
    plan, wb = build_plan(lib, d, 1, 2000.0)
    yp = planned(lib, d, 1, 2000.0, plan, wb, X)
    assert not np.array_equal(yp, screened(lib, d, X, 1, 2000.0))
    k = oracle.Kernel(pts, kind, l, nugget=0.5)
    rows = np.arange(0, n, 61)
    ref = np.array([k.matvec(x, mode=1, row_off=int(i), rows=1)[0] for i in rows])
    mass = np.array([k.matvec(np.abs(x), mode=1, row_off=int(i), rows=1)[0] for i in rows])
    assert np.all(np.abs(yp[rows] - ref) <= 2e-5 * mass)


def test_planned_product_unordered_points_and_errors(lib, oracle):
    n = 3000
    pts = oracle.uniform(3 * n, 12, 0).reshape(3, n)  # no spatial order: nothing can be dropped
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, 0, 0.05, nugget=1.0)
    x = oracle.uniform(n, 12, 1, -1, 1)
    X = dev(x)
    plan, wb = build_plan(lib, d, 0, 30.0)
    y = planned(lib, d, 0, 30.0, plan, wb, X)
    assert rel_err(y, oracle.Kernel(pts, 0, 0.05, nugget=1.0).matvec(x, mode=1)) <= 1e-7
    assert lib.mplak_kernel_matvec_planned(C.byref(d), 0, 30.0, plan.data_ptr(), wb - 8, X.data_ptr(),
                                           X.data_ptr(), None, stream()) == _native.MPLAK_ERR_WORKSPACE
    rect = kernel_desc(P, T, 0, 0.05, rows=100, cols=200)
    assert lib.mplak_screen_plan_bytes(C.byref(rect)) == 0


def plan_counts_numpy(pts, l, drop):
    """The plan's screen restated in numpy on the same FP32-cast scaled
    coordinates (Gaussian): boxes of 16-point slabs, 64-point groups and
    128-point wave target sets per 512-point tile (padding clamped to the last
    point), squared box gaps with FP32 rounding (fma emulated in FP64, exact
    for these products), compact tiles (half-diagonal^2 <= 49), and per tile
    pair I <= J the kept 128 x 64 blocks and kept 16 x 16 tiles.  Returns the
    five plan counts."""
    dim, n = pts.shape
    scale = math.sqrt(math.log2(math.e)) / (math.sqrt(2.0) * l)
    f = np.float32
    nt = (n + 511) // 512
    P = (pts * scale).astype(f)[:, np.minimum(np.arange(nt * 512), n - 1)]

    def boxes(size):
        Q = P.reshape(dim, -1, size)
        return Q.min(axis=2).T, Q.max(axis=2).T  # (count, dim)

    def gap2(alo, ahi, blo, bhi):
        g = np.maximum(f(0), np.maximum(blo - ahi, alo - bhi)).astype(f)
        d2 = (g[..., 0].astype(np.float64) ** 2).astype(f)
        for d in range(1, dim):
            d2 = (g[..., d].astype(np.float64) ** 2 + d2.astype(np.float64)).astype(f)
        return d2

    slo, shi = boxes(16)
    glo, ghi = boxes(64)
    wlo, whi = boxes(128)
    tlo, thi = boxes(512)
    half = (f(0.5) * (thi - tlo)).astype(f)
    r2 = (half[:, 0].astype(np.float64) ** 2).astype(f)
    for d in range(1, dim):
        r2 = (half[:, d].astype(np.float64) ** 2 + r2.astype(np.float64)).astype(f)
    compact = r2 <= f(49.0)
    dt = gap2(slo[:, None, :], shi[:, None, :], slo[None, :, :], shi[None, :, :]) <= f(drop)  # slab x slab
    db = gap2(wlo[:, None, :], whi[:, None, :], glo[None, :, :], ghi[None, :, :]) <= f(drop)  # wave x group
    c = [0, 0, 0, 0, 0]
    for J in range(nt):
        for I in range(J + 1):
            blocks = int(db[4 * I:4 * I + 4, 8 * J:8 * J + 8].sum())
            tiles = int(dt[32 * I:32 * I + 32, 32 * J:32 * J + 32].sum())
            if (tiles if compact[I] else blocks) == 0:
                continue
            c[0] += 1
            c[1] += blocks
            c[2] += blocks if I == J else 0
            c[4 if I == J else 3] += 256 * tiles if compact[I] else 8192 * blocks
    return c


def test_plan_counts(lib, oracle):
    import torch
    n = 3000  # 6 tiles, 21 pairs
    pts = oracle.uniform(3 * n, 12, 0).reshape(3, n)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, 0, 0.05, nugget=1.0)
    counts = torch.zeros(5, dtype=torch.int64, device="cuda")
    plan, wb = build_plan(lib, d, 1, 20.0)  # unordered points: every block is kept, no tile is compact
    ck(lib.mplak_screen_plan_counts(C.byref(d), plan.data_ptr(), wb, counts.data_ptr(), stream()))
    assert host(counts).tolist() == [21, 21 * 32, 6 * 32, 8192 * 15 * 32, 8192 * 6 * 32]
    for dim, l, n in ((3, 0.05, 20000), (2, 0.02, 9000), (1, 0.01, 5000)):
        po = ordered_points(lib, oracle, n, dim, 5)
        Po = dev(po)
        To = point_table(Po)
        do = kernel_desc(Po, To, 0, l, nugget=1.0)
        plan, wb = build_plan(lib, do, 1, 20.0)
        ck(lib.mplak_screen_plan_counts(C.byref(do), plan.data_ptr(), wb, counts.data_ptr(), stream()))
        got = host(counts).tolist()
        assert got == plan_counts_numpy(po, l, 20.0), (dim, got)
        nt = (n + 511) // 512
        assert nt <= got[0] < nt * (nt + 1) // 2 and got[3] + got[4] < 8192 * got[1]
