"""Spatial ordering and the screened product (no reference counterpart).

Bars: the permutation equals an independent numpy computation of the same
Morton keys (exact integer work: bitwise); gather/scatter are exact inverses;
the screened product with nothing to drop is bitwise the exact symmetric
product; with a drop level of 2^-30 on ordered points every row differs from
the oracle's exact FP64 product by at most the dropped mass bound
n * 2^-30 * max|x| plus the FP32 product tolerance; the planned product
(screening plan built once, all-dropped tile pairs never launched) agrees,
in FP64, with the screened product within n * 2^-drop * max|x| (the entries
only one of the two screens drops) plus 1e-12 of sum_j |K_ij x_j|, also
when the plan is reused for new vectors, and repeats bitwise.  In FP32 the planned product takes the squared distances of compact
tiles from the matrix cores (tile-centred form), so it agrees with the
screened product row by row within 2e-5 of sum_j |K_ij x_j| and with the FP64
oracle within the same bar, and it is bitwise reproducible.  In FP64 the
planned product screens per 16 x 16 tile as well, with the entries from FP64
coordinate differences: with nothing dropped it is the exact product to 1e-12
of sum_j |K_ij x_j| row by row, and it is bitwise reproducible.
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


def hilbert_perm_numpy(pts):
    """Hilbert-curve order over 2^(30/dim) cells per axis of the bounding box,
    ties by index: Skilling's axes-to-transpose map (AIP Conf. Proc. 707,
    2004), then the bits of X[0], X[1], ... interleaved from the top."""
    dim, n = pts.shape
    bits = 30 // dim
    cells = float(1 << bits)
    X = []
    for d in range(dim):
        lo, hi = pts[d].min(), pts[d].max()
        ext = hi - lo
        c = np.floor((pts[d] - lo) / ext * cells) if ext > 0 else np.zeros(n)
        X.append(np.clip(c, 0.0, cells - 1.0).astype(np.int64))
    M = 1 << (bits - 1)
    Q = M
    while Q > 1:
        P = Q - 1
        for i in range(dim):
            inv = (X[i] & Q) != 0
            t = np.where(inv, 0, (X[0] ^ X[i]) & P)
            X[0] = np.where(inv, X[0] ^ P, X[0] ^ t)
            if i > 0:
                X[i] = X[i] ^ t
        Q >>= 1
    for i in range(1, dim):
        X[i] = X[i] ^ X[i - 1]
    t = np.zeros(n, dtype=np.int64)
    Q = M
    while Q > 1:
        t = np.where((X[dim - 1] & Q) != 0, t ^ (Q - 1), t)
        Q >>= 1
    X = [x ^ t for x in X]
    code = np.zeros(n, dtype=np.uint64)
    for b in range(bits - 1, -1, -1):
        for i in range(dim):
            code = (code << np.uint64(1)) | ((X[i] >> b) & 1).astype(np.uint64)
    return np.argsort(code, kind="stable")


def spatial_order(lib, T, dim, n, offset=0):
    import torch
    perm = torch.empty(n, dtype=torch.int64, device="cuda")
    wb = lib.mplak_spatial_order_workspace_bytes(n, dim)
    assert wb > 0
    w = workspace(wb)
    ck(lib.mplak_spatial_order(T.data_ptr(), dim, offset, n, perm.data_ptr(), w.data_ptr(), wb, stream()))
    return perm


@pytest.mark.parametrize("dim,n", [(1, 1), (1, 1000), (2, 4097), (3, 20000), (4, 777), (7, 300)])
def test_spatial_order_matches_numpy_hilbert(lib, oracle, dim, n):
    pts = oracle.uniform(dim * n, 300 + dim, 0, -2.0, 3.0).reshape(dim, n)
    pts[:, n // 2] = pts[:, 0]  # duplicate point: tie broken by index
    P = dev(pts)
    T = point_table(P)
    perm = host(spatial_order(lib, T, dim, n))
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.array_equal(perm, hilbert_perm_numpy(pts))
    if dim == 1:  # the one-dimensional Hilbert curve is the line itself
        assert np.array_equal(perm, np.argsort(pts[0], kind="stable"))


@pytest.mark.parametrize("dim,m", [(2, 16), (3, 8)])
def test_spatial_order_is_a_continuous_curve(lib, dim, m):
    """Points at the centres of an m^dim lattice of dyadic blocks of the grid
    (plus one point at the far corner that fixes the bounding box): the order
    visits every lattice point once, each step to a lattice neighbour (one
    axis, distance one), as a Hilbert curve must; the Z-order jumps."""
    bits = 30 // dim
    step = (1 << bits) // m
    axes = np.meshgrid(*[np.arange(m)] * dim, indexing="ij")
    lat = np.stack([a.ravel() for a in axes]).astype(np.float64)
    pts = np.concatenate([lat * step + step // 2, np.full((dim, 1), float(1 << bits))], axis=1)
    order = rng_permutation(pts.shape[1], 77)
    pts = pts[:, order].copy()  # the curve, not the input order, decides
    P = dev(pts)
    T = point_table(P)
    perm = host(spatial_order(lib, T, dim, pts.shape[1]))
    corner = pts.shape[1] - 1
    seq = [int(i) for i in perm if order[i] != corner]
    cells = np.rint((pts[:, seq] - step // 2) / step).astype(np.int64)
    steps = np.abs(np.diff(cells, axis=1)).sum(0)
    assert len(seq) == m ** dim and np.all(steps == 1)
    mz = (pts[:, morton_perm_numpy(pts)] - step // 2) / step
    assert np.abs(np.diff(np.rint(mz[:, :-1]), axis=1)).sum(0).max() > 1


def rng_permutation(n, seed):
    return np.random.default_rng(seed).permutation(n)


def test_spatial_order_offset_and_degenerate_axis(lib, oracle):
    n, off = 3000, 500
    pts = oracle.uniform(3 * (n + off), 9, 0).reshape(3, n + off)
    pts[2, :] = 0.25  # flat axis: every point in cell 0 along z
    P = dev(pts)
    T = point_table(P)
    perm = host(spatial_order(lib, T, 3, n, offset=off))
    assert np.array_equal(perm, hilbert_perm_numpy(pts[:, off:]))


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
    """Points in Z-order: spatially ordered, with a few tiles that straddle
    octant boundaries (non-compact), so the products' vector-unit form for
    such tiles runs beside the matrix-core form."""
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


def _plan_offsets(n):
    """Byte offsets of the pair count, the pair list and the unsorted keys
    in a screening plan of n points (the plan's fixed layout: 512-point
    tiles, 256-byte aligned regions)."""
    a = lambda v: -(-v // 256) * 256
    nt = -(-n // 512)
    np_ = nt * (nt + 1) // 2
    tinfo = 256
    boxes = tinfo + a(16 * nt)
    sboxes = boxes + a(8 * 6 * 8 * nt)
    lst = sboxes + a(8 * 6 * 32 * nt)
    bmask = lst + a(4 * np_)
    tmask = bmask + a(4 * np_)
    keys_in = tmask + a(4 * 32 * np_)  # 32 tile-mask words per pair
    return np_, 0, lst, keys_in


@pytest.mark.parametrize("n,drop,prec", [(20000, 17.0, 1), (65536, 15.0, 1), (65536, 6.0, 1), (3000, 6.0, 1),
                                         (511, 20.0, 1), (13000, 55.0, 0)])
def test_plan_list_is_the_stable_class_order(lib, oracle, n, drop, prec):
    """White-box check of a screening plan's pair list: every per-pair key
    carries its pair index, and the list holds the pair indices in ascending
    key class (bits 32-38: heaviest pairs first, pairs with nothing kept
    last), equal classes in index order; the count is the number of pairs
    before the dropped class."""
    pts = ordered_points(lib, oracle, n, 3, 151)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, 0, 0.05, nugget=1.0)
    np_, off_count, off_list, off_keys = _plan_offsets(n)
    plan, wb = build_plan(lib, d, prec, drop)
    raw = host(plan).view(np.uint8)
    count = int(raw[off_count:off_count + 8].view(np.int64)[0])
    lst = raw[off_list:off_list + 4 * np_].view(np.uint32)
    keys = raw[off_keys:off_keys + 8 * np_].view(np.uint64)
    assert np.array_equal(keys & np.uint64(0xffffffff), np.arange(np_, dtype=np.uint64))
    cls = (keys >> np.uint64(32)).astype(np.int64)
    assert np.all(cls <= 127)
    assert np.array_equal(lst, np.argsort(cls, kind="stable").astype(np.uint32))
    assert count == int((cls != 127).sum())
    assert 0 < count <= np_


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
                # FP64: 16 x 16 tile screen against 128 x 64 blocks; the
                # entries only one of them drops are below 2^-drop
                mass = screened(lib, d, dev(np.abs(x)), 0, drop)
                assert np.all(np.abs(yp - ys) <= n * 2.0 ** -drop * np.abs(x).max() + 1e-12 * mass)
                assert np.array_equal(planned(lib, d, prec, drop, plan, wb, X), yp)
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
    pair I <= J the kept 128 x 64 blocks and kept 16 x 16 tiles (on a
    diagonal pair only those with target slab <= source slab: the FP32
    matrix-core form adds each strictly upper tile to its rows and its
    columns).  Returns the five plan counts."""
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
            block = dt[32 * I:32 * I + 32, 32 * J:32 * J + 32]
            tiles = int((np.triu(block) if I == J else block).sum())
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


def planned_dots(lib, desc, prec, drop, plan, wb, X, y, dots, skip=None):
    ck(lib.mplak_kernel_matvec_planned_dots(C.byref(desc), prec, drop, plan.data_ptr(), wb, X.data_ptr(),
                                            y.data_ptr(), dots.data_ptr(), skip, stream()))


@pytest.mark.parametrize("kind,l,dim,n", [(0, 0.05, 3, 20000), (1, 0.1, 3, 7000), (0, 0.02, 2, 9000),
                                          (0, 0.05, 3, 511)])
@pytest.mark.parametrize("prec", [0, 1])
def test_planned_dots(lib, oracle, kind, l, dim, n, prec):
    """The planned product that also leaves the partials of <x,x> and <x,y>:
    y bitwise that of the planned product; the partials (one pair per 64
    rows) add up to the two inner products of the returned y within 1e-13
    (FP64 sums of the same terms in another order)."""
    pts = ordered_points(lib, oracle, n, dim, 90 + dim)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, kind, l, nugget=0.5)
    plan, wb = build_plan(lib, d, prec, 20.0)
    nd = lib.mplak_planned_dots_count(n)
    assert nd == (n + 63) // 64
    assert lib.mplak_planned_dots_count(0) == 0
    x = oracle.uniform(n, 91, 1, -1, 1)
    X = dev(x)
    y = empty(n)
    dots = empty(2 * nd)
    planned_dots(lib, d, prec, 20.0, plan, wb, X, y, dots)
    yh, dh = host(y), host(dots)
    assert np.array_equal(yh, planned(lib, d, prec, 20.0, plan, wb, X))
    xx, xy = math.fsum(x * x), math.fsum(x * yh)
    assert abs(math.fsum(dh[:nd]) - xx) <= 1e-13 * xx
    assert abs(math.fsum(dh[nd:]) - xy) <= 1e-13 * math.fsum(np.abs(x * yh))
    # a set skip flag leaves y and the partials untouched
    flag = dev(np.array([1], dtype=np.int32), dtype=torch_int32())
    y2, d2 = dev(np.full(n, 7.0)), dev(np.full(2 * nd, 7.0))
    planned_dots(lib, d, prec, 20.0, plan, wb, X, y2, d2, flag.data_ptr())
    assert np.all(host(y2) == 7.0) and np.all(host(d2) == 7.0)
    assert lib.mplak_kernel_matvec_planned_dots(C.byref(d), prec, 20.0, plan.data_ptr(), wb, X.data_ptr(),
                                                y.data_ptr(), None, None, stream()) == _native.MPLAK_ERR_ARG


def torch_int32():
    import torch
    return torch.int32


def cg1_solve(lib, d, prec, drop, b, eps, iter_max=2001):
    """Single-reduction CG through the C ABI: planned product on r with its
    dot partials, then mplak_cg1_iterate, until the state is done."""
    import torch
    n = d.row_count
    plan, wb = build_plan(lib, d, prec, drop)
    nd = lib.mplak_planned_dots_count(n)
    B = dev(b)
    x, r, p, s, w, z = (torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(6))
    dots = empty(2 * nd)
    ssz = C.sizeof(_native.mplak_cg_state)
    st = torch.zeros(2 * ssz // 8, dtype=torch.float64, device="cuda")
    cw = lib.mplak_cg_workspace_bytes(n)
    cwork = workspace(cw)
    ck(lib.mplak_cg_start(n, B.data_ptr(), z.data_ptr(), r.data_ptr(), p.data_ptr(), st.data_ptr(),
                          _native.MPLAK_STOP_RELATIVE, eps, iter_max, float(n), cwork.data_ptr(), cw, stream()))
    products = 0
    for it in range(iter_max + 1):
        sin, sout = st.data_ptr() + (it & 1) * ssz, st.data_ptr() + (1 - (it & 1)) * ssz
        planned_dots(lib, d, prec, drop, plan, wb, r, w, dots, sin + _native.mplak_cg_state.done.offset)
        ck(lib.mplak_cg1_iterate(n, x.data_ptr(), r.data_ptr(), p.data_ptr(), s.data_ptr(), w.data_ptr(),
                                 dots.data_ptr(), nd, sin, sout, stream()))
        products += 1
        state = _native.mplak_cg_state.from_buffer_copy(host(st).tobytes()[(1 - (it & 1)) * ssz:][:ssz])
        if state.done != _native.MPLAK_RUNNING:
            return host(x), state, products
    raise AssertionError("single-reduction CG did not stop")


@pytest.mark.parametrize("kind,l", [(0, 0.05), (1, 0.1)])
def test_cg1_matches_classic_cg_fp64(lib, oracle, kind, l):
    """Single-reduction CG with the FP64 planned product (nothing dropped at
    2^-2000) is classic CG rearranged: it stops within one iteration of the
    oracle's CG (src/mpla.cpp:887-1001 stopping rule) on the same system,
    its true residual (oracle product) meets the tolerance, and its solution
    agrees with the oracle's."""
    n, eps = 4000, 1e-10
    pts = ordered_points(lib, oracle, n, 3, 95 + kind)
    b = oracle.uniform(n, 96, 1, -1, 1)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, kind, l, nugget=1.0)
    x, state, products = cg1_solve(lib, d, 0, 2000.0, b, eps)
    assert state.done == _native.MPLAK_CONVERGED
    assert products == state.k + 1
    k = oracle.Kernel(pts, kind, l, nugget=1.0)
    xr, info = oracle.cg_relative(k.op(), b, np.zeros(n), 2001, eps)
    assert info.converged and abs(state.k - info.iterations) <= 1
    res = np.linalg.norm(b - k.matvec(x)) / np.linalg.norm(b)
    assert res <= 1.5 * eps
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)


def test_cg1_matches_classic_cg_fp32_planned(lib, oracle):
    """The mixed solver's inner form: planned FP32 operator (2^-20 screen),
    tolerance 1e-4.  Single-reduction CG and classic CG (mplak_cg_iterate) on
    the same operator stop within one iteration of each other, and both
    solutions meet the tolerance on the oracle's FP64 residual (the FP32
    operator and its 2^-20 screen move it by far less than 1e-4)."""
    import torch
    n, eps, drop = 12000, 1e-4, 20.0
    pts = ordered_points(lib, oracle, n, 3, 97)
    b = oracle.uniform(n, 98, 1, -1, 1)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, 0, 0.05, nugget=1.0)
    x1, s1, _ = cg1_solve(lib, d, 1, drop, b, eps)
    assert s1.done == _native.MPLAK_CONVERGED
    # classic CG on the same planned operator
    plan, wb = build_plan(lib, d, 1, drop)
    B = dev(b)
    x, r, dd, z = (torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(4))
    ssz = C.sizeof(_native.mplak_cg_state)
    st = torch.zeros(2 * ssz // 8, dtype=torch.float64, device="cuda")
    cw = lib.mplak_cg_workspace_bytes(n)
    cwork = workspace(cw)
    ck(lib.mplak_cg_start(n, B.data_ptr(), z.data_ptr(), r.data_ptr(), dd.data_ptr(), st.data_ptr(),
                          _native.MPLAK_STOP_RELATIVE, eps, 2001, float(n), cwork.data_ptr(), cw, stream()))
    for it in range(2001):
        sin, sout = st.data_ptr() + (it & 1) * ssz, st.data_ptr() + (1 - (it & 1)) * ssz
        ck(lib.mplak_kernel_matvec_planned(C.byref(d), 1, drop, plan.data_ptr(), wb, dd.data_ptr(), z.data_ptr(),
                                           sin + _native.mplak_cg_state.done.offset, stream()))
        ck(lib.mplak_cg_iterate(n, x.data_ptr(), r.data_ptr(), dd.data_ptr(), z.data_ptr(), sin, sout,
                                cwork.data_ptr(), cw, stream()))
        s2 = _native.mplak_cg_state.from_buffer_copy(host(st).tobytes()[(1 - (it & 1)) * ssz:][:ssz])
        if s2.done != _native.MPLAK_RUNNING:
            break
    assert s2.done == _native.MPLAK_CONVERGED
    assert abs(s1.k - s2.k) <= 1
    k = oracle.Kernel(pts, 0, 0.05, nugget=1.0)
    for sol in (x1, host(x)):
        assert np.linalg.norm(b - k.matvec(sol)) <= 1.5 * eps * np.linalg.norm(b)


def test_planned_relaxed_levels(lib, oracle):
    """Three FP32 plans at 2^-17, 2^-12, 2^-8 behind one launch: the level
    the selector value picks (<= 2^-5 for level 1, <= 2^-9 for level 2)
    gives bitwise the planned product of that level's plan, with its dot
    partials; one level is bitwise mplak_kernel_matvec_planned_dots.
    Invalid level sets are refused."""
    import torch
    n = 20000
    pts = ordered_points(lib, oracle, n, 3, 99)
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, 0, 0.05, nugget=1.0)
    drops = [17.0, 12.0, 8.0]
    built = [build_plan(lib, d, 1, dr) for dr in drops]
    wb = built[0][1]
    plans = (C.c_void_p * 3)(*[p.data_ptr() for p, _ in built])
    dl = (C.c_double * 3)(*drops)
    below = (C.c_double * 2)(2.0 ** -5, 2.0 ** -9)
    x = oracle.uniform(n, 100, 1, -1, 1)
    X = dev(x)
    nd = lib.mplak_planned_dots_count(n)
    sel = dev(np.array([1.0]))
    for value, level in ((1.0, 0), (0.04, 0), (2.0 ** -5, 1), (2.0 ** -7, 1), (2.0 ** -9, 2), (1e-6, 2),
                         (float("nan"), 0)):
        sel.fill_(value)
        y, dots = empty(n), empty(2 * nd)
        ck(lib.mplak_kernel_matvec_planned_relaxed(C.byref(d), 3, dl, plans, wb, sel.data_ptr(), below, X.data_ptr(),
                                                   y.data_ptr(), dots.data_ptr(), None, stream()))
        yr, dr_ = host(y), host(dots)
        plan, pwb = built[level]
        y1, d1 = empty(n), empty(2 * nd)
        planned_dots(lib, d, 1, drops[level], plan, pwb, X, y1, d1)
        assert np.array_equal(yr, host(y1)), (value, level)
        assert np.array_equal(dr_, host(d1)), (value, level)
    # one level: the single-plan product
    y, dots = empty(n), empty(2 * nd)
    ck(lib.mplak_kernel_matvec_planned_relaxed(C.byref(d), 1, dl, plans, wb, None, None, X.data_ptr(), y.data_ptr(),
                                               dots.data_ptr(), None, stream()))
    y1, d1 = empty(n), empty(2 * nd)
    planned_dots(lib, d, 1, drops[0], built[0][0], wb, X, y1, d1)
    assert np.array_equal(host(y), host(y1)) and np.array_equal(host(dots), host(d1))
    bad = _native.MPLAK_ERR_ARG
    for nlev, dd, bb in ((4, drops, (2.0 ** -5, 2.0 ** -9)), (0, drops, (2.0 ** -5, 2.0 ** -9)),
                         (3, [17.0, 12.0, 12.0], (2.0 ** -5, 2.0 ** -9)), (3, drops, (2.0 ** -9, 2.0 ** -5)),
                         (2, drops, (float("nan"), 0.0))):
        dl2 = (C.c_double * 4)(*(list(dd) + [1.0] * (4 - len(dd))))
        b2 = (C.c_double * 3)(*(list(bb) + [0.0]))
        pl4 = (C.c_void_p * 4)(*([p.data_ptr() for p, _ in built] + [built[0][0].data_ptr()]))
        assert lib.mplak_kernel_matvec_planned_relaxed(C.byref(d), nlev, dl2, pl4, wb, sel.data_ptr(), b2,
                                                       X.data_ptr(), y.data_ptr(), None, None, stream()) == bad
    assert lib.mplak_kernel_matvec_planned_relaxed(C.byref(d), 2, dl, plans, wb, None, below, X.data_ptr(),
                                                   y.data_ptr(), None, None, stream()) == bad


@pytest.mark.parametrize("kind,l,dim,n,order", [(0, 0.05, 3, 20000, "hilbert"), (1, 0.1, 3, 7000, "morton"),
                                                (0, 0.02, 2, 9000, "hilbert"), (0, 0.01, 1, 5000, "none"),
                                                (1, 0.1, 3, 511, "hilbert"), (0, 0.05, 3, 13000, "none")])
def test_planned_fp64_tile_form_is_the_exact_product(lib, oracle, kind, l, dim, n, order):
    """FP64 planned product with nothing dropped (2^-2000): the tile form's
    result equals the exact symmetric product to 1e-12 of sum_j |K_ij x_j|
    in every row, and the FP64 oracle on sampled rows to the same bar; on
    ordered and unordered points, dims 1-3, both kernels, ragged sizes."""
    pts = oracle.uniform(dim * n, 110 + dim, 0).reshape(dim, n)
    if order == "hilbert":
        pts = pts[:, hilbert_perm_numpy(pts)].copy()
    elif order == "morton":
        pts = pts[:, morton_perm_numpy(pts)].copy()
    P = dev(pts)
    T = point_table(P)
    d = kernel_desc(P, T, kind, l, nugget=0.5)
    x = oracle.uniform(n, 111, 1, -1, 1)
    X = dev(x)
    plan, wb = build_plan(lib, d, 0, 2000.0)
    yp = planned(lib, d, 0, 2000.0, plan, wb, X)
    ye = host(kmv_gpu(lib, d, X))
    mass = host(kmv_gpu(lib, d, dev(np.abs(x))))
    assert np.all(np.abs(yp - ye) <= 1e-12 * mass)
    k = oracle.Kernel(pts, kind, l, nugget=0.5)
    rows = np.arange(0, n, max(1, n // 97))
    ref = np.array([k.matvec(x, mode=1, row_off=int(i), rows=1)[0] for i in rows])
    assert np.all(np.abs(yp[rows] - ref) <= 1e-12 * mass[rows])
