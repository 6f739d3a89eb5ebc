"""Ragged batched kernel products through the C ABI (no reference
counterpart): both work splits, SLABS (entries evaluated per target row) and
PAIRS (each off-diagonal entry of a pair of 512-point tiles evaluated once
for both of its rows), against the oracle's product, and bitwise the same
alone and in a batch.

Bars, per row i and column r, against the oracle's FP64 product y and
a_i = sum_j |K_ij| |v_j| + nugget |v_i|: FP64 entries |y_gpu - y| <= 1e-12 a_i;
FP32 entries 2e-5 a_i (FP32 coordinates, entries and tile sums); screened
products additionally n 2^-drop max|v| (every dropped entry is below 2^-drop).
"""
import ctypes as C

import numpy as np
import pytest

from mpla_amd import _native
from tests.gpu_util import dev, empty, host, require_gpu, stream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    require_gpu()
    return _native.mplak()


def ck(status):
    _native.check(status)


class System:
    """Host description, device buffers and the planned record of one system."""

    def __init__(self, lib, n, dim, kind, nrhs, precision, seed, drop_log2=0.0, nugget=0.75, sort=False):
        import torch
        rng = np.random.default_rng(seed)
        pts = rng.random((dim, n))
        if sort:  # spatially ordered along the first axis, so screening has blocks to drop
            pts = pts[:, np.argsort(pts[0], kind="stable")]
        self.pts = np.ascontiguousarray(pts)
        self.V = rng.standard_normal((n, nrhs))
        self.n, self.dim, self.kind, self.nrhs, self.nugget = n, dim, kind, nrhs, nugget
        self.ls = 0.05 if kind == _native.MPLAK_KERNEL_GAUSSIAN else 0.1
        self.P = dev(self.pts.reshape(-1))
        self.Vd = dev(self.V.reshape(-1, order="F"))
        self.Y = empty(n * nrhs)
        s = _native.mplak_ragged_system()
        s.points, s.V, s.Y = self.P.data_ptr(), self.Vd.data_ptr(), self.Y.data_ptr()
        s.n, s.dim, s.kind, s.nrhs, s.precision = n, dim, kind, nrhs, precision
        s.length_scale, s.nugget, s.drop_log2 = self.ls, nugget, drop_log2
        nbytes = lib.mplak_ragged_plan(C.byref(s))
        assert nbytes > 0
        self.part = torch.empty(nbytes // 8 + 1, dtype=torch.float64, device="cuda")
        s.part = self.part.data_ptr()
        self.rec = s

    def reference(self, oracle):
        k = oracle.Kernel(self.pts, 0 if self.kind == _native.MPLAK_KERNEL_GAUSSIAN else 1, self.ls,
                          nugget=self.nugget)
        ys = np.stack([k.matvec(self.V[:, r]) for r in range(self.nrhs)], axis=1)
        scale = np.stack([k.matvec(np.abs(self.V[:, r])) for r in range(self.nrhs)], axis=1)
        return ys, scale


def run_batch(lib, systems):
    """One product launch per (scheme, rt, dim, precision) class of the batch,
    then one finish launch; returns each system's Y as an (n, nrhs) array."""
    import torch
    recs = (_native.mplak_ragged_system * len(systems))(*[s.rec for s in systems])
    sys_d = torch.frombuffer(bytearray(bytes(recs)), dtype=torch.uint8).to("cuda")
    classes = {}
    for k, s in enumerate(systems):
        key = (s.rec.scheme, s.rec.rt, s.rec.dim, s.rec.precision)
        classes.setdefault(key, []).append(k)
    keep = []
    for (scheme, rt, dim, prec), members in sorted(classes.items()):
        items = []
        for k in members:
            cnt = lib.mplak_ragged_product_items(C.byref(systems[k].rec), None, k)
            arr = (_native.mplak_ragged_item * cnt)()
            assert lib.mplak_ragged_product_items(C.byref(systems[k].rec), arr, k) == cnt
            items.extend(arr)
        it = (_native.mplak_ragged_item * len(items))(*items)
        it_d = torch.frombuffer(bytearray(bytes(it)), dtype=torch.uint8).to("cuda")
        keep.append(it_d)
        print(f"ragged product: scheme {scheme} rt {rt} dim {dim} precision {prec} items {len(items)}", flush=True)
        ck(lib.mplak_ragged_product(sys_d.data_ptr(), it_d.data_ptr(), len(items), rt, dim, prec, scheme, stream()))
    fin = []
    for k, s in enumerate(systems):
        cnt = lib.mplak_ragged_finish_items(C.byref(s.rec), None, k)
        arr = (_native.mplak_ragged_item * cnt)()
        lib.mplak_ragged_finish_items(C.byref(s.rec), arr, k)
        fin.extend(arr)
    fa = (_native.mplak_ragged_item * len(fin))(*fin)
    fin_d = torch.frombuffer(bytearray(bytes(fa)), dtype=torch.uint8).to("cuda")
    ck(lib.mplak_ragged_finish(sys_d.data_ptr(), fin_d.data_ptr(), len(fin), stream()))
    out = [host(s.Y).reshape((s.n, s.nrhs), order="F").copy() for s in systems]
    del keep
    return out


G, M = _native.MPLAK_KERNEL_GAUSSIAN, _native.MPLAK_KERNEL_MATERN32
F64, F32 = _native.MPLAK_PREC_F64, _native.MPLAK_PREC_F32


@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("n,dim,kind,nrhs,pairs", [
    (9000, 3, M, 5, None),   # PAIRS by size: 18 tiles, last one partial
    (1000, 2, G, 3, "1"),    # PAIRS forced: two tiles, 488 points in the second
    (512, 1, M, 1, "1"),     # PAIRS forced: one tile, the diagonal pair only
    (2600, 4, G, 10, "1"),   # two column tiles (rt 8), four dimensions
    (700, 3, G, 2, None),    # SLABS by size
    (9000, 2, G, 8, "0"),    # SLABS forced
])
def test_ragged_product_vs_oracle(lib, oracle, monkeypatch, prec, n, dim, kind, nrhs, pairs):
    if pairs is None:
        monkeypatch.delenv("MPLAK_RAGGED_PAIRS", raising=False)
    else:
        monkeypatch.setenv("MPLAK_RAGGED_PAIRS", pairs)
    s = System(lib, n, dim, kind, nrhs, prec, seed=n + dim)
    want = _native.MPLAK_RAGGED_PAIRS if (pairs == "1" or (pairs is None and n >= 8192)) else _native.MPLAK_RAGGED_SLABS
    assert s.rec.scheme == want
    print(f"case n={n} dim={dim} kind={kind} nrhs={nrhs} prec={prec} scheme={s.rec.scheme}", flush=True)
    (Y,) = run_batch(lib, [s])
    ref, scale = s.reference(oracle)
    bar = 1e-12 if prec == F64 else 2e-5
    err = np.abs(Y - ref) / scale
    assert err.max() <= bar, (err.max(), np.unravel_index(err.argmax(), err.shape))


@pytest.mark.parametrize("prec", [F64, F32])
def test_ragged_pairs_screened(lib, oracle, monkeypatch, prec):
    """Screened PAIRS on spatially ordered points: some pairs drop out whole
    (flag 0) and the error stays within the dropped entries' bound."""
    monkeypatch.setenv("MPLAK_RAGGED_PAIRS", "1")
    drop = 20.0
    s = System(lib, 6000, 2, G, 3, prec, seed=41, drop_log2=drop, sort=True)
    (Y,) = run_batch(lib, [s])
    ref, scale = s.reference(oracle)
    bar = (1e-12 if prec == F64 else 2e-5) * scale + s.n * 2.0 ** -drop * np.abs(s.V).max(axis=0)
    assert np.all(np.abs(Y - ref) <= bar)
    import torch
    pairs = s.rec.tiles * (s.rec.tiles + 1) // 2
    es = 8 if prec == F64 else 4
    off = -(-2 * pairs * s.nrhs * 512 * es // 256) * 256
    flags = host(s.part.view(torch.uint8)[off:off + 4 * pairs].view(torch.int32))
    assert 0 < flags.sum() < pairs  # ordered points: far tile pairs are skipped whole


@pytest.mark.parametrize("prec", [F64, F32])
def test_ragged_pairs_alone_equals_batch(lib, monkeypatch, prec):
    """A system's product is bitwise the same alone and in a batch with others
    of its class and of other classes."""
    monkeypatch.setenv("MPLAK_RAGGED_PAIRS", "1")
    specs = [(3000, 3, M, 4), (1500, 3, M, 3), (2048, 3, G, 4), (900, 2, G, 1)]
    batch = [System(lib, n, d, k, r, prec, seed=7 + i) for i, (n, d, k, r) in enumerate(specs)]
    together = run_batch(lib, batch)
    for i, (n, d, k, r) in enumerate(specs):
        (alone,) = run_batch(lib, [System(lib, n, d, k, r, prec, seed=7 + i)])
        assert np.array_equal(alone, together[i]), f"system {i} differs alone vs in the batch"
