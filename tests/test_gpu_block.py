"""Batched CG over many right-hand sides and the block mixed-precision solver
(no reference counterpart) against per-column FP64 oracles.

Bars: batched FP64 CG iteration counts equal to the single-column oracle
restatement within +-1 per column; every column's true FP64 relative residual
below epsilon (recomputed on the host with the oracle's product); solution
error against the FP64 oracle solution below 1e-6 on the stated columns.
"""
import ctypes as C

import numpy as np
import pytest

from mpla_amd import _native, mpla
from tests.dropin_util import Points, Session, from_device, to_device, DeviceBuffer
from tests.gpu_util import dev, empty, host, point_table, require_gpu, stream, workspace

pytestmark = pytest.mark.gpu


def kdesc(T, dim, n, kind, ell, nugget):
    d = _native.mplak_kernel_desc()
    d.points, d.dim, d.kind, d.length_scale, d.nugget, d.coincident_shift = T.data_ptr(), dim, kind, ell, nugget, 0.0
    d.row_offset, d.row_count, d.col_offset, d.col_count = 0, n, 0, n
    return d


def test_batched_cg_c_abi_matches_per_column_oracle(oracle):
    require_gpu()
    import torch
    lib = _native.mplak()
    n, R = 1200, 7
    pts = oracle.uniform(3 * n, 21, 0).reshape(3, n)
    Bh = oracle.uniform(n * R, 21, 1, -1, 1).reshape((n, R), order="F")
    Bh[:, 3] *= 1e-3  # columns with different scales converge at different iterations
    P = dev(pts)
    T = point_table(P)
    d = kdesc(T, 3, n, 1, 0.1, 0.5)
    B = dev(Bh.ravel(order="F"))
    X = dev(np.zeros(n * R))
    Rr, D, Z = empty(n * R), empty(n * R), empty(n * R)
    ssz = C.sizeof(_native.mplak_cg_state)
    st = torch.zeros((2 * R * ssz) // 8 + 1, dtype=torch.float64, device="cuda")
    grp = torch.zeros(4, dtype=torch.float64, device="cuda")
    s0, s1 = st.data_ptr(), st.data_ptr() + R * ssz
    wb = lib.mplak_bcg_workspace_bytes(n, R)
    w = workspace(wb)
    pw = lib.mplak_kernel_mrhs_workspace_bytes(C.byref(d), 0, R)
    pwork = workspace(pw)
    s = stream()
    Z.zero_()
    _native.check(lib.mplak_bcg_start(n, R, B.data_ptr(), Z.data_ptr(), Rr.data_ptr(), D.data_ptr(), s0,
                                      grp.data_ptr(), 1, 1e-10, 2001, float(n), w.data_ptr(), wb, s))
    all_done = grp.data_ptr()
    for j in range(2000):
        sin, sout = (s0, s1) if j % 2 == 0 else (s1, s0)
        _native.check(lib.mplak_kernel_mrhs(C.byref(d), 0, R, D.data_ptr(), n, Z.data_ptr(), n, pwork.data_ptr(), pw,
                                            all_done, s))
        _native.check(lib.mplak_bcg_iterate(n, R, X.data_ptr(), Rr.data_ptr(), D.data_ptr(), Z.data_ptr(), sin, sout,
                                            grp.data_ptr(), w.data_ptr(), wb, s))
        if j % 50 == 49:
            torch.cuda.synchronize()
            if int(grp.view(torch.int32)[0].item()) == 1:
                break
    torch.cuda.synchronize()
    final = s1 if (j + 1) % 2 == 1 else s0
    raw = st.cpu().numpy().tobytes()
    off = final - s0
    states = [_native.mplak_cg_state.from_buffer_copy(raw[off + c * ssz: off + (c + 1) * ssz]) for c in range(R)]
    Xh = host(X).reshape((n, R), order="F")
    k = oracle.Kernel(pts, 1, 0.1, nugget=0.5)
    for c in range(R):
        xo, info = oracle.cg_relative(k.op(), Bh[:, c], np.zeros(n), 2001, 1e-10)
        assert states[c].done == _native.MPLAK_CONVERGED
        assert abs(states[c].k - info.iterations) <= 1, (c, states[c].k, info.iterations)
        assert np.linalg.norm(Xh[:, c] - xo) <= 1e-8 * np.linalg.norm(xo)


_ORACLE_X = {}


@pytest.mark.parametrize("prec", [mpla.MPLA_PREC_BF16X3, mpla.MPLA_PREC_BF16, mpla.MPLA_PREC_F32,
                                  mpla.MPLA_PREC_F64])
def test_block_mixed_solver_c3_shape(oracle, prec):
    """Matern-3/2 multi-RHS system at N = 4096: l = 0.1, nugget 0.5, 16 RHS ~ N(0,1), 1e-8."""
    require_gpu()
    s = Session(verbosity=0)
    try:
        n, R = 4096, 16
        pts = oracle.uniform(3 * n, 1235, 0).reshape(3, n)
        Bh = oracle.normal(n * R, 1235, 1).reshape((n, R), order="F")
        Pp = Points(pts)
        kd = mpla.mpla_kernel_matrix_data(Pp.ptr, 3, 0, mpla.MPLA_KERNEL_MATERN32, mpla.MPLA_PREC_F64, 0.1, 0.5, 0.0)
        A = s.generic(n, n, C.addressof(kd))
        Bd = DeviceBuffer(8 * n * R)
        Xd = DeviceBuffer(8 * n * R)
        to_device(Bd.ptr, Bh.ravel(order="F"))
        to_device(Xd.ptr, np.zeros(n * R))
        rel = (C.c_double * R)()
        s.lib.mpla_kernel_block_conjugate_gradient_mixed(Bd.ptr, Xd.ptr, R, C.byref(A), 20000, 1e-8, prec, rel, s.I)
        Xh = from_device(Xd.ptr, n * R).reshape((n, R), order="F")
        assert s.inst.last_converged == 1
        assert max(rel) < 1e-8
        k = oracle.Kernel(pts, 1, 0.1, nugget=0.5)
        for c in (0, 5, 10, 15):
            true_rel = np.linalg.norm(Bh[:, c] - k.matvec(Xh[:, c])) / np.linalg.norm(Bh[:, c])
            assert true_rel < 1e-8 * 1.01
            assert abs(true_rel - rel[c]) <= 1e-2 * true_rel + 1e-16
            if c not in _ORACLE_X:
                _ORACLE_X[c] = oracle.cg_relative(k.op(), Bh[:, c], np.zeros(n), 20000, 1e-12)[0]
            xo = _ORACLE_X[c]
            assert np.linalg.norm(Xh[:, c] - xo) <= 1e-6 * np.linalg.norm(xo)
    finally:
        s.close()
