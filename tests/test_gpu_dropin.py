"""The drop-in mpla.h API on the GPU, exercised the way the reference's own
drivers exercise MPLA (src/test_library.cpp, src/test_cg.cpp,
src/test_kernel_cg.cpp, src/benchmark_generic_system_cg.cpp,
src/generate_quadrature_weights.cpp), checked against the CPU oracle on the
same inputs.

Bars: exact for the integer patterns; 1e-12 relative for products; solvers
reach the same stopping test with iteration counts within +-1 (or 2 %) of the
oracle's restatement and solutions within the conditioning-scaled tolerance
stated per test.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from mpla_amd import mpla
from tests.dropin_util import Points, Session, from_device
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture()
def s():
    require_gpu()
    sess = Session(verbosity=0)
    yield sess
    sess.close()


def iters_close(a, b):
    return abs(a - b) <= max(1, int(0.02 * max(a, b)))


# --------------------------------------------------------------------------
# src/test_library.cpp: dgemv + ddot through mpla_generic_dgemv, exact


@pytest.mark.parametrize("m,n", [(100, 100), (1000, 700), (4096, 4096)])
def test_test_library_generic_dgemv_and_ddot(s, m, n):
    lib = s.lib
    A = s.matrix(m, n, np.arange(m * n, dtype=np.float64).reshape((m, n), order="F"))
    x = s.vector(n, np.arange(n))
    y = s.vector(n, np.arange(n) + 1)
    Ax = s.vector(m)
    core = mpla.core("mpla_generic_dgemv_core_cublas")
    xy = C.c_double()
    for _ in range(1, 100):
        lib.mpla_generic_dgemv(C.byref(Ax), s.as_generic(A), C.byref(x), core, s.I)
        lib.mpla_ddot(C.byref(xy), C.byref(x), C.byref(y), s.I)
    i = np.arange(m, dtype=np.float64)
    expect = m * (n - 1) * n * (2 * n - 1) / 6 + i * n * (n - 1) / 2
    assert np.array_equal(s.get(Ax), expect)
    assert xy.value == (n - 1) * n * (n + 1) / 3


def test_mpla_dgemv_stored_random(s, oracle):
    m, n = 3000, 2500
    Ah = oracle.uniform(m * n, 1234, 1, -1, 1).reshape((m, n), order="F")
    xh = oracle.uniform(n, 1234, 2, -1, 1)
    A = s.matrix(m, n, Ah)
    x = s.vector(n, xh)
    b = s.vector(m)
    s.lib.mpla_dgemv(C.byref(b), C.byref(A), C.byref(x), s.I)
    ref = oracle.dgemv(Ah, xh, mode=1)
    assert np.all(np.abs(s.get(b) - ref) <= 1e-12 * (np.abs(Ah) @ np.abs(xh)))


def test_daxpy_set_zero_redistribute(s):
    n = 10001
    x = s.vector(n, np.arange(n))
    y = s.vector(n, np.ones(n))
    s.lib.mpla_daxpy(C.byref(y), -0.5, C.byref(x), s.I)
    assert np.array_equal(s.get(y), 1.0 - 0.5 * np.arange(n))
    s.lib.mpla_vector_set_zero(C.byref(y), s.I)
    assert np.all(s.get(y) == 0.0)
    A = s.matrix(n, n)
    xr = mpla.mpla_vector()
    s.lib.mpla_init_vector_for_block_rows(C.byref(xr), s.I, n)
    s._vectors.append(xr)
    s.lib.mpla_redistribute_vector_for_dgesv(C.byref(xr), C.byref(x), C.byref(A), s.I)
    assert np.array_equal(s.get(xr), np.arange(n))


def test_instance_grid_and_distribution_fields(s):
    inst = s.inst
    assert inst.proc_count == 1 and inst.proc_rows == 1 and inst.proc_cols == 1 and inst.is_parent
    A = s.matrix(300, 200)
    assert (A.cur_proc_row_count, A.cur_proc_col_count, A.cur_proc_row_offset, A.cur_proc_col_offset) == (300, 200, 0, 0)
    assert A.proc_row_count[0][0] == 300 and A.proc_col_count[0][0] == 200
    v = s.vector(77)
    assert v.cur_proc_row_count == 77 and v.cur_proc_row_offset == 0


# --------------------------------------------------------------------------
# src/test_cg.cpp: 1-D Laplacian, manufactured RHS, random x0, RMS rule 1e-10


def laplacian(n):
    return np.diag(2.0 * np.ones(n)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)


@pytest.mark.parametrize("n", [64, 500])
def test_test_cg_laplacian(s, oracle, n):
    lib = s.lib
    Ah = laplacian(n)
    A = s.matrix(n, n, Ah)
    xt = oracle.uniform(n, 1235, 0)
    x = s.vector(n, xt)
    b = s.vector(n)
    core = mpla.core("mpla_generic_dgemv_core_cublas")
    lib.mpla_generic_dgemv(C.byref(b), s.as_generic(A), C.byref(x), core, s.I)
    x0 = oracle.uniform(n, 1235, 1)
    s.set(x, x0)
    lib.mpla_generic_conjugate_gradient(C.byref(b), s.as_generic(A), C.byref(x), 1000, 1e-10, core, s.I)
    xs = s.get(x)
    bh = s.get(b)
    assert np.sqrt(np.mean((bh - Ah @ xs) ** 2)) < 1e-10 * 1.001
    xo, info = oracle.cg_rms(oracle.dense_op(Ah), Ah @ xt, x0, 1000, 1e-10)
    assert s.inst.last_converged == 1 and iters_close(s.inst.last_iterations, info.iterations)
    assert np.linalg.norm(xs - xt) <= 1e-6 * np.linalg.norm(xt) * n


def dense_spd(oracle, n, seed=1235):
    B = oracle.uniform(n * n, seed, 10, -1, 1).reshape((n, n), order="F")
    return B @ B.T / n + np.eye(n)


@pytest.mark.parametrize("n", [512, 4096])
def test_checkpoint_cg_dense_spd(s, oracle, n):
    """Dense SPD system: A = B B^T/N + I, b ~ U[-1,1], x0 = 0, relative 1e-10, max 2000 iterations."""
    lib = s.lib
    Ah = dense_spd(oracle, n)
    bh = oracle.uniform(n, 1235, 11, -1, 1)
    A = s.matrix(n, n, Ah)
    b = s.vector(n, bh)
    x = s.vector(n, np.zeros(n))
    core = mpla.core("mpla_generic_dgemv_core_dense")
    lib.mpla_generic_conjugate_gradient_with_checkpoint_restart(C.byref(b), s.as_generic(A), C.byref(x), 2001,
                                                                1e-10, core, 0, False, s.I)
    xs = s.get(x)
    xo, info = oracle.cg_relative(oracle.dense_op(Ah), bh, np.zeros(n), 2001, 1e-10)
    assert s.inst.last_converged == 1 and info.converged == 1
    assert iters_close(s.inst.last_iterations, info.iterations)
    assert np.linalg.norm(bh - Ah @ xs) / np.linalg.norm(bh) < 1e-10 * 1.001
    assert np.linalg.norm(xs - xo) <= 1e-9 * np.linalg.norm(xo)


def test_checkpoint_restart_files_and_resume(s, oracle, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    lib = s.lib
    n = 400
    Ah = laplacian(n) + 0.05 * np.eye(n)
    bh = oracle.uniform(n, 7, 0, -1, 1)
    A = s.matrix(n, n, Ah)
    b = s.vector(n, bh)
    x = s.vector(n, np.zeros(n))
    core = mpla.core("mpla_generic_dgemv_core_dense")
    # 20 iterations with a checkpoint every 5 (odd chunks exercise the state copy)
    lib.mpla_generic_conjugate_gradient_with_checkpoint_restart(C.byref(b), s.as_generic(A), C.byref(x), 21, 1e-12,
                                                                core, 5, False, s.I)
    assert s.inst.last_converged == 0 and s.inst.last_iterations == 21
    odir = tmp_path / "oracle"
    odir.mkdir()
    oracle.cg_relative(oracle.dense_op(Ah), bh, np.zeros(n), 21, 1e-12, ckpt_every=5, directory=str(odir))
    for name in ("x.dat", "r.dat", "d.dat", "z.dat"):
        raw = (tmp_path / name).read_bytes()
        assert len(raw) == 4 + 8 * n and int.from_bytes(raw[:4], "little") == n
        g = np.frombuffer(raw[4:], dtype=np.float64)
        o = np.frombuffer((odir / name).read_bytes()[4:], dtype=np.float64)
        assert np.linalg.norm(g - o) <= 1e-9 * np.linalg.norm(o), name
    # z.dat holds the new direction, as the reference writes it
    assert (tmp_path / "z.dat").read_bytes() == (tmp_path / "d.dat").read_bytes()
    # resume from the files: total work matches an uninterrupted run
    s.set(x, np.zeros(n))
    lib.mpla_generic_conjugate_gradient_with_checkpoint_restart(C.byref(b), s.as_generic(A), C.byref(x), 2000, 1e-10,
                                                                core, 0, True, s.I)
    resumed = s.inst.last_iterations
    x_res = s.get(x)
    xo, info = oracle.cg_relative(oracle.dense_op(Ah), bh, np.zeros(n), 2000, 1e-10)
    assert abs(20 + resumed - info.iterations) <= 2
    assert np.linalg.norm(bh - Ah @ x_res) / np.linalg.norm(bh) < 1e-10 * 1.001


def test_load_vector_length_mismatch_exits(tmp_path):
    """The loader prints and exits(1) on a wrong length (src/mpla.cpp:695-700);
    run in a child process."""
    import subprocess
    import sys
    require_gpu()
    path = tmp_path / "v.dat"
    path.write_bytes((5).to_bytes(4, "little") + np.zeros(5).tobytes())
    code = f"""
import ctypes as C, sys
sys.path.insert(0, {repr(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))})
from tests.dropin_util import Session
s = Session()
v = s.vector(4)
s.lib.mpla_load_vector(C.byref(v), {repr(str(path))}.encode(), s.I)
print("not reached")
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "does not fit" in r.stdout and "not reached" not in r.stdout


# --------------------------------------------------------------------------
# BiCGSTAB (src/mpla.cpp:1003-1183), diagonally dominant dense system at reduced N


def nonsym(oracle, n, seed=1235):
    R = oracle.uniform(n * n, seed, 20, -1, 1).reshape((n, n), order="F") / math.sqrt(n)
    return R + np.diag(2.0 + oracle.uniform(n, seed, 21, 0, 1))


@pytest.mark.parametrize("n", [300, 2048])
def test_bicgstab_dense(s, oracle, n):
    lib = s.lib
    Ah = nonsym(oracle, n)
    bh = oracle.uniform(n, 1235, 22, -1, 1)
    A = s.matrix(n, n, Ah)
    b = s.vector(n, bh)
    x = s.vector(n, np.zeros(n))
    core = mpla.core("mpla_generic_dgemv_core_dense")
    lib.mpla_generic_BiCGSTAB_with_checkpoint_restart(C.byref(b), s.as_generic(A), C.byref(x), 1000, 1e-10, core,
                                                      0, False, s.I)
    xs = s.get(x)
    xo, info = oracle.bicgstab(oracle.dense_op(Ah), bh, np.zeros(n), 1000, 1e-10)
    assert s.inst.last_converged == 1 and info.converged == 1
    assert iters_close(s.inst.last_iterations, info.iterations)
    assert s.inst.last_matvecs == 1 + 3 * s.inst.last_iterations
    assert np.linalg.norm(bh - Ah @ xs) / np.linalg.norm(bh) < 1e-10 * 1.001
    assert np.linalg.norm(xs - xo) <= 1e-8 * np.linalg.norm(xo)


def test_bicgstab_checkpoint_restart(s, oracle, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    lib = s.lib
    n = 600
    Ah = nonsym(oracle, n, seed=9)
    bh = oracle.uniform(n, 9, 1, -1, 1)
    A = s.matrix(n, n, Ah)
    b = s.vector(n, bh)
    x = s.vector(n, np.zeros(n))
    core = mpla.core("mpla_generic_dgemv_core_dense")
    lib.mpla_generic_BiCGSTAB_with_checkpoint_restart(C.byref(b), s.as_generic(A), C.byref(x), 5, 1e-14, core, 2,
                                                      False, s.I)
    for name in ("x", "r", "r_hat", "z", "zz", "s", "p"):
        assert (tmp_path / f"{name}.dat").stat().st_size == 4 + 8 * n
    s.set(x, np.zeros(n))
    lib.mpla_generic_BiCGSTAB_with_checkpoint_restart(C.byref(b), s.as_generic(A), C.byref(x), 1000, 1e-10, core, 0,
                                                      True, s.I)
    assert s.inst.last_converged == 1
    xs = s.get(x)
    assert np.linalg.norm(bh - Ah @ xs) / np.linalg.norm(bh) < 1e-10 * 1.001


# --------------------------------------------------------------------------
# kernel plugin (src/kernel.cu, src/test_kernel_cg.cpp)


def test_kernel_core_generic_dgemv_matches_oracle(s, oracle):
    n = 3000
    pts = oracle.uniform(3 * n, 1235, 0).reshape(3, n)
    P = Points(pts)
    kd = mpla.kernel_matrix_data(P.ptr, 3, 123)
    A = s.generic(n, n, C.addressof(kd))
    xh = oracle.uniform(n, 1235, 1)
    x = s.vector(n, xh)
    b = s.vector(n)
    for name in ("mpla_dgemv_core_kernel_cublas", "mpla_dgemv_core_kernel_streamed_cublas"):
        s.lib.mpla_generic_dgemv(C.byref(b), C.byref(A), C.byref(x), mpla.core(name), s.I)
        ref = oracle.Kernel(pts, oracle.GAUSSIAN, 1 / math.sqrt(2.0), coincident_shift=1e-8).matvec(xh, mode=1)
        assert np.linalg.norm(s.get(b) - ref) <= 1e-12 * np.linalg.norm(ref)
    s.lib.mpla_set_gaussian_kernel_rhs(C.byref(b), C.byref(A), s.I)
    assert np.allclose(s.get(b), oracle.Kernel(pts).gaussian_rhs(), rtol=1e-14)


def test_test_kernel_cg_manufactured(s, oracle):
    """src/test_kernel_cg.cpp: manufactured b = K x_rand, random x0, RMS 1e-10."""
    n = 2000
    pts = oracle.uniform(3 * n, 1235, 0).reshape(3, n)
    P = Points(pts)
    kd = mpla.kernel_matrix_data(P.ptr, 3, 0)
    A = s.generic(n, n, C.addressof(kd))
    core = mpla.core("mpla_dgemv_core_kernel_streamed_cublas")
    xt = oracle.uniform(n, 1235, 1)
    x = s.vector(n, xt)
    b = s.vector(n)
    s.lib.mpla_generic_dgemv(C.byref(b), C.byref(A), C.byref(x), core, s.I)
    x0 = oracle.uniform(n, 1235, 2)
    s.set(x, x0)
    s.lib.mpla_generic_conjugate_gradient(C.byref(b), C.byref(A), C.byref(x), 1000, 1e-10, core, s.I)
    k = oracle.Kernel(pts, oracle.GAUSSIAN, 1 / math.sqrt(2.0), coincident_shift=1e-8)
    bh = s.get(b)
    xs = s.get(x)
    _, info = oracle.cg_rms(k.op(), bh, x0, 1000, 1e-10)
    # K + 1e-8 I is badly conditioned (kappa ~ 1e11): both runs must agree on
    # the outcome; iteration counts within 10 %
    assert s.inst.last_converged == info.converged
    if info.converged:
        assert np.sqrt(np.mean((bh - k.matvec(xs)) ** 2)) < 1e-10 * 1.05
        assert abs(s.inst.last_iterations - info.iterations) <= max(3, info.iterations // 10)


def gaussian_system(oracle, n, ell=0.05, nugget=1.0, kind=0, seed=1235):
    pts = oracle.uniform(3 * n, seed, 0).reshape(3, n)
    bh = oracle.uniform(n, seed, 1, -1, 1)
    return pts, bh, oracle.Kernel(pts, kind, ell, nugget=nugget)


def test_fused_core_plain_fp64_cg(s, oracle):
    """Gaussian operator at N=8192: (K + 1.0 I) x = b, l=0.05, relative 1e-8."""
    n = 8192
    pts, bh, k = gaussian_system(oracle, n)
    P = Points(pts)
    kd = mpla.mpla_kernel_matrix_data(P.ptr, 3, 0, mpla.MPLA_KERNEL_GAUSSIAN, mpla.MPLA_PREC_F64, 0.05, 1.0, 0.0)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n, bh)
    x = s.vector(n, np.zeros(n))
    s.lib.mpla_generic_conjugate_gradient_with_checkpoint_restart(
        C.byref(b), C.byref(A), C.byref(x), 2001, 1e-8, mpla.core("mpla_dgemv_core_kernel_fused"), 0, False, s.I)
    xs = s.get(x)
    xo, info = oracle.cg_relative(k.op(), bh, np.zeros(n), 2001, 1e-8)
    assert s.inst.last_converged == 1 and iters_close(s.inst.last_iterations, info.iterations)
    assert np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh) < 1e-8 * 1.001
    assert np.linalg.norm(xs - xo) <= 1e-7 * np.linalg.norm(xo)


@pytest.mark.parametrize("n,dim,kind,ell,nugget", [(8192, 3, 0, 0.05, 1.0), (5000, 2, 1, 0.1, 0.5),
                                                   (300, 1, 0, 0.05, 1.0)])
def test_ordered_plain_fp64_cg(s, oracle, n, dim, kind, ell, nugget):
    """mpla_kernel_conjugate_gradient_mixed with MPLA_PREC_F64: plain FP64 CG
    (src/mpla.cpp:887-1001) on Hilbert-ordered points with the FP64 product
    screened at 2^-(53 + log2 n).  Same iterations as the oracle's CG on the
    exact operator (+-2 %), true residual below the tolerance, solution within
    1e-7 of the oracle's, in the caller's order."""
    pts = oracle.uniform(dim * n, 1300 + n, 0).reshape(dim, n)
    bh = oracle.uniform(n, 1300 + n, 1, -1, 1)
    k = oracle.Kernel(pts, kind, ell, nugget=nugget)
    P = Points(pts)
    kd = mpla.mpla_kernel_matrix_data(P.ptr, dim, 0, kind, mpla.MPLA_PREC_F64, ell, nugget, 0.0)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n, bh)
    x = s.vector(n, np.zeros(n))
    s.lib.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), 2001, 1e-8, mpla.MPLA_PREC_F64,
                                               s.I)
    xs = s.get(x)
    xo, info = oracle.cg_relative(k.op(), bh, np.zeros(n), 2001, 1e-8)
    assert s.inst.last_converged == 1 and iters_close(s.inst.last_iterations, info.iterations)
    assert np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh) < 1e-8 * 1.001
    assert np.linalg.norm(xs - xo) <= 1e-7 * np.linalg.norm(xo)


# The FP64 residual products drop entries below 2^-(log2 n + log2(1/nugget) +
# log2(1/eps) + 12): the smaller nugget moves that screen by 3.3 bits, and the
# reported residual must still be the oracle's to 1e-3 of itself.
@pytest.mark.parametrize("kind,ell,nugget", [(0, 0.05, 1.0), (1, 0.1, 0.5), (0, 0.05, 0.1)])
def test_mixed_precision_ir_reaches_fp64_residual(s, oracle, kind, ell, nugget):
    n = 8192
    pts, bh, k = gaussian_system(oracle, n, ell, nugget, kind)
    P = Points(pts)
    kd = mpla.mpla_kernel_matrix_data(P.ptr, 3, 0, kind, mpla.MPLA_PREC_F64, ell, nugget, 0.0)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n, bh)
    x = s.vector(n, np.zeros(n))
    s.lib.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), 5000, 1e-8, mpla.MPLA_PREC_F32,
                                               s.I)
    xs = s.get(x)
    true_rel = np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh)
    assert s.inst.last_converged == 1
    assert true_rel < 1e-8 * 1.01
    assert abs(true_rel - s.inst.last_rel_res) <= 1e-3 * true_rel + 1e-15
    xo, _ = oracle.cg_relative(k.op(), bh, np.zeros(n), 5000, 1e-12)
    assert np.linalg.norm(xs - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("shift", [1e-8, -1e-8])
def test_mixed_precision_with_coincidence_shift(s, oracle, shift):
    """A non-zero coincidence shift (the reference core's +1e-8 where points
    coincide, src/kernel.cu:45-48, here also negative) has no planned inner
    product: the mixed solve takes the unordered path and still meets the true
    FP64 residual of the shifted operator, checked by the oracle."""
    n = 4096
    pts, bh, _ = gaussian_system(oracle, n)
    pts[:, 1::97] = pts[:, 0:1]  # coincident points, so the shift applies off the diagonal too
    k = oracle.Kernel(pts, 0, 0.05, nugget=1.0, coincident_shift=shift)
    P = Points(pts)
    kd = mpla.mpla_kernel_matrix_data(P.ptr, 3, 0, mpla.MPLA_KERNEL_GAUSSIAN, mpla.MPLA_PREC_F64, 0.05, 1.0, shift)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n, bh)
    x = s.vector(n, np.zeros(n))
    s.lib.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), 5000, 1e-8, mpla.MPLA_PREC_F32,
                                               s.I)
    xs = s.get(x)
    true_rel = np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh)
    assert s.inst.last_converged == 1
    assert true_rel < 1e-8 * 1.01
    assert abs(true_rel - s.inst.last_rel_res) <= 1e-3 * true_rel + 1e-15
    xo, _ = oracle.cg_relative(k.op(), bh, np.zeros(n), 5000, 1e-12)
    assert np.linalg.norm(xs - xo) <= 1e-6 * np.linalg.norm(xo)


# Typical interpolation nuggets: the residual screen's ||x|| <= ||b||/nugget
# bound is loosest here and the FP32 inner operator may lose definiteness
# (the inner solve then stops at its breakdown, or refinement stagnates and
# the solve finishes in FP64).  Either way the reported residual is the true
# one and meets the tolerance.
@pytest.mark.parametrize("nugget", [1e-3, 1e-5])
def test_mixed_precision_small_nugget(s, oracle, nugget):
    n = 8192
    pts, bh, k = gaussian_system(oracle, n, 0.05, nugget)
    P = Points(pts)
    kd = mpla.mpla_kernel_matrix_data(P.ptr, 3, 0, mpla.MPLA_KERNEL_GAUSSIAN, mpla.MPLA_PREC_F64, 0.05, nugget, 0.0)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n, bh)
    x = s.vector(n, np.zeros(n))
    s.lib.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), 20000, 1e-8,
                                               mpla.MPLA_PREC_F32, s.I)
    xs = s.get(x)
    assert np.all(np.isfinite(xs))
    true_rel = np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh)
    assert s.inst.last_converged == 1
    assert true_rel < 1e-8 * 1.01
    assert abs(true_rel - s.inst.last_rel_res) <= 1e-2 * true_rel + 1e-15


@pytest.mark.parametrize("n,dim,kind,ell,nugget", [(1, 3, 0, 0.05, 1.0), (2, 1, 1, 0.3, 0.5), (300, 2, 0, 0.05, 1.0),
                                                   (513, 3, 1, 0.1, 0.5), (3000, 1, 0, 0.01, 1.0),
                                                   (5000, 2, 1, 0.05, 0.75)])
def test_mixed_precision_edge_sizes_and_dims(s, oracle, n, dim, kind, ell, nugget):
    """The ordered, screened mixed path at sizes below one tile, ragged last
    tiles, dims 1-3, both kernels and a non-zero initial guess: true FP64
    residual below 1e-8 and the FP64 oracle solution to 1e-6."""
    pts = oracle.uniform(dim * n, 900 + n, 0).reshape(dim, n)
    bh = oracle.uniform(n, 900 + n, 1, -1, 1)
    x0 = oracle.uniform(n, 900 + n, 2, -0.1, 0.1)
    k = oracle.Kernel(pts, kind, ell, nugget=nugget)
    P = Points(pts)
    kd = mpla.mpla_kernel_matrix_data(P.ptr, dim, 0, kind, mpla.MPLA_PREC_F64, ell, nugget, 0.0)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n, bh)
    x = s.vector(n, x0)
    s.lib.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), 5000, 1e-8, mpla.MPLA_PREC_F32,
                                               s.I)
    xs = s.get(x)
    assert s.inst.last_converged == 1
    assert np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh) < 1e-8 * 1.01
    xo, _ = oracle.cg_relative(k.op(), bh, np.zeros(n), 5000, 1e-12)
    assert np.linalg.norm(xs - xo) <= 1e-6 * np.linalg.norm(xo)


def test_mixed_precision_ordered_screened_vs_exact_inner(s, oracle, monkeypatch):
    """The mixed solve runs on spatially ordered points with FP32 inner
    products that drop entries below 2^-30; with MPLA_IR_DROP_LOG2=0 it keeps
    every entry in the original order.  Both reach the FP64 bar, agree to the
    solution tolerance, and each is bitwise repeatable."""
    n = 13000  # above the symmetric-product size threshold, ragged last tile
    pts, bh, k = gaussian_system(oracle, n, 0.05, 1.0, 0)
    P = Points(pts)
    kd = mpla.mpla_kernel_matrix_data(P.ptr, 3, 0, 0, mpla.MPLA_PREC_F64, 0.05, 1.0, 0.0)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n, bh)
    x = s.vector(n)
    sols = {}
    for drop in ("30", "0"):
        monkeypatch.setenv("MPLA_IR_DROP_LOG2", drop)
        runs = []
        for _ in range(2):
            s.set(x, np.zeros(n))
            s.lib.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), 5000, 1e-8,
                                                       mpla.MPLA_PREC_F32, s.I)
            assert s.inst.last_converged == 1
            runs.append(s.get(x))
        assert np.array_equal(runs[0], runs[1])
        xs = runs[0]
        assert np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh) < 1e-8 * 1.01
        sols[drop] = xs
    assert np.linalg.norm(sols["30"] - sols["0"]) <= 1e-7 * np.linalg.norm(sols["0"])


@pytest.mark.parametrize("relax,drop", [("1", None), ("0", None), ("0", "7")])
def test_mixed_precision_relaxed_and_safeguarded_screens(s, oracle, monkeypatch, relax, drop):
    """Default screen (3 bits below the inner tolerance) with the inner
    products relaxed as the inner residual falls (MPLA_IR_RELAX=1), without
    relaxation, and a screen so loose (2^-7) that the refinement safeguard
    has to tighten it: every run reaches the FP64 bar on the oracle's
    residual and the FP64 oracle solution to 1e-6, bitwise repeatably."""
    n = 13000
    pts, bh, k = gaussian_system(oracle, n, 0.05, 1.0, 0)
    monkeypatch.setenv("MPLA_IR_RELAX", relax)
    if drop is not None:
        monkeypatch.setenv("MPLA_IR_DROP_LOG2", drop)
    P = Points(pts)
    kd = mpla.mpla_kernel_matrix_data(P.ptr, 3, 0, 0, mpla.MPLA_PREC_F64, 0.05, 1.0, 0.0)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n, bh)
    x = s.vector(n)
    runs = []
    for _ in range(2):
        s.set(x, np.zeros(n))
        s.lib.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), 5000, 1e-8,
                                                   mpla.MPLA_PREC_F32, s.I)
        assert s.inst.last_converged == 1
        runs.append(s.get(x))
    assert np.array_equal(runs[0], runs[1])
    xs = runs[0]
    assert np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh) < 1e-8 * 1.01
    xo, _ = oracle.cg_relative(k.op(), bh, np.zeros(n), 5000, 1e-12)
    assert np.linalg.norm(xs - xo) <= 1e-6 * np.linalg.norm(xo)


def test_mixed_precision_reuses_its_graph_for_new_points(s, oracle):
    """Two mixed solves of the same size and kernel in one instance, on
    different points and right-hand sides: the second replays the chunk the
    first captured (same kernel arguments, new contents behind them).  Both
    reach the FP64 bar on their own system, and a third solve of the first
    system repeats the first bitwise."""
    n = 6000
    sols = []
    systems = []
    for seed in (1301, 1302, 1301):
        pts, bh, k = gaussian_system(oracle, n, 0.05, 1.0, 0, seed=seed)
        P = Points(pts)
        kd = mpla.mpla_kernel_matrix_data(P.ptr, 3, 0, 0, mpla.MPLA_PREC_F64, 0.05, 1.0, 0.0)
        A = s.generic(n, n, C.addressof(kd))
        b = s.vector(n, bh)
        x = s.vector(n, np.zeros(n))
        s.lib.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), 5000, 1e-8,
                                                   mpla.MPLA_PREC_F32, s.I)
        assert s.inst.last_converged == 1
        xs = s.get(x)
        assert np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh) < 1e-8 * 1.01
        sols.append(xs)
        systems.append((P, kd, A, b, x))
    assert np.array_equal(sols[0], sols[2])
    assert not np.array_equal(sols[0], sols[1])


def test_mixed_precision_replays_its_setup_on_new_contents(s, oracle):
    """Three mixed solves through the same point, right-hand-side and
    solution buffers, whose contents are overwritten in place between the
    solves (system 1, system 2, system 1 again): the later solves replay the
    captured inner-iteration chunk with unchanged arguments, so everything
    behind them (order, plans, vectors) must come from the new contents.
    Each solve reaches the FP64 bar on the system it was given, and the third
    repeats the first bitwise."""
    from tests.dropin_util import to_device
    n = 6000
    systems = [gaussian_system(oracle, n, 0.05, 1.0, 0, seed=seed) for seed in (1311, 1312)]
    P = Points(systems[0][0])
    kd = mpla.mpla_kernel_matrix_data(P.ptr, 3, 0, 0, mpla.MPLA_PREC_F64, 0.05, 1.0, 0.0)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n)
    x = s.vector(n)
    sols = []
    for which in (0, 1, 0):
        pts, bh, k = systems[which]
        to_device(P.buf.ptr, np.ascontiguousarray(pts, dtype=np.float64).ravel())
        s.set(b, bh)
        s.set(x, np.zeros(n))
        s.lib.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), 5000, 1e-8,
                                                   mpla.MPLA_PREC_F32, s.I)
        assert s.inst.last_converged == 1
        xs = s.get(x)
        assert np.linalg.norm(bh - k.matvec(xs)) / np.linalg.norm(bh) < 1e-8 * 1.01, f"system {which}"
        sols.append(xs)
    assert np.array_equal(sols[0], sols[2])
    assert not np.array_equal(sols[0], sols[1])


def test_graph_replay_is_deterministic(s, oracle):
    n = 4096
    pts, bh, k = gaussian_system(oracle, n)
    P = Points(pts)
    kd = mpla.mpla_kernel_matrix_data(P.ptr, 3, 0, 0, mpla.MPLA_PREC_F64, 0.05, 1.0, 0.0)
    A = s.generic(n, n, C.addressof(kd))
    b = s.vector(n, bh)
    x = s.vector(n)
    outs = []
    for _ in range(3):
        s.set(x, np.zeros(n))
        s.lib.mpla_generic_conjugate_gradient_with_checkpoint_restart(
            C.byref(b), C.byref(A), C.byref(x), 2001, 1e-8, mpla.core("mpla_dgemv_core_kernel_fused"), 0, False, s.I)
        outs.append(s.get(x))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


# --------------------------------------------------------------------------
# virtual system assembler (src/generic_system_adapter.cu, src/kernel_system_assembler.cu)


def test_generic_system_adapter(s, oracle):
    from tests.dropin_util import DeviceBuffer, hip
    n = 1500
    pts = oracle.uniform(3 * n, 5, 0).reshape(3, n)
    P = Points(pts)
    holder = DeviceBuffer(8)  # device pointer-to-pointer, as the reference passes it
    s.lib.create_gaussian_kernel_system_assembler_object(holder.ptr, P.ptr, 1000, 3)
    assem = C.c_void_p()
    assert hip().hipMemcpy(C.byref(assem), holder.ptr, 8, 2) == 0
    A = s.generic(n, n, assem.value)
    xh = oracle.uniform(n, 5, 1, -1, 1)
    x = s.vector(n, xh)
    b = s.vector(n)
    ref = oracle.Kernel(pts, oracle.GAUSSIAN, 1 / math.sqrt(2.0)).matvec(xh, mode=1)  # regularization 0
    for name in ("mpla_dgemv_core_generic_system_matrix_cublas",
                 "mpla_dgemv_core_generic_system_matrix_streamed_cublas"):
        s.lib.mpla_generic_dgemv(C.byref(b), C.byref(A), C.byref(x), mpla.core(name), s.I)
        assert np.linalg.norm(s.get(b) - ref) <= 1e-12 * np.linalg.norm(ref)
    s.lib.mpla_set_generic_system_rhs(C.byref(b), C.byref(A), s.I)
    assert np.allclose(s.get(b), oracle.Kernel(pts).gaussian_rhs(), rtol=1e-14)
    # a short CG through the virtual core (src/benchmark_generic_system_cg.cpp)
    s.set(x, np.zeros(n))
    bh = oracle.uniform(n, 5, 2, -1, 1)
    s.set(b, bh)
    s.lib.mpla_generic_conjugate_gradient(C.byref(b), C.byref(A), C.byref(x), 8, 1e-15,
                                          mpla.core("mpla_dgemv_core_generic_system_matrix_streamed_cublas"), s.I)
    xo, info = oracle.cg_rms(oracle.Kernel(pts).op(), bh, np.zeros(n), 8, 1e-15)
    assert s.inst.last_iterations == info.iterations == 8
    assert np.linalg.norm(s.get(x) - xo) <= 1e-6 * np.linalg.norm(xo)
    s.lib.destroy_gaussian_kernel_system_assembler_object(holder.ptr)
