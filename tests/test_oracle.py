"""Pins the CPU oracle against golden vectors, closed forms and independent
implementations (numpy / scipy), so that it can be trusted as the checker.

Each test names the reference behaviour it pins (paths relative to the MPLA
source tree).  Runs on CPU only.
"""
import math
import os
import struct

import numpy as np
import pytest
from scipy import integrate, sparse
from scipy.sparse import linalg as spla
from scipy.spatial.distance import cdist

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --------------------------------------------------------------------------
# Input generator


def test_splitmix64_published_vector(oracle):
    # SplitMix64 seeded with 1234567 (the reference test vector of the
    # generator's published implementation).
    expected = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                4593380528125082431, 16408922859458223821]
    got = [oracle.lib().oracle_rng_u64(1234567, i) for i in range(5)]
    assert got == expected


def test_uniform_moments_and_range(oracle):
    u = oracle.uniform(200_000, 1234, 7, -1.0, 1.0)
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(u.mean()) < 0.01
    assert abs(u.var() - 1.0 / 3.0) < 0.01
    # independent streams differ, same stream is reproducible
    assert not np.array_equal(u[:100], oracle.uniform(100, 1234, 8, -1.0, 1.0))
    assert np.array_equal(u[10:20], oracle.uniform(10, 1234, 7, -1.0, 1.0, first=10))


def test_normal_moments(oracle):
    z = oracle.normal(200_000, 1235, 3)
    assert abs(z.mean()) < 0.01
    assert abs(z.std() - 1.0) < 0.01
    assert abs(np.mean(z ** 4) - 3.0) < 0.05  # Gaussian kurtosis


# --------------------------------------------------------------------------
# Block distribution (src/mpla.cpp:157-209)


def test_block_counts_hand_cases(oracle):
    c, o = oracle.block_counts(10, 3)
    assert list(c) == [4, 3, 3] and list(o) == [0, 4, 7]
    c, o = oracle.block_counts(4096, 1)
    assert list(c) == [4096] and list(o) == [0]
    c, o = oracle.block_counts(7, 7)
    assert list(c) == [1] * 7 and list(o) == list(range(7))
    # floor(M/P) == 0: the reference prints and exits (src/mpla.cpp:160-164)
    assert oracle.block_counts(2, 3) == (None, None)


@pytest.mark.parametrize("M,P", [(131072, 4), (65537, 8), (1000, 7), (33, 32)])
def test_block_counts_invariants(oracle, M, P):
    c, o = oracle.block_counts(M, P)
    assert c.sum() == M
    assert c.max() - c.min() <= 1
    assert np.all(np.diff(c) <= 0)  # the remainder goes to the first blocks
    assert o[0] == 0 and np.array_equal(o[1:], np.cumsum(c)[:-1])


# --------------------------------------------------------------------------
# Dense level 1/2: src/test_library.cpp closed forms


@pytest.mark.parametrize("m,n", [(1, 1), (7, 5), (300, 200), (1000, 1000)])
@pytest.mark.parametrize("mode", [0, 1])
def test_dgemv_test_library_closed_form(oracle, m, n, mode):
    A = np.arange(m * n, dtype=np.float64).reshape((m, n), order="F")  # A[idx] = idx, idx = j*m+i
    x = np.arange(n, dtype=np.float64)
    y = oracle.dgemv(A, x, mode=mode)
    i = np.arange(m, dtype=np.float64)
    expect = m * (n - 1) * n * (2 * n - 1) / 6 + i * n * (n - 1) / 2
    assert np.array_equal(y, expect)


@pytest.mark.parametrize("n", [1, 10, 1000, 100000])
def test_ddot_closed_form(oracle, n):
    x = np.arange(n, dtype=np.float64)
    y = x + 1
    assert oracle.ddot(x, y, 1) == (n - 1) * n * (n + 1) / 3
    assert oracle.ddot(x, y, 0) == (n - 1) * n * (n + 1) / 3


def test_dgemv_random_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, (300, 257))
    x = rng.uniform(-1, 1, 257)
    y = oracle.dgemv(A, x, mode=1)
    assert np.allclose(y, A @ x, rtol=1e-13, atol=1e-13)


# --------------------------------------------------------------------------
# Kernel entries (src/kernel.cu:7-15, 39-50; src/kernel_system_assembler.h)


def _pair(r, dim=3):
    p = np.zeros((dim, 2))
    p[0, 1] = r
    return p


def test_reference_unscaled_gaussian_values(oracle):
    # l = 1/sqrt(2) is the reference's exp(-r*r)
    k = oracle.Kernel(_pair(1.0), oracle.GAUSSIAN, 1 / math.sqrt(2.0))
    assert k.entry(0, 1) == pytest.approx(0.36787944117144233, rel=1e-15)  # e^-1
    k = oracle.Kernel(_pair(2.0), oracle.GAUSSIAN, 1 / math.sqrt(2.0))
    assert k.entry(0, 1) == pytest.approx(0.01831563888873418, rel=1e-14)  # e^-4


def test_reference_unscaled_matern_values(oracle):
    # l = sqrt(3) is the reference's (1 + r) exp(-r)
    k = oracle.Kernel(_pair(1.0), oracle.MATERN32, math.sqrt(3.0))
    assert k.entry(0, 1) == pytest.approx(0.7357588823428847, rel=1e-14)  # 2/e
    k = oracle.Kernel(_pair(0.5), oracle.MATERN32, math.sqrt(3.0))
    assert k.entry(0, 1) == pytest.approx(0.9097959895689501, rel=1e-14)  # 1.5 e^-0.5


def test_length_scale_convention(oracle):
    # Gaussian: value e^-1/2 at r = l; Matern-3/2: (1+sqrt3) e^-sqrt3 at r = l
    for l in (0.05, 0.1, 2.0):
        g = oracle.Kernel(_pair(l), oracle.GAUSSIAN, l)
        assert g.entry(0, 1) == pytest.approx(0.6065306597126334, rel=1e-14)
        m = oracle.Kernel(_pair(l), oracle.MATERN32, l)
        assert m.entry(0, 1) == pytest.approx(0.4833577245965077, rel=1e-14)


def test_diagonal_rules(oracle):
    p = np.zeros((3, 3))
    p[:, 1] = 0.25
    p[:, 2] = 0.25  # duplicate of point 1
    k = oracle.Kernel(p, oracle.GAUSSIAN, 0.1, nugget=0.5, coincident_shift=1e-8)
    assert k.entry(0, 0) == 1.0 + 1e-8 + 0.5
    assert k.entry(1, 2) == 1.0 + 1e-8          # coincident, off-diagonal: shift only
    k = oracle.Kernel(p, oracle.GAUSSIAN, 0.1, nugget=0.5)
    assert k.entry(1, 2) == 1.0 and k.entry(2, 2) == 1.5


@pytest.mark.parametrize("kind", [0, 1])
def test_kernel_matrix_is_spd(oracle, kind):
    pts = oracle.uniform(3 * 200, 1235, 1).reshape(3, 200)
    k = oracle.Kernel(pts, kind, 0.2)
    K = np.array([[k.entry(i, j) for j in range(200)] for i in range(200)])
    assert np.array_equal(K, K.T)
    assert np.linalg.eigvalsh(K).min() > -1e-10


@pytest.mark.parametrize("kind,l", [(0, 0.05), (0, 0.3), (1, 0.1)])
@pytest.mark.parametrize("dim", [1, 2, 3, 5])
def test_kernel_matvec_brute_force_scipy(oracle, kind, l, dim):
    n = 300
    pts = oracle.uniform(dim * n, 99, dim).reshape(dim, n)
    x = oracle.uniform(n, 99, 100, -1.0, 1.0)
    D = cdist(pts.T, pts.T)
    if kind == 0:
        K = np.exp(-D ** 2 / (2 * l * l))
    else:
        s = math.sqrt(3.0) * D / l
        K = (1 + s) * np.exp(-s)
    K += 0.7 * np.eye(n)
    k = oracle.Kernel(pts, kind, l, nugget=0.7)
    for mode in (0, 1):
        y = k.matvec(x, mode=mode)
        assert np.linalg.norm(y - K @ x) <= 1e-13 * np.linalg.norm(K @ x)
    # sub-block (row/column offsets)
    y = k.matvec(x[50:170], row_off=20, rows=100, col_off=50, cols=120)
    assert np.allclose(y, K[20:120, 50:170] @ x[50:170], rtol=1e-12, atol=1e-13)


def test_kernel_matvec_symmetry(oracle):
    n = 500
    pts = oracle.uniform(3 * n, 5, 0).reshape(3, n)
    k = oracle.Kernel(pts, oracle.MATERN32, 0.1, nugget=0.5)
    x = oracle.uniform(n, 5, 1, -1, 1)
    y = oracle.uniform(n, 5, 2, -1, 1)
    assert np.dot(y, k.matvec(x)) == pytest.approx(np.dot(x, k.matvec(y)), rel=1e-13)


@pytest.mark.parametrize("dim", [1, 2])
def test_gaussian_rhs_integral_identity(oracle, dim):
    pts = oracle.uniform(dim * 5, 3, 0).reshape(dim, 5)
    k = oracle.Kernel(pts, oracle.GAUSSIAN, 1 / math.sqrt(2.0))
    b = k.gaussian_rhs()
    for i in range(5):
        if dim == 1:
            val, _ = integrate.quad(lambda y: math.exp(-(pts[0, i] - y) ** 2), 0, 1, epsabs=1e-14)
        else:
            val, _ = integrate.dblquad(
                lambda y1, y0: math.exp(-(pts[0, i] - y0) ** 2 - (pts[1, i] - y1) ** 2),
                0, 1, 0, 1, epsabs=1e-13)
        assert b[i] == pytest.approx(val, rel=1e-11)


# --------------------------------------------------------------------------
# CG (src/mpla.cpp:817-1001) and BiCGSTAB (:1003-1183)


def _laplacian(n):
    return (np.diag(2.0 * np.ones(n)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1))


def test_cg_manufactured_laplacian(oracle):
    # src/test_cg.cpp:57-113: tridiag(-1,2,-1), b = A x_rand, random x0, eps 1e-10 (RMS)
    n = 200
    A = _laplacian(n)
    xt = oracle.uniform(n, 1235, 0)
    b = A @ xt
    x0 = oracle.uniform(n, 1235, 1)
    x, info = oracle.cg_rms(oracle.dense_op(A), b, x0, 1000, 1e-10)
    assert info.converged == 1
    assert np.sqrt(np.mean((b - A @ x) ** 2)) < 1e-10
    # exact arithmetic converges in <= n iterations; scipy agrees on the solution
    assert info.iterations <= n + 5
    xs, rc = spla.cg(sparse.csr_matrix(A), b, x0=x0, rtol=1e-14, atol=0, maxiter=5000)
    assert np.linalg.norm(x - xs) < 1e-6 * np.linalg.norm(xs)


def test_cg_scaled_identity_one_iteration(oracle):
    n = 50
    A = 3.5 * np.eye(n)
    b = oracle.uniform(n, 1, 0, -1, 1)
    x, info = oracle.cg_relative(oracle.dense_op(A), b, np.zeros(n), 100, 1e-12)
    assert info.iterations == 1 and info.converged == 1
    assert np.allclose(x, b / 3.5, rtol=1e-15)


def test_cg_k_distinct_eigenvalues(oracle):
    n = 60
    rng = np.random.default_rng(1)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    eig = np.repeat([1.0, 2.0, 5.0, 9.0], n // 4)  # 4 distinct eigenvalues
    A = (Q * eig) @ Q.T
    A = 0.5 * (A + A.T)
    b = rng.standard_normal(n)
    x, info = oracle.cg_relative(oracle.dense_op(A), b, np.zeros(n), 100, 1e-10)
    assert info.converged == 1 and info.iterations <= 5


def test_cg_iteration_limit_accounting(oracle):
    # loop k = 1 .. iter_max-1 (src/mpla.cpp:937); exhausting it reports k = iter_max
    n = 100
    A = _laplacian(n)
    b = np.ones(n)
    x, info = oracle.cg_relative(oracle.dense_op(A), b, np.zeros(n), 6, 1e-14)
    assert info.converged == 0 and info.iterations == 6 and info.matvecs == 6


def test_cg_rms_vs_relative_rules(oracle):
    n = 100
    A = _laplacian(n) + np.eye(n)
    b = 1e-3 * np.ones(n)
    # RMS rule stops when sqrt(rr/N) < eps, relative rule when ||r||/||b|| < eps
    _, i_rms = oracle.cg_rms(oracle.dense_op(A), b, np.zeros(n), 500, 1e-6)
    _, i_rel = oracle.cg_relative(oracle.dense_op(A), b, np.zeros(n), 500, 1e-6)
    assert i_rms.final_rel < 1e-6 and i_rel.final_rel < 1e-6
    assert i_rms.iterations < i_rel.iterations  # ||b||/sqrt(N) = 1e-3 makes RMS looser


def test_bicgstab_manufactured(oracle):
    n = 300
    rng_u = oracle.uniform(n * n, 1235, 4, -1, 1).reshape((n, n), order="F")
    A = rng_u / math.sqrt(n) + np.diag(2.0 + oracle.uniform(n, 1235, 5))
    xt = oracle.uniform(n, 1235, 6, -1, 1)
    b = A @ xt
    x, info = oracle.bicgstab(oracle.dense_op(A), b, np.zeros(n), 500, 1e-10)
    assert info.converged == 1
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-9)
    assert info.matvecs == 1 + 3 * info.iterations


# --------------------------------------------------------------------------
# Vector files and checkpoint/restart (src/mpla.cpp:605-713, 922-933, 977-991)


def test_vector_file_fixture(oracle, tmp_path):
    fixture = os.path.join(GOLDEN, "vector_n3.bin")
    rc, x = oracle.load_vector(fixture, 3)
    assert rc == 0 and list(x) == [1.0, -2.5, 3.25]
    out = tmp_path / "v.bin"
    assert oracle.save_vector(out, x) == 0
    assert out.read_bytes() == open(fixture, "rb").read()
    rc, _ = oracle.load_vector(fixture, 4)
    assert rc == -2  # length mismatch (the reference prints and exits)
    with open(fixture, "rb") as f:
        assert struct.unpack("<i", f.read(4))[0] == 3


def test_cg_checkpoint_restart_continues(oracle, tmp_path):
    n = 150
    A = _laplacian(n) + 0.01 * np.eye(n)
    b = oracle.uniform(n, 7, 0, -1, 1)
    op = oracle.dense_op(A)
    full, info_full = oracle.cg_relative(op, b, np.zeros(n), 1000, 1e-10)
    # run 20 iterations with a checkpoint every 10, then restart from the files
    _, info_a = oracle.cg_relative(op, b, np.zeros(n), 21, 1e-10, ckpt_every=10, directory=str(tmp_path))
    assert info_a.iterations == 21 and info_a.converged == 0
    for name in ("x.dat", "r.dat", "d.dat", "z.dat"):
        assert (tmp_path / name).stat().st_size == 4 + 8 * n
    x, info_b = oracle.cg_relative(op, b, np.zeros(n), 1000, 1e-10, restart=True, directory=str(tmp_path))
    assert info_b.converged == 1
    # restarted run resumes at iteration 20: total work matches the full run
    assert abs((20 + info_b.iterations) - info_full.iterations) <= 2
    assert np.linalg.norm(x - full) < 1e-8 * np.linalg.norm(full)
