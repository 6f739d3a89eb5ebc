"""Halton nodes of the quadrature driver (src/generate_quadrature_weights.cpp:
107-117 uses GSL's gsl_qrng_halton, absent here).

CPU bars: the oracle's restatement equals the exact rational radical inverse
(Python fractions) bitwise in base 2 and within 4 ulp in every base, and
returns the known first points; the product's host generator
(mpla_halton_points, libmpla) equals the oracle bitwise.  GPU bar: the driver
writes weights.dat whose FP64 residual (recomputed by the oracle's product with
the reference kernel exp(-r^2) + 1e-8 on coincident points) meets the
tolerance, and whose weights integrate the constant 1 over [0,1]^dim to 1e-3;
a restart from its checkpoint files converges at once.
"""
import ctypes as C
import math
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53]


def radical_inverse(k, b):
    r, w = Fraction(0), Fraction(1, b)
    while k > 0:
        r += w * (k % b)
        k //= b
        w /= b
    return r


def test_oracle_halton_exact_radical_inverse(oracle):
    n, dim = 1500, 6
    pts = oracle.halton(n, dim)
    for d in range(dim):
        b = PRIMES[d]
        exact = np.array([float(radical_inverse(k, b)) for k in range(1, n + 1)])
        if b == 2:
            assert np.array_equal(pts[d], exact)
        ulp = np.spacing(np.maximum(exact, 1e-300))
        assert np.all(np.abs(pts[d] - exact) <= 4 * ulp), b
    assert pts[:, 0].tolist() == [0.5, float(Fraction(1, 3)), 0.2, float(Fraction(1, 7)), float(Fraction(1, 11)),
                                  float(Fraction(1, 13))]
    assert pts[0, :7].tolist() == [0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875]
    # index offset: points first+1 .. continue the same sequence
    assert np.array_equal(oracle.halton(100, 3, first=1000), oracle.halton(1100, 3)[:, 1000:])


@pytest.mark.parametrize("dim,n,first", [(1, 1, 0), (3, 1000, 0), (16, 777, 12345), (64, 50, 0)])
def test_product_halton_equals_oracle(oracle, dim, n, first):
    from mpla_amd import mpla
    lib = mpla.lib()
    out = np.zeros((dim, n + 5))
    assert lib.mpla_halton_points(out.ctypes.data, dim, n, first, n + 5) == 0
    assert np.array_equal(out[:, :n], oracle.halton(n, dim, first=first))
    assert lib.mpla_halton_points(out.ctypes.data, 65, n, 0, n + 5) == -1
    assert lib.mpla_halton_points(out.ctypes.data, dim, n, 0, n - 1) == -1


DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mpla_amd", "bin",
                      "generate_quadrature_weights")


def run_driver(cwd, *args):
    assert os.path.exists(DRIVER), "driver not built"
    out = subprocess.run([DRIVER, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=100)
    assert out.returncode == 0, out.stderr
    return out.stdout


def read_weights(path):
    raw = open(path, "rb").read()
    n = int.from_bytes(raw[:4], "little", signed=True)
    w = np.frombuffer(raw[4:], dtype=np.float64)
    assert len(w) == n
    return w


@pytest.mark.gpu
@pytest.mark.parametrize("n,dim,tol", [(300, 2, 1e-6), (1000, 3, 1e-7)])
def test_quadrature_driver_weights(oracle, tmp_path, n, dim, tol):
    from tests.gpu_util import require_gpu
    require_gpu()
    out = run_driver(tmp_path, n, dim, tol, 0, 0)
    assert "Time spent:" in out and "MPLA CG stopped after" in out
    w = read_weights(tmp_path / "weights.dat")
    pts = oracle.halton(n, dim)
    k = oracle.Kernel(pts, 0, 1 / math.sqrt(2.0), coincident_shift=1e-8)
    b = k.gaussian_rhs()
    assert np.linalg.norm(b - k.matvec(w)) / np.linalg.norm(b) < tol * 1.01
    assert abs(w.sum() - 1.0) < 1e-3  # the weights integrate the constant 1
    # the oracle's CG from the same start reaches the same tolerance
    x0 = oracle.uniform(n, 1235, 0)
    xo, info = oracle.cg_relative(k.op(), b, x0, 100000, tol)
    assert info.converged == 1
    assert abs(w.sum() - xo.sum()) < 1e-3


@pytest.mark.gpu
def test_quadrature_driver_restart(oracle, tmp_path):
    from tests.gpu_util import require_gpu
    require_gpu()
    n, dim, tol = 500, 2, 1e-6
    out = run_driver(tmp_path, n, dim, tol, 0, 10)
    first = read_weights(tmp_path / "weights.dat")
    for name in ("x.dat", "r.dat", "d.dat", "z.dat"):
        assert (tmp_path / name).stat().st_size == 4 + 8 * n
    out = run_driver(tmp_path, n, dim, tol, 1, 10)
    again = read_weights(tmp_path / "weights.dat")
    its = int(out.split("MPLA CG stopped after ")[1].split()[0])
    assert its <= 10  # resumed from the last checkpoint, at most one checkpoint interval before the end
    pts = oracle.halton(n, dim)
    k = oracle.Kernel(pts, 0, 1 / math.sqrt(2.0), coincident_shift=1e-8)
    b = k.gaussian_rhs()
    assert np.linalg.norm(b - k.matvec(again)) / np.linalg.norm(b) < tol * 1.01
    assert abs(again.sum() - first.sum()) < 1e-4
