"""ctypes binding of the CPU oracle ``oracle/liboracle.so``.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of ``bench.py`` as the checker.  The product never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

GAUSSIAN = 0
MATERN32 = 1


class oracle_kernel(C.Structure):
    _fields_ = [
        ("points", C.c_void_p),
        ("n_points", C.c_int64),
        ("dim", C.c_int),
        ("kind", C.c_int),
        ("length_scale", C.c_double),
        ("nugget", C.c_double),
        ("coincident_shift", C.c_double),
    ]


class oracle_op(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("n", C.c_int64),
        ("A", C.c_void_p),
        ("lda", C.c_int64),
        ("kernel", oracle_kernel),
        ("mode", C.c_int),
    ]


class oracle_solve_info(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("converged", C.c_int),
        ("final_rel", C.c_double),
        ("matvecs", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (make -C oracle)")
        L = C.CDLL(LIB)
        vp, i64, u64, dbl = C.c_void_p, C.c_int64, C.c_uint64, C.c_double
        sig = {
            "oracle_rng_key": (u64, [u64, u64]),
            "oracle_rng_u64": (u64, [u64, u64]),
            "oracle_fill_uniform": (None, [vp, i64, u64, u64, u64, dbl, dbl]),
            "oracle_fill_normal": (None, [vp, i64, u64, u64, u64]),
            "oracle_block_counts": (C.c_int, [i64, C.c_int, vp, vp]),
            "oracle_dgemv": (None, [i64, i64, vp, i64, vp, vp, C.c_int]),
            "oracle_ddot": (dbl, [i64, vp, vp, C.c_int]),
            "oracle_daxpy": (None, [i64, dbl, vp, vp]),
            "oracle_kernel_entry": (dbl, [C.POINTER(oracle_kernel), i64, i64]),
            "oracle_kernel_matvec": (None, [C.POINTER(oracle_kernel), i64, i64, i64, i64, vp, vp, C.c_int]),
            "oracle_gaussian_rhs": (None, [C.POINTER(oracle_kernel), i64, i64, vp]),
            "oracle_apply": (None, [C.POINTER(oracle_op), vp, vp]),
            "oracle_cg_rms": (None, [C.POINTER(oracle_op), vp, vp, C.c_int, dbl, C.POINTER(oracle_solve_info)]),
            "oracle_cg_relative": (None, [C.POINTER(oracle_op), vp, vp, C.c_int, dbl, C.c_int, C.c_int,
                                          C.c_char_p, C.POINTER(oracle_solve_info)]),
            "oracle_bicgstab": (None, [C.POINTER(oracle_op), vp, vp, C.c_int, dbl, C.c_int, C.c_int,
                                       C.c_char_p, C.POINTER(oracle_solve_info)]),
            "oracle_save_vector": (C.c_int, [C.c_char_p, vp, C.c_int32]),
            "oracle_load_vector": (C.c_int, [C.c_char_p, vp, C.c_int32]),
            "oracle_threads": (C.c_int, []),
            "oracle_halton": (C.c_int, [vp, C.c_int, i64, i64, i64]),
            "oracle_hm_create": (vp, [C.POINTER(oracle_kernel), dbl, C.c_int, C.c_int, dbl]),
            "oracle_hm_free": (None, [vp]),
            "oracle_hm_perm": (None, [vp, vp]),
            "oracle_hm_counts": (None, [vp, vp]),
            "oracle_hm_cluster": (C.c_int, [vp, C.c_int, vp]),
            "oracle_hm_block": (C.c_int, [vp, C.c_int, C.c_int, vp]),
            "oracle_hm_matvec": (None, [vp, vp, vp]),
            "oracle_aca": (C.c_int, [vp, i64, i64, i64, C.c_int, dbl, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data


def uniform(n, seed, stream, lo=0.0, hi=1.0, first=0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().oracle_fill_uniform(_ptr(out), n, seed, stream, first, lo, hi)
    return out


def normal(n, seed, stream, first=0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().oracle_fill_normal(_ptr(out), n, seed, stream, first)
    return out


def block_counts(M, P):
    c = np.zeros(P, dtype=np.int64)
    o = np.zeros(P, dtype=np.int64)
    rc = lib().oracle_block_counts(M, P, _ptr(c), _ptr(o))
    return (None, None) if rc != 0 else (c, o)


def dgemv(A: np.ndarray, x: np.ndarray, mode=1) -> np.ndarray:
    """y = A x for a column-major (Fortran-ordered) A."""
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    y = np.empty(m)
    lib().oracle_dgemv(m, n, _ptr(A), m, _ptr(np.ascontiguousarray(x, dtype=np.float64)), _ptr(y), mode)
    return y


def ddot(x, y, mode=1) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().oracle_ddot(len(x), _ptr(x), _ptr(y), mode)


class Kernel:
    """Kernel-matrix description on host points of shape (dim, n)."""

    def __init__(self, points, kind=GAUSSIAN, length_scale=1.0 / np.sqrt(2.0), nugget=0.0,
                 coincident_shift=0.0):
        self.points = np.ascontiguousarray(points, dtype=np.float64)
        self.dim, self.n = self.points.shape
        self.s = oracle_kernel(_ptr(self.points), self.n, self.dim, kind, length_scale, nugget,
                               coincident_shift)

    def entry(self, i, j) -> float:
        return lib().oracle_kernel_entry(C.byref(self.s), i, j)

    def matvec(self, x, mode=1, row_off=0, rows=None, col_off=0, cols=None) -> np.ndarray:
        rows = self.n if rows is None else rows
        cols = self.n if cols is None else cols
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(rows)
        lib().oracle_kernel_matvec(C.byref(self.s), row_off, rows, col_off, cols, _ptr(x), _ptr(y), mode)
        return y

    def gaussian_rhs(self, row_off=0, rows=None) -> np.ndarray:
        rows = self.n if rows is None else rows
        b = np.empty(rows)
        lib().oracle_gaussian_rhs(C.byref(self.s), row_off, rows, _ptr(b))
        return b

    def op(self, mode=1):
        o = oracle_op()
        o.type = 1
        o.n = self.n
        o.kernel = self.s
        o.mode = mode
        o._keep = self
        return o


def dense_op(A: np.ndarray, mode=1):
    A = np.asfortranarray(A, dtype=np.float64)
    o = oracle_op()
    o.type = 0
    o.n = A.shape[0]
    o.A = _ptr(A)
    o.lda = A.shape[0]
    o.mode = mode
    o._keep = A
    return o


def _solve(fn, op, b, x0, *args):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.array(x0, dtype=np.float64, copy=True)
    info = oracle_solve_info()
    fn(C.byref(op), _ptr(b), _ptr(x), *args, C.byref(info))
    return x, info


def cg_rms(op, b, x0, iter_max, eps):
    return _solve(lib().oracle_cg_rms, op, b, x0, iter_max, eps)


def cg_relative(op, b, x0, iter_max, eps, ckpt_every=0, restart=False, directory="."):
    return _solve(lib().oracle_cg_relative, op, b, x0, iter_max, eps, ckpt_every, int(restart),
                  directory.encode())


def bicgstab(op, b, x0, iter_max, eps, ckpt_every=0, restart=False, directory="."):
    return _solve(lib().oracle_bicgstab, op, b, x0, iter_max, eps, ckpt_every, int(restart),
                  directory.encode())


def apply(op, x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(op.n)
    lib().oracle_apply(C.byref(op), _ptr(x), _ptr(y))
    return y


def save_vector(path, x) -> int:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().oracle_save_vector(str(path).encode(), _ptr(x), len(x))


def load_vector(path, n):
    x = np.zeros(n)
    rc = lib().oracle_load_vector(str(path).encode(), _ptr(x), n)
    return rc, x


def halton(n, dim, first=0) -> np.ndarray:
    """Halton points k = first+1 .. first+n as a (dim, n) array (GSL's sequence)."""
    out = np.empty((dim, n), dtype=np.float64)
    if lib().oracle_halton(_ptr(out), dim, n, first, n) != 0:
        raise ValueError(f"unsupported dim {dim}")
    return out


def threads() -> int:
    return lib().oracle_threads()


class HMatrix:
    """The oracle's H-matrix (cluster tree, block partition, ACA) of a kernel."""

    def __init__(self, kernel: "Kernel", eta, c_leaf, kmax, eps=0.0):
        self.kernel = kernel
        self.n = kernel.n
        self.dim = kernel.dim
        self.h = lib().oracle_hm_create(C.byref(kernel.s), eta, c_leaf, kmax, eps)
        if not self.h:
            raise ValueError("invalid H-matrix arguments")

    def close(self):
        if self.h:
            lib().oracle_hm_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def perm(self) -> np.ndarray:
        p = np.empty(self.n, dtype=np.int64)
        lib().oracle_hm_perm(self.h, _ptr(p))
        return p

    def counts(self) -> dict:
        c = np.zeros(7, dtype=np.int64)
        lib().oracle_hm_counts(self.h, _ptr(c))
        keys = ("leaves", "near_blocks", "far_blocks", "near_entries", "far_entries", "stored_terms", "max_rank")
        return dict(zip(keys, (int(v) for v in c)))

    def clusters(self):
        out = np.zeros(4 + 2 * self.dim)
        count = lib().oracle_hm_cluster(self.h, 0, _ptr(out))
        res = []
        for c in range(count):
            lib().oracle_hm_cluster(self.h, c, _ptr(out))
            res.append((int(out[0]), int(out[1]), int(out[2]), int(out[3]), out[4:4 + self.dim].copy(),
                        out[4 + self.dim:].copy()))
        return res

    def blocks(self, far):
        rc = np.zeros(2, dtype=np.int32)
        count = lib().oracle_hm_block(self.h, int(far), 0, _ptr(rc))
        out = []
        for b in range(count):
            lib().oracle_hm_block(self.h, int(far), b, _ptr(rc))
            out.append((int(rc[0]), int(rc[1])))
        return out

    def matvec(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        lib().oracle_hm_matvec(self.h, _ptr(x), _ptr(y))
        return y


def aca(A: np.ndarray, kmax, eps=0.0):
    """ACA of an explicit matrix: (rank, U, V) with A ~ U V^T."""
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    U = np.zeros((m, kmax), order="F")
    V = np.zeros((n, kmax), order="F")
    r = lib().oracle_aca(_ptr(A), m, m, n, kmax, eps, _ptr(U), _ptr(V))
    return r, U[:, :r], V[:, :r]
