"""Helpers to drive the drop-in mpla.h API from Python tests, the way the
reference's C++ drivers do (HIP memcpy in place of cudaMemcpy)."""
import ctypes as C

import numpy as np

from mpla_amd import mpla

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy.restype = C.c_int
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMalloc.restype = C.c_int
        h.hipFree.argtypes = [C.c_void_p]
        h.hipFree.restype = C.c_int
        h.hipDeviceSynchronize.restype = C.c_int
        _hip = h
    return _hip


def to_device(dst_ptr, arr):
    a = np.ascontiguousarray(arr, dtype=np.float64)
    assert hip().hipDeviceSynchronize() == 0
    assert hip().hipMemcpy(dst_ptr, a.ctypes.data, a.nbytes, H2D) == 0


def from_device(src_ptr, n):
    out = np.empty(n, dtype=np.float64)
    assert hip().hipDeviceSynchronize() == 0
    assert hip().hipMemcpy(out.ctypes.data, src_ptr, out.nbytes, D2H) == 0
    return out


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(int(nbytes), 8)) == 0

    def __del__(self):
        if self.ptr:
            hip().hipFree(self.ptr)


class Points:
    """SoA points on the device plus the device array of per-dimension pointers
    (the layout of kernel_matrix_data.points)."""

    def __init__(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        self.dim, self.n = pts.shape
        self.buf = DeviceBuffer(pts.nbytes)
        to_device(self.buf.ptr, pts.ravel())
        base = self.buf.ptr.value
        table = np.array([base + d * self.n * 8 for d in range(self.dim)], dtype=np.uint64)
        self.table = DeviceBuffer(table.nbytes)
        assert hip().hipMemcpy(self.table.ptr, table.ctypes.data, table.nbytes, H2D) == 0

    @property
    def ptr(self):
        return self.table.ptr.value


class Session:
    def __init__(self, verbosity=0):
        self.lib = mpla.lib()
        self.inst = mpla.mpla_instance()
        self.lib.mpla_init_instance(C.byref(self.inst), mpla.MPI_COMM_WORLD)
        self.inst.verbosity = verbosity
        self._vectors = []
        self._matrices = []

    def close(self):
        for v in self._vectors:
            self.lib.mpla_free_vector(C.byref(v), C.byref(self.inst))
        for m in self._matrices:
            if isinstance(m, mpla.mpla_generic_matrix):
                self.lib.mpla_free_generic_matrix(C.byref(m), C.byref(self.inst))
            else:
                self.lib.mpla_free_matrix(C.byref(m), C.byref(self.inst))
        self.lib.mpla_destroy_instance(C.byref(self.inst))

    def vector(self, n, values=None):
        v = mpla.mpla_vector()
        self.lib.mpla_init_vector(C.byref(v), C.byref(self.inst), n)
        self._vectors.append(v)
        if values is not None:
            self.set(v, values)
        return v

    def set(self, v, values):
        to_device(v.data, np.asarray(values, dtype=np.float64))

    def get(self, v):
        out = np.empty(v.vec_row_count, dtype=np.float64)
        self.lib.mpla_copy_distributed_vector_to_cpu(out.ctypes.data, C.byref(v), C.byref(self.inst))
        return out

    def matrix(self, m, n, values_colmajor=None):
        A = mpla.mpla_matrix()
        self.lib.mpla_init_matrix(C.byref(A), C.byref(self.inst), m, n)
        self._matrices.append(A)
        if values_colmajor is not None:
            to_device(A.data, np.asarray(values_colmajor, dtype=np.float64).ravel(order="F"))
        return A

    def generic(self, m, n, payload_ptr):
        A = mpla.mpla_generic_matrix()
        self.lib.mpla_init_generic_matrix(C.byref(A), C.byref(self.inst), m, n)
        A.data = payload_ptr
        self._matrices.append(A)
        return A

    def as_generic(self, A):
        """The reference drivers pass an mpla_matrix as mpla_generic_matrix
        (src/test_cg.cpp:107); the layouts are identical."""
        return C.cast(C.pointer(A), C.POINTER(mpla.mpla_generic_matrix))

    def ref(self, obj):
        return C.byref(obj)

    @property
    def I(self):
        return C.byref(self.inst)
