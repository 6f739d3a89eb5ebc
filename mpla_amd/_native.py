"""ctypes bindings of the native libraries built into ``mpla_amd/lib``.

``libmplak.so`` is the gfx950 kernel library (C ABI of ``include/mplak.h``);
``libmpla.so`` is the drop-in host library (``include/mpla.h`` and friends).
Both are built in-tree by ``__graft_entry__.build()`` (``make -C
mpla_amd/csrc``).  Loading fails with an explicit error when a library is
missing: there is no fallback implementation of any operation.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

MPLAK_OK = 0
MPLAK_ERR_ARG = 1
MPLAK_ERR_HIP = 2
MPLAK_ERR_WORKSPACE = 3
MPLAK_ERR_UNSUPPORTED = 4
MPLAK_ERR_TOO_LARGE = 5
MPLAK_HM_MAX_RANK = 32
MPLAK_KERNEL_GAUSSIAN = 0
MPLAK_KERNEL_MATERN32 = 1
MPLAK_PREC_F64 = 0
MPLAK_PREC_F32 = 1
MPLAK_PREC_BF16 = 2
MPLAK_PREC_BF16X3 = 3
MPLAK_STOP_RMS = 0
MPLAK_STOP_RELATIVE = 1
MPLAK_RUNNING = 0
MPLAK_CONVERGED = 1
MPLAK_ITER_LIMIT = 2
MPLAK_BREAKDOWN = 3
MPLAK_RAGGED_START = 0
MPLAK_RAGGED_START_FINAL = 1
MPLAK_RAGGED_GROUP = 2
MPLAK_RAGGED_DZ = 3
MPLAK_RAGGED_UPDATE = 4
MPLAK_RAGGED_DIRECT = 5
MPLAK_RAGGED_AXPY = 6
MPLAK_RAGGED_SLABS = 0
MPLAK_RAGGED_PAIRS = 1
MPLAK_RAGGED_FILL = 7


class NativeLibraryMissing(RuntimeError):
    """The in-tree native library has not been built."""


class MplakError(RuntimeError):
    """A kernel-library call returned a non-zero status."""


class mplak_kernel_desc(C.Structure):
    _fields_ = [
        ("points", C.c_void_p),
        ("dim", C.c_int32),
        ("kind", C.c_int32),
        ("length_scale", C.c_double),
        ("nugget", C.c_double),
        ("coincident_shift", C.c_double),
        ("row_offset", C.c_int64),
        ("row_count", C.c_int64),
        ("col_offset", C.c_int64),
        ("col_count", C.c_int64),
    ]


class mplak_cg_state(C.Structure):
    _fields_ = [
        ("rr", C.c_double),
        ("bnorm2", C.c_double),
        ("rel", C.c_double),
        ("eps", C.c_double),
        ("n_global", C.c_double),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("k", C.c_int32),
        ("done", C.c_int32),
        ("iter_max", C.c_int32),
        ("rule", C.c_int32),
    ]


class mplak_bcg_group(C.Structure):
    _fields_ = [("all_done", C.c_int32), ("active", C.c_int32), ("pad", C.c_int32 * 2)]


class mplak_ragged_system(C.Structure):
    _fields_ = [
        ("points", C.c_void_p),
        ("V", C.c_void_p),
        ("Y", C.c_void_p),
        ("part", C.c_void_p),
        ("skip", C.c_void_p),
        ("n", C.c_int64),
        ("slab_cols", C.c_int64),
        ("length_scale", C.c_double),
        ("nugget", C.c_double),
        ("scale", C.c_double),
        ("dim", C.c_int32),
        ("kind", C.c_int32),
        ("nrhs", C.c_int32),
        ("precision", C.c_int32),
        ("rt", C.c_int32),
        ("ctiles", C.c_int32),
        ("tiles", C.c_int32),
        ("slabs", C.c_int32),
        ("drop_log2", C.c_double),
        ("drop_s", C.c_double),
        ("scheme", C.c_int32),
        ("reserved", C.c_int32),
    ]


class mplak_ragged_item(C.Structure):
    _fields_ = [("sys", C.c_int32), ("tile", C.c_int32), ("slab", C.c_int32), ("ctile", C.c_int32)]


class mplak_bicgstab_state(C.Structure):
    _fields_ = [
        ("rho", C.c_double),
        ("bnorm2", C.c_double),
        ("rel", C.c_double),
        ("eps", C.c_double),
        ("alpha", C.c_double),
        ("omega", C.c_double),
        ("beta", C.c_double),
        ("k", C.c_int32),
        ("done", C.c_int32),
        ("iter_max", C.c_int32),
        ("true_residual", C.c_int32),
    ]


_i64, _u64, _i32, _dbl, _vp, _sz = C.c_int64, C.c_uint64, C.c_int32, C.c_double, C.c_void_p, C.c_size_t

# name -> (restype, argtypes)
MPLAK_SIGNATURES = {
    "mplak_status_string": (C.c_char_p, [C.c_int]),
    "mplak_version": (C.c_int, []),
    "mplak_fill_uniform_f64": (C.c_int, [_vp, _i64, _u64, _u64, _u64, _dbl, _dbl, _vp]),
    "mplak_fill_f64": (C.c_int, [_vp, _i64, _dbl, _vp]),
    "mplak_copy_f64": (C.c_int, [_vp, _vp, _i64, _vp]),
    "mplak_dgemv_workspace_bytes": (_sz, [_i64, _i64]),
    "mplak_dgemv_f64": (C.c_int, [_i64, _i64, _vp, _i64, _vp, _vp, _vp, _sz, _vp, _vp]),
    "mplak_dot_workspace_bytes": (_sz, [_i64]),
    "mplak_ddot_f64": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mplak_daxpy_f64": (C.c_int, [_i64, _dbl, _vp, _vp, _vp]),
    "mplak_dwaxpby_f64": (C.c_int, [_i64, _dbl, _vp, _dbl, _vp, _vp, _vp]),
    "mplak_kernel_matvec_workspace_bytes": (_sz, [C.POINTER(mplak_kernel_desc)]),
    "mplak_kernel_matvec": (C.c_int, [C.POINTER(mplak_kernel_desc), C.c_int, _vp, _vp, _vp, _sz, _vp, _vp]),
    "mplak_kernel_matvec_noalloc_max_bytes": (_i64, []),
    "mplak_kernel_matvec_noalloc": (C.c_int, [C.POINTER(mplak_kernel_desc), C.c_int, _vp, _vp, _vp, _vp]),
    "mplak_kernel_matvec_screened_workspace_bytes": (_sz, [C.POINTER(mplak_kernel_desc)]),
    "mplak_kernel_matvec_screened": (C.c_int, [C.POINTER(mplak_kernel_desc), C.c_int, _dbl, _vp, _vp, _vp, _sz,
                                               _vp, _vp]),
    "mplak_screen_plan_bytes": (_sz, [C.POINTER(mplak_kernel_desc)]),
    "mplak_screen_plan_build": (C.c_int, [C.POINTER(mplak_kernel_desc), C.c_int, _dbl, _vp, _sz, _vp]),
    "mplak_kernel_matvec_planned": (C.c_int, [C.POINTER(mplak_kernel_desc), C.c_int, _dbl, _vp, _sz, _vp, _vp, _vp,
                                              _vp]),
    "mplak_planned_dots_count": (_i64, [_i64]),
    "mplak_kernel_matvec_planned_dots": (C.c_int, [C.POINTER(mplak_kernel_desc), C.c_int, _dbl, _vp, _sz, _vp, _vp,
                                                   _vp, _vp, _vp]),
    "mplak_kernel_matvec_planned_relaxed": (C.c_int, [C.POINTER(mplak_kernel_desc), C.c_int, _vp, _vp, _sz, _vp, _vp,
                                                      _vp, _vp, _vp, _vp, _vp]),
    "mplak_screen_plan_counts": (C.c_int, [C.POINTER(mplak_kernel_desc), _vp, _sz, _vp, _vp]),
    "mplak_spatial_order_workspace_bytes": (_sz, [_i64, _i32]),
    "mplak_spatial_order": (C.c_int, [_vp, _i32, _i64, _i64, _vp, _vp, _sz, _vp]),
    "mplak_gather_points": (C.c_int, [_vp, _i32, _i64, _vp, _i64, _vp, _i64, _vp]),
    "mplak_gather_f64": (C.c_int, [_vp, _vp, _vp, _i64, _vp]),
    "mplak_scatter_f64": (C.c_int, [_vp, _vp, _vp, _i64, _vp]),
    "mplak_gaussian_kernel_rhs": (C.c_int, [_vp, _i32, _i64, _i64, _vp, _vp]),
    "mplak_hm_gather_scaled": (C.c_int, [_vp, _i32, _i64, _vp, _i64, _i32, _dbl, _vp, _i64, _vp]),
    "mplak_hm_aca": (C.c_int, [_vp, _vp, _i64, _i32, _dbl, _vp, _vp, _vp, _vp]),
    "mplak_hm_far_partials": (C.c_int, [_vp, _vp, _i64, _vp, _i32, _vp, _vp, _vp, _vp]),
    "mplak_hm_far_sum": (C.c_int, [_vp, _i64, _vp, _i32, _vp, _vp, _vp]),
    "mplak_hm_apply": (C.c_int, [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    "mplak_hm_near_pairs": (C.c_int, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "mplak_hm_apply_pairs": (C.c_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp,
                                       _vp]),
    "mplak_hm_apply_stored": (C.c_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp,
                                        _vp, _vp]),
    "mplak_hm_scatter_nugget": (C.c_int, [_vp, _vp, _i64, _dbl, _vp, _i64, _i64, _vp, _vp]),
    "mplak_kernel_mrhs_workspace_bytes": (_sz, [C.POINTER(mplak_kernel_desc), C.c_int, _i64]),
    "mplak_kernel_mrhs": (C.c_int, [C.POINTER(mplak_kernel_desc), C.c_int, _i64, _vp, _i64, _vp, _i64, _vp, _sz, _vp,
                                    _vp]),
    "mplak_cg_workspace_bytes": (_sz, [_i64]),
    "mplak_cg_start": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, C.c_int, _dbl, _i32, _dbl, _vp, _sz, _vp]),
    "mplak_cg_iterate": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mplak_cg1_iterate": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "mplak_cg_refresh_rr": (C.c_int, [_i64, _vp, _vp, _vp, _sz, _vp]),
    "mplak_cg_copy_state": (C.c_int, [_vp, _vp, _vp]),
    "mplak_bcg_workspace_bytes": (_sz, [_i64, _i64]),
    "mplak_bcg_start": (C.c_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _dbl, _i32, _dbl, _vp, _sz, _vp]),
    "mplak_bcg_iterate": (C.c_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mplak_bcg_copy_states": (C.c_int, [_vp, _vp, _i64, _vp, _vp]),
    "mplak_ragged_plan": (_sz, [_vp]),
    "mplak_ragged_product_items": (_i64, [_vp, _vp, _i32]),
    "mplak_ragged_finish_items": (_i64, [_vp, _vp, _i32]),
    "mplak_ragged_product": (C.c_int, [_vp, _vp, _i64, C.c_int, C.c_int, C.c_int, C.c_int, _vp]),
    "mplak_ragged_finish": (C.c_int, [_vp, _vp, _i64, _vp]),
    "mplak_ragged_cg_blocks": (_i32, [_i64]),
    "mplak_ragged_cg_part_bytes": (_sz, [_i64, _i32]),
    "mplak_ragged_cg_items": (_i64, [_vp, _vp, _i32]),
    "mplak_ragged_cg_op": (C.c_int, [C.c_int, _vp, _vp, _i64, C.c_int, _dbl, _vp]),
    "mplak_bicgstab_workspace_bytes": (_sz, [_i64]),
    "mplak_bicgstab_start": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _dbl, _i32, _i32, _vp, _sz, _vp]),
    "mplak_bicgstab_step_s": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mplak_bicgstab_step_x": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mplak_bicgstab_step_true": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mplak_bicgstab_refresh_rho": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _sz, _vp]),
    "mplak_bicgstab_copy_state": (C.c_int, [_vp, _vp, _vp]),
}

_libs: dict[str, C.CDLL] = {}


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def _load(name: str, signatures: dict) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no non-native fallback)")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    for fname, (res, args) in signatures.items():
        fn = getattr(lib, fname)
        fn.restype = res
        fn.argtypes = args
    _libs[name] = lib
    return lib


def mplak() -> C.CDLL:
    """The kernel library (raises NativeLibraryMissing if it is not built)."""
    return _load("libmplak.so", MPLAK_SIGNATURES)


def check(status: int, what: str = "mplak call") -> None:
    if status != MPLAK_OK:
        msg = mplak().mplak_status_string(status).decode()
        raise MplakError(f"{what} failed with status {status} ({msg})")
