"""ctypes mirror of the drop-in API (include/mpla.h, kernel.h, mpla_kernel.h,
generic_system_adapter.h, kernel_system_assembler.h, hmglib_adapter.h)
exported by libmpla.so.

This is the binding a Python caller of MPLA would use; the struct layouts are
checked against the C headers by tests/test_abi.py.  Device memory that the
library does not own (points, payloads) is passed as raw device pointers.
"""
from __future__ import annotations

import ctypes as C

from . import _native

MPI_COMM_WORLD = 0x44000000

MPLA_KERNEL_GAUSSIAN = 0
MPLA_KERNEL_MATERN32 = 1
MPLA_PREC_F64 = 0
MPLA_PREC_F32 = 1
MPLA_PREC_BF16 = 2
MPLA_PREC_BF16X3 = 3


class mpla_instance(C.Structure):
    _fields_ = [
        ("comm_storage", C.c_uint64),
        ("proc_count", C.c_int),
        ("proc_rows", C.c_int),
        ("proc_cols", C.c_int),
        ("cur_proc_coord", C.c_int * 2),
        ("cur_proc_row", C.c_int),
        ("cur_proc_col", C.c_int),
        ("cur_proc_rank", C.c_int),
        ("is_parent", C.c_bool),
        ("stream", C.c_void_p),
        ("runtime", C.c_void_p),
        ("skip_flag", C.c_void_p),
        ("last_iterations", C.c_int),
        ("last_converged", C.c_int),
        ("last_rel_res", C.c_double),
        ("last_matvecs", C.c_int),
        ("verbosity", C.c_int),
    ]


_IPP = C.POINTER(C.POINTER(C.c_int))


class mpla_matrix(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("mat_row_count", C.c_int),
        ("mat_col_count", C.c_int),
        ("cur_proc_row_count", C.c_int),
        ("cur_proc_col_count", C.c_int),
        ("cur_proc_row_offset", C.c_int),
        ("cur_proc_col_offset", C.c_int),
        ("proc_row_count", _IPP),
        ("proc_col_count", _IPP),
        ("proc_row_offset", _IPP),
        ("proc_col_offset", _IPP),
    ]


class mpla_generic_matrix(mpla_matrix):
    """Same layout as mpla_matrix; `data` is the plugin payload."""


class mpla_vector(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("vec_row_count", C.c_int),
        ("vec_col_count", C.c_int),
        ("cur_proc_row_count", C.c_int),
        ("cur_proc_row_offset", C.c_int),
        ("proc_row_count", _IPP),
        ("proc_row_offset", _IPP),
    ]


class kernel_matrix_data(C.Structure):
    _fields_ = [
        ("points", C.c_void_p),
        ("dim", C.c_int),
        ("max_row_count_per_dgemv", C.c_int),
    ]


class mpla_kernel_matrix_data(C.Structure):
    _fields_ = [
        ("points", C.c_void_p),
        ("dim", C.c_int),
        ("max_row_count_per_dgemv", C.c_int),
        ("kernel", C.c_int),
        ("precision", C.c_int),
        ("length_scale", C.c_double),
        ("nugget", C.c_double),
        ("coincident_shift", C.c_double),
    ]


_P = C.POINTER
_I, _V, _M, _G = _P(mpla_instance), _P(mpla_vector), _P(mpla_matrix), _P(mpla_generic_matrix)
_vp, _dbl, _int = C.c_void_p, C.c_double, C.c_int

# cores are passed as raw function addresses (c_void_p)
MPLA_SIGNATURES = {
    "info": (None, []),
    "mpla_init_instance": (None, [_I, C.c_int]),
    "mpla_init_instance_block_rows": (None, [_I, C.c_int]),
    "mpla_destroy_instance": (None, [_I]),
    "mpla_init_matrix": (None, [_M, _I, _int, _int]),
    "mpla_init_generic_matrix": (None, [_G, _I, _int, _int]),
    "mpla_init_vector": (None, [_V, _I, _int]),
    "mpla_init_vector_for_block_rows": (None, [_V, _I, _int]),
    "mpla_redistribute_vector_for_dgesv": (None, [_V, _V, _M, _I]),
    "mpla_save_vector": (None, [_V, C.c_char_p, _I]),
    "mpla_load_vector": (None, [_V, C.c_char_p, _I]),
    "mpla_free_vector": (None, [_V, _I]),
    "mpla_free_matrix": (None, [_M, _I]),
    "mpla_free_generic_matrix": (None, [_G, _I]),
    "mpla_dgemv": (None, [_V, _M, _V, _I]),
    "mpla_ddot": (None, [_P(C.c_double), _V, _V, _I]),
    "mpla_generic_dgemv": (None, [_V, _G, _V, _vp, _I]),
    "mpla_daxpy": (None, [_V, _dbl, _V, _I]),
    "mpla_vector_set_zero": (None, [_V, _I]),
    "mpla_generic_conjugate_gradient": (None, [_V, _G, _V, _int, _dbl, _vp, _I]),
    "mpla_generic_conjugate_gradient_with_checkpoint_restart":
        (None, [_V, _G, _V, _int, _dbl, _vp, _int, C.c_bool, _I]),
    "mpla_generic_BiCGSTAB_with_checkpoint_restart": (None, [_V, _G, _V, _int, _dbl, _vp, _int, C.c_bool, _I]),
    "mpla_copy_distributed_vector_to_cpu": (None, [_vp, _V, _I]),
    "checkCUDAError": (None, [C.c_char_p]),
    "mpla_generic_dgemv_core_dense": (None, [_V, _G, _V, _I]),
    "mpla_generic_dgemv_core_cublas": (None, [_V, _G, _V, _I]),
    "mpla_register_graph_safe_core": (None, [_vp]),
    "mpla_block_distribution": (C.c_int, [_int, _int, _P(C.c_int), _P(C.c_int)]),
    "mpla_instance_stream": (_vp, [_I]),
    # kernel.h
    "mpla_set_gaussian_kernel_rhs": (None, [_V, _G, _I]),
    "mpla_dgemv_core_kernel_cublas": (None, [_V, _G, _V, _I]),
    "mpla_dgemv_core_kernel_streamed_cublas": (None, [_V, _G, _V, _I]),
    # mpla_kernel.h
    "mpla_dgemv_core_kernel_fused": (None, [_V, _G, _V, _I]),
    "mpla_kernel_conjugate_gradient_mixed": (None, [_V, _G, _V, _int, _dbl, _int, _I]),
    "mpla_kernel_block_conjugate_gradient_mixed": (None, [_vp, _vp, _int, _G, _int, _dbl, _int, _vp, _I]),
    # generic_system_adapter.h / kernel_system_assembler.h
    "mpla_set_generic_system_rhs": (None, [_V, _G, _I]),
    "mpla_dgemv_core_generic_system_matrix_cublas": (None, [_V, _G, _V, _I]),
    "mpla_dgemv_core_generic_system_matrix_streamed_cublas": (None, [_V, _G, _V, _I]),
    "create_gaussian_kernel_system_assembler_object": (None, [_vp, _vp, _int, _int]),
    "destroy_gaussian_kernel_system_assembler_object": (None, [_vp]),
}


def lib() -> C.CDLL:
    """libmpla.so (loads libmplak.so first; raises NativeLibraryMissing if unbuilt)."""
    _native.mplak()
    return _native._load("libmpla.so", MPLA_SIGNATURES)


def core(name: str) -> int:
    """Address of an exported plugin core, for the solver/generic-dgemv calls."""
    return C.cast(getattr(lib(), name), C.c_void_p).value


# ---- mpla_server.h -----------------------------------------------------------

MPLA_SERVE_MIXED_BATCHED = 0
MPLA_SERVE_FP64_SERIAL = 1
MPLA_REQUEST_PENDING = 0
MPLA_REQUEST_CONVERGED = 1
MPLA_REQUEST_ITER_LIMIT = 2
MPLA_REQUEST_REJECTED = -1


class mpla_server_config(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_size_t),
        ("max_active", C.c_int),
        ("mode", C.c_int),
        ("iterations_per_step", C.c_int),
        ("inner_tol_floor", C.c_double),
        ("max_iterations", C.c_int),
    ]


class mpla_solve_request(C.Structure):
    _fields_ = [
        ("points", C.c_void_p),
        ("rhs", C.c_void_p),
        ("solution", C.c_void_p),
        ("n", C.c_int64),
        ("dim", C.c_int),
        ("kind", C.c_int),
        ("length_scale", C.c_double),
        ("nugget", C.c_double),
        ("nrhs", C.c_int),
        ("tolerance", C.c_double),
        ("arrival", C.c_double),
    ]


class mpla_solve_result(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("iterations", C.c_int),
        ("refinement_steps", C.c_int),
        ("max_rel_residual", C.c_double),
        ("t_arrival", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
    ]


MPLA_SIGNATURES.update({
    "mpla_server_default_config": (None, [_P(mpla_server_config)]),
    "mpla_server_create": (_vp, [_P(mpla_server_config)]),
    "mpla_server_destroy": (None, [_vp]),
    "mpla_server_request_bytes": (C.c_size_t, [C.c_int64, _int, _int]),
    "mpla_server_submit": (_int, [_vp, _P(mpla_solve_request)]),
    "mpla_server_step": (_int, [_vp]),
    "mpla_server_result": (_int, [_vp, _int, _P(mpla_solve_result)]),
    "mpla_server_run_trace": (_int, [_vp, _P(mpla_solve_request), _int, _P(mpla_solve_result)]),
    "mpla_server_clock": (C.c_double, [_vp]),
    "mpla_server_peak_bytes": (C.c_size_t, [_vp]),
    "mpla_halton_points": (C.c_int, [_vp, C.c_int32, C.c_int64, C.c_int64, C.c_int64]),
})


# ---- hmglib_adapter.h ----------------------------------------------------------


class h_matrix_data(C.Structure):
    _fields_ = [
        ("point_count", C.c_int * 2),
        ("dim", C.c_int),
        ("bits", C.c_int),
        ("eta", C.c_double),
        ("max_level", C.c_int),
        ("c_leaf", C.c_int),
        ("k", C.c_int),
        ("epsilon", C.c_double),
        ("root_level_set_1", C.c_char),
        ("root_level_set_2", C.c_char),
        ("max_batched_dense_size", C.c_int),
        ("max_batched_aca_size", C.c_int),
        ("assem", C.c_void_p),
        ("coords_d", C.c_void_p * 2),
        ("point_ids_d", C.c_void_p * 2),
        ("impl", C.c_void_p),
    ]


class mpla_hmatrix_info(C.Structure):
    _fields_ = [
        ("row_leaves", C.c_int64),
        ("col_leaves", C.c_int64),
        ("near_blocks", C.c_int64),
        ("far_blocks", C.c_int64),
        ("near_entries", C.c_int64),
        ("far_entries", C.c_int64),
        ("stored_terms", C.c_int64),
        ("max_rank", C.c_int32),
        ("k", C.c_int32),
        ("mean_rank", C.c_double),
        ("device_bytes", C.c_int64),
        ("near_pairs", C.c_int64),
    ]


MPLA_SIGNATURES.update({
    # all_coords[2] / all_point_ids[2] / global_point_count[2] are passed as pointers to 2-element arrays
    "mpla_init_hmglib": (None, [_G, _vp, _vp, _vp, _vp, _dbl, _int, _int, _int, _int, C.c_char, C.c_char, _int, _int,
                                _I]),
    "mpla_init_hmglib_coords_from_host": (None, [_G, _vp, _vp, _vp, _dbl, _int, _int, _int, _int, _int, _int, _I]),
    "mpla_destroy_hmglib": (None, [_G, _I]),
    "mpla_dgemv_core_hmglib_h_matrix": (None, [_V, _G, _V, _I]),
    "mpla_dgemv_core_hmglib_full_matrix": (None, [_V, _G, _V, _I]),
    "set_gaussian_kernel_rhs": (None, [_vp, _P(h_matrix_data)]),
    "mpla_init_hmatrix_kernel": (None, [_G, _vp, _int, _int, _dbl, _dbl, _dbl, _int, _int, _dbl, _I]),
    "mpla_hmatrix_get_info": (None, [_G, _P(mpla_hmatrix_info)]),
    "mpla_hmatrix_get_permutation": (None, [_G, _int, _P(C.c_int64)]),
    "mpla_hmatrix_partition": (_int, [_vp, C.c_int64, _int, _dbl, _int, _P(C.c_int64), _P(C.c_int64)]),
})
