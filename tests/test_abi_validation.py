"""Empty inputs and argument errors of the kernel C ABI (include/mplak.h).

These calls return before any device work, so they run without a GPU: empty
problems succeed, malformed ones return MPLAK_ERR_ARG (the reference exits on
size errors, src/mpla.cpp:160-175; the C ABI reports them instead), and the
workspace queries of empty problems are 0.
"""
import ctypes as C

import pytest

from mpla_amd import _native

OK, ARG, WS = _native.MPLAK_OK, _native.MPLAK_ERR_ARG, _native.MPLAK_ERR_WORKSPACE


@pytest.fixture(scope="module")
def lib():
    return _native.mplak()


def desc(rows=10, cols=10, dim=3, kind=0, ell=0.5, row_off=0, col_off=0):
    d = _native.mplak_kernel_desc()
    d.points = 0x1000  # never dereferenced by the paths under test
    d.dim, d.kind, d.length_scale = dim, kind, ell
    d.nugget, d.coincident_shift = 0.0, 0.0
    d.row_offset, d.row_count, d.col_offset, d.col_count = row_off, rows, col_off, cols
    return d


def test_status_strings(lib):
    for code in (OK, ARG, _native.MPLAK_ERR_HIP, WS, _native.MPLAK_ERR_UNSUPPORTED):
        assert lib.mplak_status_string(code)
    assert lib.mplak_status_string(12345)


def test_empty_vectors_succeed(lib):
    assert lib.mplak_fill_f64(None, 0, 1.0, None) == OK
    assert lib.mplak_copy_f64(None, None, 0, None) == OK
    assert lib.mplak_daxpy_f64(0, 2.0, None, None, None) == OK
    assert lib.mplak_dwaxpby_f64(0, 1.0, None, 1.0, None, None, None) == OK
    assert lib.mplak_fill_uniform_f64(None, 0, 1, 0, 0, 0.0, 1.0, None) == OK
    assert lib.mplak_gather_f64(None, None, None, 0, None) == OK
    assert lib.mplak_scatter_f64(None, None, None, 0, None) == OK


def test_negative_sizes_and_null_pointers_rejected(lib):
    assert lib.mplak_fill_f64(None, -1, 1.0, None) == ARG
    assert lib.mplak_fill_f64(None, 5, 1.0, None) == ARG
    assert lib.mplak_copy_f64(None, None, 3, None) == ARG
    assert lib.mplak_daxpy_f64(-1, 1.0, None, None, None) == ARG
    assert lib.mplak_ddot_f64(-1, None, None, None, None, 0, None) == ARG
    assert lib.mplak_gather_f64(None, None, None, -2, None) == ARG
    assert lib.mplak_cg_start(0, None, None, None, None, None, 1, 1e-8, 10, 1.0, None, 0, None) == ARG
    assert lib.mplak_cg_iterate(0, None, None, None, None, None, None, None, 0, None) == ARG
    assert lib.mplak_bicgstab_start(0, None, None, None, None, None, None, 1e-8, 10, 1, None, 0, None) == ARG


def test_dgemv_shapes(lib):
    assert lib.mplak_dgemv_workspace_bytes(0, 10) == 0
    assert lib.mplak_dgemv_workspace_bytes(10, 0) == 0
    assert lib.mplak_dgemv_f64(0, 10, None, 1, None, None, None, 0, None, None) == OK
    assert lib.mplak_dgemv_f64(10, 10, None, 5, None, None, None, 0, None, None) == ARG  # lda < m
    assert lib.mplak_dgemv_f64(-1, 10, None, 1, None, None, None, 0, None, None) == ARG


def test_kernel_desc_validation(lib):
    ok_empty = desc(rows=0)
    assert lib.mplak_kernel_matvec(C.byref(ok_empty), 0, None, None, None, 0, None, None) == OK
    assert lib.mplak_kernel_matvec_workspace_bytes(C.byref(ok_empty)) == 0
    for bad in (desc(dim=0), desc(dim=17), desc(kind=7), desc(ell=0.0), desc(ell=-1.0), desc(rows=-1),
                desc(row_off=-3)):
        assert lib.mplak_kernel_matvec(C.byref(bad), 0, None, None, None, 0, None, None) == ARG
    d = desc()
    # precision: the single-vector product takes FP64 or FP32 only
    assert lib.mplak_kernel_matvec(C.byref(d), _native.MPLAK_PREC_BF16, None, None, None, 0, None, None) == ARG
    # missing output / input / workspace
    assert lib.mplak_kernel_matvec(C.byref(d), 0, None, None, None, 0, None, None) == ARG
    y = (C.c_double * 10)()
    x = (C.c_double * 10)()
    assert lib.mplak_kernel_matvec(C.byref(d), 0, C.addressof(x), C.addressof(y), None, 0, None, None) == WS
    assert lib.mplak_kernel_matvec(None, 0, None, None, None, 0, None, None) == ARG


def test_screened_and_planned_validation(lib):
    d = desc()
    for drop in (0.0, -5.0, 3000.0, float("nan")):
        assert lib.mplak_kernel_matvec_screened(C.byref(d), 0, drop, None, None, None, 0, None, None) == ARG
    rect = desc(rows=10, cols=20)
    assert lib.mplak_screen_plan_bytes(C.byref(rect)) == 0
    assert lib.mplak_screen_plan_build(C.byref(rect), 1, 20.0, None, 0, None) == ARG
    assert lib.mplak_kernel_matvec_planned(C.byref(rect), 1, 20.0, None, 0, None, None, None, None) == ARG
    four_d = desc(dim=4)
    assert lib.mplak_screen_plan_bytes(C.byref(four_d)) == 0  # dim <= 3 only


def test_spatial_order_validation(lib):
    assert lib.mplak_spatial_order_workspace_bytes(0, 3) == 0
    assert lib.mplak_spatial_order_workspace_bytes(10, 0) == 0
    assert lib.mplak_spatial_order_workspace_bytes(10, 17) == 0
    assert lib.mplak_spatial_order(0x1000, 3, 0, 0, None, None, 0, None) == OK
    assert lib.mplak_spatial_order(0x1000, 3, 0, -1, None, None, 0, None) == ARG
    assert lib.mplak_spatial_order(None, 3, 0, 10, None, None, 0, None) == ARG
    assert lib.mplak_spatial_order(0x1000, 3, -1, 10, None, None, 0, None) == ARG
    assert lib.mplak_gather_points(0x1000, 3, 0, None, 10, None, 5, None) == ARG  # ld < n


def test_multi_rhs_validation(lib):
    d = desc()
    assert lib.mplak_kernel_mrhs_workspace_bytes(C.byref(d), 0, 0) == 0
    assert lib.mplak_kernel_mrhs(C.byref(d), 0, 0, None, 10, None, 10, None, 0, None, None) == OK
    assert lib.mplak_bcg_workspace_bytes(10, 0) >= 0
    assert lib.mplak_bcg_start(10, 0, None, None, None, None, None, None, 1, 1e-8, 10, 1.0, None, 0, None) == ARG


def test_hmatrix_kernels_validate_before_launch(lib):
    """The H-matrix kernels (mplak_hm_*) refuse malformed arguments and accept
    empty work without touching the device."""
    assert lib.mplak_hm_aca(None, None, 0, 20, 0.0, None, None, None, None) == ARG  # no point description
    assert lib.mplak_hm_far_partials(None, None, 0, None, 20, None, None, None, None) == OK
    assert lib.mplak_hm_far_partials(None, None, -1, None, 20, None, None, None, None) == ARG
    assert lib.mplak_hm_far_partials(None, None, 1, None, 33, None, None, None, None) == ARG  # rank > 32
    assert lib.mplak_hm_far_sum(None, 0, None, 20, None, None, None) == OK
    assert lib.mplak_hm_far_sum(None, 3, None, 0, None, None, None) == ARG
    assert lib.mplak_hm_apply_stored(None, 0, None, None, None, None, None, None, None, None, None, 20, None, None,
                                     None, None, None) == OK
    assert lib.mplak_hm_apply_stored(None, 2, None, None, None, None, None, None, None, None, None, 20, None, None,
                                     None, None, None) == ARG
    assert lib.mplak_hm_scatter_nugget(None, None, 0, 0.0, None, 0, 0, None, None) == OK
    assert lib.mplak_hm_scatter_nugget(None, None, 4, 1.0, None, 0, 4, None, None) == ARG
    assert lib.mplak_hm_gather_scaled(None, 3, 0, None, 10, 0, 0.5, None, 10, None) == ARG
    # symmetric near field
    assert lib.mplak_hm_near_pairs(None, None, 0, None, None, None) == ARG  # no point description
    assert lib.mplak_hm_apply_pairs(None, 0, None, None, None, None, None, None, None, None, None, 20, None, None,
                                    None, None) == OK
    assert lib.mplak_hm_apply_pairs(None, 2, None, None, None, None, None, None, None, None, None, 20, None, None,
                                    None, None) == ARG
    assert lib.mplak_hm_apply_pairs(None, 2, None, None, None, None, None, None, None, None, None, 33, None, None,
                                    None, None) == ARG


def test_ragged_product_validates_before_launch(lib):
    """The served products refuse an unknown work split, the tile-pair split
    without a single dimension, and missing arrays; empty launches succeed."""
    fake = C.c_void_p(0x1000)  # never dereferenced by the paths under test
    pairs, slabs = _native.MPLAK_RAGGED_PAIRS, _native.MPLAK_RAGGED_SLABS
    assert lib.mplak_ragged_product(None, None, 0, 8, 3, 0, pairs, None) == OK
    assert lib.mplak_ragged_product(None, None, 1, 8, 3, 0, pairs, None) == ARG
    assert lib.mplak_ragged_product(fake, fake, 1, 8, 0, 0, pairs, None) == ARG  # PAIRS needs one dimension
    assert lib.mplak_ragged_product(fake, fake, 1, 8, 3, 0, 7, None) == ARG  # unknown split
    assert lib.mplak_ragged_product(fake, fake, 1, 8, 5, 0, slabs, None) == ARG  # dimension above 4
    assert lib.mplak_ragged_finish(None, None, 0, None) == OK
    assert lib.mplak_ragged_finish(None, None, 3, None) == ARG


def test_noalloc_refuses_malformed_descriptors(lib):
    d = desc(rows=-1)
    assert lib.mplak_kernel_matvec_noalloc(C.byref(d), 0, None, None, None, None) == ARG
    assert lib.mplak_kernel_matvec_noalloc_max_bytes() == 80 << 20
