"""mpla_amd -- MI355X-native (gfx950) implementation of MPLA's solver hot path.

The product is native code: ``lib/libmplak.so`` (hand-written HIP kernels
behind the C ABI of ``include/mplak.h``) and ``lib/libmpla.so`` (the drop-in
``mpla.h`` host library).  This Python package only locates and binds them
(``_native``) and offers thin helpers used by the tests and the benchmark.
"""
from ._native import (  # noqa: F401
    LIB_DIR,
    MplakError,
    NativeLibraryMissing,
    check,
    mplak,
)
