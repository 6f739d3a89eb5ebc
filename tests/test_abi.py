"""The C-ABI libraries load on a CPU-only host and export every function
declared in include/*.h (no compute call is made here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")

# header -> library that must export its functions
HEADERS = {
    "mplak.h": "libmplak.so",
}

DECL = re.compile(r"^\s*(?:extern\s+)?[A-Za-z_][\w\s\*]*?[\s\*]((?:mplak|mpla)_\w+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in DECL.finditer(text):
        line = text[m.start():m.end()]
        if "typedef" in line or "#define" in line:
            continue
        names.add(m.group(1))
    return sorted(names)


@pytest.mark.parametrize("header,libname", sorted(HEADERS.items()))
def test_library_exports_header(header, libname):
    from mpla_amd import _native
    path = _native.lib_path(libname)
    assert os.path.exists(path), f"{path} not built"
    lib = C.CDLL(path)
    names = declared(header)
    assert len(names) > 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{libname} lacks {missing}"


def test_bindings_cover_header():
    from mpla_amd import _native
    names = declared("mplak.h")
    assert sorted(_native.MPLAK_SIGNATURES) == names


def test_status_strings_without_gpu():
    from mpla_amd import _native
    lib = _native.mplak()
    assert lib.mplak_status_string(0) == b"ok"
    assert lib.mplak_version() >= 100
    # argument validation happens before any device work
    assert lib.mplak_daxpy_f64(-1, 1.0, None, None, None) == 1
