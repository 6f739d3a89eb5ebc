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
    "mpla.h": "libmpla.so",
    "kernel.h": "libmpla.so",
    "mpla_kernel.h": "libmpla.so",
    "generic_system_adapter.h": "libmpla.so",
    "kernel_system_assembler.h": "libmpla.so",
    "mpla_server.h": "libmpla.so",
    "quadrature.h": "libmpla.so",
    "hmglib_adapter.h": "libmpla.so",
}

DECL = re.compile(r"^\s*(?:extern\s+)?[A-Za-z_][\w\s\*]*?[\s\*]((?:mplak|mpla)_\w+|info|checkCUDAError|"
                  r"(?:create|destroy)_gaussian_kernel_system_assembler_object|set_gaussian_kernel_rhs)\s*\(",
                  re.M)


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
    assert len(names) >= 1
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{libname} lacks {missing}"


def test_bindings_cover_header():
    from mpla_amd import _native
    names = declared("mplak.h")
    assert sorted(_native.MPLAK_SIGNATURES) == names


def test_binding_constants_match_header():
    from mpla_amd import _native
    text = open(os.path.join(INCLUDE, "mplak.h")).read()
    defines = dict(re.findall(r"#define\s+(MPLAK_[A-Z0-9_]+)\s+(-?\d+)\b", text))
    assert len(defines) >= 15
    for name, value in defines.items():
        if name == "MPLAK_VERSION":
            continue
        assert getattr(_native, name) == int(value), name


def test_status_strings_without_gpu():
    from mpla_amd import _native
    lib = _native.mplak()
    assert lib.mplak_status_string(0) == b"ok"
    assert lib.mplak_version() >= 100
    # argument validation happens before any device work
    assert lib.mplak_daxpy_f64(-1, 1.0, None, None, None) == 1


def test_python_bindings_cover_headers():
    from mpla_amd import mpla
    names = set()
    for h in ("mpla.h", "kernel.h", "mpla_kernel.h", "generic_system_adapter.h", "kernel_system_assembler.h",
              "mpla_server.h", "quadrature.h", "hmglib_adapter.h"):
        names |= set(declared(h))
    assert sorted(names) == sorted(mpla.MPLA_SIGNATURES)


STRUCTS = {
    "mpla_instance": ["comm_storage", "proc_count", "proc_rows", "proc_cols", "cur_proc_coord", "cur_proc_row",
                      "cur_proc_col", "cur_proc_rank", "is_parent", "stream", "runtime", "skip_flag",
                      "last_iterations", "last_converged", "last_rel_res", "last_matvecs", "verbosity"],
    "mpla_matrix": None,
    "mpla_generic_matrix": None,
    "mpla_vector": None,
    "kernel_matrix_data": None,
    "mpla_kernel_matrix_data": None,
    "mpla_server_config": None,
    "mpla_solve_request": None,
    "mpla_solve_result": None,
    "h_matrix_data": None,
    "mpla_hmatrix_info": None,
}


def test_struct_layout_matches_headers(tmp_path):
    """Compile the C headers with gcc and compare sizeof/offsetof with ctypes."""
    import subprocess
    from mpla_amd import mpla
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mpla.h"', '#include "kernel.h"',
             '#include "mpla_kernel.h"', '#include "mpla_server.h"', '#include "hmglib_adapter.h"',
             "int main(void) {"]
    for sname in STRUCTS:
        cls = getattr(mpla, sname)
        tag = "" if sname.startswith(("mpla_server", "mpla_solve")) else "struct "
        lines.append(f'printf("{sname} size %zu\\n", sizeof({tag}{sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{sname}.{fname} %zu\\n", offsetof({tag}{sname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for sname in STRUCTS:
        cls = getattr(mpla, sname)
        assert int(got[f"{sname} size"]) == C.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert int(got[f"{sname}.{fname}"]) == getattr(cls, fname).offset, (sname, fname)


MPLAK_STRUCTS = ["mplak_kernel_desc", "mplak_cg_state", "mplak_bcg_group", "mplak_bicgstab_state",
                 "mplak_ragged_system", "mplak_ragged_item"]


def test_mplak_struct_layout_matches_header(tmp_path):
    """The kernel library's records (ctypes in mpla_amd._native) against gcc's
    sizeof/offsetof of include/mplak.h."""
    import subprocess
    from mpla_amd import _native
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mplak.h"', "int main(void) {"]
    for sname in MPLAK_STRUCTS:
        cls = getattr(_native, sname)
        lines.append(f'printf("{sname} size %zu\\n", sizeof({sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{sname}.{fname} %zu\\n", offsetof({sname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout_k.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout_k"
    subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for sname in MPLAK_STRUCTS:
        cls = getattr(_native, sname)
        assert int(got[f"{sname} size"]) == C.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert int(got[f"{sname}.{fname}"]) == getattr(cls, fname).offset, (sname, fname)


def test_block_distribution_host_only():
    """The product's distribution arithmetic (no GPU needed) against hand cases
    of src/mpla.cpp:157-209."""
    from mpla_amd import mpla
    lib = mpla.lib()
    for M, P, counts in [(10, 3, [4, 3, 3]), (131072, 4, [32768] * 4), (7, 7, [1] * 7), (65537, 2, [32769, 32768])]:
        c = (C.c_int * P)()
        o = (C.c_int * P)()
        assert lib.mpla_block_distribution(M, P, c, o) == 0
        assert list(c) == counts
        assert list(o) == [sum(counts[:p]) for p in range(P)]
    c = (C.c_int * 3)()
    assert lib.mpla_block_distribution(2, 3, c, c) == -1


def test_server_request_bytes_host_only():
    """Admission arithmetic is host-only: device block = points + 6 vectors +
    product partials + CG partials, rounded to 1 MiB.  Below 8192 points the
    partials are FP64 slab sums; from 8192 on, FP64 tile-pair sums (two
    512-row blocks per pair of 512-point tiles and right-hand side)."""
    from mpla_amd import mpla
    lib = mpla.lib()
    small = lib.mpla_server_request_bytes(1024, 2, 1)
    big = lib.mpla_server_request_bytes(65536, 3, 8)
    assert small % (1 << 20) == 0 and small >= (1 << 20)
    assert small < 4 * (1 << 20)
    pairs = 128 * 129 // 2
    assert big >= 8 * (3 * 65536 + 6 * 65536 * 8) + 2 * pairs * 8 * 512 * 8
    assert big < 8 * (3 * 65536 + 6 * 65536 * 8) + 2 * pairs * 8 * 512 * 8 + (8 << 20)
    assert lib.mpla_server_request_bytes(65536, 3, 4) < big
    assert lib.mpla_server_request_bytes(0, 3, 1) == 0


def test_ragged_plan_and_items_host_only(monkeypatch):
    """The served products' work split is host arithmetic (no GPU): slabs
    below 8192 points, tile pairs from 8192 on (or as MPLAK_RAGGED_PAIRS
    says), with the partial-buffer size and the item list of each split."""
    from mpla_amd import _native
    lib = _native.mplak()

    def plan(n, nrhs, prec):
        s = _native.mplak_ragged_system()
        s.n, s.dim, s.kind, s.nrhs, s.precision, s.length_scale = n, 3, _native.MPLAK_KERNEL_GAUSSIAN, nrhs, prec, 0.1
        nbytes = lib.mplak_ragged_plan(C.byref(s))
        cnt = lib.mplak_ragged_product_items(C.byref(s), None, 5)
        items = (_native.mplak_ragged_item * cnt)()
        assert lib.mplak_ragged_product_items(C.byref(s), items, 5) == cnt
        return s, nbytes, list(items)

    def align(v):
        return -(-v // 256) * 256

    monkeypatch.delenv("MPLAK_RAGGED_PAIRS", raising=False)
    s, nbytes, items = plan(9000, 5, _native.MPLAK_PREC_F64)
    tiles = -(-9000 // 512)
    pairs = tiles * (tiles + 1) // 2
    assert s.scheme == _native.MPLAK_RAGGED_PAIRS and s.tiles == tiles and s.rt == 8 and s.ctiles == 1
    assert nbytes == align(2 * pairs * 5 * 512 * 8) + align(pairs * 4)
    assert [(it.sys, it.tile, it.slab, it.ctile) for it in items] == [(5, p, 0, 0) for p in range(pairs)]
    s32, nbytes32, _ = plan(9000, 5, _native.MPLAK_PREC_F32)
    assert nbytes32 == align(2 * pairs * 5 * 512 * 4) + align(pairs * 4)
    s, nbytes, items = plan(9000, 10, _native.MPLAK_PREC_F64)  # two column tiles
    assert s.ctiles == 2 and len(items) == 2 * pairs and items[-1].ctile == 1
    s, nbytes, items = plan(5000, 3, _native.MPLAK_PREC_F64)
    assert s.scheme == _native.MPLAK_RAGGED_SLABS and nbytes == s.slabs * 3 * 5000 * 8
    assert len(items) == s.tiles * s.slabs * s.ctiles
    monkeypatch.setenv("MPLAK_RAGGED_PAIRS", "1")
    assert plan(5000, 3, _native.MPLAK_PREC_F64)[0].scheme == _native.MPLAK_RAGGED_PAIRS
    monkeypatch.setenv("MPLAK_RAGGED_PAIRS", "0")
    assert plan(65536, 8, _native.MPLAK_PREC_F64)[0].scheme == _native.MPLAK_RAGGED_SLABS
