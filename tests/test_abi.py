"""The C ABI library loads and exports every symbol include/lap_gfx950.h
and include/lap_io.h declare (no device calls: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("lap_gfx950.h", "lap_io.h")]
LIB = os.path.join(ROOT, "lap_amd", "liblapgfx.so")


def declared_functions():
    text = "\n".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:[\w\s\*]+?)\b(\w+)\s*\(", text, flags=re.M)
    return sorted({n for n in names if n not in ("if", "while", "for", "return", "sizeof")})


def test_header_parses():
    names = declared_functions()
    for must in ("hip_csr_matvec", "hip_dot", "initialize_ichol", "hip_ichol", "cg",
                 "mallocForDevice", "memcpyDevice", "freeDevice", "lap_pcg", "lap_mm_load",
                 "lap_coo_to_csr", "read_mm_file", "coo_to_csr", "create_L_and_split"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="liblapgfx.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="liblapgfx.so not built")
def test_library_has_no_vendor_sparse_blas():
    out = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "rocsparse" not in out and "rocblas" not in out and "hipblas" not in out


@pytest.mark.skipif(not os.path.exists(LIB), reason="liblapgfx.so not built")
def test_binding_signatures_cover_header():
    from lap_amd import lib as L
    declared = set(declared_functions())
    assert set(L.EXPORTED) <= declared
