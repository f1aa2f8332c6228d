"""CPU-side checks of the C ABI: the library builds, loads without a GPU and
exports every symbol include/fall3d_amd.h declares; host-side argument
validation that needs no device."""
import ctypes
import os
import re

import pytest

from fall3d_amd import lib as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fall3d_amd.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([a-z][a-z0-9_]*)\s*\(", txt)
    skip = {"if", "for", "while", "return", "sizeof"}
    return sorted({n for n in names if n not in skip and (n.startswith("f3d_") or n.endswith("_"))})


def test_header_and_binding_agree():
    decl = set(declared_symbols())
    bound = set(F.DROPIN_SYMBOLS) | set(F.CTX_SYMBOLS)
    assert decl == bound, (decl - bound, bound - decl)


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(F.LIB_PATH):
        pytest.skip("libfall3d_amd.so not built (run __graft_entry__.build())")
    return F.lib()


def test_library_exports_every_declared_symbol(L):
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_dropin_rejects_parallel_decomposition(L):
    # a z-slab decomposition (nz_loc < nz) is not supported: outputs untouched, error set
    import numpy as np
    nx, ny, nz = 4, 3, 5
    c = np.ones((nx + 2, ny + 2, nz + 2), order="F")
    v = np.ones((nx, ny, nz), order="F")
    i = lambda x: ctypes.byref(ctypes.c_int(x))  # noqa: E731
    d = lambda x: ctypes.byref(ctypes.c_double(x))  # noqa: E731
    dp = ctypes.POINTER(ctypes.c_double)
    L.advctx_(c.ctypes.data_as(dp), i(nx), i(ny), i(nz), i(2), i(1), v.ctypes.data_as(dp),
              d(1000.0), d(1.0), None, i(1))
    assert (c == 1.0).all()
    assert "serial decomposition" in F.last_error()
