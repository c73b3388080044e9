"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import glob
import os
import re

import pytest

from wtomo3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
            if m.group(1) not in ("if", "sizeof", "return"):
                names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "wt_assemble" in names and "wt_graph" in names and "wt_asmble" in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_closed_form_nnz_from_library():
    L = _lib.lib()
    assert L.wt_graph_nnz(96, 96, 64) == 9 * 286 * 286 * 190
    assert L.wt_graph_nnz(0, 4, 4) == 0


def test_gemm_entry_rejects_unknown_modes_before_any_device_call():
    import ctypes
    L = _lib.lib()
    ms = ctypes.c_double()
    for mode in (-1, 4, 7):
        rc = L.wt_gemm_host_var(32, -1, 4, 4, 4, None, 4, None, 4, None, 4, mode, None, 1, ctypes.byref(ms))
        assert rc == 1  # WT_ERR_ARG


def test_parallel_interface_is_the_empty_stub():
    """strumpack_parallel_finter: the reference's MPI entry is an empty stub
    (strumpack_finter.c:123-126); here it does nothing and returns 0."""
    L = _lib.lib()
    L.strumpack_parallel_finter.argtypes = []
    L.strumpack_parallel_finter.restype = __import__("ctypes").c_int
    assert L.strumpack_parallel_finter() == 0
