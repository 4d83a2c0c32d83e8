"""The C-ABI library loads on a host without a GPU and exports every entry
point declared in include/qfx.h (no compute calls are made here)."""
import ctypes
import os
import re

from qflex_amd import capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "qfx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qfx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "qfx_reorder_c64" in names and "qfx_cgemm_c64" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_functions()) == set(capi.SIGNATURES)


def test_error_string_without_gpu_call():
    lib = capi.load()
    assert lib.qfx_last_error() is not None


def test_index_map_layout():
    # dims (4, 2, 8): strides 16, 8, 1 -> log2 4, 3, 0; digits least significant first
    m = capi.index_map([4, 2, 8], [0, 2])
    assert m.ndigits == 2
    assert list(m.bits[:2]) == [3, 2] and list(m.shift[:2]) == [0, 4]
    # unit axes contribute no digit
    assert capi.index_map([1, 16], [0, 1]).ndigits == 1
