"""CPU checks of the C ABI: the library loads and exports every symbol that
include/*.h declares, and the ctypes binding knows each of them."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("kmhodlr.h", "kmhodlr_testing.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(kmh_[a-z_0-9]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    import kmhodlr
    from kmhodlr import _lib
    lib = kmhodlr.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(lib.dll, n), "missing export " + n
    assert names == set(_lib.SIGNATURES) | set(_lib.TEST_SIGNATURES)


def test_struct_layouts_match_header():
    from kmhodlr import _lib as L
    assert ctypes.sizeof(L.kmh_kernel) == 4 + 4 + 8 * 4 + 8
    assert ctypes.sizeof(L.kmh_build_opts) == 4 * 4 + 8 + 4 * 4
    assert ctypes.sizeof(L.kmh_info) == 8 + 4 * 4 + 8


def test_no_device_reports_zero_or_more():
    import kmhodlr
    lib = kmhodlr.load()
    assert lib.device_count() >= 0


def test_missing_library_fails_loudly(tmp_path):
    from kmhodlr import _lib
    with pytest.raises(_lib.LibraryMissing):
        _lib.Lib(str(tmp_path / "nope.so"))
