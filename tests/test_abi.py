"""CPU: libqtn loads and exports every symbol include/qtn.h declares; the
host package refuses to run without it (no CPU fallback)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "qtn.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(qtn_[a-z0-9_]+)\s*\(", hdr)))


def test_header_lists_the_binding_exports():
    from qaintensor import _lib
    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    from qaintensor import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libqtn.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    lib.qtn_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.qtn_version()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from qaintensor import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libqtn.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.QtnLibraryError):
        _lib.load()
