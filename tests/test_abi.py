"""CPU checks of the C ABI: the library loads and exports every declared symbol,
and the ctypes mirror declares exactly those symbols.  No compute calls."""
import ctypes
import os
import re

from conftest import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "qrochet_hip.h")
LIB = os.path.join(PKG, "libqrochet_hip.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qh_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_api():
    names = declared()
    assert "qh_evolve2" in names and "qh_canonize" in names and "qh_expect" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_mirror_matches_header():
    from qrochet_hip import _lib

    assert sorted(_lib.SIGNATURES) == declared()
    lib = _lib.load()
    assert lib is _lib.load()
