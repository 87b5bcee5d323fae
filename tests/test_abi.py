"""The C-ABI library loads and exports every symbol include/pychem_hip.h
declares; the ctypes table matches the header.  No device calls."""
import os
import re

from pychem_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "pychem_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pq_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_ctypes_table_covers_header():
    assert set(header_functions()) == set(_lib.SIGNATURES)


def test_error_path_without_device():
    L = _lib.lib()
    assert L.pq_version() == 1
    assert L.pq_eri_nclass() == 21
    rc = L.pq_boys(4, None, 99, None, None)
    assert rc != 0 and b"mmax" in L.pq_last_error()
