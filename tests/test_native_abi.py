"""The C-ABI library loads and exports every entry point of include/*.h
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from samoa_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|size_t|void|double)\s+\**(swe\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_native.HIP_LIB)
    names = declared(os.path.join(ROOT, "include", "swe_mi355x.h"))
    assert "swe_step" in names and "swe_create" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(_native.SIGNATURES)


def test_headers_all_covered():
    headers = glob.glob(os.path.join(ROOT, "include", "*.h"))
    assert headers
