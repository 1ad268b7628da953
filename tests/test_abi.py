"""The C ABI (CPU; no GPU call is made): libnnkifmm loads and exports every
entry point that include/nnkifmm.h declares, and the ctypes table agrees."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols(nnk):
    with open(os.path.join(ROOT, "include", "nnkifmm.h")) as fh:
        hdr = fh.read()
    names = set(re.findall(r"\b(nnk_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 20
    lib = ctypes.CDLL(nnk.library_path())
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    from importlib import import_module
    sig = import_module("stokes-nnkifmm_amd._lib").SIGNATURES
    assert names == set(sig), names ^ set(sig)
    assert lib.nnk_version() >= 100
